"""Python mirror of the reference's hot-path interface over the C ABI.

The reference exposes the explicit stage update as Fortran module procedures
(``calc_face_value``, ``calc_face_flux``, ``update_state``,
``calc_timestep``, ``set_global_timestep``, ``advance_explicit``,
``exchange_messages``) on module-global block arrays.  :class:`Session` keeps
those names, argument meanings and argument order, and calls the gfx950
library ``libbats_hip.so`` through its C ABI (include/bats_hip.h).  A
non-zero status raises :class:`StopMpi`, the analogue of the reference's
``stop_mpi`` abort.

There is no CPU fallback: if the library is missing or no gfx950 device is
present, construction fails.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .grid import BlockGrid

__all__ = ["StopMpi", "BatsConfig", "Session", "library_path", "load_library",
           "LIMITERS", "FLUXES", "ABI_FUNCTIONS"]

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

LIMITERS = {"minmod": 1, "mc": 2, "mc3": 3, "koren": 3, "beta": 4, "no": 5}
FLUXES = {"rusanov": 1, "linde": 2, "hlld": 3}

STATUS = {0: "BATS_OK", 1: "BATS_EINVAL", 2: "BATS_ENEGFAST2", 3: "BATS_EHLLD",
          4: "BATS_ENAN", 5: "BATS_EBC", 6: "BATS_EHIP", 7: "BATS_ENOMEM",
          8: "BATS_ENODEV", 9: "BATS_ELIMIT"}


class StopMpi(RuntimeError):
    """Raised where the reference would call stop_mpi."""

    def __init__(self, code, where):
        self.code = code
        name = STATUS.get(code, str(code))
        msg = _strerror(code)
        super().__init__(f"{where}: {name}: {msg}")


class BatsConfig(C.Structure):
    _fields_ = [("nI", C.c_int), ("nJ", C.c_int), ("nK", C.c_int),
                ("nG", C.c_int), ("nDim", C.c_int), ("nBlock", C.c_int),
                ("nOrder", C.c_int), ("iLimiter", C.c_int),
                ("LimiterBeta", C.c_double), ("iFlux", C.c_int),
                ("nStage", C.c_int), ("UseHalfStep", C.c_int),
                ("Cfl", C.c_double), ("Gamma", C.c_double),
                ("pMin", C.c_double), ("UseDbTrick", C.c_int),
                ("UseDtFixed", C.c_int), ("DtFixed", C.c_double),
                ("cTiny", C.c_double), ("iDevice", C.c_int),
                ("UseDivbSource", C.c_int), ("StatePrecision", C.c_int),
                ("UseAdaptiveLowOrder", C.c_int), ("nLowOrder", C.c_int),
                ("pCritLow", C.c_double), ("pCritHigh", C.c_double),
                ("VelCrit", C.c_double),
                ("BetaLimiterResChange", C.c_double),
                ("nFaceLimiterResChange", C.c_int),
                ("UseOutflowPressure", C.c_int), ("pOutflow", C.c_double),
                ("IsSteadyState", C.c_int), ("UseTvdResChange", C.c_int)]


_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)
_hp = C.c_void_p

# name -> (restype, argtypes); every entry point of include/bats_hip.h
ABI_FUNCTIONS = {
    "bats_init": (C.c_int, [C.POINTER(BatsConfig), C.POINTER(_hp)]),
    "bats_finalize": (C.c_int, [_hp]),
    "bats_set_grid": (C.c_int, [_hp, _ip, _dp]),
    "bats_set_amr_grid": (C.c_int, [_hp, _ip, _dp, C.c_int, _ip]),
    "bats_set_low_order_face": (C.c_int, [_hp]),
    "bats_set_cell_boundary": (C.c_int, [_hp, _ip, _dp]),
    "bats_set_b0": (C.c_int, [_hp, _dp, C.c_int]),
    "bats_set_inflow_table": (C.c_int, [_hp, C.c_int, _dp, _dp]),
    "bats_set_body": (C.c_int, [_hp, C.POINTER(C.c_ubyte), _dp, C.c_double,
                                C.c_double, C.c_double]),
    "bats_set_step_graph": (C.c_int, [_hp, C.c_int]),
    "bats_get_low_order_crit": (C.c_int, [_hp, C.c_int, _dp, _dp, _dp]),
    "bats_upload_state": (C.c_int, [_hp, _dp]),
    "bats_download_state": (C.c_int, [_hp, _dp]),
    "bats_upload_state_soa": (C.c_int, [_hp, _dp]),
    "bats_download_state_soa": (C.c_int, [_hp, _dp]),
    "bats_exchange_messages": (C.c_int, [_hp]),
    "bats_set_stage": (C.c_int, [_hp, C.c_int]),
    "bats_calc_face_value": (C.c_int, [_hp, C.c_int, C.c_int, C.c_int]),
    "bats_calc_face_flux": (C.c_int, [_hp, C.c_int, C.c_int]),
    "bats_save_cons_flux": (C.c_int, [_hp, C.c_int]),
    "bats_message_pass_face": (C.c_int, [_hp]),
    "bats_apply_cons_flux": (C.c_int, [_hp, C.c_int]),
    "bats_calc_source": (C.c_int, [_hp, C.c_int]),
    "bats_update_state": (C.c_int, [_hp, C.c_int]),
    "bats_calc_timestep": (C.c_int, [_hp, C.c_int, C.c_int]),
    "bats_get_face_values": (C.c_int, [_hp] + [_dp] * 6),
    "bats_get_face_fluxes": (C.c_int, [_hp] + [_dp] * 3),
    "bats_get_dtmax": (C.c_int, [_hp, _dp, _dp]),
    "bats_set_global_timestep": (C.c_int, [_hp, C.c_double, _dp]),
    "bats_advance_explicit": (C.c_int, [_hp, C.c_int]),
    "bats_advance_stage": (C.c_int, [_hp, C.c_int, C.c_int]),
    "bats_advance": (C.c_int, [_hp, C.c_int, C.c_double]),
    "bats_init_timestep": (C.c_int, [_hp]),
    "bats_get_time": (C.c_int, [_hp, _dp, _dp, _ip]),
    "bats_synchronize": (C.c_int, [_hp]),
    "bats_set_runs": (C.c_int, [_hp, C.c_int, _ip]),
    "bats_get_run_times": (C.c_int, [_hp, _dp, _dp, _ip]),
    "bats_set_persistent_grid": (C.c_int, [_hp, C.c_int]),
    "bats_set_stage_shape": (C.c_int, [_hp, C.c_int, C.c_int]),
    "bats_mlp_create": (C.c_int, [C.POINTER(_hp)]),
    "bats_mlp_free": (C.c_int, [_hp]),
    "bats_mlp_set_params": (C.c_int, [_hp, _fp]),
    "bats_mlp_get_params": (C.c_int, [_hp, _fp, _fp]),
    "bats_mlp_forward": (C.c_int, [_hp, _fp, _fp, C.c_int]),
    "bats_mlp_forward_noalloc": (C.c_int, [_hp, _fp, _fp, C.c_int]),
    "bats_mlp_loss_grad": (C.c_int, [_hp, _fp, _fp, C.c_int, _dp, _fp]),
    "bats_mlp_set_data": (C.c_int, [_hp, _fp, _fp, C.c_long]),
    "bats_mlp_get_data": (C.c_int, [_hp, C.POINTER(C.c_long), _fp, _fp,
                                    _fp]),
    "bats_mlp_harvest": (C.c_int, [_hp, _hp, C.c_double, C.c_double,
                                   C.c_double, C.c_ulonglong,
                                   C.POINTER(C.c_long)]),
    "bats_mlp_train": (C.c_int, [_hp, C.c_int, C.c_int, C.c_double,
                                 C.c_double, C.c_int, C.c_ulonglong, _dp,
                                 _dp]),
    "bats_set_surrogate": (C.c_int, [_hp, _hp, C.c_int, C.c_double,
                                     C.c_double]),
    "bats_surrogate_counts": (C.c_int, [_hp, C.POINTER(C.c_long),
                                        C.POINTER(C.c_long)]),
    "bats_check_nan": (C.c_int, [_hp]),
    "bats_integrate_domain": (C.c_int, [_hp, _dp]),
    "bats_last_stage_kernel_ms": (C.c_int, [_hp, _dp, _ip]),
    "bats_strerror": (C.c_char_p, [C.c_int]),
}


def library_path():
    return os.path.join(_HERE, "libbats_hip.so")


def load_library():
    """Load libbats_hip.so (built by ``__graft_entry__.build()``)."""
    global _LIB
    if _LIB is None:
        path = library_path()
        if not os.path.exists(path):
            raise RuntimeError(
                f"{path} is missing: build it with __graft_entry__.build()")
        lib = C.CDLL(path)
        for name, (res, args) in ABI_FUNCTIONS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = lib
    return _LIB


def _strerror(code):
    try:
        return load_library().bats_strerror(code).decode()
    except Exception:  # library not loadable: keep the numeric code
        return "status %d" % code


def _ptr(a, ct=C.c_double):
    return a.ctypes.data_as(C.POINTER(ct))


SCHEME_DEFAULTS = dict(nOrder=2, limiter="mc", beta=1.5, flux="rusanov",
                       nStage=2, UseHalfStep=True, cfl=0.8, gamma=5.0 / 3.0,
                       pMin=-1.0, UseDbTrick=False, UseDtFixed=False,
                       DtFixed=0.0, cTiny=1e-6, UseDivbSource=False,
                       precision="f64", UseAdaptiveLowOrder=False,
                       nLowOrder=1, pCritLow=2.0, pCritHigh=1.5, VelCrit=2.0,
                       BetaLimiterResChange=2.0, nFaceLimiterResChange=2,
                       UseOutflowPressure=False, pOutflow=0.0,
                       IsSteadyState=False, UseTvdResChange=False)


# iTypeCellBc values of the outer boundaries (src/ModCellBoundary.f90:23-41)
BC_TYPES = {"periodic": 4, "float": 5, "outflow": 6, "reflect": 7,
            "fixed": 9, "inflow": 10, "vary": 11}


def make_config(grid: BlockGrid, device=0, **scheme):
    s = dict(SCHEME_DEFAULTS)
    s.update(scheme)
    c = BatsConfig()
    c.nI, c.nJ, c.nK = grid.nI, grid.nJ, grid.nK
    c.nG, c.nDim, c.nBlock = grid.nG, grid.nDim, grid.nBlock
    c.nOrder = s["nOrder"]
    lim = s["limiter"]
    c.iLimiter = LIMITERS[lim] if isinstance(lim, str) else int(lim)
    c.LimiterBeta = s["beta"]
    fl = s["flux"]
    c.iFlux = FLUXES[fl] if isinstance(fl, str) else int(fl)
    c.nStage = s["nStage"]
    c.UseHalfStep = int(s["UseHalfStep"])
    c.Cfl = s["cfl"]
    c.Gamma = s["gamma"]
    c.pMin = s["pMin"]
    c.UseDbTrick = int(s["UseDbTrick"])
    c.UseDtFixed = int(s["UseDtFixed"])
    c.DtFixed = s["DtFixed"]
    c.cTiny = s["cTiny"]
    c.iDevice = device
    c.UseDivbSource = int(s["UseDivbSource"])
    prec = s["precision"]
    if prec not in ("f64", "f32"):
        raise ValueError("precision must be 'f64' or 'f32'")
    c.StatePrecision = 8 if prec == "f64" else 4
    c.UseAdaptiveLowOrder = int(s["UseAdaptiveLowOrder"])
    c.nLowOrder = int(s["nLowOrder"])
    c.pCritLow, c.pCritHigh = s["pCritLow"], s["pCritHigh"]
    c.VelCrit = s["VelCrit"]
    c.BetaLimiterResChange = s["BetaLimiterResChange"]
    c.nFaceLimiterResChange = int(s["nFaceLimiterResChange"])
    c.UseOutflowPressure = int(s["UseOutflowPressure"])
    c.pOutflow = s["pOutflow"]
    c.IsSteadyState = int(s["IsSteadyState"])
    c.UseTvdResChange = int(s["UseTvdResChange"])
    return c


class Session:
    """One GPU-resident solution: the reference's State_VGB plus the
    hot-path routines, with reference names and argument order."""

    def __init__(self, grid: BlockGrid, device=0, default_grid=False,
                 **scheme):
        """default_grid: keep bats_init's default grid (each block its own
        domain of unit cells, six outer sides) instead of setting grid's
        neighbours and cell sizes; grid then only gives the shapes."""
        self.lib = load_library()
        self.grid = grid
        self.cfg = make_config(grid, device, **scheme)
        h = _hp()
        self._check(self.lib.bats_init(C.byref(self.cfg), C.byref(h)),
                    "bats_init")
        self.h = h
        nei = np.ascontiguousarray(grid.iNeighbor_EB, dtype=np.int32)
        cs = np.ascontiguousarray(grid.CellSize_DB, dtype=np.float64)
        slabs = getattr(grid, "iSlab_IS", None)
        if default_grid:
            pass
        elif slabs is not None and len(slabs):
            sl = np.ascontiguousarray(slabs, dtype=np.int32)
            self._check(self.lib.bats_set_amr_grid(
                self.h, _ptr(nei, C.c_int), _ptr(cs), len(sl),
                _ptr(sl, C.c_int)), "bats_set_amr_grid")
        else:
            self._check(self.lib.bats_set_grid(self.h, _ptr(nei, C.c_int),
                                               _ptr(cs)), "bats_set_grid")
        if scheme.get("TypeCellBc") is not None:
            # applied to the grid just set (before any state is uploaded)
            self.set_cell_boundary(scheme["TypeCellBc"],
                                   scheme.get("CellState_VI"))

    # ------------------------------------------------------------------
    def _check(self, code, where):
        if code != 0:
            raise StopMpi(code, where)

    def close(self):
        if getattr(self, "h", None):
            self.lib.bats_finalize(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- state transfer (sync_cpu_gpu) ---------------------------------
    def upload_state(self, State_VGB):
        S = np.ascontiguousarray(State_VGB, dtype=np.float64)
        self._check(self.lib.bats_upload_state(self.h, _ptr(S)),
                    "upload_state")

    def download_state(self, out=None):
        g = self.grid
        gI, gJ, gK = g.ghost
        shp = (g.nBlock, g.nK + 2 * gK, g.nJ + 2 * gJ, g.nI + 2 * gI, 8)
        S = np.zeros(shp) if out is None else out
        self._check(self.lib.bats_download_state(self.h, _ptr(S)),
                    "download_state")
        return S

    def upload_soa(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64)
        self._check(self.lib.bats_upload_state_soa(self.h, _ptr(U)),
                    "upload_state_soa")

    def download_soa(self):
        g = self.grid
        U = np.zeros((8, g.nBlock, g.nK, g.nJ, g.nI))
        self._check(self.lib.bats_download_state_soa(self.h, _ptr(U)),
                    "download_state_soa")
        return U

    # -- reference-named per-block routines ----------------------------
    def exchange_messages(self):
        self._check(self.lib.bats_exchange_messages(self.h),
                    "exchange_messages")

    def set_stage(self, iStage):
        self._check(self.lib.bats_set_stage(self.h, iStage), "set_stage")

    def calc_face_value(self, iBlock, DoResChangeOnly=False,
                        DoMonotoneRestrict=True):
        self._check(self.lib.bats_calc_face_value(
            self.h, iBlock, int(DoResChangeOnly), int(DoMonotoneRestrict)),
            "calc_face_value")

    def set_cell_boundary(self, TypeCellBc, CellState_VI=None):
        """Outer boundary types of the sides x-, x+, y-, y+, z-, z+ (names
        of BC_TYPES or iTypeCellBc codes) and the (6, 8) ghost states of the
        fixed/inflow sides (rho, rho u, B, p).  Reallocates the state when
        a side is fixed: upload the state afterwards."""
        t = np.array([BC_TYPES[x] if isinstance(x, str) else int(x)
                      for x in TypeCellBc], dtype=np.int32)
        if t.shape != (6,):
            raise ValueError("TypeCellBc needs 6 sides")
        st = None
        if CellState_VI is not None:
            st = np.ascontiguousarray(CellState_VI, dtype=np.float64)
            if st.shape != (6, 8):
                raise ValueError("CellState_VI must have shape (6, 8)")
        self._check(self.lib.bats_set_cell_boundary(
            self.h, _ptr(t, C.c_int), None if st is None else _ptr(st)),
            "set_cell_boundary")

    def set_inflow_table(self, Time_I, State_VI=None):
        """Time-varying inflow of the x+ side (type 'vary'): table times
        (nPoint,) and primitive states (nPoint, 8); None drops it."""
        if Time_I is None:
            self._check(self.lib.bats_set_inflow_table(self.h, 0, None, None),
                        "set_inflow_table")
            return
        t = np.ascontiguousarray(Time_I, dtype=np.float64)
        st = np.ascontiguousarray(State_VI, dtype=np.float64)
        assert st.shape == (len(t), 8), st.shape
        self._check(self.lib.bats_set_inflow_table(self.h, len(t), _ptr(t),
                                                   _ptr(st)),
                    "set_inflow_table")

    def set_b0(self, B0_DGB, UseB0Source=True):
        """B0 splitting: B0_DGB of shape (nBlock, nKg, nJg, nIg, 3), the
        static field at the cell centres and ghost cells (None drops it).
        The state then holds B1 = B - B0."""
        if B0_DGB is None:
            self._check(self.lib.bats_set_b0(self.h, None, 0), "set_b0")
            return
        b = np.ascontiguousarray(B0_DGB, dtype=np.float64)
        g = self.grid
        assert b.shape == (g.nBlock, g.nK + 2 * g.nG, g.nJ + 2 * g.nG,
                           g.nI + 2 * g.nG, 3), b.shape
        self._check(self.lib.bats_set_b0(self.h, _ptr(b), int(UseB0Source)),
                    "set_b0")

    def set_body(self, Used_GB, XyzStart_DB=None, BodyRho=1.0, BodyP=1.0,
                 B1rCoef=-1.0):
        """Inner body: Used_GB of shape (nBlock, nKg, nJg, nIg), True at the
        true cells of every block and its ghost cells, XyzStart_DB
        (nBlock, 3) the first cell centres relative to the body centre, the
        body density and pressure of the ghost face state and B1rCoef
        (None drops the body)."""
        if Used_GB is None:
            self._check(self.lib.bats_set_body(self.h, None, None, 0.0, 0.0,
                                               0.0), "set_body")
            return
        g = self.grid
        u = np.ascontiguousarray(Used_GB, dtype=np.uint8)
        assert u.shape == (g.nBlock, g.nK + 2 * g.nG, g.nJ + 2 * g.nG,
                           g.nI + 2 * g.nG), u.shape
        x = np.ascontiguousarray(XyzStart_DB, dtype=np.float64)
        assert x.shape == (g.nBlock, 3), x.shape
        self._check(self.lib.bats_set_body(
            self.h, _ptr(u, C.c_ubyte), _ptr(x), float(BodyRho), float(BodyP),
            float(B1rCoef)), "set_body")

    def set_step_graph(self, use=True):
        """bats_advance replays one captured hipGraph per time step."""
        self._check(self.lib.bats_set_step_graph(self.h, int(bool(use))),
                    "set_step_graph")

    def set_low_order_face(self):
        """set_low_order_face on the last exchanged state (all blocks)."""
        self._check(self.lib.bats_set_low_order_face(self.h),
                    "set_low_order_face")

    def get_low_order_crit(self, iBlock):
        """LowOrderCrit_[XYZ]B of block iBlock (1-based) in C order:
        (nK, nJ, nI+1), (nK, nJ+1, nI), (nK+1, nJ, nI)."""
        out = [np.zeros(s) for s in self.face_shapes()]
        self._check(self.lib.bats_get_low_order_crit(
            self.h, iBlock, *[_ptr(a) for a in out]), "get_low_order_crit")
        return out

    def calc_face_flux(self, DoResChangeOnly, iBlock):
        self._check(self.lib.bats_calc_face_flux(
            self.h, int(DoResChangeOnly), iBlock), "calc_face_flux")

    def save_cons_flux(self, iBlock):
        self._check(self.lib.bats_save_cons_flux(self.h, iBlock),
                    "save_cons_flux")

    def message_pass_face(self):
        self._check(self.lib.bats_message_pass_face(self.h),
                    "message_pass_face")

    def apply_cons_flux(self, iBlock):
        self._check(self.lib.bats_apply_cons_flux(self.h, iBlock),
                    "apply_cons_flux")

    def calc_source(self, iBlock):
        self._check(self.lib.bats_calc_source(self.h, iBlock), "calc_source")

    def update_state(self, iBlock):
        self._check(self.lib.bats_update_state(self.h, iBlock),
                    "update_state")

    def calc_timestep(self, iBlock, IsPartLocal=False):
        self._check(self.lib.bats_calc_timestep(self.h, iBlock,
                                                int(IsPartLocal)),
                    "calc_timestep")

    def face_shapes(self):
        g = self.grid
        return ((g.nK, g.nJ, g.nI + 1), (g.nK, g.nJ + 1, g.nI),
                (g.nK + 1, g.nJ, g.nI))

    def get_face_values(self):
        sx, sy, sz = self.face_shapes()
        arrs = [np.zeros(s + (8,)) for s in (sx, sx, sy, sy, sz, sz)]
        self._check(self.lib.bats_get_face_values(
            self.h, *[_ptr(a) for a in arrs]), "get_face_values")
        return arrs

    def get_face_fluxes(self):
        fl = [np.zeros(s + (14,)) for s in self.face_shapes()]
        self._check(self.lib.bats_get_face_fluxes(
            self.h, *[_ptr(a) for a in fl]), "get_face_fluxes")
        return fl

    def get_dtmax(self):
        g = self.grid
        cb = np.zeros((g.nBlock, g.nK, g.nJ, g.nI))
        b = np.zeros(g.nBlock)
        self._check(self.lib.bats_get_dtmax(self.h, _ptr(cb), _ptr(b)),
                    "get_dtmax")
        return cb, b

    def get_dtmax_b(self):
        b = np.zeros(self.grid.nBlock)
        self._check(self.lib.bats_get_dtmax(self.h, None, _ptr(b)),
                    "get_dtmax")
        return b

    # -- production path ------------------------------------------------
    def set_global_timestep(self, TimeSimulationLimit=-1.0):
        dt = C.c_double()
        self._check(self.lib.bats_set_global_timestep(
            self.h, TimeSimulationLimit, C.byref(dt)), "set_global_timestep")
        return dt.value

    def advance_explicit(self, DoCalcTimestep=True):
        self._check(self.lib.bats_advance_explicit(self.h,
                                                   int(DoCalcTimestep)),
                    "advance_explicit")

    def advance_stage(self, iStage, DoCalcTimestep=True):
        self._check(self.lib.bats_advance_stage(self.h, iStage,
                                                int(DoCalcTimestep)),
                    "advance_stage")

    def advance(self, nStep, TimeSimulationLimit=-1.0):
        """BATS_advance for nStep time steps."""
        self._check(self.lib.bats_advance(self.h, nStep, TimeSimulationLimit),
                    "advance")

    def init_timestep(self):
        self._check(self.lib.bats_init_timestep(self.h), "init_timestep")

    def synchronize(self):
        self._check(self.lib.bats_synchronize(self.h), "synchronize")

    def check_nan(self):
        self._check(self.lib.bats_check_nan(self.h), "check_nan")

    def time(self):
        dt, t, n = C.c_double(), C.c_double(), C.c_int()
        self._check(self.lib.bats_get_time(self.h, C.byref(dt), C.byref(t),
                                           C.byref(n)), "get_time")
        return dt.value, t.value, n.value

    def set_runs(self, iBlockFirst_R):
        """Independent runs over contiguous block ranges (0-based offsets,
        nRun+1 entries)."""
        f = np.ascontiguousarray(iBlockFirst_R, dtype=np.int32)
        self._check(self.lib.bats_set_runs(self.h, len(f) - 1,
                                           _ptr(f, C.c_int)), "set_runs")
        self.nRun = len(f) - 1

    def run_times(self):
        n = getattr(self, "nRun", 1)
        dt, t = np.zeros(n), np.zeros(n)
        ns = np.zeros(n, dtype=np.int32)
        self._check(self.lib.bats_get_run_times(self.h, _ptr(dt), _ptr(t),
                                                _ptr(ns, C.c_int)),
                    "get_run_times")
        return dt, t, ns

    def set_persistent_grid(self, nWorkgroup):
        self._check(self.lib.bats_set_persistent_grid(self.h, nWorkgroup),
                    "set_persistent_grid")

    def set_stage_shape(self, CellsPerThread=None, UseXcdOrder=False):
        """Cells per thread of the stage kernel (NB, NB/2 or NB/4; None:
        NB) and the XCD-contiguous work order (the same bits; a split
        rounds its part-boundary faces within an ulp)."""
        cpt = self.grid.nI if CellsPerThread is None else int(CellsPerThread)
        self._check(self.lib.bats_set_stage_shape(self.h, cpt,
                                                  int(bool(UseXcdOrder))),
                    "set_stage_shape")

    def set_surrogate(self, model, mode="network", BetaMin=0.0, JumpTol=0.05):
        """Gated face fluxes on an FP32 handle: mode 'network' (the trained
        model on smooth faces), 'rusanov' (the same gate, Rusanov flux) or
        'off'."""
        code = {"off": 0, "network": 1, "rusanov": 2}[mode]
        self._check(self.lib.bats_set_surrogate(
            self.h, model.h if model is not None else None, code,
            float(BetaMin), float(JumpTol)), "set_surrogate")
        self._surrogate = model

    def surrogate_counts(self):
        nf, nn = C.c_long(), C.c_long()
        self._check(self.lib.bats_surrogate_counts(self.h, C.byref(nf),
                                                   C.byref(nn)),
                    "surrogate_counts")
        return nf.value, nn.value

    def integrate_domain(self):
        s = np.zeros(8)
        self._check(self.lib.bats_integrate_domain(self.h, _ptr(s)),
                    "integrate_domain")
        return s

    def last_stage_kernel_ms(self):
        ms, n = C.c_double(), C.c_int()
        self._check(self.lib.bats_last_stage_kernel_ms(self.h, C.byref(ms),
                                                       C.byref(n)),
                    "last_stage_kernel_ms")
        return ms.value, n.value


class Mlp:
    """The face-flux surrogate network on the GPU (bats_mlp_* entry points):
    FP32 master weights, BF16 matrix-core forward/backward, Adam."""

    NPARAM = 27464

    def __init__(self, params=None):
        self.lib = load_library()
        h = _hp()
        st = self.lib.bats_mlp_create(C.byref(h))
        if st != 0:
            raise StopMpi(st, "bats_mlp_create")
        self.h = h
        if params is not None:
            self.set_params(params)

    def _check(self, st, what):
        if st != 0:
            raise StopMpi(st, what)

    def close(self):
        if self.h:
            self.lib.bats_mlp_free(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_params(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        assert w.size == self.NPARAM
        self._check(self.lib.bats_mlp_set_params(self.h, _ptr(w, C.c_float)),
                    "mlp_set_params")

    def params(self):
        w = np.zeros(self.NPARAM, np.float32)
        g = np.zeros(self.NPARAM, np.float32)
        self._check(self.lib.bats_mlp_get_params(self.h, _ptr(w, C.c_float),
                                                 _ptr(g, C.c_float)),
                    "mlp_get_params")
        return w, g

    def forward(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.zeros((X.shape[0], 8), np.float32)
        self._check(self.lib.bats_mlp_forward(self.h, _ptr(X, C.c_float),
                                              _ptr(Y, C.c_float), X.shape[0]),
                    "mlp_forward")
        return Y

    def forward_noalloc(self, X):
        """forward() through the entry point that allocates no device
        memory (rows staged through the library's fixed array)."""
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.zeros((X.shape[0], 8), np.float32)
        self._check(self.lib.bats_mlp_forward_noalloc(
            self.h, _ptr(X, C.c_float), _ptr(Y, C.c_float), X.shape[0]),
            "mlp_forward_noalloc")
        return Y

    def loss_grad(self, X, T):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        g = np.zeros(self.NPARAM, np.float32)
        loss = C.c_double()
        self._check(self.lib.bats_mlp_loss_grad(
            self.h, _ptr(X, C.c_float), _ptr(T, C.c_float), X.shape[0],
            C.byref(loss), _ptr(g, C.c_float)), "mlp_loss_grad")
        return loss.value, g

    def set_data(self, X, T):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        self._check(self.lib.bats_mlp_set_data(self.h, _ptr(X, C.c_float),
                                               _ptr(T, C.c_float), X.shape[0]),
                    "mlp_set_data")

    def data(self):
        """Samples held on the device: inputs, targets, baseline fluxes."""
        n = C.c_long()
        self._check(self.lib.bats_mlp_get_data(self.h, C.byref(n), None, None,
                                               None), "mlp_get_data")
        X = np.zeros((n.value, 16), np.float32)
        T = np.zeros((n.value, 8), np.float32)
        B = np.zeros((n.value, 8), np.float32)
        self._check(self.lib.bats_mlp_get_data(self.h, C.byref(n),
                                               _ptr(X, C.c_float),
                                               _ptr(T, C.c_float),
                                               _ptr(B, C.c_float)),
                    "mlp_get_data")
        return X, T, B

    def harvest(self, session, fraction, BetaMin=0.0, JumpTol=0.05, seed=1):
        n = C.c_long()
        self._check(self.lib.bats_mlp_harvest(self.h, session.h,
                                              float(fraction), float(BetaMin),
                                              float(JumpTol), int(seed),
                                              C.byref(n)),
                    "mlp_harvest")
        return n.value

    def train(self, nEpoch, batch=65536, lr=1e-3, hold=0.1, seed=1,
              standardize=False, jump_weight=False, graph=True):
        """Returns the last minibatch loss and the held-out relative L2
        errors (18: flux overall + 8 outputs, correction overall + 8)."""
        loss = C.c_double()
        err = np.zeros(18)
        self._check(self.lib.bats_mlp_train(self.h, int(nEpoch), int(batch),
                                            float(lr), float(hold),
                                            int(standardize) |
                                            2 * int(jump_weight) |
                                            4 * int(not graph), int(seed),
                                            C.byref(loss), _ptr(err)),
                    "mlp_train")
        return loss.value, err
