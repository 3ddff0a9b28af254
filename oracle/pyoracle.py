"""ctypes wrapper of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Loads ``oracle/liboracle.so`` (built by ``make -C oracle``).  Only tests/,
``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg import this
module; the product never does.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

LIMITERS = {"minmod": 1, "mc": 2, "mc3": 3, "koren": 3, "beta": 4, "no": 5}
FLUXES = {"rusanov": 1, "linde": 2, "hlld": 3}


class ObConfig(C.Structure):
    _fields_ = [("nI", C.c_int), ("nJ", C.c_int), ("nK", C.c_int),
                ("nG", C.c_int), ("nDim", C.c_int), ("nOrder", C.c_int),
                ("iLimiter", C.c_int), ("LimiterBeta", C.c_double),
                ("iFlux", C.c_int), ("nStage", C.c_int),
                ("UseHalfStep", C.c_int), ("Cfl", C.c_double),
                ("Gamma", C.c_double), ("pMin", C.c_double),
                ("UseDbTrick", C.c_int), ("UseDtFixed", C.c_int),
                ("DtFixed", C.c_double), ("cTiny", C.c_double),
                ("UseDivbSource", C.c_int),
                ("UseAdaptiveLowOrder", C.c_int), ("nLowOrder", C.c_int),
                ("pCritLow", C.c_double), ("pCritHigh", C.c_double),
                ("VelCrit", C.c_double),
                ("iTypeCellBc_I", C.c_int * 6),
                ("CellState_VI", C.c_double * 48),
                ("BetaLimiterResChange", C.c_double),
                ("nFaceLimiterResChange", C.c_int),
                ("UseOutflowPressure", C.c_int), ("pOutflow", C.c_double),
                ("IsSteadyState", C.c_int), ("UseTvdResChange", C.c_int)]

# iTypeCellBc values (src/ModCellBoundary.f90:23-41)
BC_TYPES = {"periodic": 4, "float": 5, "outflow": 6, "reflect": 7,
            "fixed": 9, "inflow": 10, "vary": 11}


class ObSession(C.Structure):
    _fields_ = [("c", ObConfig), ("nBlock", C.c_int),
                ("iNeighbor_EB", C.POINTER(C.c_int)),
                ("CellSize_DB", C.POINTER(C.c_double)),
                ("State", C.POINTER(C.c_double)),
                ("StateOld", C.POINTER(C.c_double)),
                ("DtMax_CB", C.POINTER(C.c_double)),
                ("DtMax_B", C.POINTER(C.c_double)),
                ("Dt", C.c_double), ("tSimulation", C.c_double),
                ("nStep", C.c_int), ("nThread", C.c_int),
                ("nSlab", C.c_int), ("iSlab_IS", C.POINTER(C.c_int)),
                ("Corr", C.POINTER(C.c_double)),
                ("LowOrderCrit", C.POINTER(C.c_double)),
                ("B0", C.POINTER(C.c_double)), ("UseB0Source", C.c_int),
                ("Used", C.POINTER(C.c_ubyte)),
                ("XyzStart_DB", C.POINTER(C.c_double)),
                ("BodyRho", C.c_double), ("BodyP", C.c_double),
                ("B1rCoef", C.c_double), ("nVary", C.c_int),
                ("VaryTime_I", C.POINTER(C.c_double)),
                ("VaryState_VI", C.POINTER(C.c_double))]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(path)
        dp = C.POINTER(C.c_double)
        L.ob_limiter.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, dp,
                                 C.c_int, dp, dp]
        L.ob_physical_flux.argtypes = [C.POINTER(ObConfig), C.c_int, dp, dp,
                                       dp, dp]
        L.ob_speed_max.argtypes = [C.POINTER(ObConfig), C.c_int, dp, dp, dp,
                                   dp]
        L.ob_speed_max.restype = C.c_int
        L.ob_numerical_flux.argtypes = [C.POINTER(ObConfig), C.c_int,
                                        C.c_double, dp, dp, dp]
        L.ob_numerical_flux.restype = C.c_int
        L.ob_calc_face_value.argtypes = [C.POINTER(ObConfig), dp] + [dp] * 6
        L.ob_calc_face_value_lo.argtypes = [C.POINTER(ObConfig), dp] + \
            [dp] * 9
        L.ob_calc_face_value_rc.argtypes = [C.POINTER(ObConfig), dp, dp,
                                            dp, dp, C.POINTER(C.c_int)] + \
            [dp] * 6
        L.ob_set_low_order_face.argtypes = [C.POINTER(ObConfig), dp, dp, dp,
                                            dp]
        L.ob_calc_face_flux.argtypes = [C.POINTER(ObConfig), dp] + [dp] * 9
        L.ob_calc_face_flux.restype = C.c_int
        L.ob_update_state.argtypes = [C.POINTER(ObConfig), C.c_int,
                                      C.c_double, dp, dp, dp, dp, dp, dp, dp,
                                      dp]
        L.ob_calc_source.argtypes = [C.POINTER(ObConfig), dp, dp] + [dp] * 7
        L.ob_calc_timestep.argtypes = [C.POINTER(ObConfig), dp, dp, dp, dp,
                                       dp, dp]
        L.ob_exchange_messages.argtypes = [C.POINTER(ObConfig), C.c_int,
                                           C.POINTER(C.c_int), dp]
        L.ob_set_global_timestep.argtypes = [C.POINTER(ObSession), C.c_double]
        L.ob_advance_explicit.argtypes = [C.POINTER(ObSession), C.c_int]
        L.ob_advance_explicit.restype = C.c_int
        L.ob_advance_stage.argtypes = [C.POINTER(ObSession), C.c_int, C.c_int]
        L.ob_advance_stage.restype = C.c_int
        L.ob_advance_step.argtypes = [C.POINTER(ObSession), C.c_double]
        L.ob_advance_step.restype = C.c_int
        L.ob_session_sources.argtypes = [C.POINTER(ObSession), dp]
        L.ob_session_sources.restype = C.c_int
        L.ob_session_b0.argtypes = [C.POINTER(ObSession), C.c_int, dp, dp,
                                    dp, dp, dp]
        L.ob_session_b0.restype = C.c_int
        L.ob_init_timestep.argtypes = [C.POINTER(ObSession)]
        L.ob_init_timestep.restype = C.c_int
        L.ob_exchange_amr.argtypes = [C.POINTER(ObConfig), C.c_int, C.c_int,
                                      C.POINTER(C.c_int), dp]
        L.ob_physical_flux_b0.argtypes = [C.POINTER(ObConfig), C.c_int, dp,
                                          dp, dp, dp, dp]
        L.ob_speed_max_b0.argtypes = [C.POINTER(ObConfig), C.c_int, dp, dp,
                                      dp, dp, dp]
        L.ob_speed_max_b0.restype = C.c_int
        L.ob_numerical_flux_b0.argtypes = [C.POINTER(ObConfig), C.c_int,
                                           C.c_double, dp, dp, dp, dp]
        L.ob_numerical_flux_b0.restype = C.c_int
        L.ob_calc_face_flux_b0.argtypes = [C.POINTER(ObConfig), dp] + \
            [dp] * 12
        L.ob_calc_face_flux_b0.restype = C.c_int
        L.ob_set_b0_face.argtypes = [C.POINTER(ObConfig), dp, dp, dp, dp]
        L.ob_set_b0_source.argtypes = [C.POINTER(ObConfig), dp, dp, dp, dp,
                                       dp, dp]
        L.ob_calc_source_b0.argtypes = [C.POINTER(ObConfig), dp, dp] + \
            [dp] * 6 + [dp] * 6 + [C.c_int, dp]
        L.ob_body_face_state.argtypes = [dp, dp, C.c_double, C.c_double,
                                         C.c_double, dp]
        L.ob_tvd_reschange.argtypes = [C.POINTER(ObConfig)] + [dp] * 7
        L.ob_vary_state.argtypes = [C.c_int, dp, dp, C.c_double, dp]
        L.ob_session_exchange.argtypes = [C.POINTER(ObSession)]
        L.ob_correct_monotone_restrict.argtypes = [dp, dp]
        ub = C.POINTER(C.c_ubyte)
        L.ob_session_face_values.argtypes = [C.POINTER(ObSession), C.c_int] \
            + [dp] * 6 + [ub] * 3
        L.ob_pressure_to_energy_cell.argtypes = [C.POINTER(ObConfig), dp]
        L.ob_energy_to_pressure_cell.argtypes = [C.POINTER(ObConfig), dp]
        _LIB = L
    return _LIB


def _dp(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def make_config(nI=8, nJ=8, nK=8, nG=2, nDim=3, nOrder=2, limiter="mc",
                beta=1.5, flux="rusanov", nStage=2, UseHalfStep=True, cfl=0.8,
                gamma=5.0 / 3.0, pMin=-1.0, UseDbTrick=False,
                UseDtFixed=False, DtFixed=0.0, cTiny=1e-6,
                UseDivbSource=False, UseAdaptiveLowOrder=False, nLowOrder=1,
                pCritLow=2.0, pCritHigh=1.5, VelCrit=2.0, TypeCellBc=None,
                CellState_VI=None, BetaLimiterResChange=2.0,
                nFaceLimiterResChange=2, UseOutflowPressure=False,
                pOutflow=0.0, IsSteadyState=False, UseTvdResChange=False):
    """TypeCellBc: six names or iTypeCellBc codes (x-, x+, y-, y+, z-, z+;
    default float); CellState_VI: (6, 8) ghost states of the fixed sides
    in conservative variables (rho, rho u, B, p)."""
    c = ObConfig()
    c.nI, c.nJ, c.nK, c.nG, c.nDim, c.nOrder = nI, nJ, nK, nG, nDim, nOrder
    c.iLimiter = LIMITERS[limiter] if isinstance(limiter, str) else limiter
    c.LimiterBeta = beta
    c.iFlux = FLUXES[flux] if isinstance(flux, str) else flux
    c.nStage, c.UseHalfStep, c.Cfl, c.Gamma = nStage, int(UseHalfStep), cfl, gamma
    c.pMin, c.UseDbTrick = pMin, int(UseDbTrick)
    c.UseDtFixed, c.DtFixed, c.cTiny = int(UseDtFixed), DtFixed, cTiny
    if UseDtFixed:  # src/ModSetParameters.f90:4314-4321
        c.Cfl = 1.0
    c.UseDivbSource = int(UseDivbSource)
    c.UseAdaptiveLowOrder, c.nLowOrder = int(UseAdaptiveLowOrder), nLowOrder
    c.pCritLow, c.pCritHigh, c.VelCrit = pCritLow, pCritHigh, VelCrit
    c.UseOutflowPressure, c.pOutflow = int(UseOutflowPressure), pOutflow
    c.IsSteadyState = int(IsSteadyState)
    c.UseTvdResChange = int(UseTvdResChange)
    c.BetaLimiterResChange = BetaLimiterResChange
    c.nFaceLimiterResChange = nFaceLimiterResChange
    for side, t in enumerate(TypeCellBc or ["float"] * 6):
        c.iTypeCellBc_I[side] = BC_TYPES[t] if isinstance(t, str) else int(t)
    if CellState_VI is not None:
        st = np.asarray(CellState_VI, dtype=np.float64).reshape(48)
        for t in range(48):
            c.CellState_VI[t] = st[t]
    return c


def config_from_dict(d, grid):
    return make_config(nI=grid.nI, nJ=grid.nJ, nK=grid.nK, nG=grid.nG,
                       nDim=grid.nDim, **d)


# ----------------------------------------------------------------------
# point routines
def limiter(iLimiter, P, beta):
    """Slopes for a pencil P of shape (n, 8) holding l = lMin-2 .. lMax+1;
    returns dL, dR of the same shape with rows lMin-1 .. lMax valid."""
    P = np.ascontiguousarray(P, dtype=np.float64)
    n = P.shape[0]
    lMin, lMax = 2, n - 2
    dL = np.zeros_like(P)
    dR = np.zeros_like(P)
    lib().ob_limiter(LIMITERS.get(iLimiter, iLimiter), lMin, lMax, beta,
                     _dp(P), 0, _dp(dL), _dp(dR))
    return dL, dR


def tvd_reschange(cfg, C2, C1, F1, F2):
    """tvd_reschange on primitive states: C2, C1 (8,), F1, F2 (4, 8) with
    subface f = i2 + 2 j2.  Returns (CoarseToFineF, FineToCoarseF, FineF),
    each (4, 8)."""
    a = [np.ascontiguousarray(x, dtype=np.float64) for x in (C2, C1, F1, F2)]
    out = [np.zeros((4, 8)) for _ in range(3)]
    lib().ob_tvd_reschange(C.byref(cfg), *[_dp(x) for x in a],
                           *[_dp(x) for x in out])
    return tuple(out)


def vary_state(Time_I, State_VI, t):
    """The time-varying inflow's conservative state at time t from a table
    of primitive states State_VI (nPoint, 8)."""
    T = np.ascontiguousarray(Time_I, dtype=np.float64)
    S = np.ascontiguousarray(State_VI, dtype=np.float64)
    out = np.zeros(8)
    lib().ob_vary_state(len(T), _dp(T), _dp(S), float(t), _dp(out))
    return out


def correct_monotone_restrict(Cell, Ghost):
    """The ghost's primitive state moved to the coarse cell distance."""
    c = np.ascontiguousarray(Cell, dtype=np.float64)
    g = np.array(Ghost, dtype=np.float64)
    lib().ob_correct_monotone_restrict(_dp(c), _dp(g))
    return g


def physical_flux(cfg, iDim, state):
    s = np.ascontiguousarray(state, dtype=np.float64)
    cons = np.zeros(9)
    flux = np.zeros(9)
    un = C.c_double()
    lib().ob_physical_flux(C.byref(cfg), iDim, _dp(s), _dp(cons), _dp(flux),
                           C.byref(un))
    return cons, flux, un.value


def speed_max(cfg, iDim, state):
    s = np.ascontiguousarray(state, dtype=np.float64)
    a, b, c = C.c_double(), C.c_double(), C.c_double()
    err = lib().ob_speed_max(C.byref(cfg), iDim, _dp(s), C.byref(a),
                             C.byref(b), C.byref(c))
    return err, a.value, b.value, c.value


def numerical_flux(cfg, iDim, area, left, right):
    L = np.ascontiguousarray(left, dtype=np.float64)
    R = np.ascontiguousarray(right, dtype=np.float64)
    F = np.zeros(14)
    err = lib().ob_numerical_flux(C.byref(cfg), iDim, area, _dp(L), _dp(R),
                                  _dp(F))
    return err, F


def physical_flux_b0(cfg, iDim, state, b0):
    s = np.ascontiguousarray(state, dtype=np.float64)
    b = np.ascontiguousarray(b0, dtype=np.float64)
    cons = np.zeros(9)
    flux = np.zeros(9)
    un = C.c_double()
    lib().ob_physical_flux_b0(C.byref(cfg), iDim, _dp(s), _dp(b), _dp(cons),
                              _dp(flux), C.byref(un))
    return cons, flux, un.value


def speed_max_b0(cfg, iDim, state, b0):
    s = np.ascontiguousarray(state, dtype=np.float64)
    b = np.ascontiguousarray(b0, dtype=np.float64)
    a, l, r = C.c_double(), C.c_double(), C.c_double()
    err = lib().ob_speed_max_b0(C.byref(cfg), iDim, _dp(s), _dp(b),
                                C.byref(a), C.byref(l), C.byref(r))
    return err, a.value, l.value, r.value


def numerical_flux_b0(cfg, iDim, area, left, right, b0):
    L = np.ascontiguousarray(left, dtype=np.float64)
    R = np.ascontiguousarray(right, dtype=np.float64)
    b = np.ascontiguousarray(b0, dtype=np.float64)
    F = np.zeros(14)
    err = lib().ob_numerical_flux_b0(C.byref(cfg), iDim, area, _dp(L),
                                     _dp(R), _dp(b), _dp(F))
    return err, F


def set_b0_face(cfg, b0_block):
    """Face B0 (B0X (nK, nJ, nI+1, 3), B0Y, B0Z) of one block from its
    ghosted B0_DGB (nKg, nJg, nIg, 3)."""
    blk = np.ascontiguousarray(b0_block, dtype=np.float64)
    out = [np.zeros(sh + (3,)) for sh in face_shapes(cfg)]
    lib().ob_set_b0_face(C.byref(cfg), _dp(blk), *[_dp(a) for a in out])
    return out


def set_b0_source(cfg, cell_size, b0_faces):
    """DivB0 (nK, nJ, nI) and CurlB0 (nK, nJ, nI, 3) from face B0."""
    cs = np.ascontiguousarray(cell_size, dtype=np.float64)
    div = np.zeros((cfg.nK, cfg.nJ, cfg.nI))
    curl = np.zeros((cfg.nK, cfg.nJ, cfg.nI, 3))
    lib().ob_set_b0_source(C.byref(cfg), _dp(cs),
                           *[_dp(np.ascontiguousarray(a)) for a in b0_faces],
                           _dp(div), _dp(curl))
    return div, curl


def calc_face_flux_b0(cfg, cell_size, faces, b0_faces):
    sx, sy, sz = face_shapes(cfg)
    fl = [np.zeros(s + (14,)) for s in (sx, sy, sz)]
    cs = np.ascontiguousarray(cell_size, dtype=np.float64)
    err = lib().ob_calc_face_flux_b0(
        C.byref(cfg), _dp(cs), *[_dp(a) for a in faces],
        *[_dp(np.ascontiguousarray(a)) for a in b0_faces],
        *[_dp(a) for a in fl])
    return err, fl


def calc_source_b0(cfg, cell_size, block_state, faces, b0_block, b0_faces,
                   div_b0, curl_b0, UseB0Source=True):
    cs = np.ascontiguousarray(cell_size, dtype=np.float64)
    blk = np.ascontiguousarray(block_state, dtype=np.float64)
    src = np.zeros((cfg.nK, cfg.nJ, cfg.nI, 9))
    lib().ob_calc_source_b0(
        C.byref(cfg), _dp(cs), _dp(blk),
        *[_dp(np.ascontiguousarray(a)) for a in faces],
        _dp(np.ascontiguousarray(b0_block, dtype=np.float64)),
        *[_dp(np.ascontiguousarray(a)) for a in b0_faces],
        _dp(np.ascontiguousarray(div_b0)), _dp(np.ascontiguousarray(curl_b0)),
        int(UseB0Source), _dp(src))
    return src


def body_face_state(true_state, xyz_face, BodyRho, BodyP, B1rCoef):
    t = np.ascontiguousarray(true_state, dtype=np.float64)
    x = np.ascontiguousarray(xyz_face, dtype=np.float64)
    g = np.zeros(8)
    lib().ob_body_face_state(_dp(t), _dp(x), BodyRho, BodyP, B1rCoef, _dp(g))
    return g


def energy_round_trip(cfg, state):
    s = np.array(state, dtype=np.float64)
    lib().ob_pressure_to_energy_cell(C.byref(cfg), _dp(s))
    e = s.copy()
    lib().ob_energy_to_pressure_cell(C.byref(cfg), _dp(s))
    return e, s


# ----------------------------------------------------------------------
# per-block routines
def face_shapes(cfg):
    nI, nJ, nK = cfg.nI, cfg.nJ, cfg.nK
    return ((nK, nJ, nI + 1), (nK, nJ + 1, nI), (nK + 1, nJ, nI))


def calc_face_value(cfg, block_state, crit=None, res_change=None):
    """Face states (LX, RX, LY, RY, LZ, RZ); crit = the block's low order
    weights (X, Y, Z) from set_low_order_face, or None; res_change = six
    flags (x-, x+, y-, y+, z-, z+) of the sides with DiLevel_EB /= 0, or
    None."""
    blk = np.ascontiguousarray(block_state, dtype=np.float64)
    sx, sy, sz = face_shapes(cfg)
    arrs = [np.zeros(s + (8,)) for s in (sx, sx, sy, sy, sz, sz)]
    if crit is None and res_change is None:
        lib().ob_calc_face_value(C.byref(cfg), _dp(blk),
                                 *[_dp(a) for a in arrs])
    else:
        dp = C.POINTER(C.c_double)
        cr = [None] * 3 if crit is None else \
            [np.ascontiguousarray(a, dtype=np.float64) for a in crit]
        rc = None if res_change is None else \
            np.ascontiguousarray(res_change, dtype=np.int32)
        lib().ob_calc_face_value_rc(
            C.byref(cfg), _dp(blk),
            *[dp() if a is None else _dp(a) for a in cr],
            None if rc is None else rc.ctypes.data_as(C.POINTER(C.c_int)),
            *[_dp(a) for a in arrs])
    return arrs


def set_low_order_face(cfg, block_state):
    """Low order weights of the block's faces, shapes X (nK, nJ, nI+1),
    Y (nK, nJ+1, nI), Z (nK+1, nJ, nI)."""
    blk = np.ascontiguousarray(block_state, dtype=np.float64)
    out = [np.zeros(s) for s in face_shapes(cfg)]
    lib().ob_set_low_order_face(C.byref(cfg), _dp(blk),
                                *[_dp(a) for a in out])
    return out


def calc_face_flux(cfg, cell_size, faces):
    sx, sy, sz = face_shapes(cfg)
    fl = [np.zeros(s + (14,)) for s in (sx, sy, sz)]
    cs = np.ascontiguousarray(cell_size, dtype=np.float64)
    err = lib().ob_calc_face_flux(C.byref(cfg), _dp(cs),
                                  *[_dp(a) for a in faces],
                                  *[_dp(a) for a in fl])
    return err, fl


def calc_source(cfg, cell_size, block_state, faces):
    """Source_VC of shape (nK, nJ, nI, 9) (zero unless UseDivbSource)."""
    cs = np.ascontiguousarray(cell_size, dtype=np.float64)
    blk = np.ascontiguousarray(block_state, dtype=np.float64)
    src = np.zeros((cfg.nK, cfg.nJ, cfg.nI, 9))
    lib().ob_calc_source(C.byref(cfg), _dp(cs), _dp(blk),
                         *[_dp(np.ascontiguousarray(a)) for a in faces],
                         _dp(src))
    return src


def update_state(cfg, iStage, dt, cell_size, dtmax_cb, fluxes, block_state,
                 block_old, source=None):
    cs = np.ascontiguousarray(cell_size, dtype=np.float64)
    src = None if source is None else np.ascontiguousarray(source)
    lib().ob_update_state(C.byref(cfg), iStage, dt, _dp(cs), _dp(dtmax_cb),
                          *[_dp(f) for f in fluxes],
                          None if src is None else _dp(src),
                          _dp(block_state), _dp(block_old))


def calc_timestep(cfg, cell_size, fluxes):
    cs = np.ascontiguousarray(cell_size, dtype=np.float64)
    dtcb = np.zeros((cfg.nK, cfg.nJ, cfg.nI))
    dtb = C.c_double()
    lib().ob_calc_timestep(C.byref(cfg), _dp(cs), *[_dp(f) for f in fluxes],
                           _dp(dtcb), C.byref(dtb))
    return dtcb, dtb.value


# ----------------------------------------------------------------------
class Session:
    """Whole-grid oracle run (BATS_advance-shaped time loop)."""

    def __init__(self, grid, cfg, state_vgb, nThread=0, B0_DGB=None,
                 UseB0Source=True, Used_GB=None, XyzStart_DB=None,
                 BodyRho=1.0, BodyP=1.0, B1rCoef=-1.0, VaryTable=None):
        """B0_DGB: B0 splitting, the static field (nBlock, nKg, nJg, nIg, 3)
        at the cell centres and ghost cells.  Used_GB:
        an inner body, the true cells (nBlock, nKg, nJg, nIg) of bool;
        XyzStart_DB (nBlock, 3) the centre of each block's first cell
        relative to the body centre; BodyRho, BodyP, B1rCoef the body's
        face state and radial field coefficient."""
        self.grid = grid
        self.cfg = cfg
        self.nei = np.ascontiguousarray(grid.iNeighbor_EB, dtype=np.int32)
        self.cs = np.ascontiguousarray(grid.CellSize_DB, dtype=np.float64)
        self.State = np.ascontiguousarray(state_vgb, dtype=np.float64).copy()
        self.StateOld = np.zeros_like(self.State)
        n = grid.nBlock * cfg.nI * cfg.nJ * cfg.nK
        self.DtMax_CB = np.zeros(n)
        self.DtMax_B = np.zeros(grid.nBlock)
        s = ObSession()
        s.c = cfg
        s.nBlock = grid.nBlock
        s.iNeighbor_EB = self.nei.ctypes.data_as(C.POINTER(C.c_int))
        s.CellSize_DB = _dp(self.cs)
        s.State = _dp(self.State)
        s.StateOld = _dp(self.StateOld)
        s.DtMax_CB = _dp(self.DtMax_CB)
        s.DtMax_B = _dp(self.DtMax_B)
        s.Dt = 0.0
        s.tSimulation = 0.0
        s.nStep = 0
        s.nThread = nThread
        slabs = getattr(grid, "iSlab_IS", None)
        self.nSlab = 0 if slabs is None else len(slabs)
        if self.nSlab:
            self.slabs = np.ascontiguousarray(slabs, dtype=np.int32)
            self.Corr = np.zeros(self.nSlab * cfg.nI * cfg.nJ * 4 * 14)
            s.nSlab = self.nSlab
            s.iSlab_IS = self.slabs.ctypes.data_as(C.POINTER(C.c_int))
            s.Corr = _dp(self.Corr)
        if B0_DGB is not None:
            self.B0 = np.ascontiguousarray(B0_DGB, dtype=np.float64)
            assert self.B0.shape[-1] == 3 and self.B0.shape[0] == grid.nBlock
            s.B0 = _dp(self.B0)
            s.UseB0Source = int(UseB0Source)
        if Used_GB is not None:
            self.Used = np.ascontiguousarray(Used_GB, dtype=np.uint8)
            assert self.Used.shape == self.State.shape[:-1]
            self.XyzStart = np.ascontiguousarray(XyzStart_DB,
                                                 dtype=np.float64)
            assert self.XyzStart.shape == (grid.nBlock, 3)
            s.Used = self.Used.ctypes.data_as(C.POINTER(C.c_ubyte))
            s.XyzStart_DB = _dp(self.XyzStart)
            s.BodyRho, s.BodyP, s.B1rCoef = BodyRho, BodyP, B1rCoef
        if VaryTable is not None:
            # (times, primitive states (nPoint, 8)) of the x+ vary side
            self.VaryTime = np.ascontiguousarray(VaryTable[0], np.float64)
            self.VaryState = np.ascontiguousarray(VaryTable[1], np.float64)
            assert self.VaryState.shape == (len(self.VaryTime), 8)
            s.nVary = len(self.VaryTime)
            s.VaryTime_I = _dp(self.VaryTime)
            s.VaryState_VI = _dp(self.VaryState)
        if cfg.UseAdaptiveLowOrder:
            sx, sy, sz = face_shapes(cfg)
            self.LowOrderCrit = np.zeros(
                grid.nBlock * sum(int(np.prod(a)) for a in (sx, sy, sz)))
            s.LowOrderCrit = _dp(self.LowOrderCrit)
        self.s = s

    def exchange(self):
        lib().ob_session_exchange(C.byref(self.s))

    def init_timestep(self):
        return lib().ob_init_timestep(C.byref(self.s))

    def face_values(self, b):
        """Face states (LX, RX, LY, RY, LZ, RZ) of block b (0-based) as the
        stage forms them, and the body-body face flags (SX, SY, SZ)."""
        sx, sy, sz = face_shapes(self.cfg)
        arrs = [np.zeros(s + (8,)) for s in (sx, sx, sy, sy, sz, sz)]
        skips = [np.zeros(s, np.uint8) for s in (sx, sy, sz)]
        lib().ob_session_face_values(
            C.byref(self.s), b, *[_dp(a) for a in arrs],
            *[a.ctypes.data_as(C.POINTER(C.c_ubyte)) for a in skips])
        return arrs, skips

    def sources(self):
        """Source_VC of every block (nBlock, nK, nJ, nI, 9) from the current
        state (ghost cells as last exchanged), with the resolution-change
        correction of the face fields."""
        n = self.grid.nBlock * self.cfg.nI * self.cfg.nJ * self.cfg.nK
        src = np.zeros(n * 9)
        err = lib().ob_session_sources(C.byref(self.s), _dp(src))
        assert err == 0
        return src.reshape(self.grid.nBlock, self.cfg.nK, self.cfg.nJ,
                           self.cfg.nI, 9)

    def b0(self, b):
        """Block b's face B0 as a stage uses it (B0X (nK, nJ, nI+1, 3), B0Y
        (nK, nJ+1, nI, 3), B0Z (nK+1, nJ, nI, 3)), DivB0 (nK, nJ, nI) and
        CurlB0 (nK, nJ, nI, 3)."""
        nI, nJ, nK = self.cfg.nI, self.cfg.nJ, self.cfg.nK
        X = np.zeros((nK, nJ, nI + 1, 3))
        Y = np.zeros((nK, nJ + 1, nI, 3))
        Z = np.zeros((nK + 1, nJ, nI, 3))
        D = np.zeros((nK, nJ, nI))
        Cu = np.zeros((nK, nJ, nI, 3))
        err = lib().ob_session_b0(C.byref(self.s), b, _dp(X), _dp(Y), _dp(Z),
                                  _dp(D), _dp(Cu))
        assert err == 0
        return X, Y, Z, D, Cu

    def set_global_timestep(self, tlimit=-1.0):
        lib().ob_set_global_timestep(C.byref(self.s), tlimit)
        return self.s.Dt

    def advance_explicit(self, DoCalcTimestep=True):
        return lib().ob_advance_explicit(C.byref(self.s), int(DoCalcTimestep))

    def advance_stage(self, iStage, DoCalcTimestep=True):
        return lib().ob_advance_stage(C.byref(self.s), iStage,
                                      int(DoCalcTimestep))

    def new_step(self, tlimit=-1.0):
        """nStep += 1 and set_global_timestep, as BATS_advance starts a
        step (src/ModBatsrusMethods.f90:566-572)"""
        self.s.nStep += 1
        return self.set_global_timestep(tlimit)

    def advance_step(self, tlimit=-1.0):
        return lib().ob_advance_step(C.byref(self.s), tlimit)

    @property
    def Dt(self):
        return self.s.Dt

    @property
    def tSimulation(self):
        return self.s.tSimulation

    @property
    def nStep(self):
        return self.s.nStep

    def soa(self):
        return self.grid.state_vgb_to_soa(self.State)
