"""IDL plot files (.out) and log files in the formats BATS-R-US writes, so
that runs of this library can be compared with the reference's regression
outputs by the same numeric diff.

The plot-file writer and reader of the reference (save_plot_file /
read_plot_file, module ModPlotFile of the SWMF share/Library, called from
src/ModPlotShell.f90:305-340 and src/ModBuffer.f90:344) are not part of the
reference tree; the format here is the one its committed outputs hold and
is pinned by them byte for byte:

ascii (Param/SHOCKTUBE/TestOutput/mhdhyp_1d_ref.out:1-5,
fastwave_2d_ref.out:1-5)
    line 1  StringHeader
    line 2  nStep (i10), Time (es18.10), nDim (i3, negative for a
            non-Cartesian grid), nParam (i3), nVar (i3)
    line 3  n_D (i8 each)
    line 4  Param_I (es18.10 each; absent when nParam = 0)
    line 5  NameVar: coordinate names, variable names, parameter names
    then one line per cell, first index fastest: the coordinates and
    the variables (es18.10 each)

real4 / real8 (Param/CORONA/B0local.out): Fortran sequential unformatted
records with 4-byte length markers, native (little) endian
    StringHeader (500 characters) | nStep, Time, nDim, nParam, nVar
    (int32, real, 3 x int32) | n_D (int32) | Param_I (real; absent when
    nParam = 0) | NameVar (500 characters) | Coord(n_D..., nDim) (real)
    | one record per variable, Var(n_D...) (real)
    real4 or real8 is told apart by the length of the second record
    (20 or 24 bytes), as the IDL procedures of the SWMF do.

Log files (src/ModWriteLogSatFile.f90:266-268,351-394): the line
'Volume averages, fluxes, etc', the names line ('it', ' t', then the log
variable names) and one line per output: nStep (i8), the time (es13.5)
and the variables (es14.5e3 each).
"""
from dataclasses import dataclass, field
import math
import struct

import numpy as np

_LSTRING = 500


# ----------------------------------------------------------------------
# Fortran edit descriptors
def fortran_es(v, w, d, e=None):
    """Fortran ESw.d (e=None) or ESw.dEe output of one real."""
    v = float(v)
    if math.isnan(v):
        return "NaN".rjust(w)
    if math.isinf(v):
        s = "Infinity" if v > 0 else "-Infinity"
        return s.rjust(w) if len(s) <= w else "*" * w
    mant, exp = f"{v:.{d}E}".split("E")
    ex = int(exp)
    sign = "+" if ex >= 0 else "-"
    if e is None:
        if abs(ex) <= 99:
            s = f"{mant}E{sign}{abs(ex):02d}"
        elif abs(ex) <= 999:
            s = f"{mant}{sign}{abs(ex):03d}"
        else:
            return "*" * w
    else:
        if abs(ex) >= 10 ** e:
            return "*" * w
        s = f"{mant}E{sign}{abs(ex):0{e}d}"
    return s.rjust(w) if len(s) <= w else "*" * w


def fortran_i(n, w):
    s = str(int(n))
    return s.rjust(w) if len(s) <= w else "*" * w


def _ftoken(s):
    """A real as Fortran writes it: ES without the letter for 3-digit
    exponents ('1.0-100')."""
    try:
        return float(s)
    except ValueError:
        for i in range(len(s) - 1, 0, -1):
            if s[i] in "+-" and s[i - 1] not in "eEdD":
                return float(s[:i] + "E" + s[i:])
        raise


# ----------------------------------------------------------------------
@dataclass
class PlotFile:
    """Contents of one IDL plot file. Coord has shape n_D + (nDim,) and Var
    n_D + (nVar,), first index fastest in the file."""
    StringHeader: str
    nStep: int
    Time: float
    n_D: tuple
    Coord: np.ndarray
    Var: np.ndarray
    NameVar: str
    Param: np.ndarray = field(default_factory=lambda: np.zeros(0))
    IsCartesian: bool = True
    TypeFile: str = "ascii"

    @property
    def nDim(self):
        return len(self.n_D)

    @property
    def nVar(self):
        return self.Var.shape[-1]


def _check_shapes(n_D, Coord, Var):
    n_D = tuple(int(n) for n in n_D)
    if Coord.shape != n_D + (len(n_D),):
        raise ValueError(f"Coord shape {Coord.shape} does not fit n_D {n_D}")
    if Var.shape[:-1] != n_D:
        raise ValueError(f"Var shape {Var.shape} does not fit n_D {n_D}")
    return n_D


def save_plot_file(NameFile, TypeFile="ascii", StringHeader="", nStep=0,
                   Time=0.0, Param=(), NameVar="", Coord=None, Var=None,
                   IsCartesian=True):
    """Write an IDL plot file. Coord: n_D + (nDim,); Var: n_D + (nVar,)."""
    Coord = np.asarray(Coord, dtype=np.float64)
    Var = np.asarray(Var, dtype=np.float64)
    n_D = _check_shapes(Coord.shape[:-1], Coord, Var)
    nDim, nVar = len(n_D), Var.shape[-1]
    Param = np.asarray(Param, dtype=np.float64).ravel()
    nDimOut = nDim if IsCartesian else -nDim
    ncell = int(np.prod(n_D))
    C2 = Coord.reshape(ncell, nDim, order="F")
    V2 = Var.reshape(ncell, nVar, order="F")
    if TypeFile == "ascii":
        es = lambda x: fortran_es(x, 18, 10)
        out = [StringHeader.rstrip(),
               fortran_i(nStep, 10) + es(Time) + fortran_i(nDimOut, 3)
               + fortran_i(len(Param), 3) + fortran_i(nVar, 3),
               "".join(fortran_i(n, 8) for n in n_D)]
        if len(Param):
            out.append("".join(es(p) for p in Param))
        out.append(NameVar.rstrip())
        rows = np.concatenate([C2, V2], axis=1)
        out.extend("".join(es(x) for x in r) for r in rows)
        with open(NameFile, "w") as f:
            f.write("\n".join(out) + "\n")
        return
    if TypeFile not in ("real4", "real8"):
        raise ValueError(f"TypeFile {TypeFile!r}: ascii, real4 or real8")
    rt = "<f4" if TypeFile == "real4" else "<f8"
    rc = "f" if TypeFile == "real4" else "d"

    def rec(b):
        n = struct.pack("<i", len(b))
        return n + b + n

    def text(s):
        b = s.encode("ascii")
        if len(b) > _LSTRING:
            raise ValueError("header or names longer than 500 characters")
        return b.ljust(_LSTRING, b" ")

    parts = [rec(text(StringHeader)),
             rec(struct.pack("<i" + rc + "iii", nStep, Time, nDimOut,
                             len(Param), nVar)),
             rec(np.asarray(n_D, dtype="<i4").tobytes())]
    if len(Param):
        parts.append(rec(Param.astype(rt).tobytes()))
    parts.append(rec(text(NameVar)))
    parts.append(rec(C2.astype(rt).tobytes(order="F")))
    for v in range(nVar):
        parts.append(rec(V2[:, v].astype(rt).tobytes()))
    with open(NameFile, "wb") as f:
        f.write(b"".join(parts))


def _records(b):
    p, out = 0, []
    while p < len(b):
        if p + 4 > len(b):
            raise ValueError("truncated record marker")
        n = struct.unpack_from("<i", b, p)[0]
        if n < 0 or p + 8 + n > len(b) or \
                struct.unpack_from("<i", b, p + 4 + n)[0] != n:
            raise ValueError(f"bad record at byte {p}")
        out.append(b[p + 4:p + 4 + n])
        p += 8 + n
    return out


def read_plot_file(NameFile):
    """Read an ascii, real4 or real8 IDL plot file."""
    with open(NameFile, "rb") as f:
        b = f.read()
    if len(b) >= 4 and struct.unpack_from("<i", b, 0)[0] == _LSTRING:
        r = _records(b)
        if len(r[1]) == 20:
            TypeFile, rt, rc = "real4", "<f4", "f"
        elif len(r[1]) == 24:
            TypeFile, rt, rc = "real8", "<f8", "d"
        else:
            raise ValueError("second record is neither 20 nor 24 bytes")
        nStep, Time, nDimOut, nParam, nVar = struct.unpack("<i" + rc + "iii",
                                                           r[1])
        n_D = tuple(int(n) for n in np.frombuffer(r[2], dtype="<i4"))
        if len(n_D) != abs(nDimOut):
            raise ValueError("n_D record does not match nDim")
        i = 3
        Param = np.zeros(0)
        if nParam > 0:
            Param = np.frombuffer(r[3], dtype=rt).astype(np.float64)
            i = 4
        NameVar = r[i].decode("ascii").rstrip()
        ncell = int(np.prod(n_D))
        C2 = np.frombuffer(r[i + 1], dtype=rt).astype(np.float64)
        Coord = C2.reshape(n_D + (len(n_D),), order="F")
        if len(r) != i + 2 + nVar:
            raise ValueError(f"{len(r) - i - 2} variable records, nVar {nVar}")
        V2 = np.stack([np.frombuffer(r[i + 2 + v], dtype=rt)
                       for v in range(nVar)], axis=1).astype(np.float64)
        if V2.shape[0] != ncell:
            raise ValueError("variable record length does not match n_D")
        Var = V2.reshape(n_D + (nVar,), order="F")
        return PlotFile(r[0].decode("ascii").rstrip(), nStep, float(Time),
                        n_D, Coord, Var, NameVar, Param, nDimOut > 0,
                        TypeFile)
    lines = b.decode("ascii").splitlines()
    head = lines[1].split()
    nStep, Time = int(head[0]), _ftoken(head[1])
    nDimOut, nParam, nVar = (int(x) for x in head[2:5])
    n_D = tuple(int(n) for n in lines[2].split())
    if len(n_D) != abs(nDimOut):
        raise ValueError("n_D line does not match nDim")
    i = 3
    Param = np.zeros(0)
    if nParam > 0:
        Param = np.array([_ftoken(x) for x in lines[3].split()])
        i = 4
    NameVar = lines[i]
    nDim, ncell = len(n_D), int(np.prod(n_D))
    rows = lines[i + 1:i + 1 + ncell]
    if len(rows) != ncell:
        raise ValueError(f"{len(rows)} data lines, n_D gives {ncell}")
    data = np.array([[_ftoken(x) for x in ln.split()] for ln in rows])
    if data.shape[1] != nDim + nVar:
        raise ValueError("columns do not match nDim + nVar")
    Coord = data[:, :nDim].reshape(n_D + (nDim,), order="F")
    Var = data[:, nDim:].reshape(n_D + (nVar,), order="F")
    return PlotFile(lines[0], nStep, Time, n_D, Coord, Var, NameVar, Param,
                    nDimOut > 0, "ascii")


# ----------------------------------------------------------------------
class LogFile:
    """A log file as src/ModWriteLogSatFile.f90 writes it (StringTime
    'step time')."""

    HEADER = "Volume averages, fluxes, etc"

    def __init__(self, NameFile, NameLogVar, StringTime="step time"):
        self.DoStep = "step" in StringTime
        self.DoTime = "time" in StringTime
        names = ""
        if "none" not in StringTime:
            if self.DoStep:
                names = "it"
            if self.DoTime:
                names = names + " t"
        else:
            self.DoStep = self.DoTime = False
        names = names + "".join(" " + n for n in NameLogVar)
        self.NameLogVar = list(NameLogVar)
        self.f = open(NameFile, "w")
        self.f.write(self.HEADER + "\n" + names.strip() + "\n")

    def write(self, nStep, Time, LogVar):
        if len(LogVar) != len(self.NameLogVar):
            raise ValueError("one value per log variable")
        s = ""
        if self.DoStep:
            s += fortran_i(nStep, 8)
        if self.DoTime:
            s += fortran_es(Time, 13, 5)
        s += "".join(fortran_es(v, 14, 5, 3) for v in LogVar)
        self.f.write(s + "\n")
        self.f.flush()

    def close(self):
        self.f.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def read_log_file(NameFile):
    """(names, rows) of a log file; rows is a float array, one row per
    line after the names line."""
    with open(NameFile) as f:
        lines = f.read().splitlines()
    names = lines[1].split()
    rows = [[_ftoken(x) for x in ln.split()] for ln in lines[2:] if ln.strip()]
    return names, np.array(rows)


# ----------------------------------------------------------------------
# log variables of a running session
DEFAULT_LOG_VARS = ("dt", "rho", "mx", "my", "mz", "bx", "by", "bz", "e",
                    "pmin", "pmax")
_INTEGRAL_INDEX = {"rho": 0, "mx": 1, "my": 2, "mz": 3, "bx": 4, "by": 5,
                   "bz": 6, "e": 7}


def log_vars(session, NameLogVar=DEFAULT_LOG_VARS, Dt=None):
    """Values of the log variables (src/ModWriteLogSatFile.f90:590-620):
    rho, mx..mz, bx..bz and e are domain integrals over the domain volume
    (the GPU's bats_integrate_domain sums), pmin / pmax the extreme
    pressures of the interior cells, dt the time step."""
    g = session.grid
    need_p = any(n.lower() in ("pmin", "pmax") for n in NameLogVar)
    need_i = any(n.lower() in _INTEGRAL_INDEX for n in NameLogVar)
    Sum_V = session.integrate_domain() if need_i else None
    vol = float(np.sum(np.prod(np.asarray(g.CellSize_DB)[:, :3], axis=1))
                * g.nI * g.nJ * g.nK)
    P = session.download_soa()[7] if need_p else None
    out = []
    for n in NameLogVar:
        k = n.lower()
        if k in _INTEGRAL_INDEX:
            out.append(Sum_V[_INTEGRAL_INDEX[k]] / vol)
        elif k == "pmin":
            out.append(float(P.min()))
        elif k == "pmax":
            out.append(float(P.max()))
        elif k == "dt":
            if Dt is None:
                raise ValueError("log variable dt needs Dt")
            out.append(float(Dt))
        else:
            raise ValueError(f"log variable {n!r} not supported")
    return out
