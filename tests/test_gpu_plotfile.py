"""GPU run written as an IDL plot file and a log file, then compared with
the reference's regression output through the file format, as the
reference's test suite compares its own runs (DiffNum.pl -r=1e-5 -a=1e-11,
Makefile.test:319-322)."""
import gzip
import os

import numpy as np
import pytest

from batsrus_amd import ics
from batsrus_amd import plotfile as pf
from batsrus_amd.core import Session
from batsrus_amd.grid import BlockGrid

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_gpu_shocktube_idl_and_log(tmp_path):
    """Param/SHOCKTUBE/TestOutput/mhdhyp_1d_ref.out: the 1-D Brio-Wu run
    (first-order Rusanov, gamma=2, dt=0.05, 512 steps) on 128 blocks of 8^3
    along x, logged every 64 steps and saved as an ascii IDL file."""
    print("case gpu_shocktube_idl_and_log", flush=True)
    grid = BlockGrid(nRoot=(128, 1, 1), nCell=(8, 8, 8),
                     xyzMin=(-128.0, -1.0, -1.0), xyzMax=(128.0, 1.0, 1.0),
                     periodic=(False, True, True))
    U0 = ics.brio_wu(grid, noise=0.0, x0=0.0, left=(1.0, 1.0, 1.0),
                     right=(0.125, 0.1, -1.0), bx=0.75)
    log = tmp_path / "log_n000000.log"
    snaps = []
    with Session(grid, nOrder=1, flux="rusanov", nStage=1, cfl=1.0,
                 gamma=2.0, UseDtFixed=True, DtFixed=0.05) as s, \
            pf.LogFile(log, pf.DEFAULT_LOG_VARS) as lf:
        s.upload_soa(U0)
        for _ in range(8):
            s.advance(64, 25.6)
            dt, t, n = s.time()
            lf.write(n, t, pf.log_vars(s, Dt=dt))
            snaps.append(s.download_soa())
    U = snaps[-1]
    assert n == 512 and abs(t - 25.6) < 1e-12

    col = U[:, :, 0, 0, :].reshape(8, -1)
    rho = col[0]
    prim = np.stack([rho, col[1] / rho, col[2] / rho, col[3] / rho,
                     col[4], col[5], col[6], col[7]], axis=1)
    x = -128.0 + (np.arange(1024) + 0.5) * 0.25
    out = tmp_path / "mhdhyp_1d.out"
    pf.save_plot_file(out, "ascii", "normalized variables", n, t,
                      [2.0, 0.0, 0.0], "x Rho Ux Uy Uz Bx By Bz P g cuty cutz",
                      x.reshape(1024, 1), prim)

    mine = pf.read_plot_file(out)
    ref_path = tmp_path / "ref.out"
    with gzip.open(os.path.join(GOLDEN, "idl_mhdhyp_1d_ref.out.gz")) as f:
        ref_path.write_bytes(f.read())
    ref = pf.read_plot_file(ref_path)
    assert (mine.nStep, mine.Time, mine.n_D) == (ref.nStep, ref.Time, ref.n_D)
    assert np.array_equal(mine.Coord, ref.Coord)
    ref_names = ref.NameVar.split()[1:1 + ref.nVar]
    for i, nm in enumerate(mine.NameVar.split()[1:1 + mine.nVar]):
        a, r = mine.Var[:, i], ref.Var[:, ref_names.index(nm)]
        err = np.abs(a - r)
        assert np.all((err <= 1e-11) | (err <= 1e-5 * np.abs(r))), nm

    names, rows = pf.read_log_file(log)
    assert names == ["it", "t"] + list(pf.DEFAULT_LOG_VARS)
    assert rows.shape == (8, 13)
    assert np.array_equal(rows[:, 0], 64 * np.arange(1, 9))
    for k, Us in enumerate(snaps):
        p = Us[7]
        e = p / (2.0 - 1.0) + 0.5 * (Us[1] ** 2 + Us[2] ** 2 + Us[3] ** 2) \
            / Us[0] + 0.5 * (Us[4] ** 2 + Us[5] ** 2 + Us[6] ** 2)
        want = [0.05] + [Us[v].mean() for v in range(7)] + [e.mean(),
                                                            p.min(), p.max()]
        # the log prints 6 significant digits; By and Uz sum to about 0
        assert np.allclose(rows[k, 2:], want, rtol=6e-6, atol=1e-12), k


def test_gpu_shocktube_from_param_file(tmp_path):
    """The same run set up by the reference's own parameter file
    (Param/SHOCKTUBE/PARAM.in.hyp.1d) through batsrus_amd.param: grid,
    scheme, fixed step, float ends and the #SHOCKTUBE state; 64-cell 1-D
    blocks re-cut into 8^3 blocks (periodic in y and z)."""
    from batsrus_amd import param
    print("case gpu_shocktube_from_param_file", flush=True)
    with gzip.open(os.path.join(GOLDEN, "PARAM.in.hyp.1d.gz"), "rt") as f:
        p = param.read_param(f.read())
    grid = param.grid_from_param(p, nCell=(8, 8, 8))
    nStep = int(round(p.tSimulationMax / p.scheme["DtFixed"]))
    with Session(grid, **param.session_args(p, grid)) as s:
        s.upload_soa(param.shocktube_state(p, grid))
        s.advance(nStep, p.tSimulationMax)
        dt, t, n = s.time()
        U = s.download_soa()
    assert n == 512 and abs(t - 25.6) < 1e-12
    col = U[:, :, 0, 0, :].reshape(8, -1)
    ref_path = tmp_path / "ref.out"
    with gzip.open(os.path.join(GOLDEN, "idl_mhdhyp_1d_ref.out.gz")) as f:
        ref_path.write_bytes(f.read())
    ref = pf.read_plot_file(ref_path)
    names = ref.NameVar.split()[1:1 + ref.nVar]
    mine = {"Rho": col[0], "Ux": col[1] / col[0], "By": col[5],
            "Bx": col[4], "P": col[7]}
    for nm, a in mine.items():
        r = ref.Var[:, names.index(nm)]
        err = np.abs(a - r)
        assert np.all((err <= 1e-11) | (err <= 1e-5 * np.abs(r))), nm
