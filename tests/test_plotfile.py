"""IDL plot files and log files (batsrus_amd.plotfile) against the
reference's own output files (tests/golden/, copied by
make_idl_fixtures.py): each is read and written back byte for byte."""
import gzip
import os

import numpy as np
import pytest

from batsrus_amd import plotfile as pf

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _unzip(name, tmp_path):
    p = tmp_path / name[:-3]
    with gzip.open(os.path.join(GOLDEN, name), "rb") as f:
        p.write_bytes(f.read())
    return p


@pytest.mark.parametrize("name,kind,n_D,nVar,nPar,cart", [
    ("idl_mhdhyp_1d_ref.out.gz", "ascii", (1024,), 12, 3, True),
    ("idl_fastwave_2d_ref.out.gz", "ascii", (60, 60), 11, 2, False),
    ("idl_b0local_real4.out.gz", "real4", (31, 13, 8), 3, 0, True),
])
def test_plot_file_round_trip_is_byte_exact(name, kind, n_D, nVar, nPar,
                                            cart, tmp_path):
    src = _unzip(name, tmp_path)
    p = pf.read_plot_file(src)
    assert (p.TypeFile, p.n_D, p.nVar, len(p.Param), p.IsCartesian) == \
        (kind, n_D, nVar, nPar, cart)
    out = tmp_path / "out"
    pf.save_plot_file(out, p.TypeFile, p.StringHeader, p.nStep, p.Time,
                      p.Param, p.NameVar, p.Coord, p.Var, p.IsCartesian)
    assert out.read_bytes() == src.read_bytes()


def test_plot_file_layout_first_index_fastest(tmp_path):
    """The radius of the (r, lon, lat) grid of B0local grows with the first
    index and does not change with the others: the coordinate record holds
    the first index fastest, then one coordinate after the other."""
    b = pf.read_plot_file(_unzip("idl_b0local_real4.out.gz", tmp_path))
    r, lon, lat = (b.Coord[..., d] for d in range(3))
    assert np.all(np.diff(r, axis=0) > 0)
    assert np.all(np.diff(r, axis=1) == 0) and np.all(np.diff(r, axis=2) == 0)
    assert np.all(np.diff(lon, axis=1) > 0) and np.all(np.diff(lat, axis=2) > 0)


def test_plot_file_matches_npz_fixture(tmp_path):
    """The ascii reader gives the numbers of the shock-tube fixture."""
    p = pf.read_plot_file(_unzip("idl_mhdhyp_1d_ref.out.gz", tmp_path))
    ref = np.load(os.path.join(GOLDEN, "mhdhyp_1d_ref.npz"))
    data = np.concatenate([p.Coord.reshape(-1, 1), p.Var], axis=1)
    assert np.array_equal(data, ref["data"])
    assert p.nStep == int(ref["nstep"]) and p.Time == float(ref["time"])
    assert p.Param[0] == float(ref["gamma"])
    assert p.NameVar.split()[:13] == list(ref["names"])


@pytest.mark.parametrize("kind", ["ascii", "real4", "real8"])
def test_plot_file_write_read(kind, tmp_path):
    rng = np.random.default_rng(3)
    n_D = (5, 4, 3)
    Coord = np.stack(np.meshgrid(*[np.arange(n) + 0.5 for n in n_D],
                                 indexing="ij"), axis=-1)
    Var = rng.standard_normal(n_D + (8,)) * 10.0 ** rng.integers(-5, 5)
    Var[0, 0, 0, 0] = 1e-123    # three-digit exponent in ES18.10
    pf.save_plot_file(tmp_path / "a", kind, "normalized variables", 7, 1.25,
                      [5.0 / 3, 0.0], "x y z rho ux uy uz bx by bz p g c",
                      Coord, Var)
    p = pf.read_plot_file(tmp_path / "a")
    assert (p.TypeFile, p.n_D, p.nStep, p.Time) == (kind, n_D, 7, 1.25)
    assert np.array_equal(p.Coord, Coord)
    if kind == "real8":
        assert np.array_equal(p.Var, Var)
    elif kind == "real4":
        assert np.array_equal(p.Var, Var.astype(np.float32))
    else:  # 11 significant digits
        assert np.allclose(p.Var, Var, rtol=6e-11, atol=0)


def test_fortran_es_edit_descriptor():
    assert pf.fortran_es(-127.875, 18, 10) == " -1.2787500000E+02"
    assert pf.fortran_es(0.0, 18, 10) == "  0.0000000000E+00"
    assert pf.fortran_es(1e-123, 18, 10) == "  1.0000000000-123"
    assert pf.fortran_es(5.625e-1, 14, 5, 3) == "  5.62500E-001"
    assert pf.fortran_es(-8.30199e-15, 14, 5, 3) == " -8.30199E-015"
    assert pf.fortran_es(0.0, 13, 5) == "  0.00000E+00"
    assert pf._ftoken("1.0000000000-123") == 1e-123


def test_log_file_rows_are_byte_exact(tmp_path):
    """A log with StringTime 'step time' (Param/2BODYPLOT) rewritten from
    its values reproduces every line."""
    src = _unzip("log_2bodyplot.log.gz", tmp_path)
    names, rows = pf.read_log_file(src)
    assert names[:2] == ["it", "t"]
    out = tmp_path / "log"
    with pf.LogFile(out, names[2:]) as lf:
        for r in rows:
            lf.write(int(r[0]), r[1], r[2:])
    assert out.read_text() == src.read_text()


def test_log_vars_rejects_unknown_names():
    class _S:
        grid = type("G", (), {"CellSize_DB": np.ones((1, 3)), "nI": 1,
                              "nJ": 1, "nK": 1})()
    with pytest.raises(ValueError):
        pf.log_vars(_S(), ["dt"])
    with pytest.raises(ValueError):
        pf.log_vars(_S(), ["bogus"], Dt=1.0)
