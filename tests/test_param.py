"""PARAM.in reader (batsrus_amd.param) on the reference's parameter file of
the 1-D shock tube (tests/golden/PARAM.in.hyp.1d.gz) and on the command
rules of src/ModSetParameters.f90."""
import gzip
import os

import numpy as np
import pytest

from batsrus_amd import param

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _hyp1d():
    with gzip.open(os.path.join(GOLDEN, "PARAM.in.hyp.1d.gz"), "rt") as f:
        return param.read_param(f.read())


def test_reads_shocktube_param():
    p = _hyp1d()
    assert p.nCell == (64, 1, 1) and p.nRoot == (16, 1, 1) and p.nDim == 1
    assert p.xyzMin == (-128.0, -0.125, -0.125)
    assert p.xyzMax == (128.0, 0.125, 0.125)
    assert p.TypeCellBc == ["float", "float"]
    # #SCHEME after #TIMESTEPPING sets nStage = nOrder; #FIXEDTIMESTEP
    # sets Cfl = 1 (src/ModSetParameters.f90:1748-1750,4314-4319)
    # #UPDATE fast switches the dB trick off (src/ModSetParameters.f90:
    # 3873-3892)
    assert p.scheme == dict(gamma=2.0, nStage=1, cfl=1.0, UseHalfStep=True,
                            UseDtFixed=True, DtFixed=0.05, nOrder=1,
                            flux="rusanov", beta=1.0, limiter="no",
                            UseDbTrick=False)
    assert (p.MaxIteration, p.tSimulationMax, p.TypeUpdate) == (-1, 25.6,
                                                                 "fast")
    assert np.array_equal(p.ShockLeft, [1, 0, 0, 0, 0.75, 1, 0, 0, 1])
    assert np.array_equal(p.ShockRight, [0.125, 0, 0, 0, 0.75, -1, 0, 0, 0.1])
    assert "#SAVEPLOT" in p.ignored and "#SHOCKTUBE" not in p.ignored


def test_grid_and_state_from_param():
    p = _hyp1d()
    g = param.grid_from_param(p, nCell=(8, 8, 8))
    assert g.nRoot == (128, 1, 1) and g.periodic == (False, True, True)
    args = param.session_args(p, g)
    assert args["TypeCellBc"] == ["float", "float"] + ["periodic"] * 4
    U = param.shocktube_state(p, g)
    rho = U[0, :, 0, 0, :].reshape(-1)
    assert np.all(rho[:512] == 1.0) and np.all(rho[512:] == 0.125)
    assert np.all(U[5, :64] == 1.0) and np.all(U[5, 64:] == -1.0)
    assert np.all(U[4] == 0.75) and np.all(U[1:4] == 0.0)
    assert np.array_equal(U[:, :, 3], U[:, :, 0])


def test_scheme_rules():
    p = param.read_param("#TIMESTEPPING\n2\t\tnStage\n0.8\t\tCflExpl\n\n"
                         "#SCHEME\n2\t\tnOrder\nLinde\t\tTypeFlux\n"
                         "mc3\t\tTypeLimiter\n1.2\t\tLimiterBeta\n\n"
                         "#RK\n3\t\tnStage\n0.6\t\tCflExpl\n\n"
                         "#TIMEACCURATE\nF\t\tIsTimeAccurate\n\n"
                         "#ADAPTIVELOWORDER\nT\n2\n2.5\n1.5\n3.0\n")
    # #RK (after #SCHEME) keeps nStage 3 and UseHalfStep false
    assert p.scheme["flux"] == "linde" and p.scheme["limiter"] == "mc3"
    assert p.scheme["beta"] == 1.2 and p.scheme["nStage"] == 3
    assert p.scheme["UseHalfStep"] is False and p.scheme["cfl"] == 0.6
    assert p.scheme["IsSteadyState"] is True
    assert (p.scheme["nLowOrder"], p.scheme["pCritLow"], p.scheme["pCritHigh"],
            p.scheme["VelCrit"]) == (2, 2.5, 1.5, 3.0)
    q = param.read_param("#SCHEME\n2\nHLLD\nminmod\n\n#END\n#GAMMA\n1.4\n")
    assert q.scheme["limiter"] == "minmod" and q.scheme["beta"] == 1.0
    assert "gamma" not in q.scheme
    with pytest.raises(ValueError):
        param.read_param("#SCHEME\n1\nRoe\n")
    with pytest.raises(ValueError):
        param.read_param("#DIVB\nF\nF\nT\nF\n")
    with pytest.raises(ValueError):  # too few parameters
        param.read_param("#GRID\n2\n2\n")


def test_dbtrick_default_and_update_type():
    """UseDbTrick defaults to true (src/ModAdvance.f90:41), #DBTRICK sets
    it, and a TypeUpdate of 'fast' or 'slow' switches it off."""
    assert param.read_param("#GAMMA\n1.4\tGamma\n")\
        .scheme["UseDbTrick"] is True
    assert param.read_param("#DBTRICK\nF\tUseDbTrick\n")\
        .scheme["UseDbTrick"] is False
    for t, want in (("orig", True), ("slow", False), ("fast", False)):
        p = param.read_param(f"#UPDATE\n{t}\tTypeUpdate\n\n"
                             "#DBTRICK\nT\tUseDbTrick\n")
        assert p.scheme["UseDbTrick"] is want, t
