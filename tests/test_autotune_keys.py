"""Autotuner bookkeeping on the CPU: workload keys separate what changes
the kernel, candidate lists respect the grid size, the table round-trips
through its JSON file."""
import os
import tempfile

from batsrus_amd.autotune import Autotuner, candidates, workload_key
from batsrus_amd.grid import BlockGrid


def test_keys_and_candidates():
    g8 = BlockGrid(nRoot=(8, 8, 8), nCell=(8, 8, 8))
    g16 = BlockGrid(nRoot=(4, 4, 4), nCell=(16, 16, 16))
    k = workload_key(g8, dict(flux="hlld"))
    assert k != workload_key(g16, dict(flux="hlld"))
    assert k != workload_key(g8, dict(flux="hlld", precision="f32"))
    assert k != workload_key(g8, dict(flux="hlld"), nRun=3)
    assert k == workload_key(g8, dict(flux="hlld", beta=1.2))
    c = candidates(g8)                 # 128 groups: no persistent size fits
    assert {x["nWorkgroup"] for x in c} == {0}
    assert {x["graph"] for x in c} == {False, True}
    big = BlockGrid(nRoot=(16, 16, 16), nCell=(8, 8, 8))   # 1024 groups
    assert {x["nWorkgroup"] for x in candidates(big)} == {0, 256, 512}


def test_table_round_trip():
    g = BlockGrid(nRoot=(2, 2, 2), nCell=(8, 8, 8))
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "sub", "t.json")
        a = Autotuner(p)
        a.table[workload_key(g, {})] = dict(choice=dict(nWorkgroup=0,
                                                        graph=True))
        a.save()
        b = Autotuner(p)
        assert b.lookup(g, {})["choice"]["graph"] is True
