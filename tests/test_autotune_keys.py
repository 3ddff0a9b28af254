"""Autotuner bookkeeping on the CPU: workload keys separate what changes
the kernel, candidate lists respect the grid size, the table round-trips
through its JSON file."""
import os
import tempfile

from batsrus_amd.autotune import (Autotuner, candidates, domain_key,
                                  workload_key)
from batsrus_amd.grid import BlockGrid


def test_keys_and_candidates():
    g8 = BlockGrid(nRoot=(8, 8, 8), nCell=(8, 8, 8))
    g16 = BlockGrid(nRoot=(4, 4, 4), nCell=(16, 16, 16))
    k = workload_key(g8, dict(flux="hlld"))
    assert k != workload_key(g16, dict(flux="hlld"))
    assert k != workload_key(g8, dict(flux="hlld", precision="f32"))
    assert k != workload_key(g8, dict(flux="hlld"), nRun=3)
    assert k == workload_key(g8, dict(flux="hlld", beta=1.2))
    c = candidates(g8)                 # 128 groups: 128, 256, 512 items
    assert {(x["cells_per_thread"], x["nWorkgroup"]) for x in c} == {
        (8, 0), (4, 0), (2, 0), (2, 256)}
    assert {x["graph"] for x in c} == {False, True}
    assert {x["xcd"] for x in c} == {False, True}
    assert len(c) == len({tuple(sorted(x.items())) for x in c})
    big = BlockGrid(nRoot=(16, 16, 16), nCell=(8, 8, 8))   # 1024 groups
    assert {x["nWorkgroup"] for x in candidates(big)
            if x["cells_per_thread"] == 8} == {0, 256, 512}
    assert {x["cells_per_thread"] for x in candidates(g16)} == {16, 8, 4}
    assert domain_key(8 ** 6, dict(flux="hlld")) != \
        domain_key(8 ** 6, dict(flux="hlld", precision="f32"))


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
