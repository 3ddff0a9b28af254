"""Host-side checks that run without a GPU: the C ABI library loads and
exports every entry point declared in include/bats_hip.h, the grid and
ghost-fill logic match the oracle, and the seeded generator reproduces the
published Philox-4x32-10 known-answer vectors."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import pyoracle as po
from batsrus_amd import core, ics
from batsrus_amd.grid import BlockGrid
from helpers import random_mhd_soa

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_functions():
    txt = open(os.path.join(ROOT, "include", "bats_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(bats_[a-z0-9_]+)\s*\(", txt)))


def test_abi_exports_every_declared_symbol():
    lib = C.CDLL(core.library_path())
    names = _header_functions()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n
    # the Python mirror binds exactly the declared set
    assert set(names) == set(core.ABI_FUNCTIONS), \
        set(names) ^ set(core.ABI_FUNCTIONS)


def test_abi_strerror_and_bad_config():
    lib = core.load_library()
    assert lib.bats_strerror(0) == b"ok"
    assert b"HLLD" in lib.bats_strerror(3)
    grid = BlockGrid(nRoot=(1, 1, 1), nCell=(8, 8, 8))
    cfg = core.make_config(grid)
    cfg.nI = 12  # unsupported block size is rejected before any device use
    h = C.c_void_p()
    assert lib.bats_init(C.byref(cfg), C.byref(h)) == 1
    assert lib.bats_finalize(None) == 1


def test_fortran_bridge_declares_every_entry_point():
    """The ISO_C_BINDING module binds each C entry point by its name."""
    src = open(os.path.join(ROOT, "batsrus_amd", "fortran",
                            "ModHipBridge.f90")).read().lower()
    for n in _header_functions():
        if n == "bats_strerror":
            continue
        assert f'bind(c, name="{n}")' in src, n


def test_philox_known_answers():
    """Random123 known-answer vectors for philox4x32-10."""
    kat = [((0, 0, 0, 0), (0, 0),
            (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
           ((0xffffffff,) * 4, (0xffffffff, 0xffffffff),
            (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
           ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344),
            (0xa4093822, 0x299f31d0),
            (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1))]
    for ctr, key, out in kat:
        got = ics.philox4x32_10(np.array([ctr], dtype=np.uint64), key)[0]
        assert tuple(int(v) for v in got) == out


def test_ics_independent_of_block_layout():
    """The same physical cell gets the same value for any block size or
    block order."""
    g1 = BlockGrid(nRoot=(4, 4, 4), nCell=(8, 8, 8))
    g2 = BlockGrid(nRoot=(2, 2, 2), nCell=(16, 16, 16), order="morton")

    def to_global(g, U):
        nb = g.nRoot
        out = np.zeros((8, nb[2] * g.nK, nb[1] * g.nJ, nb[0] * g.nI))
        for b in range(g.nBlock):
            ib, jb, kb = g.iRoot_DB[b]
            out[:, kb * g.nK:(kb + 1) * g.nK, jb * g.nJ:(jb + 1) * g.nJ,
                ib * g.nI:(ib + 1) * g.nI] = U[:, b]
        return out
    for gen in (ics.brio_wu, ics.orszag_tang_z, ics.kelvin_helmholtz):
        a = to_global(g1, gen(g1))
        b = to_global(g2, gen(g2))
        assert np.array_equal(a, b), gen.__name__


def test_neighbour_table():
    g = BlockGrid(nRoot=(3, 2, 2), nCell=(8, 8, 8),
                  periodic=(False, True, True), order="morton")
    for b in range(g.nBlock):
        ib, jb, kb = g.iRoot_DB[b]
        for side in range(6):
            d, hi = divmod(side, 2)
            q = [ib, jb, kb]
            q[d] += 1 if hi else -1
            n = g.iNeighbor_EB[b, side]
            if d == 0 and not (0 <= q[0] < 3):
                assert n == -1
                continue
            q[d] %= g.nRoot[d]
            assert tuple(g.iRoot_DB[n]) == tuple(q)


def test_fill_ghosts_matches_oracle_exchange():
    g = BlockGrid(nRoot=(3, 2, 2), nCell=(8, 8, 8),
                  periodic=(False, True, True))
    U = random_mhd_soa(g, seed=11)
    S1 = g.soa_to_state_vgb(U)
    g.fill_ghosts(S1)
    cfg = po.config_from_dict({}, g)
    s = po.Session(g, cfg, g.soa_to_state_vgb(U))
    s.exchange()
    assert np.array_equal(S1, s.State)


@pytest.mark.parametrize("field,value", [
    ("nI", 12), ("nDim", 2), ("nG", 1), ("nBlock", 0), ("nOrder", 3),
    ("iLimiter", 9), ("iFlux", 0), ("nStage", 3), ("Gamma", 1.0),
    ("Cfl", 0.0), ("UseDivbSource", 2), ("StatePrecision", 2),
    ("UseAdaptiveLowOrder", 2)])
def test_bad_config_rejected_before_device_use(field, value):
    """Configurations the reference stops on (or this path cannot run) give
    BATS_EINVAL from bats_init without touching a device."""
    lib = core.load_library()
    cfg = core.make_config(BlockGrid(nRoot=(1, 1, 1), nCell=(8, 8, 8)))
    setattr(cfg, field, value)
    h = C.c_void_p()
    assert lib.bats_init(C.byref(cfg), C.byref(h)) == 1, field
    assert not h.value


@pytest.mark.parametrize("kw", [
    dict(nLowOrder=3), dict(pCritLow=1.2, pCritHigh=1.5),
    dict(VelCrit=-1.0)])
def test_bad_low_order_config_rejected(kw):
    """#ADAPTIVELOWORDER checks (src/ModFaceValue.f90:222-225)."""
    lib = core.load_library()
    cfg = core.make_config(BlockGrid(nRoot=(1, 1, 1), nCell=(8, 8, 8)),
                           UseAdaptiveLowOrder=True, **kw)
    h = C.c_void_p()
    assert lib.bats_init(C.byref(cfg), C.byref(h)) == 1


def test_null_handle_rejected():
    lib = core.load_library()
    d = C.c_double()
    assert lib.bats_advance(None, 1, -1.0) == 1
    assert lib.bats_advance_stage(None, 1, 1) == 1
    assert lib.bats_advance_explicit(None, 1) == 1
    assert lib.bats_set_global_timestep(None, -1.0, C.byref(d)) == 1
    assert lib.bats_calc_face_value(None, 1, 0, 1) == 1
    assert lib.bats_set_low_order_face(None) == 1
    assert lib.bats_set_step_graph(None, 1) == 1
    assert lib.bats_set_runs(None, 1, None) == 1
