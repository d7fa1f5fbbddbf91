"""Batched parameter sweeps: the seeded sweep table, and the host-side merge
of member operators into block-diagonal systems (no device work)."""
import json
import os

import numpy as np
import pytest

from higflow_amd import cases, native

HERE = os.path.dirname(os.path.abspath(__file__))


def test_pcg32_known_answer():
    # pcg_basic.c's demo (pcg32_srandom_r(&rng, 42u, 54u)), first six outputs
    g = cases.Pcg32(42, 54)
    assert [g.next() for _ in range(6)] == [0xa15c02b7, 0x7b47f409, 0xba1d3330, 0x83d2f293, 0xbfa4784b,
                                            0xcbed606e]


def test_pcg32_bounded_is_unbiased_and_in_range():
    g = cases.Pcg32(7, 3)
    v = [g.bounded(7) for _ in range(7000)]
    assert min(v) == 0 and max(v) == 6
    counts = np.bincount(v, minlength=7)
    assert counts.min() > 850 and counts.max() < 1150


def test_sweep_table_matches_fixture():
    fx = json.load(open(os.path.join(HERE, "golden", "sweep64_seed0.json")))["rows"]
    t = cases.sweep_table(64, 0)
    assert [(r["N"], r["Re"], r["Wi"]) for r in fx] == t
    assert all(n in cases.SWEEP_SIDES for n, _, _ in t)
    assert all(1.0 <= re <= 100.0 and 0.1 <= wi <= 2.0 for _, re, wi in t)
    # about 64 x E[N^2] = 64 x 19346 cells
    assert abs(sum(n * n for n, _, _ in t) / (64 * 19346.3) - 1) < 0.1


@pytest.fixture
def host_only(monkeypatch):
    monkeypatch.setenv("HFH_HOST_ONLY", "1")


def _members():
    return [cases.ve_cavity(n, Re=re, Wi=wi, beta=0.5, dt=5e-4, tempdiscrtype=cases.SEMI_IMPLICIT_EULER)
            for n, re, wi in ((16, 1.0, 0.5), (24, 10.0, 1.0), (8, 50.0, 2.0))]


@pytest.mark.parametrize("which", [0, 1, 2])
def test_batch_systems_are_member_blocks(host_only, which):
    cs = _members()
    singles = [native.Flow(c, host_only=True) for c in cs]
    b = native.Batch(cs)
    rp, ci, v, sym = b.operator(which)
    off = 0
    for s in singles:
        r2, c2, v2, s2 = s.operator(which)
        n = len(r2) - 1
        assert np.array_equal(rp[off:off + n + 1] - rp[off], r2)
        assert np.array_equal(ci[rp[off]:rp[off + n]] - off, c2)
        assert np.array_equal(v[rp[off]:rp[off + n]], v2)
        # padding rows up to the next multiple of 1024 are empty
        pad = (n + 1023) // 1024 * 1024
        assert np.all(np.diff(rp[off + n:off + pad + 1]) == 0)
        assert sym == s2 or not s2
        off += pad
    assert off == len(rp) - 1


def test_batch_rejects_members_with_other_schemes(host_only):
    cs = _members()
    cs[1] = cases.ve_cavity(24, Re=10.0, Wi=1.0, beta=0.5, dt=1e-4, tempdiscrtype=cases.SEMI_IMPLICIT_EULER)
    with pytest.raises(native.HipError, match="differ"):
        native.Batch(cs)


def test_batch_entry_points_reject_bad_arguments(host_only):
    L = native.load()
    C = native.C
    assert L.hfh_create_batch(2, None, 0) is None
    assert b"member" in L.hfh_last_error()
    m2 = native.Flow(_members()[0], host_only=True)
    m3 = native.Flow(cases.cavity(4, dim=3), host_only=True)
    arr = (C.c_void_p * 2)(m2.h, m3.h)
    assert L.hfh_create_batch(2, arr, 2) is None
    assert b"dimension" in L.hfh_last_error()
    # a solver that is not set up
    raw = L.hfh_create_host(2)
    arr = (C.c_void_p * 1)(raw)
    assert L.hfh_create_batch(2, arr, 1) is None
    L.hfh_destroy(raw)
    # batch accessors on a single run fail cleanly
    assert L.hfh_batch_size(m2.h) == 0
    out = np.zeros(16)
    assert L.hfh_batch_get_field(m2.h, 0, 1, 0, out.ctypes.data_as(C.POINTER(C.c_double))) < 0
