"""The C-ABI library loads, exports every entry point include/higflow_hip.h
declares, and (host-only mode, no device work) builds the same point
stencils as the oracle."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import hfo
from higflow_amd import cases, native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "higflow_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hfh_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = native.load()
    syms = declared_symbols()
    assert len(syms) >= 40
    for s in syms:
        assert hasattr(L, s), s


def test_error_paths_without_compute():
    L = native.load()
    assert L.hfh_create(5) is None
    assert b"dim" in L.hfh_last_error()
    assert L.hfh_version() >= 100
    assert L.hfh_device_count() >= 0


@pytest.fixture
def host_only(monkeypatch):
    monkeypatch.setenv("HFH_HOST_ONLY", "1")


@pytest.mark.parametrize("make", [lambda: cases.cavity(16), lambda: cases.newt_channel(nx=40, ny=10),
                                  lambda: cases.cavity(4, dim=3),
                                  # 8^3: more candidates than maxpts with many equal distances, so
                                  # the candidate heap fills and the bounded tree walk prunes
                                  lambda: cases.cavity(8, dim=3),
                                  lambda: cases.wall_layer_cavity(4, dim=3, levels=1),
                                  lambda: cases.contraction(levels=2, h0=0.5, length=8.0),
                                  lambda: cases.newt_channel(nx=40, ny=10, tempdiscrtype=3)])
def test_host_stencils_match_oracle(host_only, make):
    case = make()
    o = hfo.Oracle(case)
    if True:
        o.step()  # fill the oracle's point caches in step order, as the recorder does
    h = native.Flow(case)
    dim = case["dim"]
    assert h.pinned_cell() == o.pinned_cell()
    for f, c in [(0, d) for d in range(dim)] + [(1, 0)]:
        assert np.array_equal(h.centers(f, c), o.centers(f, c))
    cen = o.centers(1)
    lo, hi = cen.min(0), cen.max(0)
    dx = (hi - lo) / max(1, round(len(cen) ** (1.0 / dim)) - 1)
    rng = np.random.default_rng(1)
    nq = 0
    for c in cen[rng.choice(len(cen), size=min(25, len(cen)), replace=False)]:
        for off in rng.choice([-1.0, -0.5, 0.0, 0.5, 1.0], size=(6, dim)):
            x = c + off * dx
            for dom in (0, 1, 2):
                for d in (range(dim) if dom == 1 else (0,)):
                    ia, wa, ra = o.stencil(dom, d, c, x)
                    ib, wb, rb = h.stencil(dom, d, c, x)
                    assert np.array_equal(ia, ib)
                    assert np.allclose(wa, wb, rtol=0, atol=1e-14)
                    assert abs(ra - rb) < 1e-14
                    nq += 1
    assert nq > 100


def test_every_binding_has_argtypes():
    # a ctypes call without argtypes truncates pointers to 32 bits
    L = native.load()
    for s in declared_symbols():
        fn = getattr(L, s)
        assert fn.argtypes is not None, s


def test_host_only_calls(host_only):
    h = native.Flow(cases.cavity(8))
    h.set_timing(True)
    h.set_timing(False)
    h.reset_timing()
    assert set(h.stage_bytes()) == set(native.STAGES)
    assert h.probe(0, 1) is None  # no device: refused, not crashed
    with pytest.raises(native.HipError):
        h.step()


def test_malformed_boundary_is_an_error_not_an_abort():
    # a Dirichlet boundary tree reaching past the domain: a ghost point near
    # the corner projects onto it outside every domain cell.  The reference
    # asserts there (domain.c:939-940); the library reports it through the
    # entry point and the caller's process goes on.
    from higflow_amd import cases
    c = cases.ve_cavity(8)
    c["bcs"][1]["amr"] = cases.amr_boundary([1.0, 0.0], [1.0, 2.0], [1, 16])
    with pytest.raises(native.HipError, match="boundary projection outside the domain"):
        native.Flow(c, host_only=True)
    # the library still works afterwards
    native.Flow(cases.ve_cavity(8), host_only=True)
