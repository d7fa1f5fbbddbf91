"""GPU parity on the full-size contraction meshes against fixtures of the CPU
oracle (tests/golden/make_contraction_fixture.py writes them; the oracle
needs minutes per step on these meshes):

* the 3-level mesh of the 4:1 contraction (47,800 cells), one explicit step
  from a seeded state, and the literal 4-level mesh (163,000 cells), the
  same: u*, the velocity gradient, the new log-conformation and the
  polymeric tensor, every value the step computes before a linear solve,
  to 1e-12 of each field's scale;
* the 3-level mesh after 3 semi-implicit steps from rest at dt 1e-4.  Here
  the fields pass through the pressure and velocity solves, and on this
  mesh an FP64 solve cannot push the true relative residual much below
  1e-11 (the oracle's BiCGStab+ILU(0) and the library's AMG-BiCGStab both
  stop there), so each field may differ by 1e-10 of its scale or by 10x the
  oracle's own solver-noise floor (its difference to a run with 10x looser
  solver tolerances), whichever is larger.

The sampled values and the L2 norms are compared; velocity components use
the velocity scale.
"""
import os
import sys

import numpy as np
import pytest

from higflow_amd import native

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_contraction_fixture as fxm  # noqa: E402

pytestmark = pytest.mark.gpu


def check(sim, fx, fields, tol=1e-10):
    vel = [name for name, f, c in fields if f in (0, 3)]
    uscale = max(float(fx[n + "_max"]) for n in vel) if vel else 0.0
    worst = {}
    for name, f, c in fields:
        v = sim.get(f, c)
        assert len(v) == int(fx[name + "_n"]), name
        scale = uscale if f in (0, 3) else float(fx[name + "_max"])
        d = np.abs(v[fx[name + "_idx"]] - fx[name + "_val"]).max() / scale
        l2 = abs(np.linalg.norm(v) - float(fx[name + "_l2"])) / max(float(fx[name + "_l2"]), 1e-300)
        allowed = tol
        if name + "_noise" in fx:
            allowed = max(tol, 10.0 * float(fx[name + "_noise"]) / scale)
        worst[name] = (d, l2, allowed)
        assert d <= allowed and l2 <= allowed, (name, d, l2, allowed)
    return worst


@pytest.mark.parametrize("which,ncells", [("c2seed", 47800), ("c3", 163000)])
def test_contraction_first_step_before_the_solves(which, ncells):
    fx = np.load(os.path.join(HERE, "golden", fxm.SPEC[which]["out"]))
    h = native.Flow(fxm.case(which), maxit=50000)
    assert h.size(1) == ncells
    fxm.seeded_state(h)
    h.step()
    # the pressure solve itself runs to its attainable floor (true relative
    # residual about 3e-10 on the 3-level mesh, 3e-9 on the literal 4-level
    # one, where the oracle's BiCGStab+ILU(0) makes no progress at all)
    st = h.solve_stats(0)
    assert st["resid"] < 1e-8, st
    print(st, check(h, fx, fxm.SPEC[which]["fields"], tol=1e-12))


def test_contraction_3lev_three_steps():
    fx = np.load(os.path.join(HERE, "golden", "contraction_3lev_3steps.npz"))
    h = native.Flow(fxm.case("c2"))
    for _ in range(int(fx["steps"])):
        h.step()
        for w in range(3):
            assert h.solve_stats(w)["resid"] < 1e-9, (w, h.solve_stats(w))
    print(check(h, fx, fxm.SPEC["c2"]["fields"]))
