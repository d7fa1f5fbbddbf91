"""Write the CPU oracle's results on the full-size contraction meshes as
fixtures for the GPU parity tests (the oracle needs minutes for them, too
long for a test):

* ``contraction_3lev_3steps.npz`` -- the 2D Oldroyd-B 4:1 contraction on its
  3-level mesh (47,800 cells: base 0.1, three levels within 1.0 of the
  re-entrant corners; Re 0.01, Wi 1, beta 1/9, semi-implicit Euler,
  dt 1e-4) after 3 steps from rest, every linear solve converged to
  rtol 1e-13 / atol 1e-15.
  The fixture also keeps, per field, the oracle's own solver-noise floor:
  the largest difference between this run and a second oracle run whose
  solves stop at rtol 1e-12 / atol 1e-14.  On this mesh FP64 solves cannot
  reach a true relative residual much below 1e-11 (b - A x itself carries
  eps |A| |x| of rounding), so two correct solvers leave differences of that
  size, amplified where the log-conformation has nearly equal eigenvalues
  (the 1/(lambda_j - lambda_i + 1e-12) of hig-flow-step-viscoelastic.c:
  1667-1680).
* ``contraction_3lev_1step_seeded.npz`` -- the 3-level mesh, one explicit
  Euler step at dt 1e-4 from the seeded state of ``seeded_state``: the
  fields computed before any linear solve.
* ``contraction_4lev_1step.npz`` -- the same contraction on its literal
  4-level mesh (163,000 cells), one explicit Euler step at dt 1e-5 from the
  seeded state of ``seeded_state`` below.  The fields kept are the ones the
  step computes before any linear solve (u*, the velocity gradient, the new
  log-conformation and the polymeric tensor), so they do not depend on the
  pressure solve, which BiCGStab+ILU(0) does not converge on this mesh.

Per field the fixture keeps the L2 norm, the largest magnitude and the
values at 2,048 seeded sample indices.  Run from the repository root:
python tests/golden/make_contraction_fixture.py [c2|c3]
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
HERE = os.path.dirname(os.path.abspath(__file__))
SAMPLES = 2048


def case(which, loose=False):
    from higflow_amd import cases
    if which == "c2":
        tol = dict(rtol=1e-12, atol=1e-14) if loose else dict(rtol=1e-13, atol=1e-15)
        return cases.contraction(levels=3, dt=1e-4, maxit=50000, **tol)
    if which == "c2seed":
        return cases.contraction(levels=3, dt=1e-4, tempdiscrtype=cases.EXPLICIT_EULER, rtol=1e-13, atol=1e-15,
                                 maxit=1)
    return cases.contraction(levels=4, dt=1e-5, tempdiscrtype=cases.EXPLICIT_EULER, rtol=1e-13, atol=1e-15,
                             maxit=1)


def seeded_state(sim):
    """Inflow-like velocity with seeded noise, a seeded pressure and a
    seeded symmetric log-conformation (same values on any machine)."""
    rng = np.random.default_rng(2024)
    for d in range(2):
        x = sim.centers(0, d)
        base = 1.5 * (1.0 - (x[:, 1] / 4.0) ** 2) if d == 0 else 0.0
        sim.set(0, d, base + 0.05 * rng.standard_normal(len(x)))
    sim.set(1, 0, 0.1 * rng.standard_normal(sim.size(1)))
    for m in (0, 1, 3):
        v = 0.1 * rng.standard_normal(sim.size(6, m))
        sim.set(6, m, v)
        if m == 1:
            sim.set(6, 2, v)


PRESOLVE = [("us0", 3, 0), ("us1", 3, 1), ("D00", 8, 0), ("D01", 8, 1), ("D10", 8, 2), ("D11", 8, 3),
            ("K00", 6, 0), ("K01", 6, 1), ("K11", 6, 3), ("S00", 7, 0), ("S01", 7, 1), ("S10", 7, 2), ("S11", 7, 3)]
SPEC = {
    "c2": dict(out="contraction_3lev_3steps.npz", steps=3, seeded=False, noise=True,
               fields=[("u0", 0, 0), ("u1", 0, 1), ("p", 1, 0), ("K00", 6, 0), ("K01", 6, 1), ("K11", 6, 3),
                       ("S00", 7, 0), ("S01", 7, 1), ("S10", 7, 2), ("S11", 7, 3)]),
    "c2seed": dict(out="contraction_3lev_1step_seeded.npz", steps=1, seeded=True, noise=False, fields=PRESOLVE),
    "c3": dict(out="contraction_4lev_1step.npz", steps=1, seeded=True, noise=False, fields=PRESOLVE),
}


def run(which, loose=False):
    import hfo
    spec = SPEC[which]
    o = hfo.Oracle(case(which, loose))
    if spec["seeded"]:
        seeded_state(o)
    for k in range(spec["steps"]):
        t0 = time.time()
        rc = o.step()
        print("step %d rc %d (%.1f s) last pressure solve %s" % (k, rc, time.time() - t0, o.last_solve()),
              flush=True)
        if rc != 0 and not spec["seeded"]:
            raise SystemExit("a solve did not converge")
    return o


def main(which):
    spec = SPEC[which]
    o = run(which)
    o2 = run(which, loose=True) if spec["noise"] else None
    rng = np.random.default_rng(0)
    data = {"steps": np.array(spec["steps"])}
    for name, f, c in spec["fields"]:
        v = o.get(f, c)
        if o2 is not None:
            # the solver-noise floor, on the scale the test compares with
            data[name + "_noise"] = np.array(np.abs(o2.get(f, c) - v).max())
        idx = np.sort(rng.choice(len(v), size=min(SAMPLES, len(v)), replace=False))
        data[name + "_n"] = np.array(len(v))
        data[name + "_idx"] = idx.astype(np.int32)
        data[name + "_val"] = v[idx]
        data[name + "_l2"] = np.array(np.linalg.norm(v))
        data[name + "_max"] = np.array(np.abs(v).max())
    out = os.path.join(HERE, spec["out"])
    np.savez_compressed(out, **data)
    print("wrote", out)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "c2")
