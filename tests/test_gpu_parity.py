"""GPU parity of the HIP projection step against the CPU oracle on the same
seeded inputs (FP64 form: 1e-10 relative), through the C ABI."""
import numpy as np
import pytest

import hfo
from higflow_amd import cases, native

pytestmark = pytest.mark.gpu

TIGHT = dict(rtol=1e-13, atol=1e-15, maxit=50000)


def rel(a, b):
    s = max(np.abs(b).max(), 1e-300)
    return np.abs(a - b).max() / s


def fields(sim, dim, ve=False):
    out = {("u", d): sim.get(0, d) for d in range(dim)}
    out["p"] = sim.get(1)
    if ve:
        for m in range(dim * dim):
            out[("K", m)] = sim.get(6, m)
            out[("S", m)] = sim.get(7, m)
    return out


def compare(case, steps, tol=1e-10, init=None, floor=False):
    """HIP against the oracle after `steps` steps, every field within `tol`
    of its scale.  floor=True also runs the oracle with its solver
    tolerances 10x looser and allows each field 10x that self-difference
    where it exceeds `tol`: where the step amplifies solver round-off (the
    log-conformation's 1/(lambda_j - lambda_i + 1e-12) near equal
    eigenvalues, hig-flow-step-viscoelastic.c:1667-1680), two correct
    solvers differ by more than 1e-10 and the oracle disagrees with itself
    by as much."""
    o = hfo.Oracle(case)
    h = native.Flow(case)
    dim = case["dim"]
    runs = [o, h]
    if floor:
        loose = dict(case, params=dict(case["params"], rtol=case["params"]["rtol"] * 10,
                                       atol=case["params"]["atol"] * 10))
        o2 = hfo.Oracle(loose)
        runs.append(o2)
    if init:
        for sim in runs:
            init(sim)
    ve = case["params"]["flowtype"] == 3
    worst = {}
    for k in range(steps):
        for sim in runs:
            sim.step()
    fo, fh = fields(o, dim, ve), fields(h, dim, ve)
    f2 = fields(o2, dim, ve) if floor else None
    # relative to the field's scale; velocity components share the velocity
    # scale (a near-zero transverse component is compared against |u|)
    uscale = max(np.abs(fo[("u", d)]).max() for d in range(dim))
    bad = {}
    for key in fo:
        scale = uscale if key[0] == "u" else np.abs(fo[key]).max()
        rel = lambda a: np.abs(a - fo[key]).max() / scale if scale > 0 else np.abs(a).max()
        worst[key] = rel(fh[key])
        allowed = max(tol, 10.0 * rel(f2[key])) if floor else tol
        if not worst[key] <= allowed:
            bad[key] = (worst[key], allowed)
    assert not bad, bad
    return worst


def test_cavity_small_parity():
    w = compare(cases.cavity(16, **TIGHT), 5)
    print(w)


def test_velocity_range_log(monkeypatch, capfd):
    # the Vmin/Vmax line of the reference's final-velocity step
    # (hig-flow-step.c:302-306): a device reduction over the unknown facets
    monkeypatch.setenv("HFH_VERBOSE", "1")
    case = cases.cavity(16, **TIGHT)
    o, h = hfo.Oracle(case), native.Flow(case)
    for _ in range(3):
        o.step()
        h.step()
    lines = [ln for ln in capfd.readouterr().out.splitlines() if ln.startswith("===>")]
    assert len(lines) == 3 * 2
    for d in range(2):
        u = h.get(0, d)
        assert h.range(0, d) == (u.min(), u.max())
        uo = o.get(0, d)
        scale = max(np.abs(uo).max(), 1e-300)
        mn, mx = h.range(0, d)
        assert abs(mn - uo.min()) <= 1e-10 * scale and abs(mx - uo.max()) <= 1e-10 * scale
        assert lines[-2 + d] == "===> %d: Vmin = %15.10f <===> Vmax = %15.10f <===" % (d, mn, mx)


def test_cavity_32_parity_20_steps():
    compare(cases.cavity(32, **TIGHT), 20)


@pytest.mark.slow
def test_cavity_64_re100_100_steps():
    # lid-driven cavity 64x64, Re=100, dt=1e-3, 100 explicit Euler steps
    compare(cases.cavity(64, Re=100.0, dt=1e-3, **TIGHT), 100)


def test_cavity_first_order_and_central():
    compare(cases.cavity(16, convecdiscrtype=cases.FIRST_ORDER, **TIGHT), 4)
    compare(cases.cavity(16, convecdiscrtype=cases.CENTRAL, **TIGHT), 4)


def test_channel_from_reference_fixture():
    import os
    fx = np.load(os.path.join(os.path.dirname(__file__), "golden", "newt_poiseuille.npz"))

    def init(sim):
        for name, f, c in (("u0", 0, 0), ("u1", 0, 1), ("p", 1, 0)):
            sim.set(f, c, fx[name][:, 2])
    compare(cases.newt_channel(**TIGHT), 2, init=init)


def test_cavity_3d_parity():
    compare(cases.cavity(8, dim=3, Re=100.0, **TIGHT), 3)


def test_wall_layer_cavity_3d_parity():
    # 3D lid-driven cavity, base 6^3 with 2 levels in a 0.05 wall layer
    # (6,376 cells; the shape of the 128^3 wall-layer configuration), Re=400
    compare(cases.wall_layer_cavity(6, dim=3, levels=2, Re=400.0, dt=2e-4, **TIGHT), 2)


@pytest.mark.parametrize("force32", ["0", "1"])
def test_wall_layer_cavity_3d_mixed_error(force32, monkeypatch):
    # FP32 Poisson inner iterations in FP64 refinement to 1e-9 against the
    # FP64 step on the same mesh (FP32 refinement forced on every solve, or
    # chosen by time)
    monkeypatch.setenv("HFH_MIXED_FP32", force32)
    case = cases.wall_layer_cavity(8, dim=3, levels=2, Re=400.0, dt=2e-4, rtol=1e-9, atol=1e-9, maxit=50000)
    a = native.Flow(case, precision=0)
    b = native.Flow(case, precision=1, inner_rtol=1e-5)
    for _ in range(5):
        a.step()
        b.step()
    assert b.solve_stats(0)["converged"] == 1
    uscale = max(np.abs(a.get(0, d)).max() for d in range(3))
    for d in range(3):
        assert np.abs(b.get(0, d) - a.get(0, d)).max() / uscale < 1e-6
    assert np.linalg.norm(b.get(1) - a.get(1)) / np.linalg.norm(a.get(1)) < 1e-6


@pytest.mark.parametrize("model", [cases.OLDROYD_B, cases.GIESEKUS])
def test_viscoelastic_cavity_3d_semi_implicit(model):
    # 3D log-conformation (3x3 Jacobi eigendecomposition), semi-implicit Euler
    case = cases.cavity(6, dim=3, Re=1.0, dt=5e-4, flowtype=cases.VISCOELASTIC, ve_model=model, kernel_transform=1,
                        De=1.0, beta=0.5, tempdiscrtype=cases.SEMI_IMPLICIT_EULER, **TIGHT)
    compare(case, 3)


def test_cavity_3d_semi_implicit():
    compare(cases.cavity(8, dim=3, Re=10.0, dt=1e-3, tempdiscrtype=cases.SEMI_IMPLICIT_EULER, **TIGHT), 3)


def test_nonincremental_projection():
    compare(cases.cavity(16, projtype=cases.NON_INCREMENTAL, **TIGHT), 3)


@pytest.mark.parametrize("model,convec", [(cases.OLDROYD_B, cases.CELL_CUBISTA),
                                          (cases.GIESEKUS, cases.CELL_CUBISTA),
                                          (cases.OLDROYD_B, cases.CELL_CENTRAL)])
def test_viscoelastic_cavity_parity(model, convec):
    case = cases.ve_cavity(16, Re=1.0, Wi=1.0, beta=0.5, dt=2e-4, ve_model=model, ve_convec=convec, **TIGHT)
    compare(case, 5)


def test_mixed_precision_error_bound(monkeypatch):
    # FP32 refinement on every solve (not only where it is the faster form)
    monkeypatch.setenv("HFH_MIXED_FP32", "1")
    case = cases.cavity(64, Re=100.0, **TIGHT)
    h64 = native.Flow(case)
    hmx = native.Flow(case, precision=1, inner_rtol=1e-5)
    for _ in range(20):
        h64.step()
        hmx.step()
    for d in range(2):
        a, b = hmx.get(0, d), h64.get(0, d)
        assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-6
    a, b = hmx.get(1), h64.get(1)
    assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-6


def test_semi_implicit_cavity():
    compare(cases.cavity(32, Re=10.0, dt=2e-3, tempdiscrtype=cases.SEMI_IMPLICIT_EULER, **TIGHT), 5)


def test_semi_implicit_channel_fixture():
    # the reference fixture was written by semi-implicit Euler
    import os
    fx = np.load(os.path.join(os.path.dirname(__file__), "golden", "newt_poiseuille.npz"))

    def init(sim):
        for name, f, c in (("u0", 0, 0), ("u1", 0, 1), ("p", 1, 0)):
            sim.set(f, c, fx[name][:, 2])
    compare(cases.newt_channel(tempdiscrtype=cases.SEMI_IMPLICIT_EULER, **TIGHT), 2, init=init)


def test_amr_contraction_newtonian():
    case = cases.contraction(levels=2, h0=0.5, length=8.0, radius=1.0, viscoelastic=False, Re=1.0,
                             dt=1e-3, **TIGHT)
    compare(case, 3)


def test_amr_contraction_oldroyd_b():
    case = cases.contraction(levels=2, h0=0.5, length=8.0, radius=1.0, Re=0.01, Wi=1.0, beta=1.0 / 9.0,
                             dt=1e-4, **TIGHT)
    compare(case, 3)


def test_ve_cavity_semi_implicit_low_re():
    case = cases.ve_cavity(24, Re=1.0, Wi=0.5, beta=0.5, dt=5e-4, tempdiscrtype=cases.SEMI_IMPLICIT_EULER,
                           **TIGHT)
    compare(case, 4)


def _agree(a, b, dim, ve, tol, noise=None):
    # noise: a third run whose difference to `b` is the solver-noise floor;
    # each field may then differ by 10x that floor when it exceeds tol
    fa, fb = fields(a, dim, ve), fields(b, dim, ve)
    fn = fields(noise, dim, ve) if noise is not None else None
    uscale = max(np.abs(fb[("u", d)]).max() for d in range(dim))
    for key in fb:
        scale = uscale if key[0] == "u" else max(np.abs(fb[key]).max(), 1e-300)
        allowed = tol
        if fn is not None:
            allowed = max(tol, 10.0 * np.abs(fn[key] - fb[key]).max() / scale)
        assert np.abs(fa[key] - fb[key]).max() / scale <= allowed, (key, allowed)


@pytest.mark.parametrize("which", ["cavity", "contraction"])
def test_fused_and_classic_krylov_agree(which, monkeypatch):
    # fused iterations (2 launches per PCG / 3 per BiCGStab iteration, device
    # stopping test) against one kernel per operation
    if which == "cavity":
        case = cases.cavity(32, **TIGHT)
    else:
        case = cases.contraction(levels=2, h0=0.5, length=8.0, radius=1.0, Re=0.01, Wi=1.0, beta=1.0 / 9.0,
                                 dt=1e-4, **TIGHT)
    monkeypatch.setenv("HFH_KRYLOV", "classic")
    hc = native.Flow(case)
    # the classic solver at 10x looser tolerances: on the AMR contraction an
    # FP64 solve stops near a true relative residual of 1e-11 whatever rtol
    # asks, and two Krylov histories differ by that floor (amplified in p
    # where the pin and the WLS rows meet), not by 1e-10 of the field scale
    loose = dict(case)
    loose["params"] = dict(case["params"], rtol=1e-12, atol=1e-14)
    hn = native.Flow(loose) if which == "contraction" else None
    monkeypatch.delenv("HFH_KRYLOV")
    hf = native.Flow(case)
    for _ in range(3):
        hc.step()
        hf.step()
        if hn is not None:
            hn.step()
    _agree(hf, hc, case["dim"], case["params"]["flowtype"] == 3, 1e-10, noise=hn)


def test_amg_fused_prolongation_agrees(monkeypatch):
    # the V-cycle's prolongation and post-smoothing in one launch through
    # A P against the two-launch form: the same converged step
    case = cases.contraction(levels=2, h0=0.5, length=8.0, radius=1.0, Re=0.01, Wi=1.0, beta=1.0 / 9.0,
                             dt=1e-4, rtol=1e-12, atol=1e-14, maxit=50000)
    monkeypatch.setenv("HFH_AMG_FUSE", "0")
    a = native.Flow(case)
    monkeypatch.setenv("HFH_AMG_FUSE", "1")
    b = native.Flow(case)
    for _ in range(3):
        a.step()
        b.step()
    _agree(b, a, 2, True, 1e-10)
    assert b.solve_stats(0)["converged"] == 1


@pytest.mark.parametrize("force32", ["0", "1"])
def test_mixed_precision_amr_contraction(force32, monkeypatch):
    # nonsymmetric systems: FP32 BiCGStab corrections inside FP64 refinement,
    # forced on every solve or chosen by timing
    monkeypatch.setenv("HFH_MIXED_FP32", force32)
    case = cases.contraction(levels=2, h0=0.5, length=8.0, radius=1.0, Re=0.01, Wi=1.0, beta=1.0 / 9.0,
                             dt=1e-4, rtol=1e-10, atol=1e-10, maxit=50000)
    h64 = native.Flow(case)
    hmx = native.Flow(case, precision=1, inner_rtol=1e-5)
    for _ in range(5):
        h64.step()
        hmx.step()
    for d in range(2):
        a, b = hmx.get(0, d), h64.get(0, d)
        assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-6
    a, b = hmx.get(1), h64.get(1)
    assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-6


@pytest.mark.parametrize("which", ["contraction", "ve_cavity"])
def test_mixed_fp32_step_kernels_error(which, monkeypatch):
    # momentum, velocity-gradient and constitutive evaluations in FP32
    # arithmetic (eigendecompositions, divergence and projection in FP64):
    # velocity within 1e-6 of the FP64 step, pressure within 1e-5 (1.3e-6
    # measured on the contraction), log-conformation within 1e-4 (it starts
    # from zero, so early steps are relatively noisy)
    monkeypatch.setenv("HFH_MIXED_KERNELS", "1")
    if which == "contraction":
        case = cases.contraction(levels=2, h0=0.5, length=8.0, radius=1.0, Re=0.01, Wi=1.0, beta=1.0 / 9.0,
                                 dt=1e-4, rtol=1e-10, atol=1e-10, maxit=50000)
    else:
        case = cases.ve_cavity(32, Re=1.0, Wi=1.0, beta=0.5, dt=5e-4, tempdiscrtype=cases.SEMI_IMPLICIT_EULER,
                               rtol=1e-10, atol=1e-10, maxit=50000)
    h64 = native.Flow(case)
    hmx = native.Flow(case, precision=1, inner_rtol=1e-5)
    for _ in range(10):
        h64.step()
        hmx.step()
    for d in range(2):
        a, b = hmx.get(0, d), h64.get(0, d)
        assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-6
    a, b = hmx.get(1), h64.get(1)
    assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-5
    for m in (0, 1, 3):
        a, b = hmx.get(6, m), h64.get(6, m)
        nb = np.linalg.norm(b)
        assert np.linalg.norm(a - b) <= 1e-4 * max(nb, 1e-30), (m, np.linalg.norm(a - b) / max(nb, 1e-30))
    # the same runs through the FP64 kernels differ at solver-tolerance level
    assert np.abs(hmx.get(0, 0) - h64.get(0, 0)).max() > 0


def test_reference_example_inputs():
    # the reference example's own input files (higflow/example2d_Newt/input),
    # loaded as a case, through the HIP step against the oracle
    import os
    from higflow_amd import io
    ex = os.path.join(os.path.dirname(__file__), "golden", "example2d_newt")

    def inflow(bid, kind, x):
        return 1.5 * (1.0 - x[1] * x[1]) if (bid == 0 and kind == 1) else 0.0
    compare(io.load_reference_inputs(ex, "input/example-2d.load", bc_value=inflow, **TIGHT), 3)
