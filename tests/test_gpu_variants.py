"""GPU parity of the step variants the configurations do not exercise, against
the CPU oracle on the same inputs (FP64 form, 1e-10 relative):

* the conformation form of the constitutive equation (identity kernel
  callbacks, the form the reference's shipped viscoelastic drivers register,
  higflow/example2d_Oldroyd/ns-example-2d.c:21-57), explicit and
  semi-implicit;
* QUICK and MCU convection (hig-flow-terms.c:122-227, hig-flow-discret.c:
  122-129);
* non-zero cell and facet source terms F, FU (hig-flow-step.c:343-399);
* boundary values changed between steps (sb_set_value of the time-dependent
  boundary refresh, hig-flow-step.c:402-606).
"""
import numpy as np
import pytest

import hfo
from higflow_amd import cases, native

from test_gpu_parity import TIGHT, compare

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("semi", [False, True])
def test_conformation_form_oldroyd_b(semi):
    kw = dict(tempdiscrtype=cases.SEMI_IMPLICIT_EULER) if semi else {}
    case = cases.ve_cavity(16, Re=1.0, Wi=1.0, beta=0.5, dt=2e-4, kernel_transform=0, **kw, **TIGHT)

    def init(sim):
        # the conformation form starts from A = I
        for m in (0, 3):
            sim.set(6, m, np.ones(sim.size(6, m)))
    compare(case, 5, init=init)


def test_conformation_form_giesekus_3d():
    case = cases.cavity(6, dim=3, Re=1.0, dt=5e-4, flowtype=cases.VISCOELASTIC, ve_model=cases.GIESEKUS,
                        kernel_transform=0, De=1.0, beta=0.5, tempdiscrtype=cases.SEMI_IMPLICIT_EULER, **TIGHT)

    def init(sim):
        for m in (0, 4, 8):
            sim.set(6, m, np.ones(sim.size(6, m)))
    compare(case, 3, init=init)


@pytest.mark.parametrize("scheme", [cases.QUICK, cases.MCU])
def test_quick_and_mcu_convection(scheme):
    compare(cases.cavity(16, Re=100.0, secondconvecdiscrtype=scheme, **TIGHT), 6)
    compare(cases.cavity(6, dim=3, Re=100.0, secondconvecdiscrtype=scheme, **TIGHT), 3)


def _sources(sim, dim):
    rng = np.random.default_rng(7)
    xc = sim.centers(1)
    sim.set(4, 0, np.sin(3.0 * xc[:, 0]) * np.cos(2.0 * xc[:, 1]) + 0.1 * rng.standard_normal(len(xc)))
    for d in range(dim):
        xf = sim.centers(0, d)
        sim.set(5, d, 0.5 * xf[:, (d + 1) % dim] - 0.2 * xf[:, d] ** 2)


@pytest.mark.parametrize("ve", [False, True])
def test_nonzero_source_terms(ve):
    if ve:
        case = cases.ve_cavity(16, Re=1.0, Wi=0.5, beta=0.5, dt=2e-4, tempdiscrtype=cases.SEMI_IMPLICIT_EULER,
                               **TIGHT)
    else:
        case = cases.cavity(16, Re=10.0, dt=1e-3, **TIGHT)
    w = compare(case, 4, init=lambda sim: _sources(sim, 2))
    print(w)


def test_sources_move_the_flow():
    # the sources are not ignored: with F, FU the state differs from the
    # unforced run by far more than the parity tolerance
    case = cases.cavity(16, Re=10.0, dt=1e-3, **TIGHT)
    a, b = native.Flow(case), native.Flow(case)
    _sources(b, 2)
    for _ in range(3):
        a.step()
        b.step()
    assert np.abs(a.get(0, 0) - b.get(0, 0)).max() > 1e-4


def test_boundary_values_changed_between_steps():
    case = cases.cavity(16, Re=100.0, **TIGHT)
    o, h = hfo.Oracle(case), native.Flow(case)
    for _ in range(3):
        o.step()
        h.step()
    # new lid profile and a moving bottom wall: boundary trees 3 (y=1) and 2
    # (y=0) in the cavity's wall order
    for sim in (o, h):
        xl = sim.bc_centers(1, 3)
        sim.set_bc(1, 3, 0.5 + np.sin(np.pi * xl[:, 0]))
        xb = sim.bc_centers(1, 2)
        sim.set_bc(1, 2, -0.25 * xb[:, 0])
        sim.set_bc(2, 3, 0.1 * np.sin(2 * np.pi * xl[:, 0]))
    for _ in range(4):
        o.step()
        h.step()
    uscale = max(np.abs(o.get(0, d)).max() for d in range(2))
    for d in range(2):
        assert np.abs(h.get(0, d) - o.get(0, d)).max() / uscale <= 1e-10, d
    p = o.get(1)
    assert np.abs(h.get(1) - p).max() / np.abs(p).max() <= 1e-10


def _time_dependent(sim):
    # channel: inflow (boundary 0) grows in time, outflow pressure
    # (boundary 2) varies in time and along y; sources depend on t
    sim.set_functions(
        boundary_velocity=lambda i, x, d, t: (1.5 * (1.0 - x[1] ** 2) * (1.0 + 20.0 * t) if (i == 0 and d == 0)
                                              else 0.0),
        boundary_pressure=lambda i, x, t: 0.5 * t * (1.0 + x[1]) if i == 2 else 0.0,
        source=lambda x, t: np.sin(x[0]) * (1.0 + t),
        facet_source=lambda x, d, t: 0.2 * (x[(d + 1) % 2] - t))
    sim.set_valuetype(1, 0, 1)
    sim.set_valuetype(0, 2, 1)


@pytest.mark.parametrize("ve,proj", [(False, cases.INCREMENTAL), (False, cases.NON_INCREMENTAL),
                                     (True, cases.INCREMENTAL)])
def test_time_dependent_boundaries_and_sources(ve, proj):
    kw = dict(flowtype=cases.VISCOELASTIC, De=0.5, beta=0.5) if ve else {}
    case = cases.newt_channel(nx=40, ny=10, projtype=proj, **kw, **TIGHT)
    w = compare(case, 4, init=_time_dependent, floor=ve)
    print(w)


def test_polymeric_stress_range_log(monkeypatch, capfd):
    # the Tmin/Tmax line of the polymeric-tensor step (hig-flow-step-
    # viscoelastic.c:319-335) for i >= j, then the velocity line, each step
    monkeypatch.setenv("HFH_VERBOSE", "1")
    case = cases.ve_cavity(16, Re=1.0, Wi=1.0, beta=0.5, dt=2e-4, **TIGHT)
    o, h = hfo.Oracle(case), native.Flow(case)
    for _ in range(3):
        o.step()
        h.step()
    out = capfd.readouterr().out.splitlines()
    tl = [ln for ln in out if "Tmin" in ln]
    assert len(tl) == 3 * 3
    beta, Re = 0.5, 1.0
    for k, (i, j) in enumerate(((0, 0), (1, 0), (1, 1))):
        c = i * 2 + j
        T = h.get(9, c)
        # field 9 is S + (1 - beta) Du / Re in the reference's order
        assert np.array_equal(T, h.get(7, c) + (1.0 - beta) * h.get(8, c) / Re)
        To = o.get(7, c) + (1.0 - beta) * o.get(8, c) / Re
        scale = np.abs(To).max()
        mn, mx = h.range(9, c)
        assert (mn, mx) == (T.min(), T.max())
        assert abs(mn - To.min()) <= 1e-10 * scale and abs(mx - To.max()) <= 1e-10 * scale
        assert tl[-3 + k] == "===> %d %d: Tmin = %15.10f <===> Tmax = %15.10f <===" % (i, j, mn, mx)
