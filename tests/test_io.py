"""Reference file formats: a HiG-Flow example's input files become the same
case as the hand-built one, and checkpoint files in the reference's format
round-trip (run on the CPU oracle, which shares the field interface)."""
import os

import numpy as np
import pytest

import hfo
from higflow_amd import cases, io

EX = os.path.join(os.path.dirname(__file__), "golden", "example2d_newt")


def _inflow(bid, kind, x):
    # the example driver's get_boundary_velocity (ns-example-2d.c:46-100)
    return 1.5 * (1.0 - x[1] * x[1]) if (bid == 0 and kind == 1) else 0.0


def _nums(text):
    return [float(t) for t in text.split()]


def test_example_inputs_match_the_channel_case():
    c = io.load_reference_inputs(EX, "input/example-2d.load", bc_value=_inflow)
    ref = cases.newt_channel(tempdiscrtype=cases.SEMI_IMPLICIT_EULER)
    assert c["dim"] == 2 and len(c["domains"]) == 1
    assert _nums(c["domains"][0]) == _nums(ref["domains"][0])
    assert [b["id"] for b in c["bcs"]] == [b["id"] for b in ref["bcs"]]
    for a, b in zip(c["bcs"], ref["bcs"]):
        assert _nums(a["amr"]) == _nums(b["amr"])
        assert a["ptype"] == b["ptype"] and a["utypes"] == b["utypes"]
    for k in ("Re", "dt", "projtype", "tempdiscrtype", "convecdiscrtype", "secondconvecdiscrtype", "flowtype"):
        assert c["params"][k] == ref["params"][k], k
    assert c["numsteps"] == 1001


def test_example_inputs_step_like_the_channel_case():
    kw = dict(rtol=1e-13, atol=1e-15)
    a = hfo.Oracle(io.load_reference_inputs(EX, "input/example-2d.load", bc_value=_inflow, **kw))
    b = hfo.Oracle(cases.newt_channel(tempdiscrtype=cases.SEMI_IMPLICIT_EULER, **kw))
    a.step()
    b.step()
    for f, comp in ((0, 0), (0, 1), (1, 0)):
        assert np.array_equal(a.get(f, comp), b.get(f, comp))


def test_checkpoint_round_trip(tmp_path):
    case = cases.ve_cavity(8, Re=1.0, Wi=1.0, beta=0.5, dt=1e-4)
    a = hfo.Oracle(case)
    for _ in range(2):
        a.step()
    prefix = str(tmp_path / "ve.save")
    io.save_properties(a, prefix, dim=2, viscoelastic=True)
    # record layout: (x, y, value) FP64 in local-id order
    rec = np.fromfile(prefix + "._dpp", dtype=np.float64).reshape(-1, 3)
    assert np.array_equal(rec[:, :2], np.asarray(a.centers(1)).reshape(-1, 2))
    assert np.array_equal(rec[:, 2], a.get(1))
    assert os.path.exists(prefix + "._dpKernel01") and os.path.exists(prefix + "._dpu1")
    b = hfo.Oracle(case)
    io.load_properties(b, prefix, dim=2, viscoelastic=True)
    for f, comp in ((0, 0), (0, 1), (1, 0), (6, 0), (6, 1), (6, 3)):
        assert np.array_equal(a.get(f, comp), b.get(f, comp)), (f, comp)


def test_load_finds_records_by_point(tmp_path):
    # the reference locates each record by its point: order does not matter
    case = cases.cavity(8)
    a = hfo.Oracle(case)
    a.step()
    prefix = str(tmp_path / "c.save")
    io.save_properties(a, prefix)
    rec = np.fromfile(prefix + "._dpu0", dtype=np.float64).reshape(-1, 3)
    rec[::-1].tofile(prefix + "._dpu0")
    b = hfo.Oracle(case)
    io.load_properties(b, prefix)
    assert np.array_equal(a.get(0, 0), b.get(0, 0))


GOLD = os.path.join(os.path.dirname(__file__), "golden", "newt_poiseuille.npz")


def _vtk_section(lines, head, count):
    i = lines.index(head)
    return np.array([[float(v) for v in ln.split()] for ln in lines[i + 1:i + 1 + count]])


def test_vtk_writer_on_the_channel_fixture(tmp_path, monkeypatch):
    # higflow_print_vtk2D layout on the reference-held channel state: the
    # corner velocities interpolate the Poiseuille parabola (quadratic WLS
    # reproduces it; the fixture is steady to ~1e-7 and %e keeps 7 digits),
    # the pressure is the cell value
    monkeypatch.setenv("HFH_HOST_ONLY", "1")
    from higflow_amd import native
    fx = np.load(GOLD)
    h = native.Flow(cases.newt_channel())
    f = {(0, 0): fx["u0"][:, 2], (0, 1): fx["u1"][:, 2], (1, 0): fx["p"][:, 2]}
    name = io.print_vtk(h, str(tmp_path / "newt"), rank=0, frame=3, fields=f)
    assert name.endswith("newt_0-3.vtk")
    lines = open(name, encoding="utf-8").read().split("\n")
    n = 6400
    assert lines[:4] == ["# vtk DataFile Version 3.0", "higtree", "ASCII", "DATASET UNSTRUCTURED_GRID"]
    pts = _vtk_section(lines, "POINTS %d float" % (4 * n), 4 * n)
    cells = _vtk_section(lines, "CELLS %d %d" % (n, 5 * n), n)
    assert np.array_equal(cells[:, 0], np.full(n, 4)) and np.array_equal(cells[:, 1:].ravel(), np.arange(4 * n))
    assert lines[lines.index("CELL_TYPES %d" % n) + 1] == "7 " * n
    lo, hi = h.cell_bounds()
    assert np.allclose(0.5 * (lo + hi), h.centers(1), rtol=0, atol=1e-15)
    assert np.allclose(pts[0::4, :2], lo, rtol=1e-6, atol=1e-12) and np.allclose(pts[2::4, :2], hi, rtol=1e-6)
    vel = _vtk_section(lines, "VECTORS vel FLOAT", 4 * n)
    inner = (pts[:, 0] > 0.1) & (pts[:, 0] < 7.9)
    assert np.abs(vel[inner, 0] - 1.5 * (1 - pts[inner, 1] ** 2)).max() < 1e-5
    assert np.abs(vel[inner, 1]).max() < 1e-5
    p = _vtk_section(lines, "LOOKUP_TABLE default", n)[:, 0]
    assert np.array_equal(p, np.array(["%e" % v for v in fx["p"][:, 2]], dtype=float))


def test_vtk_writer_viscoelastic_tensor(tmp_path, monkeypatch):
    # constant S and Du: the tensor-cell stencils reproduce constants, so
    # every corner prints tau_p = S + 2 (1 - beta) 0.5 Du / Re
    monkeypatch.setenv("HFH_HOST_ONLY", "1")
    from higflow_amd import native
    h = native.Flow(cases.ve_cavity(8, Re=2.0, Wi=1.0, beta=0.25, dt=1e-4))
    n = h.size(1)
    S0, D0 = np.array([[1.5, -0.5], [0.25, 2.0]]), np.array([[0.1, 0.3], [-0.2, -0.1]])
    f = {(0, 0): np.zeros(h.size(0, 0)), (0, 1): np.zeros(h.size(0, 1)), (1, 0): np.zeros(n)}
    for m in range(4):
        f[(7, m)] = np.full(h.size(7, m), S0.flat[m])
        f[(8, m)] = np.full(h.size(8, m), D0.flat[m])
    name = io.print_vtk(h, str(tmp_path / "ve"), Re=2.0, beta=0.25, viscoelastic=True, fields=f)
    lines = open(name, encoding="utf-8").read().split("\n")
    t = _vtk_section(lines, "TENSORS τₚ FLOAT", 12 * n).reshape(4 * n, 3, 3)
    want = S0 + 2.0 * (1 - 0.25) * 0.5 * D0 / 2.0
    assert np.abs(t[:, :2, :2] - want).max() < 1e-6 and not t[:, 2, :].any() and not t[:, :, 2].any()


@pytest.mark.parametrize("viscoelastic", [False, True])
def test_vtk_writer_3d(tmp_path, monkeypatch, viscoelastic):
    # higflow_print_vtk3D / higflow_print_vtk3D_viscoelastic layout on a
    # 3D wall-layer AMR cavity with a linear velocity field: every corner,
    # coarse or fine, interpolates it exactly (quadratic WLS reproduces
    # linear fields; walls hold the field's own boundary values)
    monkeypatch.setenv("HFH_HOST_ONLY", "1")
    from higflow_amd import native
    G = np.array([[0.3, -0.2, 0.1], [0.05, 0.2, -0.4], [0.15, 0.1, -0.5]])
    lin = lambda x, d: 0.1 * (d + 1) + G[d] @ np.asarray(x)
    c = cases.wall_layer_cavity(4, dim=3, levels=1, layer=0.25)
    if viscoelastic:
        c["params"]["flowtype"] = cases.VISCOELASTIC
    for bc in c["bcs"]:
        bc["value"] = lambda kind, x: lin(x, kind - 1) if kind > 0 else 0.0
    h = native.Flow(c)
    n = h.size(1)
    f = {(0, d): np.array([lin(x, d) for x in h.centers(0, d)]) for d in range(3)}
    f[(1, 0)] = np.linspace(-1.0, 2.0, n)
    if viscoelastic:
        f[(7, 0)] = np.linspace(0.5, 0.75, n)
    name = io.print_vtk(h, str(tmp_path / "c3"), rank=1, frame=2, viscoelastic=viscoelastic, fields=f)
    assert name.endswith("c3_1-2.vtk")
    lines = open(name, encoding="utf-8").read().split("\n")
    assert lines[:5] == ["# vtk DataFile Version 3.0", "higtree", "ASCII", "DATASET UNSTRUCTURED_GRID",
                         "POINTS %d float" % (8 * n)]
    pts = _vtk_section(lines, "POINTS %d float" % (8 * n), 8 * n)
    lo, hi = h.cell_bounds()
    assert np.allclose(pts[0::8], lo, atol=1e-6) and np.allclose(pts[6::8], hi, atol=1e-6)
    cells = _vtk_section(lines, "CELLS %d %d" % (n, 9 * n), n)
    assert np.array_equal(cells[:, 0], np.full(n, 8)) and np.array_equal(cells[:, 1:].ravel(), np.arange(8 * n))
    assert lines[lines.index("CELL_TYPES %d" % n) + 1] == "12 " * n
    vel = _vtk_section(lines, "VECTORS vel FLOAT", 8 * n)
    want = np.array([[lin(p, d) for d in range(3)] for p in pts])
    # corners off the x and y walls: on those walls the boundary interpolant
    # is evaluated at the reference's unreduced projection point
    # (get_stencil_dirichlet_boundary, domain.c:1271-1280), not at the wall
    # point itself
    off = np.all((pts[:, :2] > 1e-9) & (pts[:, :2] < 1.0 - 1e-9), axis=1)
    assert off.sum() > 8 * n // 2
    assert np.abs(vel[off] - want[off]).max() < 1e-6
    key = "SCALARS S00 FLOAT" if viscoelastic else "SCALARS p FLOAT"
    i = lines.index(key)
    assert lines[i - 1] == "CELL_DATA %d" % n and lines[i + 1] == "LOOKUP_TABLE default"
    vals = np.array([float(v) for v in lines[i + 2:i + 2 + n]])
    src = f[(7, 0)] if viscoelastic else f[(1, 0)]
    assert np.array_equal(vals, np.array(["%e" % v for v in src], dtype=float))
    assert ("SCALARS p FLOAT" in lines) != viscoelastic
