"""Reference file formats: a HiG-Flow example's input files become the same
case as the hand-built one, and checkpoint files in the reference's format
round-trip (run on the CPU oracle, which shares the field interface)."""
import os

import numpy as np

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
