"""HiG-Tree restatement of the oracle pinned by the reference-held saved
state of the Newtonian channel example (higflow/example2d_Newt/output,
decoded into tests/golden/newt_poiseuille.npz): same unknowns, same order,
bit-identical coordinates."""
import os

import numpy as np

import hfo
from higflow_amd import cases

GOLD = os.path.join(os.path.dirname(__file__), "golden", "newt_poiseuille.npz")


def test_channel_unknowns_order_and_coordinates():
    fx = np.load(GOLD)
    o = hfo.Oracle(cases.newt_channel())
    for name, field, comp in (("u0", 0, 0), ("u1", 0, 1), ("p", 1, 0)):
        assert np.array_equal(o.centers(field, comp), fx[name][:, :2]), name
    assert o.pinned_cell() == -1  # a Dirichlet pressure boundary exists


def test_cavity_sizes_and_pinned_cell():
    o = hfo.Oracle(cases.cavity(64))
    assert o.size(0, 0) == 63 * 64 and o.size(0, 1) == 64 * 63 and o.size(1) == 4096
    # hig-flow-step.c:22-28: the leaf holding the tree centre (0.5, 0.5)
    # is the lower-left neighbour, cell (31, 31)
    assert o.pinned_cell() == 31 * 64 + 31


def test_cavity_3d_sizes():
    o = hfo.Oracle(cases.cavity(6, dim=3))
    assert o.size(1) == 216
    assert all(o.size(0, d) == 5 * 36 for d in range(3))


def test_stencil_closed_forms():
    o = hfo.Oracle(cases.cavity(8))
    h = 1.0 / 8
    # centre hit -> identity
    ids, w, rhs = o.stencil(0, 0, [0.5 - h / 2, 0.5 - h / 2], [0.5 - h / 2, 0.5 - h / 2])
    assert len(ids) == 1 and w[0] == 1.0 and rhs == 0.0
    # outside a Dirichlet wall (u on the bottom wall): ghost = 2*bc - u_in
    ids, w, rhs = o.stencil(1, 0, [h, h / 2], [h, -h / 2])
    assert len(ids) == 1 and w[0] == -1.0 and rhs == 0.0
    # outside a Neumann wall (pressure): ghost equals the interior value
    ids, w, rhs = o.stencil(0, 0, [h / 2, h / 2], [-h / 2, h / 2])
    assert len(ids) == 1 and w[0] == 1.0
