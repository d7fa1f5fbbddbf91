"""Synthetic case definitions for the HiG-Flow projection step.

A case is a plain dict that both the HIP library (``higflow_amd.native``) and
any other driver can consume:

``dim``      2 or 3
``domains``  list of HiG-Tree ``.amr`` texts (higtree/src/higtree-io.c:269-293)
``bcs``      list of boundary trees: ``amr`` text, user ``id``, pressure type
             ``ptype``, velocity types ``utypes`` (DIRICHLET=0, NEUMANN=1) and
             ``value(kind, x)`` giving the boundary value at a boundary-cell
             centre (kind 0 = pressure, 1+d = velocity component d), i.e. the
             drivers' get_boundary_pressure/get_boundary_velocity callbacks
             (higflow/example2d_Newt/ns-example-2d.c:46-100).
``params``   the higflow_parameters/controllers fields the step reads
             (higflow/src/hig-flow-kernel.h:108-214), reference enum values.
"""
import math

DIRICHLET, NEUMANN = 0, 1
NON_INCREMENTAL, INCREMENTAL = 0, 1
NEWTONIAN, VISCOELASTIC = 0, 3
EXPLICIT_EULER, SEMI_IMPLICIT_EULER = 0, 3
CENTRAL, FIRST_ORDER, SECOND_ORDER = 0, 1, 2
MCU, CUBISTA, QUICK = 0, 1, 2
OLDROYD_B, GIESEKUS = 0, 1
CELL_CENTRAL, CELL_CUBISTA = 0, 1


def _fmt(v):
    return repr(float(v))


def amr_uniform(lo, hi, n):
    """Single-level ``.amr`` text: box [lo, hi] split into n cells per dim."""
    dim = len(lo)
    head = " ".join("%s %s" % (_fmt(lo[d]), _fmt(hi[d])) for d in range(dim))
    delta = " ".join(_fmt((hi[d] - lo[d]) / n[d]) for d in range(dim))
    init = " ".join("1" for _ in range(dim))
    size = " ".join(str(int(n[d])) for d in range(dim))
    return "%s\n1\n%s 1\n%s %s\n" % (head, delta, init, size)


def amr_boundary(lo, hi, n):
    """Degenerate boundary tree: extent zero along the wall normal."""
    dim = len(lo)
    head = " ".join("%s %s" % (_fmt(lo[d]), _fmt(hi[d])) for d in range(dim))
    delta = " ".join(_fmt((hi[d] - lo[d]) / n[d]) for d in range(dim))
    init = " ".join("1" for _ in range(dim))
    size = " ".join(str(int(n[d])) for d in range(dim))
    return "%s\n1\n%s 1\n%s %s\n" % (head, delta, init, size)


def default_params(**kw):
    p = dict(Re=1.0, dt=1e-3, t=0.0, projtype=INCREMENTAL, flowtype=NEWTONIAN,
             tempdiscrtype=EXPLICIT_EULER, convecdiscrtype=SECOND_ORDER,
             secondconvecdiscrtype=CUBISTA, ve_model=OLDROYD_B, ve_convec=CELL_CUBISTA,
             kernel_transform=1, De=1.0, beta=0.5, kernel_tol=1e-2, giesekus_alpha=0.2,
             rtol=1e-10, atol=1e-10, maxit=20000, pin_pressure=0.0, order=2)
    p.update(kw)
    return p


def _walls(dim, lo, hi, n):
    """Boundary trees on the 2*dim walls: for each dim the low wall, then the
    high wall.  Returns (amr, axis, side)."""
    out = []
    for d in range(dim):
        for side in (0, 1):
            blo = list(lo)
            bhi = list(hi)
            pos = lo[d] if side == 0 else hi[d]
            blo[d] = pos
            bhi[d] = pos
            nn = list(n)
            nn[d] = 1
            out.append((amr_boundary(blo, bhi, nn), d, side))
    return out


def cavity(n=64, dim=2, Re=100.0, dt=1e-3, lid=1.0, lid_axis=None, **kw):
    """Lid-driven cavity on [0,1]^dim, uniform n^dim, lid velocity ``lid``
    along x on the top wall (y=1 in 2D, z=1 in 3D); no-slip elsewhere;
    pressure Neumann everywhere (so the pressure is desingularised)."""
    lo = [0.0] * dim
    hi = [1.0] * dim
    nn = [n] * dim
    top = dim - 1 if lid_axis is None else lid_axis
    bcs = []
    for i, (amr, axis, side) in enumerate(_walls(dim, lo, hi, nn)):
        is_lid = axis == top and side == 1

        def value(kind, x, is_lid=is_lid):
            if kind == 1 and is_lid:
                return lid
            return 0.0

        bcs.append(dict(id=i, amr=amr, ptype=NEUMANN, utypes=[DIRICHLET] * dim, value=value))
    params = default_params(Re=Re, dt=dt, **kw)
    return dict(dim=dim, domains=[amr_uniform(lo, hi, nn)], bcs=bcs, params=params,
                name="cavity%dd_%d" % (dim, n), ncells=n ** dim)


def ve_cavity(n=64, Re=1.0, Wi=1.0, beta=0.5, dt=5e-4, **kw):
    """2D Oldroyd-B lid-driven cavity (log-conformation, psi = 0 initially)."""
    kw.setdefault("flowtype", VISCOELASTIC)
    kw.setdefault("ve_model", OLDROYD_B)
    kw.setdefault("kernel_transform", 1)
    c = cavity(n=n, dim=2, Re=Re, dt=dt, De=Wi, beta=beta, **kw)
    c["name"] = "ve_cavity2d_%d" % n
    return c


def newt_channel(nx=160, ny=40, **kw):
    """The reference's Newtonian channel example (higflow/example2d_Newt):
    [0,8]x[-1,1], parabolic inflow u = 1.5(1-y^2) on x=0, no-slip walls,
    outflow (velocity Neumann, pressure Dirichlet 0) on x=8, Re=1, dt=1e-3.
    Boundary ids and order follow amrs/bc/ch-bc-{0..3}.amr."""
    dx = 8.0 / nx
    dy = 2.0 / ny
    bcs = [
        dict(id=0, amr="0.0 0.0 -1.0 1.0\n1\n0.0 %s 1\n1 1 1 %d\n" % (_fmt(dy), ny),
             ptype=NEUMANN, utypes=[DIRICHLET, DIRICHLET],
             value=lambda kind, x: 1.5 * (1.0 - x[1] * x[1]) if kind == 1 else 0.0),
        dict(id=1, amr="0.0 8.0 1.0 1.0\n1\n%s 0.0 1\n1 1 %d 1\n" % (_fmt(dx), nx),
             ptype=NEUMANN, utypes=[DIRICHLET, DIRICHLET], value=lambda kind, x: 0.0),
        dict(id=2, amr="8.0 8.0 -1.0 1.0\n1\n0.0 %s 1\n1 1 1 %d\n" % (_fmt(dy), ny),
             ptype=DIRICHLET, utypes=[NEUMANN, NEUMANN], value=lambda kind, x: 0.0),
        dict(id=3, amr="0.0 8.0 -1.0 -1.0\n1\n%s 0.0 1\n1 1 %d 1\n" % (_fmt(dx), nx),
             ptype=NEUMANN, utypes=[DIRICHLET, DIRICHLET], value=lambda kind, x: 0.0),
    ]
    dom = "0.0 8.0 -1.0 1.0\n1\n%s %s 1\n1 1 %d %d\n" % (_fmt(dx), _fmt(dy), nx, ny)
    params = default_params(Re=1.0, dt=1e-3, **kw)
    return dict(dim=2, domains=[dom], bcs=bcs, params=params, name="newt_channel",
                ncells=nx * ny)


def _contraction_amr(lo, hi, h0, levels, boxes):
    """.amr text of one tree: base spacing h0 on [lo, hi], then `levels`
    factor-2 refinement levels over the cells inside any of `boxes`
    ((xlo, xhi, ylo, yhi) in absolute coordinates; a callable lv -> boxes
    gives each level its own boxes).  Patches follow the reference's
    convention: level delta, 1-based first cell, cell counts."""
    nx = int(round((hi[0] - lo[0]) / h0))
    ny = int(round((hi[1] - lo[1]) / h0))
    lines = ["%s %s %s %s" % (_fmt(lo[0]), _fmt(hi[0]), _fmt(lo[1]), _fmt(hi[1])), str(1 + levels),
             "%s %s 1" % (_fmt(h0), _fmt(h0)), "1 1 %d %d" % (nx, ny)]
    for lv in range(1, levels + 1):
        h = h0 / (2 ** lv)
        patches = []
        for (bx0, bx1, by0, by1) in (boxes(lv) if callable(boxes) else boxes):
            x0, x1 = max(bx0, lo[0]), min(bx1, hi[0])
            y0, y1 = max(by0, lo[1]), min(by1, hi[1])
            if x1 <= x0 or y1 <= y0:
                continue
            i0 = int(round((x0 - lo[0]) / h))
            j0 = int(round((y0 - lo[1]) / h))
            ni = int(round((x1 - x0) / h))
            nj = int(round((y1 - y0) / h))
            patches.append("%d %d %d %d" % (i0 + 1, j0 + 1, ni, nj))
        lines.append("%s %s %d" % (_fmt(h), _fmt(h), len(patches)))
        lines.extend(patches)
    return "\n".join(lines) + "\n"


def contraction(levels=3, h0=0.1, length=20.0, H=4.0, h=1.0, radius=1.0, Re=0.01, Wi=1.0,
                beta=1.0 / 9.0, dt=1e-4, viscoelastic=True, graded=False, **kw):
    """Planar 4:1 contraction: upstream channel [-L/2, 0] x [-H, H], downstream
    [0, L/2] x [-h, h] as two HiG-Tree trees (the layout of
    higflow/example2d_Newt_contraction), base spacing h0 and `levels` levels
    of refinement over the boxes within `radius` of the re-entrant corners
    (0, +-h).  Parabolic inflow with mean 1, no-slip walls, outflow with
    velocity Neumann / pressure Dirichlet.  Oldroyd-B, log-conformation."""
    Lh = length / 2.0

    def corner_boxes(r):
        # boxes snapped outward to the base grid so every level stays aligned
        r = math.ceil(r / h0 - 1e-9) * h0
        return [(-r, r, h - r, h + r), (-r, r, -h - r, -h + r)]
    if graded:
        # level lv refines within radius (levels - lv + 1) / levels: each
        # level boundary is a 2:1 jump instead of one 2^levels:1 jump
        boxes = lambda lv: corner_boxes(radius * (levels - lv + 1) / levels)
    else:
        boxes = corner_boxes(radius)
    up = _contraction_amr((-Lh, -H), (0.0, H), h0, levels, boxes)
    dn = _contraction_amr((0.0, -h), (Lh, h), h0, levels, boxes)
    nyu, nyd, nx = int(round(2 * H / h0)), int(round(2 * h / h0)), int(round(Lh / h0))
    nw = int(round((H - h) / h0))

    def wall(amr):
        return dict(amr=amr, ptype=NEUMANN, utypes=[DIRICHLET, DIRICHLET], value=lambda kind, x: 0.0)

    def inflow(kind, x):
        return 1.5 * (1.0 - (x[1] / H) ** 2) if kind == 1 else 0.0

    f = _fmt
    bcs = [
        dict(amr="%s %s %s %s\n1\n0 %s 1\n1 1 1 %d\n" % (f(-Lh), f(-Lh), f(-H), f(H), f(h0), nyu),
             ptype=NEUMANN, utypes=[DIRICHLET, DIRICHLET], value=inflow),
        wall("%s 0.0 %s %s\n1\n%s 0 1\n1 1 %d 1\n" % (f(-Lh), f(H), f(H), f(h0), nx)),
        wall("0.0 0.0 %s %s\n1\n0 %s 1\n1 1 1 %d\n" % (f(h), f(H), f(h0), nw)),
        wall("0.0 %s %s %s\n1\n%s 0 1\n1 1 %d 1\n" % (f(Lh), f(h), f(h), f(h0), nx)),
        dict(amr="%s %s %s %s\n1\n0 %s 1\n1 1 1 %d\n" % (f(Lh), f(Lh), f(-h), f(h), f(h0), nyd),
             ptype=DIRICHLET, utypes=[NEUMANN, NEUMANN], value=lambda kind, x: 0.0),
        wall("0.0 %s %s %s\n1\n%s 0 1\n1 1 %d 1\n" % (f(Lh), f(-h), f(-h), f(h0), nx)),
        wall("0.0 0.0 %s %s\n1\n0 %s 1\n1 1 1 %d\n" % (f(-H), f(-h), f(h0), nw)),
        wall("%s 0.0 %s %s\n1\n%s 0 1\n1 1 %d 1\n" % (f(-Lh), f(-H), f(-H), f(h0), nx)),
    ]
    for i, b in enumerate(bcs):
        b["id"] = i
    kw.setdefault("tempdiscrtype", SEMI_IMPLICIT_EULER)
    if viscoelastic:
        kw.setdefault("flowtype", VISCOELASTIC)
        kw.setdefault("ve_model", OLDROYD_B)
        kw.setdefault("kernel_transform", 1)
    params = default_params(Re=Re, dt=dt, De=Wi, beta=beta, **kw)
    return dict(dim=2, domains=[up, dn], bcs=bcs, params=params,
                name="contraction_%dlev%s_h%s" % (levels, "_graded" if graded else "", f(h0)), ncells=None)
