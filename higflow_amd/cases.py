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


def amr_refined(lo, hi, h0, levels, boxes):
    """.amr text of one tree in any dimension: base spacing h0 on [lo, hi],
    then `levels` factor-2 refinement levels over the cells inside any of
    `boxes` ((blo, bhi) corner pairs in absolute coordinates; a callable
    lv -> boxes gives each level its own boxes).  Patches follow the
    reference's convention (higtree/src/higtree-io.c:97-153, 269-293): level
    delta and patch count, then per patch the 1-based first cell (odd: the
    point delta*initcell is the centre of a parent cell) and the cell
    counts.  Boxes must be aligned to the parent level's grid."""
    dim = len(lo)
    n = [int(round((hi[d] - lo[d]) / h0)) for d in range(dim)]
    lines = [" ".join("%s %s" % (_fmt(lo[d]), _fmt(hi[d])) for d in range(dim)), str(1 + levels),
             " ".join([_fmt(h0)] * dim) + " 1", " ".join(["1"] * dim + [str(v) for v in n])]
    for lv in range(1, levels + 1):
        h = h0 / (2 ** lv)
        patches = []
        for (blo, bhi) in (boxes(lv) if callable(boxes) else boxes):
            a = [max(blo[d], lo[d]) for d in range(dim)]
            b = [min(bhi[d], hi[d]) for d in range(dim)]
            if any(b[d] <= a[d] for d in range(dim)):
                continue
            first = [int(round((a[d] - lo[d]) / h)) + 1 for d in range(dim)]
            size = [int(round((b[d] - a[d]) / h)) for d in range(dim)]
            patches.append(" ".join(str(v) for v in first + size))
        lines.append(" ".join([_fmt(h)] * dim) + " %d" % len(patches))
        lines.extend(patches)
    return "\n".join(lines) + "\n"


def _contraction_amr(lo, hi, h0, levels, boxes):
    """2D form of amr_refined with boxes given as (xlo, xhi, ylo, yhi)."""
    conv = lambda bs: [((x0, y0), (x1, y1)) for (x0, x1, y0, y1) in bs]
    return amr_refined(lo, hi, h0, levels, (lambda lv: conv(boxes(lv))) if callable(boxes) else conv(boxes))


def wall_layer_cavity(n=16, dim=3, levels=2, layer=0.05, Re=400.0, dt=2e-4, **kw):
    """Lid-driven cavity on [0,1]^dim with base n^dim cells and `levels`
    extra factor-2 levels in a layer of thickness `layer` along every wall
    (each level's layer snapped outward to its parent grid).  Lid u = 1 on
    the top wall (z = 1 in 3D), no-slip elsewhere, pressure Neumann."""
    h0 = 1.0 / n

    def boxes(lv):
        hp = h0 / 2 ** (lv - 1)                      # parent spacing
        t = min(0.5, math.ceil(layer / hp - 1e-9) * hp)
        out = []
        for d in range(dim):
            for side in (0, 1):
                blo, bhi = [0.0] * dim, [1.0] * dim
                if side == 0:
                    bhi[d] = t
                else:
                    blo[d] = 1.0 - t
                out.append((blo, bhi))
        return out
    c = cavity(n=n, dim=dim, Re=Re, dt=dt, **kw)
    c["domains"] = [amr_refined([0.0] * dim, [1.0] * dim, h0, levels, boxes)]
    c["name"] = "cavity%dd_%d_wall%d" % (dim, n, levels)
    c["ncells"] = None
    return c


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


# ---------------------------------------------------------------- sweeps
class Pcg32:
    """PCG32 (O'Neill, "PCG: A Family of Simple Fast Space-Efficient
    Statistically Good Algorithms for Random Number Generation", 2014;
    pcg_basic.c): 64-bit LCG state, XSH-RR output."""
    MULT = 6364136223846793005
    MASK = (1 << 64) - 1

    def __init__(self, initstate=0, initseq=0):
        self.state = 0
        self.inc = ((initseq << 1) | 1) & self.MASK
        self.next()
        self.state = (self.state + initstate) & self.MASK
        self.next()

    def next(self):
        old = self.state
        self.state = (old * self.MULT + self.inc) & self.MASK
        xs = (((old >> 18) ^ old) >> 27) & 0xFFFFFFFF
        rot = old >> 59
        return ((xs >> rot) | (xs << ((-rot) & 31))) & 0xFFFFFFFF

    def bounded(self, bound):
        # unbiased: reject the low 2^32 mod bound values (pcg32_boundedrand_r)
        threshold = ((1 << 32) - bound) % bound
        while True:
            r = self.next()
            if r >= threshold:
                return r % bound

    def uniform(self, a, b):
        return a + (b - a) * (self.next() / 4294967296.0)


SWEEP_SIDES = (32, 48, 64, 96, 128, 192, 256)


def sweep_table(count=64, seed=0):
    """Draws, in order per simulation, the side N uniformly from
    SWEEP_SIDES, Re = exp(U[ln 1, ln 100]) and Wi = exp(U[ln 0.1, ln 2])
    with PCG32(initstate=seed, initseq=0)."""
    g = Pcg32(seed, 0)
    out = []
    for _ in range(count):
        n = SWEEP_SIDES[g.bounded(len(SWEEP_SIDES))]
        re = math.exp(g.uniform(math.log(1.0), math.log(100.0)))
        wi = math.exp(g.uniform(math.log(0.1), math.log(2.0)))
        out.append((n, re, wi))
    return out


def sweep(count=64, seed=0, beta=0.5, dt=5e-4, table=None, **kw):
    """Independent 2D Oldroyd-B lid-driven cavities (log-conformation,
    semi-implicit Euler) of a parameter sweep, as case dicts for a batch."""
    kw.setdefault("tempdiscrtype", SEMI_IMPLICIT_EULER)
    rows = table if table is not None else sweep_table(count, seed)
    return [ve_cavity(n, Re=re, Wi=wi, beta=beta, dt=dt, **kw) for (n, re, wi) in rows]
