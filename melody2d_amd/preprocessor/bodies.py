"""Body construction: node clouds, triangulation, quadrature, masses, borders.

Plays the part of the reference preprocessor's body analysis
(Preprocessor/AnalyseSimpleBody.m, AnalyseRigidBody.m,
SetUnstructuredNodesForClosedContour.m, SetRigidNodesForPolygonContour.m)
for the shapes the workloads need: discs and rectangles, deformable or
rigid.  Everything is seeded; one numpy Generator per body in creation order.

Conventions the step relies on (and that this module reproduces):
* border nodes are renumbered to be the first nodes of the body, in
  counter-clockwise order, so the right-hand normal (ty, -tx) of each
  segment points outwards (Code/Tools.h:111-114);
* a closed border lists its first node again at the end
  (SetUnstructuredNodesForClosedContour.m:174) and is its own border
  before/after (cell layout {'Closed', n+1, list, interp, 0, 0});
* three Gauss points per triangle at barycentric (2/3,1/6,1/6) with
  weight 1/3 and jacobian = triangle area (GaussCoefficientsTriangle.m,
  TriangleGaussPoints.m), so weight*jacobian = area/3;
* lumped mass = row sum of the consistent mass, rho * sum_g c0 phi_j sum_k phi_k
  (AnalyseSimpleBody.m:316-331, Code/Body.h:2753-2783).
"""
from __future__ import annotations

from dataclasses import dataclass, field
import numpy as np
from scipy.spatial import Delaunay, cKDTree

from . import shape


@dataclass
class BC:
    """One direction of a border condition, as written in STATIC_CONTROL.asc."""
    kind: str = "None"          # None | Driven | Soft | Oriented | Following
    history: list = field(default_factory=list)   # [(t, value), ...]
    params: tuple = ()           # (stiffness, damping) for Soft


@dataclass
class Border:
    nodes: np.ndarray            # node ids incl. repeated first node for closed borders
    periodicity: str = "Closed"
    interpolant: str = "Linear"
    before: int = 0
    after: int = 0
    xbc: BC = field(default_factory=BC)
    ybc: BC = field(default_factory=BC)


@dataclass
class Body:
    rigid: bool
    material: str
    X: np.ndarray                # (N,2) initial positions
    borders: list
    tri: np.ndarray              # (T,3)
    nodal_distance: float
    detection: float
    contact: float
    periodicity: str = "Simple"
    status: str = "active"
    velocity: tuple = (0.0, 0.0)   # (vx, vy), or per node an (n, 2) array (deformable bodies)
    # deformable only
    domain: np.ndarray | None = None
    mass: np.ndarray | None = None          # (N,2)
    gp_xy: np.ndarray | None = None
    gp_w: np.ndarray | None = None
    gp_jac: np.ndarray | None = None
    gp_ptr: np.ndarray | None = None
    gp_idx: np.ndarray | None = None
    gp_phi: np.ndarray | None = None
    gp_phix: np.ndarray | None = None
    gp_phiy: np.ndarray | None = None
    nregions: int = 1
    d2d_ptr: np.ndarray | None = None
    d2d_idx: np.ndarray | None = None
    d2d_phi: np.ndarray | None = None
    d2d_phix: np.ndarray | None = None
    d2d_phiy: np.ndarray | None = None
    matrix: list = field(default_factory=list)   # [(dof0, dof1, stiffness, damping)] upper triangle
    # absorbing layers (ALID section, ExportStatic.m:514-550): (xmin, xmax,
    # ymin, ymax, range, alphain, alphaout, betain, betaout, exponent),
    # x velocity history [(t, v)], y velocity history [(t, v)]
    alid: tuple | None = None
    # rigid only
    centre: tuple = (0.0, 0.0)
    rigid_mass: float = 0.0
    rigid_inertia: float = 0.0

    @property
    def n_nodes(self):
        return len(self.X)

    @property
    def n_gauss(self):
        return 0 if self.gp_w is None else len(self.gp_w)


# ----------------------------------------------------------------------------
# node clouds
# ----------------------------------------------------------------------------

def poisson_disc(rng, inside, bbox, h, exclude=None, k=30):
    """Bridson's Poisson-disc sampling with minimum distance h.

    `inside(p)` tells whether a point is admissible; `exclude` are fixed
    points (border nodes) the samples must also keep h away from.
    """
    (x0, y0), (x1, y1) = bbox
    cell = h / np.sqrt(2.0)
    nx = int(np.ceil((x1 - x0) / cell)) + 1
    ny = int(np.ceil((y1 - y0) / cell)) + 1
    grid = -np.ones((nx, ny), dtype=np.int64)
    pts = []
    fixed = [] if exclude is None else [tuple(p) for p in exclude]

    def gidx(p):
        return int((p[0] - x0) / cell), int((p[1] - y0) / cell)

    allp = []

    def add(p):
        allp.append(p)
        i, j = gidx(p)
        if 0 <= i < nx and 0 <= j < ny:
            grid[i, j] = len(allp) - 1

    def ok(p):
        i, j = gidx(p)
        for a in range(max(i - 2, 0), min(i + 3, nx)):
            for b in range(max(j - 2, 0), min(j + 3, ny)):
                q = grid[a, b]
                if q >= 0:
                    d = allp[q]
                    if (d[0] - p[0]) ** 2 + (d[1] - p[1]) ** 2 < h * h:
                        return False
        return True

    for p in fixed:
        add(p)
    # seed
    for _ in range(1000):
        p = (x0 + rng.random() * (x1 - x0), y0 + rng.random() * (y1 - y0))
        if inside(p) and ok(p):
            add(p); pts.append(p)
            break
    active = list(range(len(allp)))
    while active:
        ai = int(rng.integers(len(active)))
        c = allp[active[ai]]
        found = False
        for _ in range(k):
            rr = h * (1.0 + rng.random())
            th = 2.0 * np.pi * rng.random()
            p = (c[0] + rr * np.cos(th), c[1] + rr * np.sin(th))
            if x0 <= p[0] <= x1 and y0 <= p[1] <= y1 and inside(p) and ok(p):
                add(p); pts.append(p); active.append(len(allp) - 1)
                found = True
                break
        if not found:
            active.pop(ai)
    return np.array(pts, dtype=np.float64).reshape(-1, 2)


def disc_nodes(rng, xc, yc, R, n_border, h):
    """Border nodes equally spaced (counter-clockwise) + Poisson-disc interior."""
    # node 0 at the bottom (270 deg); with an even count a node also sits at the top
    th = 2.0 * np.pi * np.arange(n_border) / n_border - 0.5 * np.pi
    border = np.stack([xc + R * np.cos(th), yc + R * np.sin(th)], axis=1)
    margin = 0.5 * h
    inside = lambda p: (p[0] - xc) ** 2 + (p[1] - yc) ** 2 < (R - margin) ** 2
    interior = poisson_disc(rng, inside, ((xc - R, yc - R), (xc + R, yc + R)), h, exclude=border)
    return border, interior


def rectangle_border(x0, y0, x1, y1, h):
    """Counter-clockwise border nodes of a rectangle with spacing <= h."""
    pts = []
    corners = [(x0, y0), (x1, y0), (x1, y1), (x0, y1)]
    for i in range(4):
        a = np.array(corners[i]); b = np.array(corners[(i + 1) % 4])
        n = max(int(np.ceil(np.linalg.norm(b - a) / h)), 1)
        for k in range(n):
            pts.append(a + (b - a) * k / n)
    return np.array(pts)


# ----------------------------------------------------------------------------
# body analysis
# ----------------------------------------------------------------------------

GP_BARY = np.array([[2.0 / 3.0, 1.0 / 6.0, 1.0 / 6.0],
                    [1.0 / 6.0, 2.0 / 3.0, 1.0 / 6.0],
                    [1.0 / 6.0, 1.0 / 6.0, 2.0 / 3.0]])


def _triangulate(X, keep=None):
    tri = Delaunay(X).simplices.astype(np.int64)
    P = X[tri]
    area2 = (P[:, 1, 0] - P[:, 0, 0]) * (P[:, 2, 1] - P[:, 0, 1]) - (P[:, 2, 0] - P[:, 0, 0]) * (P[:, 1, 1] - P[:, 0, 1])
    flip = area2 < 0
    tri[flip] = tri[flip][:, [0, 2, 1]]
    area = 0.5 * np.abs(area2)
    good = area > 1e-14 * np.max(area)
    if keep is not None:
        c = P.mean(axis=1)
        good &= keep(c)
    return tri[good]


def closed_border(n_border, first=0, index=0):
    """Closed border through nodes first .. first+n_border-1, listing its
    first node again at the end, its own border before and after."""
    nodes = np.concatenate([np.arange(first, first + n_border), [first]]).astype(np.int64)
    return Border(nodes=nodes, before=index, after=index)


def simple_borders(corner_ids, n_total):
    """A closed contour through nodes 0 .. n_total-1 cut into 'Simple'
    borders at the given corner node ids (increasing, the first one 0).
    Border k runs from corner k to corner k+1 inclusive, so consecutive
    borders share their end node and the last one ends on node 0; each
    border's before/after are its neighbours on the contour
    (SetUnstructuredNodesForSimpleContours.m)."""
    nc = len(corner_ids)
    out = []
    for k in range(nc):
        a = corner_ids[k]
        b = corner_ids[k + 1] if k + 1 < nc else n_total
        nodes = np.arange(a, b + 1, dtype=np.int64)
        if k + 1 == nc:
            nodes[-1] = 0
        out.append(Border(nodes=nodes, periodicity="Simple", before=(k - 1) % nc, after=(k + 1) % nc))
    return out


def annulus_nodes(rng, xc, yc, R, r, n_outer, n_inner, h):
    """Outer border counter-clockwise, hole border clockwise (so both
    right-hand normals point out of the material), Poisson-disc interior
    between them."""
    th = 2.0 * np.pi * np.arange(n_outer) / n_outer - 0.5 * np.pi
    outer = np.stack([xc + R * np.cos(th), yc + R * np.sin(th)], axis=1)
    th = -2.0 * np.pi * np.arange(n_inner) / n_inner - 0.5 * np.pi
    inner = np.stack([xc + r * np.cos(th), yc + r * np.sin(th)], axis=1)
    margin = 0.5 * h
    inside = lambda p: (r + margin) ** 2 < (p[0] - xc) ** 2 + (p[1] - yc) ** 2 < (R - margin) ** 2
    border = np.concatenate([outer, inner])
    interior = poisson_disc(rng, inside, ((xc - R, yc - R), (xc + R, yc + R)), h, exclude=border)
    return outer, inner, interior


def _rpim_chunk(job):
    pts, X, nn, c = job
    k = nn.shape[1]
    ptr = np.arange(len(pts) + 1, dtype=np.int64) * k
    return shape.rpim(pts, X, ptr, nn.reshape(-1).astype(np.int64), shape_c=c)


def _rpim(points, X, nn, c, workers, device=None):
    """RPIM over fixed-size supports (rows of nn), in point chunks on worker
    processes when there are many points (the result does not depend on the
    worker count: every point is solved on its own), or batched on a torch
    device when one is named."""
    if device:
        return shape.rpim_fixed_torch(points, X, nn, shape_c=c, device=device)
    n = len(points)
    if workers is None or workers <= 1 or n < 200000:
        return _rpim_chunk((points, X, nn, c))
    import concurrent.futures as cf
    step = 50000
    jobs = [(points[a:a + step], X, nn[a:a + step], c) for a in range(0, n, step)]
    with cf.ProcessPoolExecutor(max_workers=workers) as ex:
        parts = list(ex.map(_rpim_chunk, jobs))
    return tuple(np.concatenate([p[i] for p in parts]) for i in range(3))


def deformable_body(border_xy, interior_xy, material, h, rho, radius_factor=2.5,
                    detection=None, contact=None, nregions=1, velocity=(0.0, 0.0),
                    keep_triangle=None, kind="MLS", rpim_c=1.0, rpim_support=30, workers=None,
                    borders=None, periodicity="Simple", device=None):
    """Assemble a deformable meshfree body.

    `borders` lists the body's Border records over the node numbering in
    which the `border_xy` nodes come first; by default the body has one
    closed border through all of them.
    kind "MLS": node radius R = radius_factor * sqrt(area / N) (uniform);
    kind "RPIM": supports are the `rpim_support` nearest nodes of each point
    (shape functions on `workers` processes for large bodies).
    """
    X = np.concatenate([border_xy, interior_xy], axis=0)
    nb = len(border_xy)
    tri = _triangulate(X, keep_triangle)
    P = X[tri]
    area = 0.5 * np.abs((P[:, 1, 0] - P[:, 0, 0]) * (P[:, 2, 1] - P[:, 0, 1]) - (P[:, 2, 0] - P[:, 0, 0]) * (P[:, 1, 1] - P[:, 0, 1]))
    gp = np.einsum('gi,tij->tgj', GP_BARY, P).reshape(-1, 2)
    w = np.full(len(gp), 1.0 / 3.0)
    jac = np.repeat(area, 3)
    # influence radius = radius_factor x the mean local spacing sqrt(A/N)
    spacing = float(np.sqrt(np.sum(area) / len(X)))
    domain = np.full(len(X), radius_factor * spacing)
    if kind == "MLS":
        gptr, gidx = shape.supports(gp, X, domain)
        phi, phix, phiy = shape.mls(gp, X, domain, gptr, gidx)
        dptr, didx = shape.supports(X, X, domain)
        dphi, dphix, dphiy = shape.mls(X, X, domain, dptr, didx)
    else:
        tree = cKDTree(X)
        _, nn = tree.query(gp, k=rpim_support, workers=-1)
        nn = np.sort(nn, axis=1)
        gptr = np.arange(len(gp) + 1, dtype=np.int64) * rpim_support
        gidx = nn.reshape(-1).astype(np.int64)
        phi, phix, phiy = _rpim(gp, X, nn, rpim_c, workers, device)
        _, nn2 = tree.query(X, k=rpim_support, workers=-1)
        nn2 = np.sort(nn2, axis=1)
        dptr = np.arange(len(X) + 1, dtype=np.int64) * rpim_support
        didx = nn2.reshape(-1).astype(np.int64)
        dphi, dphix, dphiy = _rpim(X, X, nn2, rpim_c, workers, device)
    # lumped mass: rho * sum_g c0 * phi_j * sum_k phi_k  (row sum of consistent mass)
    c0 = w * jac
    rowsum = np.add.reduceat(phi, gptr[:-1]) if len(phi) else np.zeros(len(gp))
    m = np.zeros(len(X))
    np.add.at(m, gidx, rho * np.repeat(c0 * rowsum, np.diff(gptr)) * phi)
    mass = np.stack([m, m], axis=1)
    b = Body(rigid=False, material=material, X=X,
             borders=borders if borders is not None else [closed_border(nb)], tri=tri,
             periodicity=periodicity,
             nodal_distance=h, detection=detection if detection is not None else 0.2 * h,
             contact=contact if contact is not None else 0.2 * h, velocity=velocity,
             domain=domain, mass=mass, gp_xy=gp, gp_w=w, gp_jac=jac, gp_ptr=gptr,
             gp_idx=gidx, gp_phi=phi, gp_phix=phix, gp_phiy=phiy, nregions=nregions,
             d2d_ptr=dptr, d2d_idx=didx, d2d_phi=dphi, d2d_phix=dphix, d2d_phiy=dphiy)
    return b


def rectangle_corners(x0, y0, x1, y1, h):
    """Node ids of the four corners in rectangle_border's numbering."""
    ids = [0]
    for a, b in (((x0, y0), (x1, y0)), ((x1, y0), (x1, y1)), ((x1, y1), (x0, y1))):
        ids.append(ids[-1] + max(int(np.ceil(np.hypot(b[0] - a[0], b[1] - a[1]) / h)), 1))
    return ids


def rigid_rectangle(x0, y0, x1, y1, material, h, rho, detection=None, contact=None,
                    xbc=None, ybc=None, periodicity="Simple", simple_sides=False):
    """Rigid rectangle: border nodes on the perimeter, triangulated for mass
    and inertia (AnalyseRigidBody.m; Code/Body.h:2877-2905 recomputes both
    from the triangulation at start-up).  One closed border, or with
    `simple_sides` four 'Simple' borders, one per side (BCs on the first)."""
    X = rectangle_border(x0, y0, x1, y1, h)
    tri = _triangulate(X)
    P = X[tri]
    area = 0.5 * np.abs((P[:, 1, 0] - P[:, 0, 0]) * (P[:, 2, 1] - P[:, 0, 1]) - (P[:, 2, 0] - P[:, 0, 0]) * (P[:, 1, 1] - P[:, 0, 1]))
    xc, yc = 0.5 * (x0 + x1), 0.5 * (y0 + y1)
    mass = rho * (x1 - x0) * (y1 - y0)
    inertia = mass * ((x1 - x0) ** 2 + (y1 - y0) ** 2) / 12.0
    brds = simple_borders(rectangle_corners(x0, y0, x1, y1, h), len(X)) if simple_sides else [closed_border(len(X))]
    brd = brds[0]
    if xbc is not None:
        brd.xbc = xbc
    if ybc is not None:
        brd.ybc = ybc
    return Body(rigid=True, material=material, X=X, borders=brds, tri=tri,
                nodal_distance=h, detection=detection if detection is not None else 0.2 * h,
                contact=contact if contact is not None else 0.2 * h, periodicity=periodicity,
                centre=(xc, yc), rigid_mass=mass, rigid_inertia=inertia)


def rigid_periodic_wall(xmin, xmax, y0, y1, material, h, rho, detection=None, contact=None,
                        xbc=None, ybc=None, bc_border=1):
    """Rigid wall spanning the periodic cell [xmin, xmax) between heights
    y0 < y1, with the two 'Periodic' borders Body::Initialize_segments
    expects (Code/Body.h:403-605): border 0 is the lower face, run left to
    right from the node at (xmin, y0), whose repeated last entry stands for
    that node one period to the right (shift +1); border 1 is the upper
    face, run right to left from the node at (xmax, y1), whose repeated last
    entry stands for that node one period to the left (shift -1).  Both
    right-hand normals point out of the wall.  The BCs are set on border
    `bc_border` (a rigid body sums its borders' BC resultants)."""
    P = xmax - xmin
    n = max(int(np.ceil(P / h)), 3)
    xs = xmin + P * np.arange(n) / n
    lower = np.stack([xs, np.full(n, y0)], axis=1)
    upper = np.stack([xmax - (xs - xmin), np.full(n, y1)], axis=1)
    X = np.concatenate([lower, upper])
    tri = _triangulate(X)
    brd = [Border(nodes=np.concatenate([np.arange(n), [0]]).astype(np.int64), periodicity="Periodic",
                  before=0, after=0),
           Border(nodes=np.concatenate([np.arange(n, 2 * n), [n]]).astype(np.int64), periodicity="Periodic",
                  before=1, after=1)]
    if xbc is not None:
        brd[bc_border].xbc = xbc
    if ybc is not None:
        brd[bc_border].ybc = ybc
    mass = rho * P * (y1 - y0)
    inertia = mass * (P ** 2 + (y1 - y0) ** 2) / 12.0
    return Body(rigid=True, material=material, X=X, borders=brd, tri=tri, nodal_distance=h,
                detection=detection if detection is not None else 0.2 * h,
                contact=contact if contact is not None else 0.2 * h, periodicity="Periodic",
                centre=(0.5 * (xmin + xmax), 0.5 * (y0 + y1)), rigid_mass=mass, rigid_inertia=inertia)
