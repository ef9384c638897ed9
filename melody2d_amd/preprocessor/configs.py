"""Seeded workload generators.

Seeds: workload k uses seed 1000 + k; inside a workload one numpy PCG64
stream per body in creation order (seed sequence spawned from the workload
seed), so a body's node cloud does not depend on how many bodies follow it.

Workloads
  tiny        two deformable discs (~30 nodes each) touching, no platens
  tiny_platen two small discs between two rigid platens (driven)
  c1          two elastic discs R = 1 mm (64 border nodes, Poisson-disc
              interior at 0.1 mm), neo-Hookean, between rigid platens; top
              platen driven down at 0.1 m/s; penalty Mohr-Coulomb mu = 0.3
  c2          biaxial compression of polydisperse deformable grains
              (R ~ U[0.8, 1.2] mm, ~200 MLS nodes each) placed by random
              sequential deposition, rigid walls, mu = 0.5, density x100
              (mass scaling expressed as density).  1000 grains cannot fit
              a 40 x 40 mm box (their area is ~3200 mm^2), so the box side
              is chosen for a 0.45 area fraction (~84 mm for 1000 grains).
  c3          a 0.1 x 0.1 m plate with random circular holes (porosity 0.3),
              ~1M RPIM nodes, linear elastic, no contact: the
              internal-force workload.
"""
from __future__ import annotations

import math
import os

import numpy as np

from .bodies import (BC, annulus_nodes, closed_border, deformable_body, disc_nodes, poisson_disc,
                     rectangle_border, rectangle_corners, rigid_periodic_wall, rigid_rectangle, simple_borders)
from .export import Model

E_GRAIN = 1.0e7
NU_GRAIN = 0.3
RHO_GRAIN = 1000.0


def _streams(seed, n):
    ss = np.random.SeedSequence(seed)
    return [np.random.Generator(np.random.PCG64(s)) for s in ss.spawn(n)]


def _critical_dt(model, kn):
    """0.2 x the smaller of the elastic (h / c) and contact (2 sqrt(m / (kn L)))
    stability limits over the deformable bodies."""
    dts = []
    for b in model.bodies:
        if b.rigid:
            continue
        mat = [m for m in model.materials if m[0] == b.material][0]
        rho, E = mat[2][0], mat[2][3]
        c = math.sqrt(E / rho)
        h = b.nodal_distance
        mmin = float(np.min(b.mass[:, 0]))
        dts.append(h / c)
        dts.append(2.0 * math.sqrt(mmin / (kn * h)))
    return 0.2 * min(dts)


def tiny(seed=1000, h=0.35e-3, R=1.0e-3, with_platens=False, nregions=2, overlap=2e-6):
    """Two small discs pressed together by `overlap` (so contact forces act
    from the first step), optionally between rigid platens that also overlap."""
    rngs = _streams(seed, 4)
    nb = 2 * max(int(round(math.pi * R / h)), 4)
    bodies = []
    for i, yc in enumerate([R, 3 * R - overlap]):
        bxy, ixy = disc_nodes(rngs[i], 0.0, yc, R, nb, h)
        bodies.append(deformable_body(bxy, ixy, "Grain", h, RHO_GRAIN, detection=0.5 * h,
                                      contact=0.5 * h, nregions=nregions,
                                      velocity=(0.0, -0.05 * i)))
    mats = [("Grain", "NeoHookean", [RHO_GRAIN, 0.0, 0.0, E_GRAIN, NU_GRAIN, 1.0])]
    laws = [("Grain", "Grain", "MohrCoulomb", "Evolutive", [1e13, 1e13, 0.3, 0.0, 0.0])]
    if with_platens:
        mats.append(("Steel", "ElasticLinear", [7800.0, 0.0, 0.0, 2e11, 0.3, 1.0]))
        laws.append(("Grain", "Steel", "MohrCoulomb", "Evolutive", [1e13, 1e13, 0.3, 0.0, 0.0]))
        t = 0.3e-3
        bodies.append(rigid_rectangle(-1.5e-3, -t + overlap, 1.5e-3, overlap, "Steel", h, 7800.0,
                                      detection=0.5 * h, contact=0.5 * h,
                                      xbc=BC("Driven", [(0.0, 0.0), (1e3, 0.0)]),
                                      ybc=BC("Driven", [(0.0, 0.0), (1e3, 0.0)])))
        top = 4 * R - 2 * overlap
        bodies.append(rigid_rectangle(-1.5e-3, top, 1.5e-3, top + t, "Steel", h, 7800.0,
                                      detection=0.5 * h, contact=0.5 * h,
                                      xbc=BC("Driven", [(0.0, 0.0), (1e3, 0.0)]),
                                      ybc=BC("Driven", [(0.0, -0.1), (1e3, -0.1)])))
    m = Model(name="tiny_platen" if with_platens else "tiny", materials=mats, contact_laws=laws,
              bodies=bodies, periodic=(-1.0, 1.0))
    m.dt = _critical_dt(m, 1e13)
    m.contact_update_period = 10 * m.dt
    m.tend = 1.0
    return m


def c1(seed=1001, h=1.0e-4, R=1.0e-3, nregions=4):
    """Two elastic discs between rigid platens, FP64 on the CPU."""
    rngs = _streams(seed, 4)
    bodies = []
    for i, yc in enumerate([R, 3 * R]):
        bxy, ixy = disc_nodes(rngs[i], 0.0, yc, R, 64, 0.8 * h)
        bodies.append(deformable_body(bxy, ixy, "Grain", h, RHO_GRAIN, detection=0.5 * h,
                                      contact=0.5 * h, nregions=nregions))
    t = 0.3e-3
    bodies.append(rigid_rectangle(-1.5e-3, -t, 1.5e-3, 0.0, "Steel", h, 7800.0, detection=0.5 * h,
                                  contact=0.5 * h, xbc=BC("Driven", [(0.0, 0.0), (1e3, 0.0)]),
                                  ybc=BC("Driven", [(0.0, 0.0), (1e3, 0.0)])))
    bodies.append(rigid_rectangle(-1.5e-3, 4 * R, 1.5e-3, 4 * R + t, "Steel", h, 7800.0,
                                  detection=0.5 * h, contact=0.5 * h,
                                  xbc=BC("Driven", [(0.0, 0.0), (1e3, 0.0)]),
                                  ybc=BC("Driven", [(0.0, -0.1), (1e3, -0.1)])))
    mats = [("Grain", "NeoHookean", [RHO_GRAIN, 0.0, 0.0, E_GRAIN, NU_GRAIN, 1.0]),
            ("Steel", "ElasticLinear", [7800.0, 0.0, 0.0, 2e11, 0.3, 1.0])]
    laws = [("Grain", "Grain", "MohrCoulomb", "Evolutive", [1e13, 1e13, 0.3, 0.0, 0.0]),
            ("Grain", "Steel", "MohrCoulomb", "Evolutive", [1e13, 1e13, 0.3, 0.0, 0.0])]
    m = Model(name="c1_two_discs", materials=mats, contact_laws=laws, bodies=bodies, periodic=(-1.0, 1.0))
    m.dt = _critical_dt(m, 1e13)
    m.contact_update_period = 10 * m.dt
    m.tend = 10000 * m.dt
    return m


def _rsa(rng, n, rmin, rmax, W, H, gap, max_tries=20000000):
    """Random sequential deposition of n discs (radius ~ U[rmin, rmax]) in a
    W x H box, rejecting overlaps (with a clearance `gap`)."""
    cell = 2 * rmax + gap
    nx, ny = int(W / cell) + 1, int(H / cell) + 1
    grid = [[[] for _ in range(ny)] for _ in range(nx)]
    out = []
    tries = 0
    while len(out) < n:
        tries += 1
        if tries > max_tries:
            raise RuntimeError("random sequential deposition stalled at %d grains" % len(out))
        r = rmin + (rmax - rmin) * rng.random()
        x = r + (W - 2 * r) * rng.random()
        y = r + (H - 2 * r) * rng.random()
        i, j = int(x / cell), int(y / cell)
        ok = True
        for a in range(max(i - 1, 0), min(i + 2, nx)):
            for b in range(max(j - 1, 0), min(j + 2, ny)):
                for (xo, yo, ro) in grid[a][b]:
                    if (x - xo) ** 2 + (y - yo) ** 2 < (r + ro + gap) ** 2:
                        ok = False
                        break
                if not ok:
                    break
            if not ok:
                break
        if ok:
            out.append((x, y, r))
            grid[i][j].append((x, y, r))
    return out


def _deposit(rng, n, rmin, rmax, W, max_rolls=400):
    """Sequential deposition under gravity (Visscher-Bolsterli): disc k,
    radius ~ U[rmin, rmax], is dropped at a random abscissa in [r, W - r]
    onto the pile, then rolls down over the disc it lands on until it rests
    on two discs (its centre between theirs), on the floor y = 0 or against
    a side wall x = 0 or x = W.  Every disc rests on the pile, so the
    packing comes out with every grain touching two others or a wall; discs
    never overlap (touching to rounding).  Returns [(x, y, r)]."""
    cell = 2.0 * rmax
    grid = {}
    xs, ys, rs = [], [], []
    cols = {}

    def near(x, y, reach):
        i0, i1 = int((x - reach) // cell), int((x + reach) // cell)
        j0, j1 = int((y - reach) // cell), int((y + reach) // cell)
        for i in range(i0, i1 + 1):
            for j in range(j0, j1 + 1):
                for k in grid.get((i, j), ()):
                    yield k

    def drop(x, r, ylim):
        # highest resting height not above ylim when falling at abscissa x
        best, sup = r, -1
        for i in range(int((x - r - rmax) // cell), int((x + r + rmax) // cell) + 1):
            for k in cols.get(i, ()):
                dx = x - xs[k]
                R = r + rs[k]
                if abs(dx) < R:
                    yy = ys[k] + math.sqrt(R * R - dx * dx)
                    if best < yy <= ylim:
                        best, sup = yy, k
        return best, sup

    def overlaps(x, y, r):
        return any(math.hypot(x - xs[k], y - ys[k]) < (r + rs[k]) * (1.0 - 1e-9) for k in near(x, y, r + rmax))

    def place():
        r = rmin + (rmax - rmin) * rng.random()
        x = r + (W - 2.0 * r) * rng.random()
        y, s1 = drop(x, r, math.inf)
        rolls, prev = 0, -1
        while s1 >= 0 and rolls < max_rolls:
            rolls += 1
            x1, y1, R1 = xs[s1], ys[s1], r + rs[s1]
            th = math.atan2(x - x1, y - y1)     # from the vertical, positive towards +x
            sgn = 1.0 if th > 0 else (-1.0 if th < 0 else (1.0 if rng.random() < 0.5 else -1.0))
            a0 = abs(th)
            best_a, best_k, kind = 0.5 * math.pi, -1, "fall"
            if y1 - r < R1:                     # floor: y1 + R1 cos a = r
                c = (r - y1) / R1
                if -1.0 <= c <= 1.0:
                    a = math.acos(c)
                    if a0 - 1e-12 < a < best_a:
                        best_a, kind = a, "floor"
            wx = (W - r) if sgn > 0 else r      # side wall: x1 + sgn R1 sin a = wx
            sv = (wx - x1) * sgn / R1
            if 0.0 <= sv <= 1.0:
                a = math.asin(sv)
                if a0 - 1e-12 < a < best_a:
                    best_a, kind = a, "wall"
            # a disc touched where the roll starts stops it at once (the
            # disc rests in the pocket), except the one it just rolled off
            for k in near(x1, y1, R1 + r + rmax):
                if k == s1 or k == prev:
                    continue
                Rk = r + rs[k]
                dxk, dyk = xs[k] - x1, ys[k] - y1
                D = math.hypot(dxk, dyk)
                if D == 0.0 or D > R1 + Rk or D < abs(R1 - Rk):
                    continue
                aa = (R1 * R1 - Rk * Rk + D * D) / (2.0 * D)
                hh = math.sqrt(max(R1 * R1 - aa * aa, 0.0))
                px, py = aa * dxk / D, aa * dyk / D
                for q in (1.0, -1.0):
                    ix, iy = px - q * hh * dyk / D, py + q * hh * dxk / D
                    a = math.atan2(ix, iy) * sgn
                    if a0 - 1e-9 < a < best_a:
                        best_a, best_k, kind = a, k, "disc"
            x, y = x1 + sgn * R1 * math.sin(best_a), y1 + R1 * math.cos(best_a)
            if kind in ("floor", "wall"):
                break
            if kind == "fall":
                y, s1 = drop(x, r, y)
                prev = -1
                continue
            if min(x1, xs[best_k]) <= x <= max(x1, xs[best_k]):
                break
            prev, s1 = s1, best_k
        return x, y, r

    for _ in range(n):
        # with equal radii a disc can roll into a pocket another disc
        # already fills; such a drop is drawn again
        x, y, r = place()
        tries = 1
        while overlaps(x, y, r):
            if tries == 1000:
                raise RuntimeError("deposition found no free resting place in 1000 drops")
            x, y, r = place()
            tries += 1
        xs.append(x)
        ys.append(y)
        rs.append(r)
        k = len(xs) - 1
        grid.setdefault((int(x // cell), int(y // cell)), []).append(k)
        cols.setdefault(int(x // cell), []).append(k)
    return list(zip(xs, ys, rs))


def _grain(job):
    """One seeded polydisperse grain: border + Poisson-disc interior, MLS."""
    ss, x, y, r, nodes_per_grain, rho, nregions = job
    rng = np.random.Generator(np.random.PCG64(ss))
    # spacing so that border + Poisson interior gives ~nodes_per_grain nodes
    h = r * 0.108 * math.sqrt(200.0 / nodes_per_grain)
    nb = int(round(2 * math.pi * r / h))
    bxy, ixy = disc_nodes(rng, x, y, r, nb, h)
    return deformable_body(bxy, ixy, "Grain", h, rho, detection=0.5 * h, contact=0.5 * h, nregions=nregions)


def _map(fn, jobs, workers):
    """Build bodies in worker processes; the result does not depend on the
    worker count (each job carries its own seed)."""
    if workers is None:
        workers = min(16, os.cpu_count() or 1)
    if workers <= 1 or len(jobs) < 64:
        return [fn(j) for j in jobs]
    import concurrent.futures as cf
    with cf.ProcessPoolExecutor(max_workers=workers) as ex:
        return list(ex.map(fn, jobs, chunksize=max(1, len(jobs) // (4 * workers))))


def c2(n_grains=1000, seed=1002, nodes_per_grain=200, area_fraction=0.45, mu=0.5,
       mass_scaling=100.0, confinement=1e5, wall_velocity=0.05, nregions=1, workers=None, packing="deposit"):
    """Biaxial compression of polydisperse deformable grains with rigid walls.

    packing "deposit" (default): the grains are deposited under gravity
    (_deposit) into a box whose width makes the pile about square (1000
    grains of radius U[0.8, 1.2] mm cover ~3200 mm^2, more than a
    40 x 40 mm box holds); every grain rests on two others or a wall, so
    contacts act from the start of the compression.  The top wall is put on
    the highest grain.  packing "rsa": random sequential addition at
    `area_fraction` in a square box, grains apart (a loose start the walls
    take thousands of steps to reach)."""
    rng_place = np.random.Generator(np.random.PCG64(np.random.SeedSequence(seed).spawn(1)[0]))
    rmin, rmax = 0.8e-3, 1.2e-3
    mean_area = math.pi * (rmin * rmin + rmin * rmax + rmax * rmax) / 3.0
    if packing == "rsa":
        side = math.sqrt(n_grains * mean_area / area_fraction)
        grains = _rsa(rng_place, n_grains, rmin, rmax, side, side, gap=0.02e-3)
        width = height = side
    else:
        width = math.sqrt(n_grains * mean_area / 0.74)
        grains = _deposit(rng_place, n_grains, rmin, rmax, width)
        height = max(y + r for _, y, r in grains)
    seeds = np.random.SeedSequence(seed + 7).spawn(n_grains)
    rho = RHO_GRAIN * mass_scaling
    jobs = [(s, x, y, r, nodes_per_grain, rho, nregions) for s, (x, y, r) in zip(seeds, grains)]
    bodies = _map(_grain, jobs, workers)
    hmax = max(b.nodal_distance for b in bodies)
    t = 2.0e-3
    hw = hmax
    # bottom and left walls fixed; top wall driven down; right wall under confinement
    fixed = lambda: BC("Driven", [(0.0, 0.0), (1e3, 0.0)])
    walls = [
        rigid_rectangle(-t, -t, width + t, 0.0, "Steel", hw, 7800.0, detection=0.5 * hw, contact=0.5 * hw,
                        xbc=fixed(), ybc=fixed()),
        rigid_rectangle(-t, 0.0, 0.0, height, "Steel", hw, 7800.0, detection=0.5 * hw, contact=0.5 * hw,
                        xbc=fixed(), ybc=fixed()),
        rigid_rectangle(-t, height, width + t, height + t, "Steel", hw, 7800.0, detection=0.5 * hw,
                        contact=0.5 * hw, xbc=fixed(),
                        ybc=BC("Driven", [(0.0, -wall_velocity), (1e3, -wall_velocity)])),
    ]
    # a rigid border takes the Neumann value at every border node times its
    # tributary length (Code/Body.h:854-858): scale so the net push equals
    # confinement x the wall's face length
    rw = rigid_rectangle(width, 0.0, width + t, height, "Steel", hw, 7800.0, detection=0.5 * hw,
                         contact=0.5 * hw, ybc=fixed())
    perim = 2 * (t + height)
    p = -confinement * height / perim
    rw.borders[0].xbc = BC("Oriented", [(0.0, p), (1e3, p)])
    walls.append(rw)
    bodies += walls
    mats = [("Grain", "NeoHookean", [rho, 0.0, 0.0, E_GRAIN, NU_GRAIN, 1.0]),
            ("Steel", "ElasticLinear", [7800.0 * mass_scaling, 0.0, 0.0, 2e11, 0.3, 1.0])]
    laws = [("Grain", "Grain", "MohrCoulomb", "Evolutive", [1e13, 1e13, mu, 0.0, 0.0]),
            ("Grain", "Steel", "MohrCoulomb", "Evolutive", [1e13, 1e13, mu, 0.0, 0.0])]
    m = Model(name="c2_biaxial_%d" % n_grains, materials=mats, contact_laws=laws, bodies=bodies,
              periodic=(-1.0, 1.0))
    m.dt = _critical_dt(m, 1e13)
    m.contact_update_period = 10 * m.dt
    m.tend = 5000 * m.dt
    return m


def _holes(rng, x0, y0, x1, y1, porosity, rmin, rmax, margin, clearance, avoid_x=None, max_tries=200000):
    """Seeded circular holes, radius U[rmin, rmax], at least `margin` from the
    box edges and `clearance` from each other (and from the line x = avoid_x
    when given), added until their area reaches `porosity` of the box."""
    holes = []
    target = porosity * (x1 - x0) * (y1 - y0)
    area = 0.0
    tries = 0
    while area < target and tries < max_tries:
        tries += 1
        r = rng.uniform(rmin, rmax)
        if x1 - x0 <= 2 * (r + margin) or y1 - y0 <= 2 * (r + margin):
            continue
        x = rng.uniform(x0 + r + margin, x1 - r - margin)
        y = rng.uniform(y0 + r + margin, y1 - r - margin)
        if avoid_x is not None and abs(x - avoid_x) < r + margin:
            continue
        if any((x - a) ** 2 + (y - b) ** 2 < (r + c + clearance) ** 2 for a, b, c in holes):
            continue
        holes.append((x, y, r))
        area += math.pi * r * r
    return holes


def _plate_body(rng, x0, y0, x1, y1, holes, h, material, rho, nregions=1, workers=None, device=None,
                hole_borders=True, simple_sides=False, left_offset=0.0):
    """One perforated rectangular RPIM body (30-node supports, Gaussian RBF
    exp(-(3 d / D)^2), D the support size, as Preprocessor/Gauss_RBF.m;
    three Gauss points per triangle).  Borders: the outer rectangle
    (counter-clockwise; one closed border, or four 'Simple' borders, one per
    side, with `simple_sides`), then one closed border per hole run
    clockwise, as Preprocessor/SetUnstructuredNodesForHoledContour.m lays out
    a holed contour (without `hole_borders` the holes are free surfaces
    without border nodes).  Interior nodes on a jittered grid of spacing h;
    triangles whose centroid falls in a hole are dropped.  `left_offset`
    moves the nodes of the left side (corners excepted) along y by that
    fraction of their spacing, so that two plates bonded side to side do not
    put nodes on each other's nodes.  Masses are lumped uniformly (rho x
    solid area / N): RPIM row sums are not all positive."""
    from .bodies import Border
    H = np.array(holes).reshape(-1, 3)

    def in_holes(p, grow=0.0):
        inside = np.zeros(len(p), dtype=bool)
        for x, y, r in H:
            inside |= (p[:, 0] - x) ** 2 + (p[:, 1] - y) ** 2 < (r + grow) ** 2
        return inside

    outer = rectangle_border(x0, y0, x1, y1, h)
    corners = rectangle_corners(x0, y0, x1, y1, h)
    if left_offset:
        n3 = len(outer) - corners[3]
        dy = (y1 - y0) / n3
        outer[corners[3] + 1:, 1] -= left_offset * dy
    parts = [outer]
    if hole_borders:
        for x, y, r in H:
            n = max(8, int(math.ceil(2 * math.pi * r / h)))
            th = -2.0 * np.pi * np.arange(n) / n - 0.5 * np.pi
            parts.append(np.stack([x + r * np.cos(th), y + r * np.sin(th)], axis=1))
    border = np.concatenate(parts)
    gxs = np.arange(x0 + h, x1 - 0.5 * h, h)
    gys = np.arange(y0 + h, y1 - 0.5 * h, h)
    gx, gy = np.meshgrid(gxs, gys, indexing="ij")
    pts = np.stack([gx.ravel(), gy.ravel()], axis=1)
    pts += rng.uniform(-0.25 * h, 0.25 * h, pts.shape)
    interior = pts[~in_holes(pts, (0.5 if hole_borders else 0.3) * h)]
    brds = simple_borders(corners, len(outer)) if simple_sides else [closed_border(len(outer))]
    off = len(outer)
    for part in parts[1:]:
        brds.append(closed_border(len(part), off, len(brds)))
        off += len(part)
    body = deformable_body(border, interior, material, h, rho, kind="RPIM", rpim_c=3.0, rpim_support=30,
                           keep_triangle=lambda c: ~in_holes(c), nregions=nregions, workers=workers,
                           detection=0.5 * h, contact=0.5 * h, borders=brds, device=device)
    solid = (x1 - x0) * (y1 - y0) - sum(math.pi * r * r for _, _, r in H)
    body.mass = np.full((body.n_nodes, 2), rho * solid / body.n_nodes)
    return body


def c3(n_nodes=1_000_000, seed=1003, side=0.1, porosity=0.3, rmin=1.0e-3, rmax=4.0e-3, E=7.0e10, nu=0.33,
       rho=2700.0, nregions=64, workers=None, device=None, hole_borders=True):
    """A side x side plate with seeded random circular holes (radius
    U[rmin, rmax], no overlap, until the hole area reaches `porosity`),
    ~n_nodes RPIM nodes, linear elastic: one body (_plate_body) whose holes
    are closed borders of their own.  The internal-force workload: no
    contact (NO_MONITORING, no contact update within a run).  Masses are
    kept as written (no UPDATE_MASS_MATRIX).  `device` names a torch device
    on which the RPIM systems are solved (the 1M-node plate has ~7M)."""
    if workers is None:
        workers = min(16, os.cpu_count() or 1)
    rng_holes, rng_nodes = _streams(seed, 2)
    holes = _holes(rng_holes, 0.0, 0.0, side, side, porosity, rmin, rmax, 0.5e-3 * side / 0.1,
                   0.2e-3 * side / 0.1)
    H = np.array(holes)
    area = sum(math.pi * r * r for _, _, r in holes)
    solid = side * side - area
    h = math.sqrt(solid / n_nodes)
    body = _plate_body(rng_nodes, 0.0, 0.0, side, side, holes, h, "Plate", rho, nregions=nregions, workers=workers,
                       device=device, hole_borders=hole_borders)
    m = Model(name="c3_plate_%d" % body.n_nodes, materials=[("Plate", "ElasticLinear", [rho, 0.0, 0.0, E, nu, 1.0])],
              contact_laws=[("Plate", "Plate", "Frictionless", "Evolutive", [1e13])], bodies=[body],
              periodic=(-1.0, 1.0))
    m.flags = [f for f in m.flags if f != "UPDATE_MASS_MATRIX"] + ["NO_MONITORING"]
    # RPIM stiffness is higher than a linear element's at the same spacing:
    # its stable step is about 0.2 h / c, so half of that is used
    m.dt = 0.1 * h / math.sqrt(E / rho)
    m.contact_update_period = 1.0e6 * m.dt
    m.tend = 2000 * m.dt
    m.holes = H
    return m


def c5_job(n_nodes, seed, cohesive, porosity=0.3, E=7.0e10, nu=0.33, rho=2700.0, strain_rate=10.0,
           peak=1.0e6, softening=1.0e-5, workers=1, device=None):
    """One serving job built like the perforated plate: ~n_nodes RPIM nodes
    at the plate's spacing, holes of radius U[1, 4] mm up to porosity 0.3
    (the side follows the size, so small plates hold few holes), pulled in uniaxial tension at `strain_rate` by
    soft Dirichlet springs on its left and right sides (the outer contour is
    four 'Simple' borders).  A cohesive job is two plates bonded along
    x = side / 2 by a CZMlinear law (peak traction `peak`, linear softening
    over `softening`, contacts closer than 2 um start intact under
    INITIALIZE_CZM: Code/Proximity.h:481-541); holes keep clear of the
    bond line.  Otherwise one plate, with a frictionless self-law it never
    uses (NO_MONITORING)."""
    rng_holes, rng_a, rng_b = _streams(seed, 3)
    h = 8.4e-5
    side = math.sqrt(n_nodes * h * h / (1.0 - porosity))
    holes = _holes(rng_holes, 0.0, 0.0, side, side, porosity, 1.0e-3, 4.0e-3, 0.5e-3, 0.2e-3,
                   avoid_x=0.5 * side if cohesive else None, max_tries=20000)
    # spring stiffness per unit length of border: a tenth of E / h
    k = 0.1 * E / h
    v = strain_rate * side
    hold = lambda: BC("Soft", [(0.0, 0.0), (1e3, 0.0)], (k, 0.0))
    pull = lambda: BC("Soft", [(0.0, v), (1e3, v)], (k, 0.0))
    if cohesive:
        mid = 0.5 * side
        ha = [c for c in holes if c[0] < mid]
        hb = [c for c in holes if c[0] >= mid]
        a = _plate_body(rng_a, 0.0, 0.0, mid, side, ha, h, "Plate", rho, workers=workers, device=device,
                        simple_sides=True)
        b = _plate_body(rng_b, mid, 0.0, side, side, hb, h, "Plate", rho, workers=workers, device=device,
                        simple_sides=True, left_offset=0.5)
        a.borders[3].xbc = hold()
        b.borders[1].xbc = pull()
        bodies = [a, b]
        gapnlim = peak / 1e13 + softening
        laws = [("Plate", "Plate", "CZMlinear", "Evolutive", [1e13, peak, gapnlim, 0.0, 1.0, 2e-6])]
    else:
        a = _plate_body(rng_a, 0.0, 0.0, side, side, holes, h, "Plate", rho, workers=workers, device=device,
                        simple_sides=True)
        a.borders[3].xbc = hold()
        a.borders[1].xbc = pull()
        bodies = [a]
        laws = [("Plate", "Plate", "Frictionless", "Evolutive", [1e13])]
    m = Model(name="c5_plate_%d%s" % (sum(b.n_nodes for b in bodies), "_czm" if cohesive else ""),
              materials=[("Plate", "ElasticLinear", [rho, 0.0, 0.0, E, nu, 1.0])], contact_laws=laws, bodies=bodies,
              periodic=(-1.0, 1.0))
    m.flags = [f for f in m.flags if f != "UPDATE_MASS_MATRIX"]
    if cohesive:
        m.flags = m.flags + ["INITIALIZE_CZM", "NO_MONITORING"]
    else:
        m.flags = m.flags + ["NO_MONITORING"]
    m.dt = 0.1 * h / math.sqrt(E / rho)
    m.contact_update_period = 10 * m.dt
    m.tend = 20000 * m.dt
    m.holes = np.array(holes).reshape(-1, 3)
    return m


def tiny_periodic(seed=1010, h=0.45e-3, R=1.0e-3, period=5.0e-3, shear=20.0, overlap=2e-6, t=0.3e-3):
    """A miniature periodic shear cell: two discs between two rigid
    'Periodic' walls spanning the cell [-period/2, period/2).  Disc A
    straddles the right end of the cell; disc B touches A across the
    periodic boundary (one period to the left of its contact partner).  The
    top wall slides at `shear` m/s (frictionless against the grains), so
    contact points walk along its segments and across its periodic seam;
    the bottom wall is fixed (Mohr-Coulomb).  Exercises the periodic shift
    ranges of Update_proximity (Code/Proximity.h:100-101, 145-147), the
    shifted segment stencils (Code/Body.h:403-605) and the walk across them
    (Code/Body.h:1402-1599)."""
    rngs = _streams(seed, 2)
    xmin, xmax = -0.5 * period, 0.5 * period
    nb = 2 * max(int(round(math.pi * R / h)), 4)
    xa = xmax - 0.5 * R
    xb = xa + 2 * R - overlap - period
    bodies = []
    for i, xc in enumerate([xa, xb]):
        bxy, ixy = disc_nodes(rngs[i], xc, R - overlap, R, nb, h)
        bodies.append(deformable_body(bxy, ixy, "Grain", h, RHO_GRAIN, detection=0.5 * h, contact=0.5 * h,
                                      nregions=2))
    fixed = lambda: BC("Driven", [(0.0, 0.0), (1e3, 0.0)])
    bodies.append(rigid_periodic_wall(xmin, xmax, -t, 0.0, "SteelB", h, 7800.0, detection=0.5 * h,
                                      contact=0.5 * h, xbc=fixed(), ybc=fixed()))
    ytop = 2 * R - 2 * overlap
    bodies.append(rigid_periodic_wall(xmin, xmax, ytop, ytop + t, "SteelT", h, 7800.0, detection=0.5 * h,
                                      contact=0.5 * h, xbc=BC("Driven", [(0.0, shear), (1e3, shear)]),
                                      ybc=fixed(), bc_border=0))
    mats = [("Grain", "NeoHookean", [RHO_GRAIN, 0.0, 0.0, E_GRAIN, NU_GRAIN, 1.0]),
            ("SteelB", "ElasticLinear", [7800.0, 0.0, 0.0, 2e11, 0.3, 1.0]),
            ("SteelT", "ElasticLinear", [7800.0, 0.0, 0.0, 2e11, 0.3, 1.0])]
    laws = [("Grain", "Grain", "MohrCoulomb", "Evolutive", [1e13, 1e13, 0.3, 0.0, 0.0]),
            ("Grain", "SteelB", "MohrCoulomb", "Evolutive", [1e13, 1e13, 0.3, 0.0, 0.0]),
            ("Grain", "SteelT", "Frictionless", "Evolutive", [1e13])]
    m = Model(name="tiny_periodic", materials=mats, contact_laws=laws, bodies=bodies, periodic=(xmin, xmax))
    m.dt = _critical_dt(m, 1e13)
    m.contact_update_period = 10 * m.dt
    m.tend = 1.0
    return m


def tiny_simple(seed=1011, h=0.4e-3, W=3.0e-3, H=1.0e-3, R=0.8e-3, overlap=2e-6, t=0.9e-3):
    """A deformable block whose contour is cut into four 'Simple' borders
    (one per side, sharing their corner nodes), on a rigid platen that also
    has four 'Simple' borders, with a disc pressed on its top face.  BCs on
    the block: soft Dirichlet on the left side, oriented Neumann on the
    right, following Neumann on the top.  Exercises the Simple branches of
    the segment stencils and tributary lengths (Code/Body.h:403-661) and the
    BC spreading over borders that share end nodes (Code/Body.h:669-901)."""
    rngs = _streams(seed, 2)
    x0, y0, x1, y1 = -0.5 * W, 0.0, 0.5 * W, H
    bxy = rectangle_border(x0, y0, x1, y1, h)
    corners = rectangle_corners(x0, y0, x1, y1, h)
    margin = 0.5 * h
    inside = lambda p: (x0 + margin < p[0] < x1 - margin) and (y0 + margin < p[1] < y1 - margin)
    ixy = poisson_disc(rngs[0], inside, ((x0, y0), (x1, y1)), h, exclude=bxy)
    brds = simple_borders(corners, len(bxy))
    brds[3].xbc = BC("Soft", [(0.0, 0.0), (1e3, 0.0)], (2e5, 1.0))
    brds[1].xbc = BC("Oriented", [(0.0, -1e4), (1e3, -2e4)])
    brds[2].ybc = BC("Following", [(0.0, -5e3), (1e3, -5e3)])
    block = deformable_body(bxy, ixy, "Grain", h, RHO_GRAIN, detection=0.5 * h, contact=0.5 * h,
                            nregions=2, borders=brds)
    nbd = 2 * max(int(round(math.pi * R / h)), 4)
    dxy, dixy = disc_nodes(rngs[1], 0.3e-3, y1 + R - overlap, R, nbd, h)
    disc = deformable_body(dxy, dixy, "Grain", h, RHO_GRAIN, detection=0.5 * h, contact=0.5 * h,
                           velocity=(0.0, -0.05))
    platen = rigid_rectangle(x0 - 0.5e-3, -t + overlap, x1 + 0.5e-3, overlap, "Steel", h, 7800.0,
                             detection=0.5 * h, contact=0.5 * h, simple_sides=True,
                             xbc=BC("Driven", [(0.0, 0.0), (1e3, 0.0)]), ybc=BC("Driven", [(0.0, 0.0), (1e3, 0.0)]))
    mats = [("Grain", "NeoHookean", [RHO_GRAIN, 0.0, 0.0, E_GRAIN, NU_GRAIN, 1.0]),
            ("Steel", "ElasticLinear", [7800.0, 0.0, 0.0, 2e11, 0.3, 1.0])]
    laws = [("Grain", "Grain", "MohrCoulomb", "Evolutive", [1e13, 1e13, 0.3, 0.0, 0.0]),
            ("Grain", "Steel", "MohrCoulomb", "Evolutive", [1e13, 1e13, 0.3, 0.0, 0.0])]
    m = Model(name="tiny_simple", materials=mats, contact_laws=laws, bodies=[block, disc, platen],
              periodic=(-1.0, 1.0))
    m.dt = _critical_dt(m, 1e13)
    m.contact_update_period = 10 * m.dt
    m.tend = 1.0
    return m


def tiny_holed(seed=1012, h=0.3e-3, R=1.5e-3, r=0.7e-3, rd=0.5e-3, hd=0.2e-3, overlap=2e-6):
    """A deformable ring (outer border counter-clockwise, hole border
    clockwise: two closed borders) with a small disc inside the hole pressed
    against the hole's lower face, and a disc resting on the ring's top.
    Exercises bodies with several borders: per-border stencils and
    before/after indices, neighbour triples naming border 1, contact
    elements with borderS / borderM = 1 (Code/Proximity.h:112-330)."""
    rngs = _streams(seed, 3)
    no = 2 * max(int(round(math.pi * R / h)), 4)
    ni = 2 * max(int(round(math.pi * r / h)), 4)
    outer, inner, ixy = annulus_nodes(rngs[0], 0.0, 0.0, R, r, no, ni, h)
    brds = [closed_border(no, 0, 0), closed_border(ni, no, 1)]
    ring = deformable_body(np.concatenate([outer, inner]), ixy, "Grain", h, RHO_GRAIN, detection=0.5 * h,
                           contact=0.5 * h, nregions=2, borders=brds,
                           keep_triangle=lambda c: c[:, 0] ** 2 + c[:, 1] ** 2 > r * r)
    nbd = 2 * max(int(round(math.pi * rd / hd)), 4)
    dxy, dixy = disc_nodes(rngs[1], 0.0, -(r - rd) + overlap, rd, nbd, hd)
    inner_disc = deformable_body(dxy, dixy, "Grain", hd, RHO_GRAIN, detection=0.5 * hd, contact=0.5 * hd,
                                 velocity=(0.02, -0.05))
    Rt = 1.0e-3
    nbt = 2 * max(int(round(math.pi * Rt / h)), 4)
    txy, tixy = disc_nodes(rngs[2], 0.2e-3, R + Rt - overlap, Rt, nbt, h)
    top = deformable_body(txy, tixy, "Grain", h, RHO_GRAIN, detection=0.5 * h, contact=0.5 * h,
                          velocity=(0.0, -0.05))
    mats = [("Grain", "NeoHookean", [RHO_GRAIN, 0.0, 0.0, E_GRAIN, NU_GRAIN, 1.0])]
    laws = [("Grain", "Grain", "MohrCoulomb", "Evolutive", [1e13, 1e13, 0.3, 0.0, 0.0])]
    m = Model(name="tiny_holed", materials=mats, contact_laws=laws, bodies=[ring, inner_disc, top],
              periodic=(-1.0, 1.0))
    m.gravity = (0.0, -9.81)
    m.dt = _critical_dt(m, 1e13)
    m.contact_update_period = 10 * m.dt
    m.tend = 1.0
    return m


def tiny_free(h=0.45e-3):
    """Two discs that do not touch within a few hundred steps, loaded by
    every force the works track except contact: a soft Dirichlet spring and
    a following pressure on one, an oriented pressure on the other, gravity,
    Rayleigh damping entries.  With no contact force, the kinetic energy
    from velocity parameters changes by exactly the sum of the works
    (Code/Body.h:2058-2066 accumulate dt (v + dt a / 2) . f for each force
    kind): the energy balance test case."""
    m = tiny(h=h, overlap=-3e-4)
    m.name = "tiny_free"
    m.materials[0] = ("Grain", "ElasticLinear", [RHO_GRAIN, 2e-6, 50.0, E_GRAIN, NU_GRAIN, 1.0])
    m.gravity = (0.3, -9.81)
    b0 = m.bodies[0]
    b0.borders[0].xbc = BC("Soft", [(0.0, 0.01), (1e3, 0.01)], (1e5, 2.0))
    b0.borders[0].ybc = BC("Following", [(0.0, -2e4), (1e3, -2e4)])
    b1 = m.bodies[1]
    b1.borders[0].xbc = BC("Oriented", [(0.0, 1e4), (1e3, 3e4)])
    n = b1.n_nodes
    b1.matrix = [(2 * i + d, 2 * i + d, 1e3 * (1 + i), 0.0) for i in range(n) for d in (0, 1)]
    return m


def _template_grain(job):
    """A grain of radius R centred at the origin with about n nodes (MLS,
    seeded by n): translated copies stand for every grain of that node count
    (shape functions depend on relative positions only)."""
    n, R, rho = job
    rng = np.random.Generator(np.random.PCG64(np.random.SeedSequence(7000 + n)))
    h = R * 0.108 * math.sqrt(200.0 / n)
    nb = int(round(2 * math.pi * R / h))
    bxy, ixy = disc_nodes(rng, 0.0, 0.0, R, nb, h)
    return deformable_body(bxy, ixy, "Grain", h, rho, detection=0.5 * h, contact=0.5 * h)


def _translated(body, dx, dy, material):
    import copy
    b = copy.copy(body)
    b.X = body.X + np.array([dx, dy])
    b.gp_xy = body.gp_xy + np.array([dx, dy])
    b.material = material
    return b


def c4_batch(n_samples=512, seed=1004, period=20.0e-3, R=1.0e-3, grains=(10, 200), nodes=(100, 300),
             mu=(0.2, 0.6), shear=0.05, gap=2.0e-3, t=0.5e-3, only=None, dt=None, workers=None):
    """Independent periodic shear cells in one model: sample s has
    U{grains} grains of radius R (C1's material and contact) deposited into a
    cell of width `period` (_deposit), a rigid 'Periodic' wall under them
    (fixed) and one on top of the highest grain, driven along x at `shear`
    m/s; its Coulomb friction is U[mu] (a grain and a wall material of its
    own carry the sample's laws).  Samples are stacked along y, `gap` apart,
    so no two samples' borders ever come within detection distance: in one
    context they run as a batch, each exactly as it would alone.  Grains
    with the same node count (U{nodes}) are translated copies of one
    template grain.  `only` = s builds sample s alone at the same place
    (with `dt` the batch's step, for comparing a sample alone and in the
    batch)."""
    ss = np.random.SeedSequence(seed).spawn(n_samples)
    plan = []
    y0 = 0.0
    for s in range(n_samples):
        rng = np.random.Generator(np.random.PCG64(ss[s]))
        ng = int(rng.integers(grains[0], grains[1] + 1))
        nn = rng.integers(nodes[0], nodes[1] + 1, ng)
        mus = float(rng.uniform(mu[0], mu[1]))
        pile = _deposit(rng, ng, R, R, period)
        H = max(y + r for _, y, r in pile)
        plan.append((y0 + t, pile, nn, mus, H))
        y0 += H + 2 * t + gap
    need = sorted({int(n) for (_, _, nn, _, _) in plan for n in nn})
    tmpl = dict(zip(need, _map(_template_grain, [(n, R, RHO_GRAIN) for n in need], workers)))
    bodies, mats, laws = [], [], []
    for s, (base, pile, nn, mus, H) in enumerate(plan):
        if only is not None and s != only:
            continue
        g, w = "G%d" % s, "W%d" % s
        mats += [(g, "NeoHookean", [RHO_GRAIN, 0.0, 0.0, E_GRAIN, NU_GRAIN, 1.0]),
                 (w, "ElasticLinear", [7800.0, 0.0, 0.0, 2e11, 0.3, 1.0])]
        laws += [(g, g, "MohrCoulomb", "Evolutive", [1e13, 1e13, mus, 0.0, 0.0]),
                 (g, w, "MohrCoulomb", "Evolutive", [1e13, 1e13, mus, 0.0, 0.0])]
        for (x, y, _), n in zip(pile, nn):
            bodies.append(_translated(tmpl[int(n)], x, base + y, g))
        hw = min(tmpl[int(n)].nodal_distance for n in nn)
        fixed = lambda: BC("Driven", [(0.0, 0.0), (1e3, 0.0)])
        bodies.append(rigid_periodic_wall(0.0, period, base - t, base, w, hw, 7800.0, detection=0.5 * hw,
                                          contact=0.5 * hw, xbc=fixed(), ybc=fixed()))
        bodies.append(rigid_periodic_wall(0.0, period, base + H, base + H + t, w, hw, 7800.0, detection=0.5 * hw,
                                          contact=0.5 * hw, xbc=BC("Driven", [(0.0, shear), (1e3, shear)]),
                                          ybc=fixed(), bc_border=0))
    m = Model(name="c4_shear_%d%s" % (n_samples, "" if only is None else "_s%d" % only), materials=mats,
              contact_laws=laws, bodies=bodies, periodic=(0.0, period))
    m.dt = dt if dt is not None else _critical_dt(m, 1e13)
    m.contact_update_period = 10 * m.dt
    m.tend = 10000 * m.dt
    m.samples = [(int(len(p[1])), int(np.sum(p[2])), p[3]) for p in plan]
    return m


def _set_law(m, typ, params, evolution="Evolutive"):
    m.contact_laws = [(a, b, typ, evolution, list(params)) for (a, b, _, _, _) in m.contact_laws]
    return m


def variants():
    """Small models that together exercise every branch of the step the
    library implements; the reference solver's outputs on them are the
    golden vectors under tests/golden."""
    v = {}
    v["tiny"] = tiny(h=0.45e-3)
    v["tiny_platen"] = tiny(h=0.45e-3, with_platens=True)
    v["tiny_frictionless"] = _set_law(tiny(h=0.45e-3), "Frictionless", [1e13])
    v["tiny_cohesion"] = _set_law(tiny(h=0.45e-3), "Cohesion", [1e13, 5e12, 2e4, 1e4])
    v["tiny_damped"] = _set_law(tiny(h=0.45e-3), "DampedMohrCoulomb", [1e13, 1e13, 0.3, 1e3, 1e3, 0.2], "Fixed")
    v["tiny_twoslopes"] = _set_law(tiny(h=0.45e-3), "TwoSlopesMohrCoulomb", [1e13, 1e13, 0.4, 0.0, 0.0, 0.5])
    m = _set_law(tiny(h=0.45e-3, overlap=-1e-6), "CZMlinear", [1e13, 2e6, 5e-6, 1e5, 1.0, 2e-6])
    m.flags = m.flags + ["INITIALIZE_CZM"]
    v["tiny_czm"] = m
    # the three-variable laws (damage and the last tractions per element)
    m = _set_law(tiny(h=0.45e-3, overlap=-1e-6), "CZMfatigue",
                 [1e13, 4e7, 5e-6, 1e5, 1.0, 2e6, 1e-8, 1e-8, 2e-6])
    m.flags = m.flags + ["INITIALIZE_CZM"]
    v["tiny_czmfatigue"] = m
    m = _set_law(tiny(h=0.45e-3, overlap=-1e-6), "BondedMohrCoulomb",
                 [1e13, 1e13, 1e12, 0.5, 4e6, 3e6, 0.3, 0.0, 0.0, 0.05, 2e-6])
    m.flags = m.flags + ["INITIALIZE_CZM"]
    v["tiny_bonded"] = m
    # linear elastic material, gravity, Neumann/Dirichlet on deformable
    # borders, Rayleigh damping entries, a free rigid body
    m = tiny(h=0.45e-3, with_platens=True)
    m.materials[0] = ("Grain", "ElasticLinear", [RHO_GRAIN, 2e-6, 50.0, E_GRAIN, NU_GRAIN, 1.0])
    m.gravity = (0.3, -9.81)
    b0 = m.bodies[0]
    b0.borders[0].xbc = BC("Soft", [(0.0, 0.01), (1e3, 0.01)], (1e5, 2.0))
    b0.borders[0].ybc = BC("Following", [(0.0, -2e4), (1e3, -2e4)])
    b1 = m.bodies[1]
    b1.borders[0].xbc = BC("Oriented", [(0.0, 1e4), (1e3, 3e4)])
    b1.borders[0].ybc = BC("Soft", [(0.0, 0.0), (1e-6, -0.02), (1e3, -0.02)], (5e4, 1.0))
    n = b1.n_nodes
    b1.matrix = [(2 * i + d, 2 * i + d, 1e3 * (1 + i), 0.0) for i in range(n) for d in (0, 1)]
    b1.matrix += [(2 * i, 2 * i + 3, -2e2, 0.0) for i in range(n - 2)]
    b1.matrix.sort()
    top = m.bodies[3]
    top.borders[0].xbc = BC("Oriented", [(0.0, 5e2), (1e3, 5e2)])
    # free rigid body (no driven direction): translates and rotates
    top.borders[0].ybc = BC("Oriented", [(0.0, -3e4), (1e3, -3e4)])
    m.flags = [f for f in m.flags if f != "NO_LOG"] + ["NO_LOG"]
    v["tiny_bc"] = m
    # absorbing layers (ALID) on the grain with the matrix: graded layers
    # along its lower and left sides, an imposed y velocity
    import copy
    m = copy.deepcopy(m)
    b1 = m.bodies[1]
    xmin, ymin = b1.X.min(axis=0)
    xmax, ymax = b1.X.max(axis=0)
    rng = 0.35 * (ymax - ymin)
    b1.alid = ((xmin, xmax, ymin, ymax, rng, 0.0, 1e-4, 0.0, 2e4, 2.0),
               [(0.0, 0.0), (1e3, 0.0)], [(0.0, 0.01), (1e3, 0.01)])
    v["tiny_alid"] = m
    # self-contact disabled (NO_MONITORING)
    m = tiny(h=0.45e-3)
    m.flags = m.flags + ["NO_MONITORING"]
    v["tiny_nomonitor"] = m
    # periodic shear cell, Simple borders, a body with two borders
    v["tiny_periodic"] = tiny_periodic()
    v["tiny_simple"] = tiny_simple()
    v["tiny_holed"] = tiny_holed()
    # deactivation rules (Update_status, Code/Proximity.h:550-1007) on the
    # discs bonded to each other and to the platens: the upper disc's damage
    # crosses its threshold while its bonds still carry forces (its last
    # sends keep reaching the lower disc and the top platen); the other
    # rules read positions, velocities, accelerations, forces of nodes and
    # bodies and never hold, and a Jacobian line the reference ignores
    m = _set_law(tiny(h=0.45e-3, with_platens=True, overlap=-1e-6), "CZMlinear", [1e13, 2e6, 5e-6, 1e5, 1.0, 2e-6])
    m.flags = m.flags + ["INITIALIZE_CZM"]
    m.deactivation = [("Position Y 0 -1", "below", 0.99e-3), ("Velocity Mag 0 5", "above", 10.0),
                      ("Force Y Contact 3 -1", "above", 1e9), ("Force X Internal 1 -1", "above", 1e12),
                      ("Acceleration R 2 -1", "above", 1e30), ("Position X 2 -1", "equal", 5.0),
                      ("Force Y Total 0 3", "below", -1e30), ("Jacobian 0 0", "below", 1.0),
                      ("Displacement Mag 1 -1", "above", 1.0), ("Damage 1", "above", 0.9)]
    v["tiny_deact"] = m
    # the same under Adaptive_Euler (Solver_step skips the stopped disc in
    # its half steps, Store and Restore, Code/Solver.h:160-330)
    m = copy.deepcopy(m)
    m.scheme = "Adaptive_Euler"
    m.target_error, m.control_parameter, m.accepted_ratio = 1e-4, 0.2, 2.0
    v["tiny_deact_adaptive"] = m
    # Monitoring rows (Code/Monitoring.h:8-559) on the model with BCs,
    # damping and a free rigid platen: every entry type, node- and
    # body-level, the energies of MONITOR_ENERGY (entries chosen away from
    # values that are rounding noise, such as a body's internal-force sum)
    m = copy.deepcopy(v["tiny_bc"])
    m.flags = m.flags + ["MONITOR_ENERGY"]
    m.monitoring = ["Position X 0 3", "Position Y 1 -1", "Position R 3 -1", "Position R 0 -1",
                    "Displacement Mag 0 -1", "Displacement R 3 -1", "Displacement Y 2 4",
                    "Velocity Y 1 5", "Velocity X 3 -1", "Velocity Mag 2 -1",
                    "Acceleration Mag 1 -1", "Acceleration R 3 -1", "Acceleration X 0 7",
                    "Force Y Total 0 3", "Force X Contact 1 -1", "Force Mag Contact 0 -1", "Force Y Body 0 -1",
                    "Force Mag Internal 0 5",
                    "Force R Neumann 3 -1", "Force X Damping 1 4", "Force Y Alid 0 2", "Force X Internal 2 1",
                    "Force Y Contact 3 -1", "Force X Total 2 5",
                    "Jacobian 0 3", "Stress VM 1 2", "Stress Spherical 1 2",
                    "Contact Gapn 0 0 13", "Contact Xsi 1 0 0", "Contact Gapn 1 0 7", "Contact Fx 0 0 13",
                    "Contact Ynorm 2 0 11", "Damage 1",
                    "Energy Kinetic 1", "Energy Kinetic 3", "Energy Kinetic -1", "Energy Deformation 0",
                    "Work Internal 0", "Work Contact -1", "Work Damping 1", "Work Neumann 3", "Work Body -1",
                    "Work Dirichlet 2", "Error X 0 1", "MassScaling X 0 1"]
    v["tiny_monitor"] = m
    # a body inactive from the start (the DYNAMIC file's status): no update,
    # no detection, and the other bodies do not see it
    m = tiny(h=0.45e-3, with_platens=True)
    m.bodies[1].status = "inactive"
    v["tiny_inactive"] = m
    # the adaptive schemes (Code/Solver.h:98-486) on the two discs between
    # platens: error-controlled time steps with rejected attempts, and the
    # two mass-scaling forms (the last one rejects on the mass ratio)
    for name, scheme, ratio in (("tiny_adaptive", "Adaptive_Euler", 2.0), ("tiny_mass_scaling", "Mass_Scaling", 2.0),
                                ("tiny_adaptive_ms", "Adaptive_Mass_Scaling", 1.1)):
        m = tiny(h=0.45e-3, with_platens=True)
        m.scheme = scheme
        m.target_error, m.control_parameter, m.accepted_ratio = 1e-4, 0.2, ratio
        m.mass_scaling = (4.0, 0.5, 0.5, 0.99)
        v[name] = m
    return v


WORKLOADS = {"tiny": tiny, "tiny_platen": lambda: tiny(with_platens=True), "c1": c1, "c2": c2, "c3": c3}
