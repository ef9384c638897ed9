"""Writers for the three model files the step reads.

Same keywords, line structure and blank-line layout as the reference's
Preprocessor/ExportStatic.m (STATIC_DATA.asc, STATIC_CONTROL.asc) and
Preprocessor/ExportDynamic.m (DYNAMIC_nnnnn.asc), so that the reference
loaders Load_static / Load_dynamic (Code/IO.h:598-1862, 1871-2343) and this
repository's loader read them identically.  Two deliberate differences:

* numbers are written with 17 significant digits (round-trip exact), where
  the MATLAB writer uses %0.16g, so a model held in memory and the same model
  read back from its files are bit-identical;
* the material line always carries the six numbers rho alpha beta E nu c that
  Code/IO.h:675-676 reads, and a Frictionless law always carries its kn
  (Code/IO.h:723-724), where the MATLAB writer may omit them.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field
import numpy as np

from .bodies import Body, BC


@dataclass
class Model:
    name: str
    materials: list          # [(name, type, [rho, alpha, beta, E, nu, c])]
    contact_laws: list       # [(mat1, mat2, type, 'Fixed'|'Evolutive', [params])]
    bodies: list
    scheme: str = "Euler"
    tini: float = 0.0
    dt: float = 1e-7
    tend: float = 1.0
    target_error: float = 1e-4
    control_parameter: float = 0.2
    accepted_ratio: float = 2.0
    # Mass_Scaling / Adaptive_Mass_Scaling: maximum factor, control
    # exponent, error factor, decrease factor (Code/IO.h:866)
    mass_scaling: tuple = (10.0, 0.5, 0.5, 0.99)
    save_period: float = 1e9
    print_period: float = 1e9
    contact_update_period: float = 1e-6
    periodic: tuple = (0.0, 1.0)
    gravity: tuple = (0.0, 0.0)
    flags: list = field(default_factory=lambda: ["UPDATE_MASS_MATRIX", "UPDATE_DAMPING_MATRIX", "NO_LOG"])
    # DEACTIVATION rules (Code/IO.h:947-968): spy-line text (the words
    # Read_spy_line reads, e.g. "Position Y 1 -1"), "below" | "equal" |
    # "above", threshold
    deactivation: list = field(default_factory=list)
    # MONITORING entries (Code/IO.h:933-945): spy-line text, e.g. "Velocity Y 1 5"
    monitoring: list = field(default_factory=list)

    @property
    def n_nodes(self):
        return sum(b.n_nodes for b in self.bodies)

    @property
    def n_gauss(self):
        return sum(b.n_gauss for b in self.bodies)


def g(x):
    return "%.17g" % float(x)


_BAR = "%%%%%%%%%%%%%%%%%%%%%%%%%%%%%%%%%%%%%%%%%%%%%%"


def _header(out):
    out += [" ", _BAR, "%%%%%%%          GENERAL DATA          %%%%%%%", _BAR, " "]


def _body_banner(out, i, rigid):
    pad = "             " if rigid else "              "
    out += [" ", _BAR, "%%%%%%%" + pad + "BODY " + str(i) + "            %%%%%%%", _BAR, " "]


def _bc_lines(direction, bc: BC):
    if bc.kind == "None":
        return ["%s None" % direction]
    if bc.kind in ("Driven", "Soft"):
        head = "%s Dirichlet Driven" % direction if bc.kind == "Driven" else \
            "%s Dirichlet Soft %s %s" % (direction, g(bc.params[0]), g(bc.params[1]))
    else:
        head = "%s Neumann %s" % (direction, bc.kind)
    lines = [head, "%d" % len(bc.history)]
    lines += ["%s %s" % (g(t), g(v)) for t, v in bc.history]
    return lines


def write_static(model: Model, d: str):
    # ---- STATIC_DATA.asc (ExportStatic.m:1-202) ----
    out = []
    _header(out)
    out += ["SIMULATION_NAME", model.name, " ", " ", "NUMBER_BODIES", "%d" % len(model.bodies), " ", " "]
    for i, b in enumerate(model.bodies):
        _body_banner(out, i, b.rigid)
        if not b.rigid:
            out += ["DEFORMABLE", "%d" % i, " ", "NODES", "%d" % b.n_nodes, " "]
            for j in range(b.n_nodes):
                m = b.mass[j]
                out.append("%d %s %s %s %s %s %s %s" % (j, g(b.X[j, 0]), g(b.X[j, 1]), g(b.domain[j]),
                                                        g(m[0]), g(m[1]), g(1.0 / m[0]), g(1.0 / m[1])))
            out.append(" ")
        else:
            out += ["RIGID", "%d" % i, "%s %s" % (g(b.centre[0]), g(b.centre[1])),
                    "%s %s %s %s" % (g(b.rigid_mass), g(b.rigid_inertia), g(1.0 / b.rigid_mass), g(1.0 / b.rigid_inertia)),
                    " ", "NODES", "%d" % b.n_nodes, " "]
            for j in range(b.n_nodes):
                out.append("%d %s %s" % (j, g(b.X[j, 0]), g(b.X[j, 1])))
            out.append(" ")
        out += ["BORDERS", "%d" % len(b.borders), " "]
        for br in b.borders:
            out.append("%s %s" % (br.periodicity, br.interpolant))
            out.append("%d %d %d" % (len(br.nodes), br.before, br.after))
            out += ["%d" % n for n in br.nodes]
            out.append(" ")
        if not b.rigid:
            out += ["INTEGRATION", "Gauss", "%d %d" % (b.n_gauss, b.nregions), " "]
            ptr = b.gp_ptr
            for gi in range(b.n_gauss):
                out.append("Gauss Point %d" % gi)
                out.append("%s %s %s %s" % (g(b.gp_xy[gi, 0]), g(b.gp_xy[gi, 1]), g(b.gp_w[gi]), g(b.gp_jac[gi])))
                s, e = ptr[gi], ptr[gi + 1]
                out.append("%d" % (e - s))
                for k in range(s, e):
                    n = b.gp_idx[k]
                    out.append("%d %d %d %s %s %s" % (n, 2 * n, 2 * n + 1, g(b.gp_phi[k]), g(b.gp_phix[k]), g(b.gp_phiy[k])))
                out.append(" ")
        out += ["TRIANGULATION", "%d" % len(b.tri)]
        out += ["%d %d %d %d" % (c, t[0], t[1], t[2]) for c, t in enumerate(b.tri)]
        out.append(" ")
        if not b.rigid:
            out += ["DOF_TO_DISP", " "]
            ptr = b.d2d_ptr
            for j in range(b.n_nodes):
                out.append("NODE %d" % j)
                s, e = ptr[j], ptr[j + 1]
                out.append("%d" % (e - s))
                for k in range(s, e):
                    out.append("%d %s %s %s" % (b.d2d_idx[k], g(b.d2d_phi[k]), g(b.d2d_phix[k]), g(b.d2d_phiy[k])))
                out.append(" ")
            out.append(" ")
            out += ["MATRICES", "%d" % len(b.matrix), " "]
            out += ["%d %d %s %s" % (d0, d1, g(k), g(dmp)) for d0, d1, k, dmp in b.matrix]
            out += [" ", " ", " "]
    with open(os.path.join(d, "STATIC_DATA.asc"), "w") as f:
        f.write("\n".join(out) + "\n")

    # ---- STATIC_CONTROL.asc (ExportStatic.m:205-695) ----
    out = []
    _header(out)
    out += ["SIMULATION_NAME", model.name, " ", " "]
    out += ["MATERIALS", "%d" % len(model.materials), " "]
    for name, typ, p in model.materials:
        out += [name, typ, " ".join(g(v) for v in p) + "  ", " "]
    out.append(" ")
    out += ["CONTACT_LAWS", "%d" % len(model.contact_laws), " "]
    for m1, m2, typ, lev, p in model.contact_laws:
        out += [m1, m2, typ, lev, " ".join(g(v) for v in p) + "  ", " "]
    out.append(" ")
    out += ["SOLVER", model.scheme, "%s %s %s" % (g(model.tini), g(model.dt), g(model.tend)),
            "%s %s %s" % (g(model.target_error), g(model.control_parameter), g(model.accepted_ratio))]
    if model.scheme in ("Mass_Scaling", "Adaptive_Mass_Scaling"):
        out.append(" ".join(g(v) for v in model.mass_scaling))
    out += ["%s %s %s" % (g(model.save_period), g(model.print_period), g(model.contact_update_period)), " ", " "]
    out += ["PERIODIC_BOUNDARIES", "%s %s" % (g(model.periodic[0]), g(model.periodic[1])), " ", " "]
    out += ["GRAVITY", "%s %s" % (g(model.gravity[0]), g(model.gravity[1])), " ", " "]
    out += ["NUMBER_BODIES", "%d" % len(model.bodies), " ", " "]
    out += ["MONITORING", "%d" % len(model.monitoring)] + list(model.monitoring) + [" ", " "]
    out += ["SPIES", "0", " ", " ", "DEACTIVATION", "%d" % len(model.deactivation)]
    out += ["%s %s %s" % (spy, cmp, g(thr)) for spy, cmp, thr in model.deactivation]
    out += [" ", " "]
    # Code/IO.h:1004-1009 reads 41 graphic flags
    out += ["GRAPHIC"] + ["0 Flag%d" % i for i in range(41)] + [" ", " "]
    for i, b in enumerate(model.bodies):
        _body_banner(out, i, b.rigid)
        out += ["RIGID" if b.rigid else "DEFORMABLE", "%d" % i, b.material, b.periodicity,
                "%s %s %s" % (g(b.nodal_distance), g(b.detection), g(b.contact)), " "]
        out += ["NODES", "%d" % b.n_nodes, " "]
        out += ["BORDERS", "%d" % len(b.borders), " "]
        for br in b.borders:
            out.append("%s %s" % (br.periodicity, br.interpolant))
            out.append("%d" % len(br.nodes))
            out += _bc_lines("X", br.xbc)
            out += _bc_lines("Y", br.ybc)
            out.append(" ")
        if b.alid is not None:
            par, xh, yh = b.alid
            out += ["ALID", " ".join(g(v) for v in [len(xh), len(yh)] + list(par))]
            for hist in (xh, yh):
                if hist:
                    out.append(" ")
                    out += ["%s %s" % (g(t), g(v)) for t, v in hist]
            out.append(" ")
        if not b.rigid:
            out += ["INTEGRATION", "%d %d" % (b.n_gauss, b.nregions), " "]
        out += ["TRIANGULATION", "%d" % len(b.tri), " ", " ", " "]
    with open(os.path.join(d, "STATIC_CONTROL.asc"), "w") as f:
        f.write("\n".join(out) + "\n")


def write_dynamic(model: Model, d: str, number_save: int = 0):
    """DYNAMIC_nnnnn.asc for the initial state (ExportDynamic.m)."""
    out = []
    _header(out)
    out += ["SIMULATION_NAME", model.name, " "]
    out += ["SOLVER", model.scheme, "%d 0 %s 0 " % (number_save, g(model.tini)), g(model.dt),
            "%s %s %s " % (g(model.tini + model.save_period), g(model.tini + model.print_period),
                           g(model.tini + model.contact_update_period))]
    out += list(model.flags)
    out.append(" ")
    for i, b in enumerate(model.bodies):
        _body_banner(out, i, b.rigid)
        vel = np.asarray(b.velocity, dtype=float)
        if not b.rigid:
            V = vel if vel.ndim == 2 else np.tile(vel, (b.n_nodes, 1))
            out += ["DEFORMABLE", b.status, "%d" % i, " ", "KINEMATICS", "%d" % b.n_nodes]
            for j in range(b.n_nodes):
                vx, vy = V[j]
                out.append("%d %s %s 0 0 %s %s 0 0 0 0 %s %s 0 0" % (j, g(b.X[j, 0]), g(b.X[j, 1]), g(vx), g(vy), g(vx), g(vy)))
            out.append(" ")
            out.append("FORCES")
            out += ["%d 0 0 0 0 0 0 0 0 0 0 0 0 0 0 0 0" % j for j in range(b.n_nodes)]
            out.append(" ")
        else:
            vx, vy = vel
            out += ["RIGID", b.status, "%d" % i, " ", "KINEMATICS",
                    "%s %s 0" % (g(b.centre[0]), g(b.centre[1])), "0 0 0", "%s %s 0" % (g(vx), g(vy)), " "]
            out += ["FORCES"] + ["0 0 0"] * 6 + [" "]
        out += ["NEIGHBOURS", "0", " "]
        if not b.rigid:
            out += ["BORDERS", " "]
            for k, br in enumerate(b.borders):
                out += ["%d" % (k + 1), "%d" % len(br.nodes)]
                out += ["0 0 0 0 0 0 0 0 0"] * len(br.nodes)
                out.append(" ")
        out += ["CONTACTS", "0", " "]
        out += ["CONTACT_PRESSURES", " "]
        for k, br in enumerate(b.borders):
            out += ["%d" % (k + 1), "%d" % len(br.nodes)] + ["0 0"] * len(br.nodes) + [" "]
        if b.rigid:
            out.append(" ")
        out += ["BC_PRESSURES", " "]
        for k, br in enumerate(b.borders):
            out += ["%d" % (k + 1), "%d" % len(br.nodes)] + ["0 0"] * len(br.nodes) + [" "]
    name = "DYNAMIC_%05d.asc" % number_save
    with open(os.path.join(d, name), "w") as f:
        f.write("\n".join(out) + "\n")


def write_model(model: Model, d: str):
    os.makedirs(d, exist_ok=True)
    write_static(model, d)
    write_dynamic(model, d, 0)
