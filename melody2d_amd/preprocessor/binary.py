"""Binary model container: the records of oracle/dump_format.hpp layout
(u32 name length, name, u8 kind 0=f64 1=i32, u64 count, payload), holding
the same content as the three ASCII model files.  Read by
mel::load_model_binary (melody2d_amd/csrc/model_io.cpp); used for models whose
ASCII files would be gigabytes."""
from __future__ import annotations

import struct

import numpy as np

from .export import Model, g

_MAT = {"ElasticLinear": 0, "NeoHookean": 1}
_LAW = {"Frictionless": 0, "Cohesion": 1, "MohrCoulomb": 2, "DampedMohrCoulomb": 3, "TwoSlopesMohrCoulomb": 4,
        "BondedMohrCoulomb": 5, "CZMlinear": 6, "CZMfatigue": 7}
_PER = {"Closed": 0, "Simple": 1, "Periodic": 2}
_BC = {"None": 0, "Driven": 1, "Soft": 2, "Oriented": 3, "Following": 4}
_FLAGS = ["KILL_VELOCITY", "MONITOR_ENERGY", "MONITOR_BOUNDARIES", "INITIALIZE_CZM", "UPDATE_MASS_MATRIX",
          "UPDATE_STIFFNESS_MATRIX", "UPDATE_DAMPING_MATRIX", "NO_LOG", "NO_MONITORING", "NO_SELF_CONTACT",
          "RESET_WORK"]


def _rec(f, name, arr):
    arr = np.asarray(arr)
    if arr.dtype.kind == "f":
        kind, a = 0, np.ascontiguousarray(arr, dtype="<f8").reshape(-1)
    else:
        kind, a = 1, np.ascontiguousarray(arr, dtype="<i4").reshape(-1)
    nb = name.encode()
    f.write(struct.pack("<I", len(nb)) + nb + struct.pack("<BQ", kind, a.size))
    f.write(a.tobytes())


_SCHEME = {"Euler": 0, "Adaptive_Euler": 1, "Mass_Scaling": 2, "Adaptive_Mass_Scaling": 3}


def write_binary(model: Model, path: str):
    mats = [m[0] for m in model.materials]
    with open(path, "wb") as f:
        m = model
        _rec(f, "model.header", np.array([m.tini, m.dt, m.tend, m.save_period, m.print_period,
                                          m.contact_update_period, m.periodic[0], m.periodic[1], m.gravity[0],
                                          m.gravity[1], m.tini, 0, 0, 0, m.tini + m.save_period,
                                          m.tini + m.print_period, m.tini + m.contact_update_period], dtype=float))
        _rec(f, "model.flags", np.array([1 if fl in m.flags else 0 for fl in _FLAGS], dtype=np.int32))
        if m.scheme != "Euler":
            _rec(f, "model.solver", np.array([_SCHEME[m.scheme], m.target_error, m.control_parameter,
                                              m.accepted_ratio] + list(m.mass_scaling), dtype=float))
        _rec(f, "model.material_type", np.array([_MAT[t] for _, t, _ in m.materials], dtype=np.int32))
        _rec(f, "model.material_raw", np.array([p for _, _, p in m.materials], dtype=float).reshape(-1))
        _rec(f, "model.law_materials", np.array([[mats.index(a), mats.index(b)] for a, b, *_ in m.contact_laws],
                                                dtype=np.int32).reshape(-1))
        _rec(f, "model.law_type", np.array([_LAW[t] for _, _, t, _, _ in m.contact_laws], dtype=np.int32))
        _rec(f, "model.law_fixed", np.array([1 if lev == "Fixed" else 0 for *_, lev, _ in m.contact_laws], dtype=np.int32))
        ptr = np.cumsum([0] + [len(p) for *_, p in m.contact_laws]).astype(np.int32)
        _rec(f, "model.law_ptr", ptr)
        _rec(f, "model.law_params", np.array([x for *_, p in m.contact_laws for x in p], dtype=float))
        if getattr(m, "monitoring", None):
            # the MONITORING section's text, one character per entry
            text = "%d\n" % len(m.monitoring) + "".join("%s\n" % q for q in m.monitoring)
            _rec(f, "model.monitoring", np.frombuffer(text.encode("ascii"), dtype=np.uint8).astype(np.int32))
        if getattr(m, "deactivation", None):
            # the DEACTIVATION section's text, one character per entry
            text = "%d\n" % len(m.deactivation) + "".join("%s %s %s\n" % (spy, cmp, g(thr))
                                                           for spy, cmp, thr in m.deactivation)
            _rec(f, "model.deactivation", np.frombuffer(text.encode("ascii"), dtype=np.uint8).astype(np.int32))
        _rec(f, "model.n_bodies", np.array([len(m.bodies)], dtype=np.int32))
        for i, b in enumerate(m.bodies):
            if b.alid is not None:
                raise ValueError("the binary container has no ALID records: write the model as ASCII files")
            p = "b%d." % i
            _rec(f, p + "info", np.array([1 if b.rigid else 0, mats.index(b.material), _PER[b.periodicity],
                                          {"active": 0, "inactive": 1, "nogravity": 2}.get(b.status, 0),
                                          b.n_nodes, b.n_gauss, b.nregions], dtype=np.int32))
            _rec(f, p + "dist", np.array([b.nodal_distance, b.detection, b.contact], dtype=float))
            _rec(f, p + "X", b.X[:, 0])
            _rec(f, p + "Y", b.X[:, 1])
            bptr = np.cumsum([0] + [len(br.nodes) for br in b.borders]).astype(np.int32)
            _rec(f, p + "border_ptr", bptr)
            _rec(f, p + "border_nodes", np.concatenate([br.nodes for br in b.borders]).astype(np.int32))
            meta, bcp, hptr, ht, hv = [], [], [0], [], []
            for br in b.borders:
                meta += [_PER[br.periodicity], br.before, br.after, _BC[br.xbc.kind], _BC[br.ybc.kind]]
                xs = br.xbc.params if br.xbc.kind == "Soft" else (0.0, 0.0)
                ys = br.ybc.params if br.ybc.kind == "Soft" else (0.0, 0.0)
                bcp += [xs[0], xs[1], ys[0], ys[1]]
                for bc in (br.xbc, br.ybc):
                    ht += [t for t, _ in bc.history]
                    hv += [v for _, v in bc.history]
                    hptr.append(len(ht))
            _rec(f, p + "border_meta", np.array(meta, dtype=np.int32))
            _rec(f, p + "border_bcparams", np.array(bcp, dtype=float))
            _rec(f, p + "border_hist_ptr", np.array(hptr, dtype=np.int32))
            _rec(f, p + "hist_t", np.array(ht, dtype=float))
            _rec(f, p + "hist_v", np.array(hv, dtype=float))
            _rec(f, p + "cells", np.asarray(b.tri, dtype=np.int32).reshape(-1))
            vel = np.asarray(b.velocity, dtype=float)
            vx, vy = (vel[:, 0], vel[:, 1]) if vel.ndim == 2 else (float(vel[0]), float(vel[1]))
            if not b.rigid:
                _rec(f, p + "domain", b.domain)
                _rec(f, p + "mx", b.mass[:, 0])
                _rec(f, p + "my", b.mass[:, 1])
                _rec(f, p + "gauss", np.stack([b.gp_xy[:, 0], b.gp_xy[:, 1], b.gp_w, b.gp_jac], axis=1))
                _rec(f, p + "gptr", b.gp_ptr)
                _rec(f, p + "gidx", b.gp_idx)
                _rec(f, p + "gphi", b.gp_phi)
                _rec(f, p + "gphix", b.gp_phix)
                _rec(f, p + "gphiy", b.gp_phiy)
                _rec(f, p + "dptr", b.d2d_ptr)
                _rec(f, p + "didx", b.d2d_idx)
                _rec(f, p + "dphi", b.d2d_phi)
                _rec(f, p + "dphix", b.d2d_phix)
                _rec(f, p + "dphiy", b.d2d_phiy)
                n = b.n_nodes
                kin = np.zeros((12, n))
                kin[2] = vx; kin[3] = vy; kin[8] = vx; kin[9] = vy
                _rec(f, p + "kin", kin)
                if b.matrix:
                    raise ValueError("damping matrices are written to the ASCII files only")
            else:
                _rec(f, p + "rigid", np.array([b.centre[0], b.centre[1], b.rigid_mass, b.rigid_inertia,
                                               1.0 / b.rigid_mass, 1.0 / b.rigid_inertia], dtype=float))
                _rec(f, p + "rkin", np.array([b.centre[0], b.centre[1], 0, 0, 0, 0, vx, vy, 0], dtype=float))
