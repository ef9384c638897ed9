"""Parity helpers: run the HIP library and the oracle on the same model and
compare their states field by field."""
import ctypes
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

from mel_dump import read_dump  # noqa: E402
from melody2d_amd import native  # noqa: E402

ORACLE_LIB = os.path.join(ROOT, "oracle", "liboracle.so")
_oracle = None


def oracle():
    global _oracle
    if _oracle is None:
        lib = ctypes.CDLL(ORACLE_LIB)
        lib.oracle_run.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_char_p,
                                   ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_int,
                                   ctypes.POINTER(ctypes.c_double)]
        lib.oracle_last_error.restype = ctypes.c_char_p
        _oracle = lib
    return _oracle


def run_oracle(model_path, steps, out_dir, save_index=0, nthreads=0):
    """Run the oracle up to max(steps), dumping the listed steps; returns
    {step: dump dict} and the wall time of the stepping loop."""
    os.makedirs(out_dir, exist_ok=True)
    arr = (ctypes.c_int * len(steps))(*steps)
    sec = ctypes.c_double()
    r = oracle().oracle_run(os.fsencode(model_path), save_index, max(steps), os.fsencode(out_dir), arr,
                            len(steps), nthreads, ctypes.byref(sec))
    if r != 0:
        raise RuntimeError(oracle().oracle_last_error().decode())
    return {s: read_dump(os.path.join(out_dir, "step_%06d.bin" % s)) for s in steps}, sec.value


def device_state(ctx, n_nodes_per_body, n_borders_per_body, border_sizes):
    """The same records as a dump, read through the C-ABI."""
    D = {"time": np.array([ctx.get_info().time])}
    si = ctx.solver_info()
    ms = si.scheme in (2, 3)
    if si.scheme != 0:
        # the adaptive schemes' step control (Code/Solver.h:285-484)
        D["dt"] = np.array([si.dt])
        D["neval"] = np.array([float(si.neval)])
        D["max_error"] = np.array([si.max_error])
        D["mean_error"] = np.array([si.mean_error])
        if ms:
            D["total_mass"] = np.array([si.total_mass])
            D["max_mass"] = np.array([si.max_mass])
    for b, n in enumerate(n_nodes_per_body):
        p = "b%d." % b
        for f in native.NODE_FIELDS:
            D[p + f] = ctx.node_field(b, f, n)
        ids, st = ctx.contacts(b)
        D[p + "contacts_i"] = ids
        D[p + "contacts_d"] = st
        D[p + "contacts_int"] = ctx.contact_internals(b)
        D[p + "neighbours"] = ctx.neighbours(b)
        r, w = ctx.body_state(b)
        D[p + "rigid"] = r
        D[p + "works"] = w
        D[p + "energy"] = ctx.body_energy(b)
        st, dmg = ctx.body_status(b)
        D[p + "status"] = np.array([float(st), dmg])
        if si.scheme != 0:
            D[p + "error"] = ctx.body_error(b)[:5 if ms else 3]
            if ms:
                D[p + "x_msf"], D[p + "y_msf"] = ctx.mass_scaling(b, n)
        for k in range(n_borders_per_body[b]):
            q = p + "border%d." % k
            ns = border_sizes[b][k]
            D[q + "length"] = ctx.border_field(b, k, 0, ns)
            D[q + "x_contact_pressure"] = ctx.border_field(b, k, 1, ns)
            D[q + "y_contact_pressure"] = ctx.border_field(b, k, 2, ns)
    return D


def layout_from_dump(D):
    nb = int(D["n_bodies"][0])
    nn = [len(D["b%d.x_current" % b]) for b in range(nb)]
    nbr, sizes = [], []
    for b in range(nb):
        k = 0
        s = []
        while ("b%d.border%d.length" % (b, k)) in D:
            s.append(len(D["b%d.border%d.length" % (b, k)]))
            k += 1
        nbr.append(k)
        sizes.append(s)
    return nn, nbr, sizes


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b)
    num = np.linalg.norm(a - b)
    if den == 0.0:
        return 0.0 if num == 0.0 else float("inf")
    return float(num / den)


def compare(dev, ref, skip=()):
    """Per-record (relative normwise error, bitwise-equal flag); integer
    records must match exactly (error 0 or inf)."""
    out = {}
    for k, r in ref.items():
        if k in ("iteration", "n_bodies") or k.endswith(".mass") or any(s in k for s in skip):
            continue
        if k not in dev:
            continue
        d = dev[k]
        if len(d) != len(r):
            out[k] = (float("inf"), False)
            continue
        if r.dtype.kind == "i":
            out[k] = (0.0 if np.array_equal(d, r) else float("inf"), bool(np.array_equal(d, r)))
        elif k.endswith(".energy") or k == "monitor":
            # kinetic energies and the boundary work, and the entries of a
            # Monitoring row, differ in scale: each against its own magnitude
            out[k] = (max(rel_err(d[i:i + 1], r[i:i + 1]) for i in range(len(r))), bool(np.array_equal(d, r)))
        else:
            out[k] = (rel_err(d, r), bool(np.array_equal(d, r)))
    return out
