"""Energy scalars, long trajectories and the force evaluation, through the
C-ABI, against the oracle and against the mathematics.

* The kinetic energies and boundary work of every body
  (mel_get_body_energy, Code/Monitoring.h:395-412 and 508-537) within
  1e-12 of the oracle, component by component, on every golden model.
* The energy balance: with no contact force the kinetic energy from velocity
  parameters changes by the sum of the works (Code/Body.h:2058-2066) to
  rounding, on the device as on the oracle.
* The two-disc platen workload (`configs.c1`, ~800 nodes) through 1000 steps
  against the oracle: contact sets and neighbour lists exactly; every
  floating-point record within 1e-12 through step 100; at step 1000
  kinematic records (positions, displacements, velocities, lengths, works,
  energies) within 1e-9 and force-like records (forces, accelerations,
  contact pressures, element forces) within 1e-6 of the body's force scale
  (the larger of the x and y record norms), and a rigid body's resultant
  moment and angular acceleration within 1e-4 of their largest: the device
  sums internal forces in another order, and differences of 1e-16 per step
  grow through the contact dynamics, most in resultants that are small
  differences of large terms (the friction resultant on a platen and its
  moment about the platen's centre).  The whole-model energy
  scalars are held to the kinematic bound.  Where the generator rebuilds
  the model bit for bit, the reference's own dumps are compared too.
* mel_eval_forces at a state gives the force fields of the next step and
  leaves the state untouched."""
import os

import numpy as np
import pytest

from conftest import extract_golden, golden_steps
from mel_compare import compare, device_state, layout_from_dump, run_oracle
from mel_dump import read_dump
from melody2d_amd import native
from melody2d_amd.preprocessor import write_model
from melody2d_amd.preprocessor.configs import tiny_free

pytestmark = pytest.mark.gpu

FORCE_FIELDS = ["x_internal_force", "y_internal_force", "x_contact_force", "y_contact_force", "x_body_force",
                "y_body_force", "x_dirichlet_force", "y_dirichlet_force", "x_neumann_force", "y_neumann_force",
                "x_damping_force", "y_damping_force", "x_force", "y_force"]


def _bad(res, tol):
    return sorted(((k, v[0]) for k, v in res.items() if not v[0] <= tol), key=lambda kv: -kv[1])[:8]


FORCE_LIKE = ("_force", "_acc", "_contact_pressure")


def _trajectory_errors(dev, ref, tol_kin, tol_force):
    """Records over their bound: kinematic ones normwise, force-like ones
    (x/y pairs) against the body's force scale, the rigid record split into
    its kinematic part (0:9) and its force part (9:)."""
    bad = []
    for k, r in ref.items():
        if k in ("iteration", "n_bodies") or k.endswith(".mass") or k not in dev:
            continue
        d = np.asarray(dev[k], dtype=np.float64) if r.dtype.kind == "f" else dev[k]
        if r.dtype.kind == "i":
            if not np.array_equal(d, r):
                bad.append((k, "ids differ"))
            continue
        if len(d) != len(r):
            bad.append((k, "length"))
            continue
        if k.endswith(".rigid"):
            # kinematics; accelerations and force resultants (max-norm against
            # the largest of them); mass and inertia
            for name, fi, tol in (("forces", [9, 10, 12, 13, 15, 16, 20, 21, 23, 24], tol_force),
                                  ("moments", [11, 14, 17, 22, 25], 100 * tol_force)):
                dr = np.abs(d[fi] - r[fi]).max()
                if dr > tol * max(np.abs(r[fi]).max(), 1e-300):
                    bad.append((k + "[%s]" % name, dr / max(np.abs(r[fi]).max(), 1e-300)))
            parts = [(d[:9], r[:9], tol_kin), (d[18:20], r[18:20], 1e-12)]
        elif any(f in k for f in FORCE_LIKE):
            other = k.replace(".x_", ".y_") if ".x_" in k else k.replace(".y_", ".x_")
            scale = max(np.linalg.norm(r), np.linalg.norm(ref.get(other, r)))
            parts = [(d, r, tol_force, scale)]
        elif k.endswith(".contacts_d"):
            parts = [(d, r, tol_force)]
        else:
            parts = [(d, r, tol_kin)]
        for p in parts:
            dd, rr, tol = p[0], p[1], p[2]
            scale = p[3] if len(p) > 3 else np.linalg.norm(rr)
            err = np.linalg.norm(dd - rr)
            if err > tol * scale:
                bad.append((k, err / max(scale, 1e-300)))
    return bad


def test_energy_balance_without_contact(tmp_path):
    write_model(tiny_free(), str(tmp_path / "model"))
    print("case energy balance tiny_free", flush=True)
    ctx = native.Context(str(tmp_path / "model"))
    t0, w0, e0 = ctx.scalars()
    prev = 0
    for s in (1, 10, 100, 400):
        ctx.euler_steps(s - prev)
        prev = s
        t, w, e = ctx.scalars()
        dk, dw = e[1] - e0[1], float(np.sum(w - w0))
        assert np.count_nonzero(w) >= 4
        assert abs(dk - dw) <= 1e-12 * abs(dk), (s, dk, dw)
        assert t > t0
    ctx.close()


@pytest.mark.parametrize("name", ["tiny_bc", "tiny_simple", "tiny_periodic", "tiny_damped"])
def test_body_energy_matches_oracle(name, golden_dir, tmp_path):
    model, _ = golden_dir(name)
    print("case energy", name, flush=True)
    steps = [0, 1, 10, 50]
    refs, _ = run_oracle(model, steps, str(tmp_path / "oracle"))
    ctx = native.Context(model)
    prev = 0
    for s in steps:
        ctx.euler_steps(s - prev)
        prev = s
        nb = int(refs[s]["n_bodies"][0])
        for b in range(nb):
            e = ctx.body_energy(b)
            r = refs[s]["b%d.energy" % b]
            for k in range(3):
                assert abs(e[k] - r[k]) <= 1e-12 * abs(r[k]) + 1e-300, (s, b, k, e, r)
    ctx.close()


def test_c1_trajectory_1000_steps(tmp_path):
    """Configuration C1 through 1000 steps: device vs oracle (and vs the
    reference's own dumps where the model rebuilds bit for bit)."""
    print("case c1 1000 steps", flush=True)
    d = str(tmp_path / "c1")
    os.makedirs(d)
    try:
        model, ref_dir = extract_golden("c1", d)
        have_ref = True
    except pytest.skip.Exception:
        from melody2d_amd.preprocessor.configs import c1
        model = os.path.join(d, "model")
        write_model(c1(), model)
        have_ref = False
    steps = [0, 1, 10, 100, 1000]
    refs, _ = run_oracle(model, steps, str(tmp_path / "oracle"))
    ctx = native.Context(model)
    prev = 0
    for s in steps:
        ctx.euler_steps(s - prev)
        prev = s
        nn, nbr, sizes = layout_from_dump(refs[s])
        dev = device_state(ctx, nn, nbr, sizes)
        tol = 1e-12 if s <= 100 else 1e-9
        if s <= 100:
            res = compare(dev, refs[s])
            assert len(res) > 50
            assert not _bad(res, tol), (s, _bad(res, tol))
        else:
            bad = _trajectory_errors(dev, refs[s], 1e-9, 1e-6)
            assert not bad, (s, bad[:8])
        if have_ref:
            R = read_dump(os.path.join(ref_dir, "step_%06d.bin" % s))
            bad = _trajectory_errors(dev, R, tol, 1e-12 if s <= 100 else 1e-6)
            assert not bad, ("reference", s, bad[:8])
        # whole-model energy balance against the oracle's scalars
        t, w, e = ctx.scalars()
        nb = len(nn)
        re = sum(refs[s]["b%d.energy" % b] for b in range(nb))
        rw = sum(refs[s]["b%d.works" % b] for b in range(nb))
        assert abs(t - refs[s]["time"][0]) <= 1e-15 * abs(refs[s]["time"][0]) + 1e-300
        for k in range(3):
            assert abs(e[k] - re[k]) <= tol * abs(re[k]) + 1e-300, (s, k, e, re)
        assert np.linalg.norm(w - rw) <= tol * max(np.linalg.norm(rw), 1e-300), (s, w, rw)
    ctx.close()


@pytest.mark.parametrize("name", ["tiny_bc", "tiny_czm", "tiny_holed"])
def test_eval_forces_gives_next_step_forces(name, golden_dir, tmp_path):
    """mel_eval_forces after k steps = the force fields the oracle's step
    k + 1 forms (deformable bodies), and a step after it gives the oracle's
    state k + 1: the evaluation left the state untouched."""
    model, _ = golden_dir(name)
    print("case eval", name, flush=True)
    refs, _ = run_oracle(model, [10, 11], str(tmp_path / "oracle"))
    nn, nbr, sizes = layout_from_dump(refs[11])
    ctx = native.Context(model)
    ctx.euler_steps(10)
    ctx.eval_forces(native.EVAL_ALL)
    for b in range(len(nn)):
        if np.all(refs[11]["b%d.x_mass" % b] == 0):
            continue   # rigid body (its nodes carry no mass): resultants are not re-formed
        for f in FORCE_FIELDS:
            got = ctx.node_field(b, f, nn[b])
            want = refs[11]["b%d.%s" % (b, f)]
            err = np.linalg.norm(got - want) / max(np.linalg.norm(want), 1e-300)
            assert err <= 1e-12 or np.linalg.norm(want) == np.linalg.norm(got) == 0.0, (b, f, err)
    # internal only: the other fields read zero
    ctx.eval_forces(native.EVAL_INTERNAL)
    assert not np.any(ctx.node_field(0, "x_contact_force", nn[0]))
    assert np.any(ctx.node_field(0, "x_internal_force", nn[0]))
    ctx.euler_steps(1)
    res = compare(device_state(ctx, nn, nbr, sizes), refs[11])
    assert not _bad(res, 1e-12), _bad(res, 1e-12)
    ctx.close()
