"""The oracle (CPU restatement) against the reference solver's own outputs.

The golden dumps were produced by the reference compiled from its sources
(tests/golden/make_golden.py).  The restatement keeps the reference's
expression and summation order, so every record must agree bit for bit."""
import os

import numpy as np
import pytest

from conftest import RESTARTS, RESTART_STEP0_SKIP, VARIANTS, extract_restart, golden_steps
from mel_compare import run_oracle
from mel_dump import read_dump

STEPS = [0, 1, 2, 10, 50, 200]


@pytest.mark.parametrize("name", VARIANTS + ["c1"])
def test_oracle_matches_reference_bitwise(name, golden_dir, tmp_path):
    model, ref = golden_dir(name)
    steps = golden_steps(ref)
    got, _ = run_oracle(model, steps, str(tmp_path / "oracle"))
    for s in steps:
        R = read_dump(os.path.join(ref, "step_%06d.bin" % s))
        G = got[s]
        assert set(R) <= set(G), set(R) - set(G)
        for k, v in R.items():
            w = G[k]
            assert v.shape == w.shape, (s, k)
            if v.dtype.kind == "f":
                same = np.array_equal(v.view(np.int64), w.view(np.int64)) or np.array_equal(v, w)
            else:
                same = np.array_equal(v, w)
            assert same, "step %d record %s differs" % (s, k)


def test_goldens_exercise_the_step():
    """The golden set is only a pin if contact, Neumann, Dirichlet, damping and
    cohesive-damage paths are actually non-zero somewhere in it."""
    import tempfile
    from conftest import extract_golden
    seen = {"contact": False, "neumann": False, "dirichlet": False, "damping": False, "czm": False,
            "rigid_rotation": False, "periodic_neighbour": False, "periodic_stencil": False,
            "slave_border_1": False, "master_border_3": False, "boundary_work": False,
            "fatigue_damage": False, "bond_broken_later": False, "alid": False, "rejected_attempt": False,
            "mass_scaled": False, "deactivated": False, "damage": False, "monitored": False}
    for name in VARIANTS:
        with tempfile.TemporaryDirectory() as d:
            _, ref = extract_golden(name, d)
            for s in golden_steps(ref):
                D = read_dump(os.path.join(ref, "step_%06d.bin" % s))
                nb = int(D["n_bodies"][0])
                for b in range(nb):
                    p = "b%d." % b
                    seen["contact"] |= bool(np.any(D[p + "x_contact_force"] != 0))
                    seen["neumann"] |= bool(np.any(D[p + "y_neumann_force"] != 0))
                    seen["dirichlet"] |= bool(np.any(D[p + "x_dirichlet_force"] != 0))
                    seen["damping"] |= bool(np.any(D[p + "x_damping_force"] != 0))
                    cd = D[p + "contacts_d"].reshape(-1, 18)
                    ci = D[p + "contacts_i"].reshape(-1, 16)
                    seen["czm"] |= bool(np.any((cd[:, 17] > 0) & (cd[:, 17] < 1)))
                    if p + "contacts_int" in D:
                        ci3 = D[p + "contacts_int"].reshape(-1, 3)
                        # damage between 0 and 1 with a last traction kept (CZMfatigue)
                        seen["fatigue_damage"] |= bool(np.any((ci3[:, 0] > 0) & (ci3[:, 0] < 1) & (ci3[:, 1] != 0)))
                        # an intact bond (damage 0, tractions kept) after step 10 (BondedMohrCoulomb)
                        seen["bond_broken_later"] |= bool(s > 10 and np.any((ci3[:, 0] == 0) & (ci3[:, 1] != 0)))
                    seen["rigid_rotation"] |= bool(D[p + "rigid"][8] != 0)
                    # periodic cells: neighbours one period away, and
                    # contact stencils that cross the master's periodic seam
                    seen["periodic_neighbour"] |= bool(np.any(D[p + "neighbours"].reshape(-1, 3)[:, 2] != 0))
                    seen["periodic_stencil"] |= bool(np.any(ci[:, 11:15] != 0))
                    # bodies with several borders
                    seen["slave_border_1"] |= bool(np.any(ci[:, 0] == 1))
                    seen["master_border_3"] |= bool(np.any(ci[:, 4] == 3))
                    seen["boundary_work"] |= bool(D[p + "energy"][2] != 0)
                    if p + "x_alid_force" in D:
                        seen["alid"] |= bool(np.any(D[p + "x_alid_force"] != 0) and D[p + "works"][6] != 0)
                    if p + "x_msf" in D:
                        seen["mass_scaled"] |= bool(np.any(D[p + "x_msf"] > 1))
                    if p + "status" in D:
                        # a body the DEACTIVATION rules stopped, a damage value
                        # between 0 and 1 (Body::Update_damage)
                        seen["deactivated"] |= bool(D[p + "status"][0] == 1)
                        seen["damage"] |= bool(0 < D[p + "status"][1] < 1)
                # a Monitoring row with entries beyond its six leading scalars
                seen["monitored"] |= bool("monitor" in D and len(D["monitor"]) > 6)
                if "neval" in D:
                    seen["rejected_attempt"] |= bool(D["neval"][0] > 1)
    assert all(seen.values()), seen


def test_oracle_energy_balance(tmp_path):
    """Pins the oracle's work bookkeeping on the mathematics: with no contact
    force, v' = v + dt a and a = f / m give 1/2 m (v'^2 - v^2) =
    dt (v + dt a / 2) . f for every node, so the kinetic energy from velocity
    parameters changes by exactly the sum of the works (internal, body,
    Dirichlet, Neumann, damping) up to rounding."""
    from melody2d_amd.preprocessor import write_model
    from melody2d_amd.preprocessor.configs import tiny_free
    write_model(tiny_free(), str(tmp_path / "model"))
    steps = [0, 1, 10, 100, 400]
    got, _ = run_oracle(str(tmp_path / "model"), steps, str(tmp_path / "oracle"))
    nb = int(got[0]["n_bodies"][0])
    ke = {s: sum(got[s]["b%d.energy" % b][1] for b in range(nb)) for s in steps}
    w = {s: sum(got[s]["b%d.works" % b] for b in range(nb)) for s in steps}
    for s in steps:
        assert all(len(got[s]["b%d.contacts_i" % b]) == 0 for b in range(nb))
    for s in steps[1:]:
        dk = ke[s] - ke[0]
        dw = float(np.sum(w[s] - w[0]))
        assert np.count_nonzero(w[s]) >= 4
        assert abs(dk - dw) <= 1e-12 * abs(dk), (s, dk, dw)


@pytest.mark.parametrize("name", RESTARTS)
def test_oracle_restart_matches_reference_bitwise(name, tmp_path):
    """Restart from a DYNAMIC file written by the library's writer: the
    reference (Load_dynamic + the initialisation block) and the oracle agree
    bit for bit at every step after the restart, and at its step 0 on every
    record that carries state."""
    model, ref = extract_restart(name, str(tmp_path / "g"))
    steps = golden_steps(ref)
    got, _ = run_oracle(model, steps, str(tmp_path / "oracle"), save_index=1)
    for s in steps:
        R = read_dump(os.path.join(ref, "step_%06d.bin" % s))
        for k, v in R.items():
            if s == 0 and any(x in k for x in RESTART_STEP0_SKIP):
                continue
            w = got[s][k]
            same = np.array_equal(v.view(np.int64), w.view(np.int64)) if v.dtype.kind == "f" else np.array_equal(v, w)
            assert same, "step %d record %s differs" % (s, k)


@pytest.mark.parametrize("name", __import__("conftest").STRESSES)
def test_oracle_nodal_stresses_match_reference(name, tmp_path, monkeypatch):
    """Body::Compute_nodal_stresses (Code/Body.h:2310-2519) restated in the
    oracle: its 19 node fields agree with the reference's bit for bit at
    steps 0, 10 and 50, for neo-Hookean and linear-elastic bodies and rigid
    ones (zeros, jacobian 1); the stresses are not trivially zero."""
    from conftest import extract_golden
    model, ref = extract_golden(name + "_stress", str(tmp_path / "g"))
    steps = golden_steps(ref)
    monkeypatch.setenv("MEL_DUMP_STRESS", "1")
    got, _ = run_oracle(model, steps, str(tmp_path / "oracle"))
    nonzero = False
    for s in steps:
        R = read_dump(os.path.join(ref, "step_%06d.bin" % s))
        for k, v in R.items():
            if not k.endswith(".stress"):
                continue
            w = got[s][k]
            assert v.shape == w.shape and np.array_equal(v.view(np.int64), w.view(np.int64)), (s, k)
            nonzero |= bool(s > 0 and np.any(v.reshape(19, -1)[13] != 0))
    assert nonzero
