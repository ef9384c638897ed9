"""Writes the golden vectors under tests/golden.

For every small model of melody2d_amd.preprocessor.configs.variants() this
script writes the three model files, runs the reference MELODY 2D solver
compiled from its own sources (oracle/_ref/melody_ref, built by
`make -C oracle ref` from /root/reference/Code), and stores the model files
together with the reference's state dumps (format of oracle/dump_format.hpp)
in <variant>.tar.gz: steps 0, 1, 2, 10, 50 and 200, and also 1000 for the
periodic shear cell, whose top wall needs that long to carry contact points
across its periodic seam, and 80 and 90 for the deactivation variant.

The two-disc platen workload (configs.c1, ~800 nodes) is too large to store
as model files: c1.tar.gz holds only the reference's dumps at steps 0, 1,
10, 100 and 1000 (a subset of the records) and the SHA-256 of the model
files the generator writes, which the tests regenerate and check before
use.  Only data is stored.

Restart sets <variant>_restart.tar.gz: the oracle runs the variant for 30
steps and writes its state with the library's DYNAMIC writer
(DYNAMIC_00001.asc); the reference restarts from that file (Load_dynamic
and the initialisation block) and its dumps at steps 0, 1, 10 and 50 are
stored beside the file.

Stress sets <variant>_stress.tar.gz: the model files and the reference's
nodal stress fields (Body::Compute_nodal_stresses) at steps 0, 10 and 50.

Run from the repository root:  python tests/golden/make_golden.py [name,...]
"""
import hashlib
import os
import subprocess
import sys
import tarfile
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from melody2d_amd.preprocessor import write_model  # noqa: E402
from melody2d_amd.preprocessor.configs import c1, variants  # noqa: E402

STEPS = [0, 1, 2, 10, 50, 200]
# the deactivation variant also at the contact update that stops the upper
# disc (step 80) and within the period after it (90), while the lower disc
# and the top platen still name it as master
LONG = {"tiny_periodic": STEPS + [1000], "tiny_deact": sorted(STEPS + [80, 90])}
C1_STEPS = [0, 1, 10, 100, 1000]
C1_FIELDS = ("x_current,y_current,x_disp_param,y_disp_param,x_vel_param,y_vel_param,x_internal_force,"
             "y_internal_force,x_contact_force,y_contact_force,x_force,y_force,rigid,works,energy,neighbours,"
             "contacts_i,contacts_d,length")
RESTART = ["tiny_czm", "tiny_bc", "tiny_holed", "tiny_periodic"]
RESTART_STEPS = [0, 1, 10, 50]
# nodal stress sets <name>_stress.tar.gz: the model files and the
# reference's Body::Compute_nodal_stresses fields (the harness's "stress"
# record, MEL_DUMP_STRESS) with the displacement parameters they come from;
# tiny_free has linear-elastic grains, the others neo-Hookean ones and rigid
# platens
STRESS = ["tiny_platen", "tiny_bc", "tiny_free"]
STRESS_STEPS = [0, 10, 50]
REF = os.path.join(ROOT, "oracle", "_ref", "melody_ref")


def model_hash(mdir):
    h = hashlib.sha256()
    for f in sorted(os.listdir(mdir)):
        h.update(f.encode())
        with open(os.path.join(mdir, f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()


def run_ref(mdir, odir, steps, env=None, save_index=0):
    subprocess.run([REF, mdir, str(max(steps)), odir, ",".join(map(str, steps)), str(save_index)], check=True,
                   stdout=subprocess.DEVNULL, env=env)


def make_restart(name, tmp):
    import ctypes
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from mel_compare import oracle
    mdir = os.path.join(tmp, "model")
    odir = os.path.join(tmp, "ref")
    os.makedirs(odir)
    write_model(variants()[name], mdir)
    lib = oracle()
    lib.oracle_write_after.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
    dyn = os.path.join(mdir, "DYNAMIC_00001.asc")
    if lib.oracle_write_after(os.fsencode(mdir), 0, 30, os.fsencode(dyn), 0) != 0:
        raise RuntimeError(lib.oracle_last_error().decode())
    run_ref(mdir, odir, RESTART_STEPS, save_index=1)
    out = os.path.join(HERE, name + "_restart.tar.gz")
    with tarfile.open(out, "w:gz", compresslevel=9) as tf:
        tf.add(dyn, arcname="DYNAMIC_00001.asc")
        for f in sorted(os.listdir(odir)):
            tf.add(os.path.join(odir, f), arcname="ref/" + f)
    print(name + "_restart", os.path.getsize(out))


def make_stress(name, tmp):
    from melody2d_amd.preprocessor.configs import tiny_free
    mdir = os.path.join(tmp, "model")
    odir = os.path.join(tmp, "ref")
    os.makedirs(odir)
    write_model(tiny_free() if name == "tiny_free" else variants()[name], mdir)
    env = dict(os.environ, MEL_DUMP_STRESS="1", MEL_DUMP_FIELDS="stress,x_disp_param,y_disp_param")
    run_ref(mdir, odir, STRESS_STEPS, env)
    out = os.path.join(HERE, name + "_stress.tar.gz")
    with tarfile.open(out, "w:gz", compresslevel=9) as tf:
        for f in sorted(os.listdir(mdir)):
            tf.add(os.path.join(mdir, f), arcname="model/" + f)
        for f in sorted(os.listdir(odir)):
            tf.add(os.path.join(odir, f), arcname="ref/" + f)
    print(name + "_stress", os.path.getsize(out))


def main(names=None):
    if not os.path.exists(REF):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    todo = dict(variants())
    todo["c1"] = None
    for name in (names or list(todo) + [r + "_restart" for r in RESTART] + [r + "_stress" for r in STRESS]):
        with tempfile.TemporaryDirectory() as tmp:
            if name.endswith("_restart"):
                make_restart(name[:-len("_restart")], tmp)
                continue
            if name.endswith("_stress"):
                make_stress(name[:-len("_stress")], tmp)
                continue
            mdir = os.path.join(tmp, "model")
            odir = os.path.join(tmp, "ref")
            os.makedirs(odir)
            out = os.path.join(HERE, name + ".tar.gz")
            if name == "c1":
                write_model(c1(), mdir)
                env = dict(os.environ, MEL_DUMP_FIELDS=C1_FIELDS)
                run_ref(mdir, odir, C1_STEPS, env)
                with open(os.path.join(tmp, "model.sha256"), "w") as f:
                    f.write(model_hash(mdir) + "\n")
                with tarfile.open(out, "w:gz", compresslevel=9) as tf:
                    tf.add(os.path.join(tmp, "model.sha256"), arcname="model.sha256")
                    for f in sorted(os.listdir(odir)):
                        tf.add(os.path.join(odir, f), arcname="ref/" + f)
            else:
                write_model(todo[name], mdir)
                run_ref(mdir, odir, LONG.get(name, STEPS))
                with tarfile.open(out, "w:gz", compresslevel=9) as tf:
                    for f in sorted(os.listdir(mdir)):
                        tf.add(os.path.join(mdir, f), arcname="model/" + f)
                    for f in sorted(os.listdir(odir)):
                        tf.add(os.path.join(odir, f), arcname="ref/" + f)
            print(name, os.path.getsize(out))


if __name__ == "__main__":
    main(sys.argv[1].split(",") if len(sys.argv) > 1 else None)
