import os
import sys
import tarfile

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

GOLDEN = os.path.join(HERE, "golden")
# golden sets that carry their model files; c1.tar.gz carries the
# reference's dumps and the hash of a model the generator rebuilds
VARIANTS = sorted(f[:-7] for f in os.listdir(GOLDEN)
                  if f.endswith(".tar.gz") and f != "c1.tar.gz" and not f.endswith("_restart.tar.gz")
                  and not f.endswith("_stress.tar.gz"))
# golden sets run with the adaptive schemes of Solver_step; the others are Euler
ADAPTIVE_VARIANTS = [v for v in VARIANTS
                     if v in ("tiny_adaptive", "tiny_mass_scaling", "tiny_adaptive_ms", "tiny_deact_adaptive")]
EULER_VARIANTS = [v for v in VARIANTS if v not in ADAPTIVE_VARIANTS]
# models the two-launch step runs (absorbing layers, deactivation rules and
# inactive bodies take the launch sequence)
TWO_LAUNCH_VARIANTS = [v for v in EULER_VARIANTS if v not in ("tiny_alid", "tiny_deact", "tiny_inactive")]
RESTARTS = sorted(f[:-15] for f in os.listdir(GOLDEN) if f.endswith("_restart.tar.gz"))
# nodal stress sets: model files and the reference's stress fields
STRESSES = sorted(f[:-14] for f in os.listdir(GOLDEN) if f.endswith("_stress.tar.gz"))
# records a restart does not carry over at its step 0: the force fields and
# pressures Load_dynamic reads back (Code/IO.h:2055-2130, 2292-2340; the
# internal force is then zeroed by Node::Sum_up_forces, Code/Node.h:212-225)
# only feed outputs, never the next step
RESTART_STEP0_SKIP = ("_force", "contact_pressure", ".energy", ".rigid")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def extract_golden(name, dest):
    """Model files under dest/model, reference dumps under dest/ref.  For
    c1 the model is rebuilt by the seeded generator and must hash to the
    value stored beside the dumps."""
    with tarfile.open(os.path.join(GOLDEN, name + ".tar.gz"), "r:gz") as tf:
        for m in tf.getmembers():
            if not m.isfile() or m.name.startswith("/") or ".." in m.name:
                raise ValueError("unexpected member %s" % m.name)
        tf.extractall(dest)
    model = os.path.join(dest, "model")
    if name == "c1":
        from golden.make_golden import model_hash
        from melody2d_amd.preprocessor import write_model
        from melody2d_amd.preprocessor.configs import c1
        write_model(c1(), model)
        with open(os.path.join(dest, "model.sha256")) as f:
            want = f.read().strip()
        if model_hash(model) != want:
            pytest.skip("the generator did not rebuild the c1 model bit for bit on this host")
    return model, os.path.join(dest, "ref")


def extract_restart(name, dest):
    """The model files of golden `name` with the restart file DYNAMIC_00001.asc
    the oracle wrote after 30 steps under dest/model, and the reference's
    dumps of the restart under dest/ref."""
    model, _ = extract_golden(name, os.path.join(dest, "base"))
    with tarfile.open(os.path.join(GOLDEN, name + "_restart.tar.gz"), "r:gz") as tf:
        for m in tf.getmembers():
            if not m.isfile() or m.name.startswith("/") or ".." in m.name:
                raise ValueError("unexpected member %s" % m.name)
        tf.extractall(os.path.join(dest, "r"))
    import shutil
    shutil.copy(os.path.join(dest, "r", "DYNAMIC_00001.asc"), os.path.join(model, "DYNAMIC_00001.asc"))
    return model, os.path.join(dest, "r", "ref")


def golden_steps(ref_dir):
    """Steps of the reference dumps in a golden set."""
    return sorted(int(f[5:11]) for f in os.listdir(ref_dir) if f.startswith("step_"))


@pytest.fixture(scope="session")
def golden_dir(tmp_path_factory):
    cache = {}

    def get(name):
        if name not in cache:
            d = tmp_path_factory.mktemp("golden_" + name)
            cache[name] = extract_golden(name, str(d))
        return cache[name]
    return get
