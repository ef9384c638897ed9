import os
import sys
import tarfile

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

GOLDEN = os.path.join(HERE, "golden")
VARIANTS = sorted(f[:-7] for f in os.listdir(GOLDEN) if f.endswith(".tar.gz"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def extract_golden(name, dest):
    """Model files under dest/model, reference dumps under dest/ref."""
    with tarfile.open(os.path.join(GOLDEN, name + ".tar.gz"), "r:gz") as tf:
        for m in tf.getmembers():
            if not m.isfile() or m.name.startswith("/") or ".." in m.name:
                raise ValueError("unexpected member %s" % m.name)
        tf.extractall(dest)
    return os.path.join(dest, "model"), os.path.join(dest, "ref")


@pytest.fixture(scope="session")
def golden_dir(tmp_path_factory):
    cache = {}

    def get(name):
        if name not in cache:
            d = tmp_path_factory.mktemp("golden_" + name)
            cache[name] = extract_golden(name, str(d))
        return cache[name]
    return get
