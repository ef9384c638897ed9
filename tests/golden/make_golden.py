"""Writes the golden vectors under tests/golden.

For every small model of melody2d_amd.preprocessor.configs.variants() this
script writes the three model files, runs the reference MELODY 2D solver
compiled from its own sources (oracle/_ref/melody_ref, built by
`make -C oracle ref` from /root/reference/Code) for 200 steps, and stores the
model files together with the reference's state dumps at steps 0, 1, 2, 10,
50 and 200 in <variant>.tar.gz.  Only data is stored: model files and FP64
state dumps (format of oracle/dump_format.hpp).

Run from the repository root:  python tests/golden/make_golden.py
"""
import io
import os
import subprocess
import sys
import tarfile
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from melody2d_amd.preprocessor import write_model  # noqa: E402
from melody2d_amd.preprocessor.configs import variants  # noqa: E402

STEPS = [0, 1, 2, 10, 50, 200]
REF = os.path.join(ROOT, "oracle", "_ref", "melody_ref")


def main():
    if not os.path.exists(REF):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    for name, model in variants().items():
        with tempfile.TemporaryDirectory() as tmp:
            mdir = os.path.join(tmp, "model")
            odir = os.path.join(tmp, "ref")
            os.makedirs(odir)
            write_model(model, mdir)
            subprocess.run([REF, mdir, str(max(STEPS)), odir, ",".join(map(str, STEPS))], check=True,
                           stdout=subprocess.DEVNULL)
            out = os.path.join(HERE, name + ".tar.gz")
            with tarfile.open(out, "w:gz", compresslevel=9) as tf:
                for f in sorted(os.listdir(mdir)):
                    tf.add(os.path.join(mdir, f), arcname="model/" + f)
                for f in sorted(os.listdir(odir)):
                    tf.add(os.path.join(odir, f), arcname="ref/" + f)
            print(name, os.path.getsize(out))


if __name__ == "__main__":
    main()
