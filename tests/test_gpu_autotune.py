"""Load-time launch options and the autotuner.

Every launch configuration the autotuner chooses between (two-launch step or
launch sequence, graph replay, register window) runs the same device routines
in the same summation orders, so a context loaded with the options it returns
is bit-identical to one loaded with the defaults; options passed in the
struct act as the MELODY2D_* environment switches of the same meaning."""
import numpy as np
import pytest

from melody2d_amd import native
from melody2d_amd.preprocessor import configs
from melody2d_amd.preprocessor.binary import write_binary

pytestmark = pytest.mark.gpu

FIELDS = ["x_disp_param", "y_disp_param", "x_vel_param", "y_vel_param", "x_contact_force", "y_contact_force"]


def _fields(ctx, nodes):
    out = {"b%d.%s" % (b, f): ctx.node_field(b, f, n) for b, n in enumerate(nodes) for f in FIELDS}
    t, works, energy = ctx.scalars()
    out["works"] = np.asarray(works)
    return out


def _same(a, b):
    return [k for k in a if not np.array_equal(a[k], b[k])]


def test_options_act_as_environment(tmp_path, monkeypatch):
    print("case options vs environment", flush=True)
    m = configs.variants()["tiny_platen"]
    path = str(tmp_path / "m.bin")
    write_binary(m, path)
    nodes = [b.n_nodes for b in m.bodies]
    opt = native.Context(path, binary=True, options={"fused_step": 0, "wide_window": 1})
    assert opt.options()["fused_step"] == 0 and opt.options()["wide_window"] == 1
    monkeypatch.setenv("MELODY2D_FUSED_STEP", "0")
    monkeypatch.setenv("MELODY2D_WIDE_LANES", "1")
    env = native.Context(path, binary=True)
    assert env.options()["fused_step"] == 0 and env.options()["wide_window"] == 1
    # the struct wins over the environment
    over = native.Context(path, binary=True, options={"fused_step": 1, "graph_replay": 1, "wide_window": 0})
    o = over.options()
    assert o["fused_step"] == 1 and o["graph_replay"] == 1 and o["wide_window"] == 0
    for s in (10, 40):
        for c in (opt, env, over):
            c.euler_steps(s - (10 if s == 40 else 0))
        a, b, c3 = _fields(opt, nodes), _fields(env, nodes), _fields(over, nodes)
        assert not _same(a, b), "step %d" % s
        assert not _same(a, c3), "step %d" % s
    for c in (opt, env, over):
        c.close()


@pytest.mark.parametrize("precision", [native.FP64, native.FP32_MIXED])
def test_autotune_choice_is_bit_identical(precision, tmp_path):
    print("case autotune", precision, flush=True)
    m = configs.c2(n_grains=60, nodes_per_grain=120, workers=1)
    path = str(tmp_path / "c2.bin")
    write_binary(m, path)
    best, ms, n = native.autotune(path, precision=precision, steps=40)
    print("autotune", best, ms, n, flush=True)
    assert n >= 3 and ms > 0
    assert best["fused_step"] in (0, 1) and best["graph_replay"] in (0, 1) and best["wide_window"] in (0, 1)
    nodes = [b.n_nodes for b in m.bodies]
    tuned = native.Context(path, binary=True, precision=precision, options=best)
    plain = native.Context(path, binary=True, precision=precision)
    for c in (tuned, plain):
        c.euler_steps(120)
    assert not _same(_fields(tuned, nodes), _fields(plain, nodes))
    tuned.close()
    plain.close()
