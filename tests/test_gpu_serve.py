"""Job server (mel_server_*): independent jobs run concurrently on one device
give exactly the state each gives alone, whatever the slicing, the arrival
order and the HBM budget; failures stay with their job."""
import os

import numpy as np
import pytest

from melody2d_amd import native
from melody2d_amd.preprocessor import configs
from melody2d_amd.preprocessor.binary import write_binary

pytestmark = pytest.mark.gpu

JOBS = ["tiny", "tiny_platen", "tiny_czm", "tiny_cohesion", "tiny_damped"]
FIELDS = ["x_disp_param", "y_disp_param", "x_vel_param", "y_vel_param", "x_current", "y_current",
          "x_internal_force", "y_internal_force", "x_contact_force", "y_contact_force"]


@pytest.fixture(scope="module")
def job_files(tmp_path_factory):
    d = tmp_path_factory.mktemp("serve")
    v = configs.variants()
    out = {}
    for name in JOBS:
        p = str(d / (name + ".bin"))
        write_binary(v[name], p)
        out[name] = (p, [b.n_nodes for b in v[name].bodies])
    return out


def _state(ctx, nodes):
    s = {}
    for b, n in enumerate(nodes):
        for f in FIELDS:
            s["b%d.%s" % (b, f)] = ctx.node_field(b, f, n)
        r, w = ctx.body_state(b)
        s["b%d.rigid" % b] = r
        s["b%d.works" % b] = w
        ids, st = ctx.contacts(b)
        s["b%d.cids" % b] = ids
        s["b%d.cstate" % b] = st
    return s


def _alone(path, nodes, steps, precision):
    with native.Context(path, binary=True, precision=precision) as c:
        c.euler_steps(steps)
        c.synchronize()
        return _state(c, nodes), c.device_bytes()


def _same(a, b):
    assert a.keys() == b.keys()
    diff = [k for k in a if not np.array_equal(np.asarray(a[k]), np.asarray(b[k]))]
    assert not diff, diff[:8]


@pytest.mark.parametrize("precision,graphs", [(native.FP64, True), (native.FP32_MIXED, True), (native.FP64, False),
                                              (native.FP32_MIXED, False)])
def test_jobs_in_the_mix_equal_jobs_alone(job_files, precision, graphs):
    steps = {name: 60 + 17 * i for i, name in enumerate(JOBS)}
    print("case serve mix", precision, graphs, flush=True)
    alone = {n: _alone(job_files[n][0], job_files[n][1], steps[n], precision)[0] for n in JOBS}
    with native.Server(hbm_budget=8 << 30, steps_per_slice=7, graph_replay=graphs) as srv:
        ids = {}
        for i, n in enumerate(JOBS):
            # two arrive later, while the others run
            ids[n] = srv.submit(job_files[n][0], steps[n], precision=precision,
                                arrival=0.0 if i < 3 else 0.02 * i, keep=True)
        assert srv.run(timeout=120.0) == 0
        for n in JOBS:
            inf = srv.job(ids[n])
            assert inf.state == native.JOB_DONE, (n, srv.job_error(ids[n]))
            assert inf.steps_done == steps[n]
            assert inf.admitted >= inf.arrival and inf.finished >= inf.admitted
            assert inf.n_nodes == sum(job_files[n][1])
            _same(alone[n], _state(srv.context(ids[n]), job_files[n][1]))
        used, budget = srv.usage()
        assert used == sum(srv.job(ids[n]).device_bytes for n in JOBS)
        for n in JOBS:
            srv.release(ids[n])
            assert srv.job(ids[n]).state == native.JOB_RELEASED
        assert srv.usage()[0] == 0


def test_budget_admits_one_job_at_a_time(job_files):
    """A budget that holds one job: jobs run one after another, in arrival
    order, each admitted only after the one before it is done."""
    names = ["tiny_platen", "tiny", "tiny_czm"]
    print("case serve budget", flush=True)
    fp = {n: _alone(job_files[n][0], job_files[n][1], 1, native.FP64)[1] for n in names}
    budget = max(fp.values()) + min(fp.values()) // 2
    with native.Server(hbm_budget=budget, steps_per_slice=25) as srv:
        ids = [srv.submit(job_files[n][0], 80, arrival=0.001 * i) for i, n in enumerate(names)]
        assert srv.run(timeout=120.0) == 0
        infos = [srv.job(j) for j in ids]
        for n, inf in zip(names, infos):
            assert inf.state == native.JOB_DONE
            assert inf.device_bytes == fp[n]
        for a, b in zip(infos, infos[1:]):
            assert b.admitted >= a.finished
        # a job that was not kept holds no context
        with pytest.raises(native.MelError):
            srv.context(ids[0])


def test_failed_job_does_not_stop_the_others(job_files, tmp_path):
    print("case serve failure", flush=True)
    bad = str(tmp_path / "garbage.bin")
    with open(bad, "wb") as f:
        f.write(os.urandom(4096))
    path, nodes = job_files["tiny"]
    ref, fp = _alone(path, nodes, 40, native.FP64)
    with native.Server(hbm_budget=8 << 30, steps_per_slice=16) as srv:
        jb = srv.submit(bad, 10)
        jg = srv.submit(path, 40, keep=True)
        # a missing container is refused at submission
        with pytest.raises(native.MelError):
            srv.submit(str(tmp_path / "missing.bin"), 10)
        assert srv.run(timeout=60.0) == 0
        assert srv.job(jb).state == native.JOB_FAILED
        assert srv.job(jb).status in (2, 3)
        assert srv.job_error(jb)
        assert srv.job(jg).state == native.JOB_DONE
        _same(ref, _state(srv.context(jg), nodes))
    with native.Server(hbm_budget=fp // 2, steps_per_slice=16) as srv:
        j = srv.submit(path, 10)
        assert srv.run(timeout=60.0) == 0
        assert srv.job(j).state == native.JOB_FAILED
        assert srv.job(j).status == 7
        assert srv.usage()[0] == 0


def test_low_estimate_waits_for_memory(job_files):
    """A job whose device footprint per container byte is above every ratio
    seen so far is estimated low; when its real footprint does not fit beside
    the running job it goes back to the queue and starts once that job is
    done, so the jobs on the device never hold more than the budget."""
    print("case serve low estimate", flush=True)
    fp = {n: _alone(job_files[n][0], job_files[n][1], 1, native.FP64)[1] for n in JOBS}
    ratio = {n: fp[n] / os.path.getsize(job_files[n][0]) for n in JOBS}
    pairs = [(a, b) for a in JOBS for b in JOBS if a != b and ratio[b] > max(4.0, ratio[a])]
    if not pairs:
        pytest.skip("no job here has a footprint ratio above the server's first estimate (%s)" % ratio)
    a, b = pairs[0]
    budget = fp[a] + fp[b] - 1
    with native.Server(hbm_budget=budget, steps_per_slice=10) as srv:
        ja = srv.submit(job_files[a][0], 600, arrival=0.0)
        jb = srv.submit(job_files[b][0], 30, arrival=0.001)
        assert srv.run(timeout=120.0) == 0
        ia, ib = srv.job(ja), srv.job(jb)
        assert ia.state == native.JOB_DONE and ib.state == native.JOB_DONE, srv.job_error(jb)
        assert ib.device_bytes == fp[b]
        assert ib.admitted >= ia.finished
