"""DEACTIVATION rules as Load_static reads them (Code/IO.h:947-968 with
Read_spy_line, Code/IO.h:11-590) and the ones the library refuses because
Update_status (Code/Proximity.h:550-1007) would read past them.  The parser
is shared by the library and the oracle; the oracle's loader reports."""
import copy
import os

import pytest

from mel_compare import run_oracle
from melody2d_amd.preprocessor import write_model
from melody2d_amd.preprocessor.configs import variants


def _with_rules(tmp_path, rules):
    m = copy.deepcopy(variants()["tiny_deact"])
    m.deactivation = rules
    d = str(tmp_path / "model")
    write_model(m, d)
    return d


@pytest.mark.parametrize("rule,message", [
    (("Position Length 0", "below", 1.0), "reads entries"),
    (("Position AllX", "below", 1.0), "reads entries"),
    (("Temperature X 0 0", "below", 1.0), "unknown quantity"),
    (("Position Y 0 -1", "under", 1.0), "unknown comparison"),
    (("Position Y 9 -1", "below", 1.0), "body out of range"),
    (("Velocity X 0 9999", "below", 1.0), "node out of range"),
    (("Force X Internal 2 3", "below", 1.0), "rigid body's nodes"),
    # body-level force: body [2] (the kind, 0: a disc) loops over nodes of body 3 (a platen)
    (("Force X Total 3 -1", "below", 1.0), "deformable nodes"),
    # kind 5 names body 5, which the model does not have
    (("Force X Neumann 0 -1", "below", 1.0), "does not exist"),
])
def test_refused_rules(tmp_path, rule, message):
    d = _with_rules(tmp_path, [rule])
    with pytest.raises(RuntimeError, match=message):
        run_oracle(d, [0, 1], str(tmp_path / "out"))


def test_ignored_and_stale_rules_run(tmp_path):
    """Quantities Update_status does not act on are read and ignored; a
    rule that holds at the first contact update stops its body there."""
    rules = [("Energy Kinetic 0", "below", -1.0), ("Work Contact 1", "above", 0.0),
             ("Stress VM 0 3", "above", 0.0), ("Contact Gapn 0 0 1", "below", 1.0),
             ("MassScaling Polar Contact_Normal 0 0.0 1.0 0.0 1.0", "below", 1.0),
             ("Position X 3 -1", "below", 1.0)]
    d = _with_rules(tmp_path, rules)
    refs, _ = run_oracle(d, [0, 9, 10], str(tmp_path / "out"))
    assert refs[9]["b3.status"][0] == 0
    assert refs[10]["b3.status"][0] == 1
    assert all(refs[10]["b%d.status" % b][0] == 0 for b in range(3))


@pytest.mark.parametrize("name", ["tiny_deact", "tiny_monitor"])
def test_binary_container_carries_rules_and_monitoring(tmp_path, name):
    """The binary container holds the DEACTIVATION and MONITORING text; a
    run from it matches the run from the ASCII files record for record."""
    import numpy as np
    from melody2d_amd.preprocessor.binary import write_binary
    v = variants()
    if name == "tiny_monitor":
        # the same entries on the platen model (the container holds no
        # damping matrices)
        m = copy.deepcopy(v["tiny_platen"])
        m.monitoring = list(v["tiny_monitor"].monitoring)
        m.flags = m.flags + ["MONITOR_ENERGY"]
    else:
        m = v[name]
    write_model(m, str(tmp_path / "m"))
    write_binary(m, str(tmp_path / "m.bin"))
    a, _ = run_oracle(str(tmp_path / "m"), [0, 100], str(tmp_path / "oa"))
    b, _ = run_oracle(str(tmp_path / "m.bin"), [0, 100], str(tmp_path / "ob"), save_index=-1)
    assert set(a[100]) == set(b[100])
    assert not [k for k in a[100] if not np.array_equal(a[100][k], b[100][k])]
