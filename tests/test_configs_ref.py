"""CPU checks of the configuration fixtures the GPU tests compare with
(tests/golden/config0_ref.json, config1_ref.json): the oracle reproduces the
first recorded trajectory, the oracle's ten trajectories of configs[0] agree
with what the reference's own susy_hmc printed for the same run
(tests/golden/config0_refbuild.json) to every printed digit, and the recorded
solve is self-consistent."""
import json
import os

import numpy as np

import measure_ref
from oracle_lib import GOLDEN, Oracle
from refcheck import check_traj, printed_ok

CFG0 = json.load(open(os.path.join(GOLDEN, "config0_ref.json")))
CFG1 = json.load(open(os.path.join(GOLDEN, "config1_ref.json")))
CFG0_REF = json.load(open(os.path.join(GOLDEN, "config0_refbuild.json")))


def test_config0_fixture_first_trajectory():
    kw = dict(CFG0["params"])
    kw["dims"] = tuple(kw["dims"])
    O = Oracle(2, **kw)
    st = O.update()
    ref = CFG0["trajectories"][0]
    assert np.array_equal(st[:23], np.array(ref["stats"]))
    assert list(O.plaquette()) == ref["plaq"]
    re_im, mod = measure_ref.wilson_line(measure_ref.Geometry(kw["dims"]), O.get_links(), 3)
    assert [re_im.real, re_im.imag, mod] == ref["ploop"]


def test_config0_oracle_fixture_matches_the_reference_build():
    """the oracle's recorded run of configs[0] (written by the oracle) against
    the numbers the reference's susy_hmc printed for the same input"""
    trajs = CFG0_REF["ref"]["trajectories"]
    assert len(trajs) == len(CFG0["trajectories"]) == 10
    for k, (tr, rec) in enumerate(zip(trajs, CFG0["trajectories"])):
        check_traj(np.array(rec["stats"]), tr)
        # GMES: Re, Im Polyakov loop, iterations, ss and st plaquettes (%.8g)
        assert printed_ok(rec["ploop"][0], tr["gmes"][0], 8), k
        assert printed_ok(rec["ploop"][1], tr["gmes"][1], 8), k
        assert printed_ok(rec["plaq"][0], tr["gmes"][3], 8), k
        assert printed_ok(rec["plaq"][1], tr["gmes"][4], 8), k


def test_config1_fixture_consistency():
    """the recorded shifts are the x^(-1/4) shifts of the oracle's Norder 15
    tables, and the shifted solutions are ordered as the shifts force them:
    |psi_j|^2 and <src, psi_j> decrease as the shift grows"""
    c = CFG1
    O = Oracle(3, dims=(2, 2, 2, 2))
    assert np.array_equal(np.array(c["shifts"]), O.rhmc()[16:31])
    n2 = np.array(c["psi_norm2"])
    d = np.array([x[0] for x in c["src_dot_psi"]])
    assert np.all(np.diff(n2) < 0) and np.all(np.diff(d) < 0)
    assert np.all(np.abs([x[1] for x in c["src_dot_psi"]]) < 1e-8 * d)
    assert c["size_r"] <= c["errormin"]
