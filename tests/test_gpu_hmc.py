"""End-to-end RHMC trajectories on the GPU against the reference's own
regression outputs (testsuite/scalar/hmc.U{2,3,4}.ref): per-term actions to
the 8 printed digits, start/end S to 12 digits, accept/reject, force monitors
and the exact CG iteration count of every trajectory; then the measurements
printed after each trajectory (control.c:85-120): GMES (Polyakov loop,
plaquettes, bosonic action per site, mean |Polyakov|) to 8 digits, FLINK and
the Wilson lines in all five directions to 6 digits."""
import json
import os
import sys

import pytest

import susy_amd
from oracle_lib import GOLDEN, config_path
from refcheck import check_traj, hmc_kwargs, printed_ok

MEAS = json.load(open(os.path.join(GOLDEN, "meas_ref.json")))

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ncol", [2, 3, 4])
def test_trajectories_reproduce_hmc_ref(ncol, hmc_ref):
    print("hmc", ncol, flush=True)
    ref = hmc_ref[f"U{ncol}"]
    lat = susy_amd.Lattice(ncol, **hmc_kwargs(ref["input"]))
    assert lat.load_config(config_path(ncol)) == 0
    ss, st = lat.plaquette()
    assert abs(ss / ref["ref"]["check_plaq"][0] - 1) < 1e-14
    for k, tr in enumerate(ref["ref"]["trajectories"]):
        stats = lat.update(susy_amd.FORM_FP64)
        print("traj", k, stats[:5], flush=True)
        sys.stdout.flush()
        check_traj(stats, tr)
        # GMES Re P, Im P, CG iterations, ss and st plaquettes, S_B / V, <|P|>
        gm = tr["gmes"]
        re, im, mod = lat.wilson_line(3)
        ss, st = lat.plaquette()
        sb = lat.actions()[0] / lat.V
        for got, want in zip((re, im, ss, st, sb, mod), (gm[0], gm[1], gm[3], gm[4], gm[5],
                                                         gm[6])):
            assert printed_ok(got, want, 8), ("gmes", k, got, want)
        assert int(stats[4]) == int(gm[2])
        ms = MEAS[f"U{ncol}"]["trajectories"][k]
        lt, ave, width = lat.link_trace()[:3]
        for got, want in zip(list(lt) + [ave, width], ms["flink"]):
            assert printed_ok(got, want, 6), ("flink", k, got, want)
        for a in range(5):
            re, im, _ = lat.wilson_line(a)
            assert printed_ok(re, ms["lines"][2 * a], 6), ("lines", k, a, re)
            assert printed_ok(im, ms["lines"][2 * a + 1], 6), ("lines", k, a, im)


def test_config_writer_round_trip(tmp_path):
    """a loaded reference configuration written back is byte-identical, and
    the written file loads with valid checksums"""
    path = config_path(2)
    raw = open(path, "rb").read()
    lat = susy_amd.Lattice(2, dims=(4, 4, 4, 4))
    assert lat.load_config(path) == 0
    out = tmp_path / "saved.U2.4444"
    lat.save_config(str(out), stamp=raw[20:84])
    assert open(out, "rb").read() == raw
    lat2 = susy_amd.Lattice(2, dims=(4, 4, 4, 4))
    assert lat2.load_config(str(out)) == 0
    assert (lat2.get_links() == lat.get_links()).all()
