"""End-to-end RHMC trajectories on the GPU against the reference's own
regression outputs (testsuite/scalar/hmc.U{2,3,4}.ref): per-term actions to
the 8 printed digits, start/end S to 12 digits, accept/reject, force monitors
and the exact CG iteration count of every trajectory."""
import sys

import pytest

import susy_amd
from oracle_lib import config_path
from refcheck import check_traj, hmc_kwargs

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
