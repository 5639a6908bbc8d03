"""The pole-split form (SUSY_FORM_SPLIT) against the same three-level
integrator driven step by step through the C oracle's pieces
(tests/split_ref.py, pinned on the CPU by tests/test_split_ref.py).

U(2) 4^4 from config.U2.4444, lambda 1, bmass 0.5, kappa_u1 0.3, G 0.05,
Nroot 2, trajectory length 1, nstep 2, nstep_gauge 2, first pole 10, two
middle steps: start S, end S and dS within 1e-10 relative (FP64; 1e-7 in
the FP32 form), the same accept/reject, CG iterations (all, and the middle
level's) exact in FP64, links within 1e-12 after each trajectory."""
import numpy as np
import pytest

import susy_amd
from oracle_lib import Oracle, config_path, rel_err
from split_ref import oracle_split_trajectory

pytestmark = pytest.mark.gpu

KW = dict(dims=(4, 4, 4, 4), PBC=-1, iseed=5, Nroot=2, Norder=15, lam=1.0, kappa_u1=0.3,
          bmass=0.5, fmass=0.0, G=0.05, niter=5000, error_per_site=1e-5, traj_length=1.0,
          nstep=2, nstep_gauge=2)
FIRST, NMID = 10, 2


def run_pair(form, weights, ntraj, tol):
    O = Oracle(2, **KW)
    assert O.load_config(config_path(2))[0] == 0
    lat = susy_amd.Lattice(2, **KW)
    assert lat.load_config(config_path(2)) == 0
    lat.set_split(FIRST, NMID, weights)
    w = O.rhmc()[1 + FIRST:16] if weights is None else np.asarray(weights)
    accepts = []
    for k in range(ntraj):
        st = lat.update(form | susy_amd.FORM_SPLIT)
        start, end, dS, accept, its, mid = oracle_split_trajectory(O, KW, FIRST, NMID, w)
        print("split", form, k, st[:5], st[23], (start, end, dS, accept, its, mid), flush=True)
        assert abs(st[0] / start - 1) < tol and abs(st[1] / end - 1) < tol, k
        assert abs(st[2] - dS) <= tol * abs(start), k
        assert bool(st[3]) == accept, k
        if form == susy_amd.FORM_FP64:
            assert int(st[4]) == its and int(st[23]) == mid, k
            assert rel_err(lat.get_links(), O.get_links()) < 1e-12, k
        else:
            assert mid > 0 and st[23] > 0, k
            assert rel_err(lat.get_links(), O.get_links()) < 1e-6, k
        accepts.append(accept)
    return accepts


def test_split_fp64_matches_oracle_driver():
    run_pair(susy_amd.FORM_FP64, None, 3, 1e-10)


def test_split_fp64_weighted_middle_level():
    """middle weights other than the rational approximation's: the outer
    level keeps amp4_j - w_j of the large-shift poles"""
    w = 0.7 * Oracle(2, **KW).rhmc()[1 + FIRST:16]
    run_pair(susy_amd.FORM_FP64, w, 2, 1e-10)


def test_split_fp32_matches_oracle_driver():
    run_pair(susy_amd.FORM_FP32, None, 2, 1e-7)


def test_split_needs_its_setup_and_the_fermions():
    lat = susy_amd.Lattice(2, **KW)
    with pytest.raises(RuntimeError):
        lat.update(susy_amd.FORM_SPLIT)
    with pytest.raises(RuntimeError):
        lat.set_split(0, 1)
    with pytest.raises(RuntimeError):
        lat.set_split(15, 1)
    with pytest.raises(RuntimeError):
        lat.set_split(10, 0)
    lat.set_split(10, 1)
    with pytest.raises(RuntimeError):
        lat.update(susy_amd.FORM_PUREGAUGE | susy_amd.FORM_SPLIT)


def test_split_settings_are_fixed_for_the_open_trajectory():
    """susy_set_split during a trajectory applies from the next one: the
    open trajectory ends bitwise as it would have"""
    def run(change):
        lat = susy_amd.Lattice(2, **KW)
        assert lat.load_config(config_path(2)) == 0
        lat.set_split(FIRST, NMID)
        lat.traj_begin(susy_amd.FORM_FP64 | susy_amd.FORM_SPLIT)
        lat.traj_step()
        if change:
            lat.set_split(12, 1, np.ones(3))
        for _ in range(KW["nstep"] - 1):
            lat.traj_step()
        st = lat.traj_end()
        return st, lat.get_links()
    st0, l0 = run(False)
    st1, l1 = run(True)
    assert np.array_equal(st0, st1) and np.array_equal(l0, l1)
