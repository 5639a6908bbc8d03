"""Trajectories after a rejection, and after a gauge file is written and read
back, against the C oracle running the same chain.

U(2) 4^4 from config.U2.4444, Nroot 2, lambda 1, bmass 0.5, G 0.05,
trajectory length 1 with nstep 2 / nstep_gauge 4 (a step large enough that
the chain rejects), seed 41.  The oracle's chain accepts trajectories 0 and
1, rejects 2, 3 and 4 and accepts 5: the reject branch (saved links restored,
momenta reversed, link fields recomputed; update_o.c:230-243) runs three
times in a row and is followed by an accepted trajectory.

  * FP64 form: every stats entry of all six trajectories against the
    oracle's (actions 1e-10 relative, dS 1e-8, iterations and accept/reject
    exact, force monitors 1e-9).
  * FP32 form: the same six trajectories (accept/reject exact; actions 1e-7
    relative, dS 1e-4; iterations within 20%).
  * save / load / continue: after trajectory 2 the links are written as a
    MILC file (single precision, as the reference writes them) and read into
    a new context whose generators restart from the seed; three more
    trajectories against the oracle restarted from the same single-precision
    links and seed.
"""
import sys

import numpy as np
import pytest

import susy_amd
from oracle_lib import Oracle, config_path

pytestmark = pytest.mark.gpu

KW = dict(dims=(4, 4, 4, 4), PBC=-1, iseed=41, Nroot=2, Norder=15, lam=1.0, kappa_u1=0.0,
          bmass=0.5, fmass=0.0, G=0.05, niter=5000, error_per_site=1e-5, traj_length=1.0,
          nstep=2, nstep_gauge=4)
NTRAJ = 6
SAVE_AFTER = 3
NCONT = 3


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def oracle_chains():
    """the oracle's six trajectories, and its continuation from the
    single-precision links after SAVE_AFTER of them"""
    O = Oracle(2, **KW)
    assert O.load_config(config_path(2))[0] == 0
    chain, saved = [], None
    for k in range(NTRAJ):
        chain.append(O.update())
        say("oracle traj", k, chain[-1][:5])
        if k + 1 == SAVE_AFTER:
            saved = O.get_links().astype(np.complex64).astype(np.complex128)
    O = Oracle(2, **KW)
    O.set_links(saved)
    cont = [O.update() for _ in range(NCONT)]
    for k, st in enumerate(cont):
        say("oracle continuation", k, st[:5])
    return chain, cont


def compare(st, want, k, fp32=False):
    assert bool(st[3]) == bool(want[3]), ("accept", k, st[2], want[2])
    if fp32:
        it, iw = int(st[4]), int(want[4])
        assert abs(it - iw) <= 0.2 * iw, ("iterations", k, it, iw)
        rel_a, abs_ds, rel_f = 1e-7, 1e-4, 1e-5
    else:
        assert int(st[4]) == int(want[4]), ("iterations", k)
        rel_a, abs_ds, rel_f = 1e-10, 1e-8, 1e-9
    for q in (0, 1) + tuple(range(11, 23)):
        assert abs(st[q] - want[q]) <= rel_a * abs(want[q]) + 1e-9, (k, q, st[q], want[q])
    assert abs(st[2] - want[2]) < abs_ds, ("dS", k, st[2], want[2])
    for q in range(5, 11):
        assert abs(st[q] - want[q]) <= rel_f * abs(want[q]) + 1e-12, ("force", k, q)


def test_oracle_chain_rejects_mid_chain(oracle_chains):
    chain, _ = oracle_chains
    assert [int(s[3]) for s in chain] == [1, 1, 0, 0, 0, 1]


@pytest.mark.parametrize("form", [susy_amd.FORM_FP64, susy_amd.FORM_FP32])
def test_chain_through_rejections(oracle_chains, form):
    chain, _ = oracle_chains
    lat = susy_amd.Lattice(2, **KW)
    assert lat.load_config(config_path(2)) == 0
    for k in range(NTRAJ):
        before = lat.get_links()
        st = lat.update(form)
        say("gpu form", form, "traj", k, st[:5])
        compare(st, chain[k], k, fp32=form == susy_amd.FORM_FP32)
        if not st[3]:
            # a rejected trajectory restores the links it started from
            assert np.array_equal(lat.get_links(), before), k
    lat.close()


def test_save_load_continue(oracle_chains, tmp_path):
    chain, cont = oracle_chains
    lat = susy_amd.Lattice(2, **KW)
    assert lat.load_config(config_path(2)) == 0
    for k in range(SAVE_AFTER):
        compare(lat.update(susy_amd.FORM_FP64), chain[k], k)
    path = str(tmp_path / "saved.U2.4444")
    lat.save_config(path)
    want_links = lat.get_links().astype(np.complex64).astype(np.complex128)
    lat.close()
    lat2 = susy_amd.Lattice(2, **KW)
    assert lat2.load_config(path) == 0
    assert np.array_equal(lat2.get_links(), want_links)
    for k in range(NCONT):
        st = lat2.update(susy_amd.FORM_FP64)
        say("gpu continuation", k, st[:5])
        compare(st, cont[k], k)
    lat2.close()


@pytest.mark.parametrize("form", [susy_amd.FORM_FP64, susy_amd.FORM_FP32,
                                  susy_amd.FORM_FP64 | susy_amd.FORM_SURROGATE])
def test_pieces_equal_the_whole_trajectory(form):
    """a trajectory driven through susy_traj_begin / susy_traj_substep /
    susy_traj_end is the trajectory of susy_update, bit for bit"""
    runs = []
    for pieces in (False, True):
        lat = susy_amd.Lattice(2, **KW)
        assert lat.load_config(config_path(2)) == 0
        lat.sg_init(seed=2, scale=0.5)
        if pieces:
            lat.traj_begin(form)
            n = 0
            while n < 2 * KW["Nroot"] * KW["nstep"]:
                n = lat.traj_substep()
            st = lat.traj_end()
        else:
            st = lat.update(form)
        runs.append((st, lat.get_links(), lat.get_mom()))
        lat.close()
    assert np.array_equal(runs[0][0], runs[1][0])
    assert np.array_equal(runs[0][1], runs[1][1]) and np.array_equal(runs[0][2], runs[1][2])
