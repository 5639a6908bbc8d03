"""Bitwise repeatability of the operator, the multi-shift CG and whole FP64
trajectories at U(3) 8^3 x 16 (regression test: stores whose data registers
were overwritten before the store read them made repeated FP64 DSq and CG
runs differ in 64-byte groups of the output; 4^4 never showed it)."""
import json
import os

import numpy as np
import pytest

import susy_amd
from oracle_lib import GOLDEN, random_links, random_tf

pytestmark = pytest.mark.gpu
CFG1 = json.load(open(os.path.join(GOLDEN, "config1_ref.json")))


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def setup():
    c = CFG1
    lat = susy_amd.Lattice(3, dims=tuple(c["dims"]))
    lat.set_links(random_links(lat.link_shape(), c["link_seed"], c["scale"]))
    src = random_tf(lat.tf_shape(), c["src_seed"])
    src = src / np.sqrt(np.vdot(src, src).real)
    h = lat.tf_alloc(18)
    lat.tf_upload(h, src)
    return lat, h


@pytest.mark.parametrize("form", [susy_amd.FORM_FP64, susy_amd.FORM_FP32])
def test_dsq_repeats_bitwise(setup, form):
    lat, h = setup
    say("dsq repeats, form", form)
    lat.dsq(h, h + 1, form=form)
    ref = lat.tf_download(h + 1).copy()
    for k in range(100):
        lat.dsq(h, h + 1, form=form)
        assert np.array_equal(lat.tf_download(h + 1), ref), k


@pytest.mark.parametrize("form", [susy_amd.FORM_FP64, susy_amd.FORM_FP32])
def test_congrad_multi_repeats_bitwise(setup, form):
    lat, h = setup
    say("cg repeats, form", form)
    shifts = np.array(CFG1["shifts"])
    runs = []
    for _ in range(3):
        it, sz = lat.congrad_multi(h, h + 2, shifts, 5000, CFG1["errormin"], form=form)
        runs.append((it, sz, lat.tf_download(h + 2), lat.tf_download(h + 16)))
    for r in runs[1:]:
        assert r[0] == runs[0][0] and r[1] == runs[0][1]
        assert np.array_equal(r[2], runs[0][2]) and np.array_equal(r[3], runs[0][3])


def test_fp64_trajectories_repeat_bitwise():
    kw = dict(dims=(8, 8, 8, 16), PBC=-1, iseed=5, Nroot=2, Norder=15, lam=1.0, kappa_u1=0.3,
              bmass=0.5, fmass=0.0, G=0.05, niter=5000, error_per_site=1e-5, traj_length=0.2,
              nstep=4, nstep_gauge=10)
    outs = []
    for rep in range(2):
        say("trajectory repeat", rep)
        lat = susy_amd.Lattice(3, **kw)
        st = lat.update(susy_amd.FORM_FP64)
        outs.append((st, lat.get_links(), lat.get_mom()))
        lat.close()
    assert np.array_equal(outs[0][0], outs[1][0])
    assert np.array_equal(outs[0][1], outs[1][1]) and np.array_equal(outs[0][2], outs[1][2])


def test_forces_and_link_update_repeat_bitwise():
    """fermion_force, gauge_force + update_u (with det_force: kappa_u1 > 0) and
    the surrogate force at U(3) 8^3 x 16, each repeated from identical inputs"""
    kw = dict(dims=(8, 8, 8, 16), PBC=-1, iseed=5, Nroot=1, Norder=15, lam=2.0, kappa_u1=0.3,
              bmass=0.4, G=0.05)
    lat = susy_amd.Lattice(3, **kw)
    U = random_links(lat.link_shape(), 11, 0.2)
    mom = random_tf(lat.link_shape(), 12)
    sol = random_tf((3,) + lat.tf_shape(), 13)
    h = lat.tf_alloc(3)
    for j in range(3):
        lat.tf_upload(h + j, sol[j])
    amps = lat.rhmc()[1:4]
    lat.sg_init(seed=4, scale=1.0)
    outs = []
    for rep in range(3):
        say("force repeat", rep)
        lat.set_links(U)
        lat.set_mom(mom)
        nf = lat.fermion_force(0.02, h, amps)
        m1 = lat.get_mom()
        ng = lat.gauge_force(0.03)
        lat.update_u(0.01)
        U1 = lat.get_links()
        S, F = lat.sg_eval(force=True)
        outs.append((nf, m1, ng, U1, S, F))
    for o in outs[1:]:
        assert o[0] == outs[0][0] and o[2] == outs[0][2] and o[4] == outs[0][4]
        for q in (1, 3, 5):
            assert np.array_equal(o[q], outs[0][q]), q
