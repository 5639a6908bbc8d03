"""What the trained surrogate gains over the method without it, at equal outer
steps: U(2) 8^4 (lambda 1, bmass 0.5, G 0.05, Nroot 1; trajectory length 1,
nstep 20, nstep_gauge 10), the network trained by
`python -m susy_amd.surrogate_run --train-traj 200 --pairs 2000 --adam 5000
--save-params tests/golden/surrogate_u2_theta.npy
--save-start tests/golden/surrogate_u2_start.milc` (the final configuration
of its generating chain and the parameters after training are committed).

A plain FP64 chain of NCONF trajectories from the committed start provides
the configurations; from each one, one trajectory of the plain FP64 form and
one of the FP64 form with the surrogate on the inner level run with the same
seed (identical momenta and pseudofermions), so the comparison is paired and
deterministic.  Measured on an MI355X: mean |dH| 0.02043 (plain) against
0.01976 (surrogate), ratio 0.967 -- the trained network (held-out relative
force error 0.50) takes 3% off the outer level's energy error at equal outer
steps, worth about 1.7% fewer outer steps at equal |dH| (|dH| ~ dt^2).  The
test asserts that gain (ratio <= GAIN).  Control: the same network with all
parameters zero has S_theta = 0, and its trajectories are the plain ones bit
for bit."""
import os

import numpy as np
import pytest

import susy_amd
from oracle_lib import GOLDEN

pytestmark = pytest.mark.gpu

THETA = os.path.join(GOLDEN, "surrogate_u2_theta.npy")
START = os.path.join(GOLDEN, "surrogate_u2_start.milc")
KW = dict(dims=(8, 8, 8, 8), PBC=-1, iseed=77, Nroot=1, Norder=15, lam=1.0, kappa_u1=0.0,
          bmass=0.5, fmass=0.0, G=0.05, niter=5000, error_per_site=1e-5, traj_length=1.0,
          nstep=20, nstep_gauge=10)
NCONF = 24
GAIN = 0.98


def paired_dH(lat, configs, form):
    out = []
    for k, U in enumerate(configs):
        lat.set_links(U)
        lat.reseed(1000 + k)
        out.append(float(lat.update(form)[2]))
    return np.array(out)


def test_trained_surrogate_lowers_the_energy_error():
    theta = np.load(THETA)
    lat = susy_amd.Lattice(2, **KW)
    assert lat.load_config(START) == 0
    configs = []
    for k in range(NCONF):
        lat.update(susy_amd.FORM_FP64)
        configs.append(lat.get_links())
    lat.sg_init(seed=1, scale=1.0)
    lat.sg_set_params(np.zeros_like(theta))
    plain = paired_dH(lat, configs, susy_amd.FORM_FP64)
    zero = paired_dH(lat, configs, susy_amd.FORM_FP64 | susy_amd.FORM_SURROGATE)
    assert np.array_equal(zero, plain)
    lat.sg_set_params(theta)
    sg = paired_dH(lat, configs, susy_amd.FORM_FP64 | susy_amd.FORM_SURROGATE)
    ratio = np.mean(np.abs(sg)) / np.mean(np.abs(plain))
    print(f"mean |dH|: plain {np.mean(np.abs(plain)):.5f}, surrogate {np.mean(np.abs(sg)):.5f}, "
          f"ratio {ratio:.3f}", flush=True)
    assert ratio <= GAIN, (ratio, np.mean(np.abs(plain)), np.mean(np.abs(sg)))
