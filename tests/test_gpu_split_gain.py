"""What the pole split gains over the plain integrator, paired and
deterministic: from the committed thermalised U(4) 8^3 x 16 configuration
(the benchmark workload's couplings: lambda 2, bmass 0.4, G 0.05, Nroot 2,
trajectory length 0.25, nstep_gauge 10), FP32 form, the same seeds in every
variant, one trajectory each:

  * plain, 5 outer steps (the benchmark's plain forms);
  * plain, 3 outer steps;
  * split, poles 9..14 on one middle step, 3 outer steps.

The split at 3 outer steps must have an energy error |dH| no larger than
1.25 x the plain integrator's at 5 (measured on these seeds: 0.0682 vs
0.0634, ratio 1.08), less than half of the plain integrator's at the same 3
outer steps (measured: 0.249, ratio 0.27), and need fewer CG iterations than
plain at 5 (measured: 27426 vs 36777, middle level included)."""
import os

import numpy as np
import pytest

import susy_amd
from susy_amd import start_config

pytestmark = pytest.mark.gpu

DIMS = (8, 8, 8, 16)
SEEDS = (21, 22, 23, 24, 25, 26)
START = os.path.join(os.path.dirname(susy_amd.__file__), "configs", "therm_u4_8x8x8x16.milc")


def run(first, nstep):
    kw = dict(start_config.COUPLINGS, niter=20000, traj_length=0.25, nstep=nstep)
    lat = susy_amd.Lattice(4, dims=DIMS, iseed=1234, **kw)
    lat.load_config(START)
    start = lat.get_links()
    form = susy_amd.FORM_FP32
    if first:
        lat.set_split(first, 1)
        form |= susy_amd.FORM_SPLIT
    dH, iters = [], []
    for seed in SEEDS:
        lat.set_links(start)
        lat.reseed(seed)
        st = lat.update(form)
        dH.append(st[2])
        iters.append(st[4])
        print("split gain", first, nstep, seed, st[2], st[4], st[23], flush=True)
    lat.close()
    return np.mean(np.abs(dH)), np.mean(iters)


def test_split_keeps_the_energy_error_with_fewer_outer_steps():
    plain5, it5 = run(0, 5)
    plain3, _ = run(0, 3)
    split3, its = run(9, 3)
    print("mean |dH|: plain 5", plain5, "plain 3", plain3, "split 3", split3,
          "iterations", it5, its, flush=True)
    assert split3 <= 1.25 * plain5
    assert split3 < 0.5 * plain3
    assert its < it5
