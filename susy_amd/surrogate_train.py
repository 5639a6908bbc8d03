"""Training of the learned surrogate for the U(4) 16^3 x 32 workload
(lambda 2, bmass 0.4, G 0.05, kappa_u1 0, Nroot 2, Norder 15).

The network is local and translation invariant (two convolution layers over
the nearest-neighbour taps of the A4* lattice on gauge-invariant plaquette
and 1x2-loop features), so parameters trained on one lattice size apply on
any other.  It is trained here on the 8^3 x 16 lattice of the workload's
start configuration, which makes hundreds of (configuration, fermion force)
pairs affordable:

  1. plain RHMC in the FP32 form from the committed thermalised configuration
     (susy_amd/configs/therm_u4_8x8x8x16.milc), trajectories of length
     `--traj-length` with `--nstep` outer steps, recording the (links,
     summed fermion force) pair of every outer kick and keeping the last
     `--pairs`;
  2. Adam (lr 1e-3, batch 2) on the force-matching loss |F_theta - F|^2/|F|^2
     over all pairs but the last `--held`, which are held out;
  3. the parameters (float32, the order of susy_sg_get_params) are saved.

With `--split-first J` the trajectories run in the pole-split form (poles
J..14 on a middle MD level of `--split-nmid` steps): the recorded force is
then the outer level's (the small-shift poles), which is what the surrogate
stands in for on the inner level when it is used together with the split.

    python -m susy_amd.surrogate_train --traj 30 --pairs 500 --adam 5000 \\
        --save susy_amd/configs/surrogate_u4_theta.npy
"""
import argparse
import json
import time

import numpy as np

import susy_amd
from susy_amd import start_config


def say(*a):
    print(*a, flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--traj", type=int, default=30)
    ap.add_argument("--traj-length", type=float, default=0.25)
    ap.add_argument("--nstep", type=int, default=20)
    ap.add_argument("--pairs", type=int, default=500)
    ap.add_argument("--held", type=int, default=50)
    ap.add_argument("--adam", type=int, default=5000)
    ap.add_argument("--seed", type=int, default=57)
    ap.add_argument("--save", default=None)
    ap.add_argument("--max-seconds", type=float, default=None,
                    help="start no further trajectory after this many seconds")
    ap.add_argument("--split-first", type=int, default=0)
    ap.add_argument("--split-nmid", type=int, default=1)
    a = ap.parse_args(argv)

    kw = dict(start_config.COUPLINGS, traj_length=a.traj_length, nstep=a.nstep)
    lat = susy_amd.Lattice(4, dims=start_config.SMALL_DIMS, iseed=a.seed, **kw)
    lat.load_config(start_config.START_FILE)
    lat.sg_init(seed=1, scale=1.0)
    form = susy_amd.FORM_FP32
    if a.split_first:
        lat.set_split(a.split_first, a.split_nmid)
        form |= susy_amd.FORM_SPLIT
    lat.sg_record(a.pairs)
    t0 = time.perf_counter()
    acc, done = 0, 0
    for k in range(a.traj):
        if a.max_seconds is not None and time.perf_counter() - t0 > a.max_seconds:
            break
        st = lat.update(form)
        acc += int(st[3])
        done += 1
        say(f"generate traj {k}: dH {st[2]:.5f} accept {int(st[3])} iters {int(st[4])} "
            f"plaq {lat.plaquette()[0]:.5f} pairs {lat.sg_pairs()}")
    lat.sg_record(0)
    t_gen = time.perf_counter() - t0
    npairs = lat.sg_pairs()
    held = min(a.held, npairs // 5)
    ntrain = npairs - held
    test = range(ntrain, npairs)
    before = float(np.mean([lat.sg_pair_loss(i) for i in test]))
    t0 = time.perf_counter()
    hist = lat.sg_train(0, ntrain, steps=a.adam, batch=2, lr=1e-3, seed=3)
    t_train = time.perf_counter() - t0
    after = float(np.mean([lat.sg_pair_loss(i) for i in test]))
    say(f"pairs {npairs} (train {ntrain}, held out {held}); held-out loss {before:.4f} -> "
        f"{after:.4f}; {a.adam} Adam steps in {t_train:.1f} s")
    if a.save:
        np.save(a.save, lat.sg_get_params().astype(np.float32))
    say(json.dumps({"dims": start_config.SMALL_DIMS, "trajectories": done,
                    "traj_length": a.traj_length, "nstep": a.nstep,
                    "acceptance": acc / max(done, 1), "generate_s": t_gen, "pairs": npairs,
                    "held_out": held, "adam_steps": a.adam, "train_s": t_train,
                    "train_loss_first_last": [float(np.mean(hist[:20])),
                                              float(np.mean(hist[-20:]))],
                    "held_out_loss": [before, after]}))


if __name__ == "__main__":
    main()
