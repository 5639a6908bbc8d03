"""The learned-surrogate workload: U(2) 8^4, lambda 1, kappa_u1 0, bmass 0.5,
fmass 0, G 0.05, Nroot 1, Norder 15, antiperiodic temporal fermion BCs.

  1. plain FP64 RHMC from a cold start (trajectory length 1, nstep 10,
     nstep_gauge 10, seed 1234) records a (links, fermion force) pair at every
     outer kick; the last `--pairs` are kept (the earliest trajectories, the
     cold-start transient, fall out of the buffer);
  2. Adam (lr 1e-3, batch 2) on the force-matching loss
     |F_theta - F|^2 / |F|^2 over all pairs but the last `--held`, which are
     held out;
  3. from the same final configuration, the same seeds and momenta, runs
     `--traj` trajectories in each form: plain FP64 and FP32 with FP64
     reliable updates at each outer step count of `--nsteps`, and FP32 with
     the surrogate on the inner MD level (the exact multi-shift-CG action
     decides acceptance) at each of `--sg-nsteps`.

Prints one progress line per trajectory and a final JSON line with the
relative force error (held out, before and after training), and per form the
acceptance, <exp(-dH)>, <|dH|>, CG iterations and seconds per trajectory.

    python -m susy_amd.surrogate_run [--train-traj 200] [--pairs 2000]
        [--adam 5000] [--traj 50] [--nsteps 20] [--sg-nsteps 20]
        [--save-params THETA.npy] [--save-start START.milc]

--save-params / --save-start write the trained parameters (float32, the
order of susy_sg_get_params) and the final configuration of the generating
chain (MILC file), the inputs of the test that measures what the trained
network gains over the plain form.
"""
import argparse
import json
import time

import numpy as np

import susy_amd

COUPLINGS = dict(dims=(8, 8, 8, 8), PBC=-1, Nroot=1, Norder=15, lam=1.0, kappa_u1=0.0,
                 bmass=0.5, fmass=0.0, G=0.05, niter=5000, error_per_site=1e-5,
                 traj_length=1.0, nstep_gauge=10)


def say(*a):
    print(*a, flush=True)


def run_form(lat, form, ntraj, seed):
    dh, acc, its, secs, plaq = [], [], [], [], []
    lat.reseed(seed)
    for k in range(ntraj):
        t0 = time.perf_counter()
        st = lat.update(form)
        secs.append(time.perf_counter() - t0)
        dh.append(float(st[2]))
        acc.append(bool(st[3]))
        its.append(int(st[4]))
        plaq.append(lat.plaquette()[0])
        say(f"form {form} traj {k} dH {st[2]:.5f} accept {int(st[3])} iters {int(st[4])} "
            f"{secs[-1]:.2f} s")
    dh = np.array(dh)
    return {"acceptance": float(np.mean(acc)), "exp_minus_dH": float(np.mean(np.exp(-dh))),
            "exp_minus_dH_err": float(np.std(np.exp(-dh)) / np.sqrt(max(len(dh) - 1, 1))),
            "abs_dH": float(np.mean(np.abs(dh))), "iters_per_traj": float(np.mean(its)),
            "s_per_traj": float(np.mean(secs)), "plaquette": float(np.mean(plaq))}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--train-traj", type=int, default=200)
    ap.add_argument("--pairs", type=int, default=2000)
    ap.add_argument("--held", type=int, default=200)
    ap.add_argument("--adam", type=int, default=5000)
    ap.add_argument("--traj", type=int, default=50)
    ap.add_argument("--nsteps", default="20", help="outer steps of the plain forms (comma list)")
    ap.add_argument("--sg-nsteps", default="20", help="outer steps of the surrogate form")
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--save-params", default=None)
    ap.add_argument("--save-start", default=None)
    a = ap.parse_args(argv)

    gen = susy_amd.Lattice(2, iseed=a.seed, nstep=10, **COUPLINGS)
    gen.sg_init(seed=1, scale=1.0)
    gen.sg_record(a.pairs)
    t0 = time.perf_counter()
    acc = 0
    for k in range(a.train_traj):
        st = gen.update(susy_amd.FORM_FP64)
        acc += int(st[3])
        say(f"generate traj {k} dH {st[2]:.5f} accept {int(st[3])} iters {int(st[4])} "
            f"pairs {gen.sg_pairs()}")
    gen.sg_record(0)
    t_gen = time.perf_counter() - t0
    npairs = gen.sg_pairs()
    held = min(a.held, npairs // 5)
    ntrain = npairs - held
    test = range(ntrain, npairs)
    before = float(np.mean([gen.sg_pair_loss(i) for i in test]))
    t0 = time.perf_counter()
    hist = gen.sg_train(0, ntrain, steps=a.adam, batch=2, lr=1e-3, seed=3)
    t_train = time.perf_counter() - t0
    after = float(np.mean([gen.sg_pair_loss(i) for i in test]))
    say(f"pairs {npairs} (train {ntrain}, held out {held}); held-out loss {before:.4f} -> "
        f"{after:.4f}; {a.adam} Adam steps in {t_train:.1f} s")
    theta = gen.sg_get_params()
    start = gen.get_links()
    if a.save_params:
        np.save(a.save_params, theta.astype(np.float32))
    if a.save_start:
        gen.save_config(a.save_start)
    gen.close()

    plain = [int(x) for x in a.nsteps.split(",") if x]
    sg = [int(x) for x in a.sg_nsteps.split(",") if x]
    out = {}
    for nstep in sorted(set(plain) | set(sg), reverse=True):
        lat = susy_amd.Lattice(2, iseed=a.seed, nstep=nstep, **COUPLINGS)
        lat.sg_init(seed=1, scale=1.0)
        lat.sg_set_params(theta)
        forms = {}
        if nstep in plain:
            forms["fp64"] = susy_amd.FORM_FP64
            forms["fp32"] = susy_amd.FORM_FP32
        if nstep in sg:
            forms["fp32_surrogate"] = susy_amd.FORM_FP32 | susy_amd.FORM_SURROGATE
        for name, form in forms.items():
            lat.set_links(start)
            out[f"{name}@nstep{nstep}"] = run_form(lat, form, a.traj, a.seed + 1)
        lat.close()
    res = {"workload": "surrogate U(2) 8^4, lambda 1, bmass 0.5, G 0.05, Nroot 1",
           "generation": {"trajectories": a.train_traj, "nstep": 10, "acceptance":
                          acc / max(a.train_traj, 1), "seconds": t_gen},
           "training": {"pairs": npairs, "held_out": held, "adam_steps": a.adam,
                        "lr": 1e-3, "batch": 2, "seconds": t_train,
                        "train_loss_first_last": [float(hist[0]), float(hist[-1])],
                        "held_out_loss": [before, after],
                        "held_out_rel_force_error": [float(np.sqrt(before)),
                                                     float(np.sqrt(after))]},
           "trajectories_per_form": a.traj, "forms": out}
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
