"""A production run of the U(4) 16^3 x 32 workload in one form: trajectories
from the benchmark's start (the thermalised 8^3 x 16 configuration tiled
2 x 2 x 2 x 2), with the benchmark's couplings and trajectory shape (length
0.25, nstep 5, nstep_gauge 10: MD step 0.05), and after every trajectory the
observables the reference's driver prints in its GMES/BACTION lines
(control.c:85-138): Polyakov loop, spatial and temporal plaquettes, bosonic
action per site.  Every form started with the same --seed draws the same
momenta and pseudofermions for its first trajectory.

    python -m susy_amd.production_run --form fp32_surrogate --traj 5 [--seed 1234]
        [--max-seconds S] [--nstep N] [--split-first J [--split-nmid M]]

--split-first J puts the poles J..14 of the rational action on a middle MD
level of M Omelyan steps per half outer step (the split form).

Prints one line per trajectory and a final JSON line (per trajectory: dH,
accept, CG iterations, seconds, observables; and the acceptance, mean
Metropolis probability min(1, exp(-dH)), <exp(-dH)> and seconds per
trajectory).
"""
import argparse
import json
import os
import time

import numpy as np

import susy_amd
from susy_amd import start_config

HERE = os.path.dirname(os.path.abspath(__file__))
DIMS = (16, 16, 16, 32)
FORMS = {"fp64": susy_amd.FORM_FP64, "fp32": susy_amd.FORM_FP32,
         "fp32_surrogate": susy_amd.FORM_FP32 | susy_amd.FORM_SURROGATE}
THETA = os.path.join(HERE, "configs", "surrogate_u4_theta.npy")


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--form", choices=sorted(FORMS), default="fp32_surrogate")
    ap.add_argument("--traj", type=int, default=5)
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--traj-length", type=float, default=0.25)
    ap.add_argument("--nstep", type=int, default=5)
    ap.add_argument("--max-seconds", type=float, default=None,
                    help="start no further trajectory after this many seconds")
    ap.add_argument("--split-first", type=int, default=0,
                    help="first pole of the middle level (0: no split)")
    ap.add_argument("--split-nmid", type=int, default=1)
    ap.add_argument("--theta", default=THETA, help="surrogate parameters (.npy)")
    a = ap.parse_args(argv)
    kw = dict(start_config.COUPLINGS, niter=20000, traj_length=a.traj_length, nstep=a.nstep)
    lat = susy_amd.Lattice(4, dims=DIMS, iseed=a.seed, **kw)
    lat.set_links(start_config.tiled_start(susy_amd, DIMS))
    form = FORMS[a.form]
    if "surrogate" in a.form:
        lat.sg_init(seed=1, scale=1.0)
        lat.sg_set_params(np.load(a.theta))
    if a.split_first:
        lat.set_split(a.split_first, a.split_nmid)
        form |= susy_amd.FORM_SPLIT
    per_traj = 2 * kw["Nroot"] * a.nstep
    rows = []
    t_start = time.perf_counter()
    for k in range(a.traj):
        if a.max_seconds is not None and time.perf_counter() - t_start > a.max_seconds:
            break
        t0 = time.perf_counter()
        # the trajectory in its pieces (the same operations as one update):
        # start, 2 halves x Nroot roots x nstep solves with their kicks, end;
        # a line after every solve (an FP64 solve takes 10-15 s here)
        lat.traj_begin(form)
        for j in range(per_traj):
            lat.traj_substep()
            print(f"{a.form} traj {k}: solve {j + 1}/{per_traj} "
                  f"{time.perf_counter() - t0:.1f} s", flush=True)
        st = lat.traj_end()
        secs = time.perf_counter() - t0
        re, im, mod = lat.wilson_line(3)
        ss, stp = lat.plaquette()
        rows.append({"dH": float(st[2]), "accept": int(st[3]), "cg_iters": int(st[4]),
                     "mid_iters": int(st[23]),
                     "s": secs, "ploop": [re, im, mod], "plaq": [ss, stp],
                     "sb_per_site": float(lat.actions()[0] / lat.V)})
        print(f"{a.form} traj {k}: dH {st[2]:.5g} accept {int(st[3])} iters {int(st[4])} "
              f"plaq {ss:.6f} {stp:.6f} sb/V {rows[-1]['sb_per_site']:.6f} "
              f"P {re:.5f} {im:.5f} {secs:.1f} s", flush=True)
    dH = np.array([r["dH"] for r in rows])
    print(json.dumps({
        "workload": "U(4) 16^3x32, lambda 2, bmass 0.4, G 0.05, kappa_u1 0, Nroot 2, "
                    f"traj length {a.traj_length}, nstep {a.nstep}, nstep_gauge 10",
        "form": a.form, "seed": a.seed,
        "split": {"first_pole": a.split_first, "nstep_mid": a.split_nmid}
        if a.split_first else None,
        "trajectories": rows,
        "acceptance": float(np.mean([r["accept"] for r in rows])) if rows else None,
        "metropolis_probability": float(np.mean(np.minimum(1.0, np.exp(-dH))))
        if rows else None,
        "exp_minus_dH": float(np.mean(np.exp(-dH))) if rows else None,
        "s_per_traj": float(np.mean([r["s"] for r in rows])) if rows else None}),
        flush=True)


if __name__ == "__main__":
    main()
