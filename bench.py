"""Benchmark of the RHMC hot path on one MI355X.

Workload: U(4) on a 16^3 x 32 lattice, lambda = 2.0, bmass = 0.4, G = 0.05,
kappa_u1 = 0, fmass = 0, Nroot = 2, Norder = 15, antiperiodic temporal
fermion BC, CG stopping at |r|^2 <= 1e-10 |src|^2 (error_per_site 1e-5).
Synthetic data: cold start (unit links); momenta and pseudofermions from the
reference's generators seeded with 1234.

One step = one full RHMC trajectory: heatbath (two multi-shift solves), the
two-level Omelyan integrator with fermion forces (two multi-shift solves per
force) and nstep_gauge gauge steps inside each, and the Metropolis test on
the action recomputed with two more solves.  The default trajectory length
is 0.25 with nstep = 5 (MD step 0.05, the step of a unit-length trajectory
with nstep = 20): one unit-length trajectory on this lattice takes several
minutes of GPU time (a thermalising configuration needs 2000-7000 CG
iterations per solve), so `--traj-length 1 --nstep 20` is available but not
the default.

Forms: the headline FP32 form (FP32 operator and multi-shift CG with FP64
reliable updates, FP64 forces, actions and dH), the plain FP64 form and the
FP32 form with the learned surrogate force on the inner MD level, all from
the same links with the same generator seed (identical momenta and
pseudofermions), so their first trajectories differ only by arithmetic and
by the surrogate.  The surrogate is trained in the run itself: the plain
FP64 trajectory records (links, fermion force) at every outer kick, the
network is trained on them with the library's own forward and backward
kernels (Adam, `--sg-steps`), and its relative force error on the last two
recorded pairs (held out) is reported with the form.  The surrogate form is
timed but is not the headline `value`: it is reported beside it.

`--kernels` runs only a short operator workload (5 field copies for the
memory-counter calibration, then 10 launches of the operator in each
precision) for counter runs under rocprofv3.

Prints one JSON line.  Only one GPU is supported.
"""
import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

DIMS = (16, 16, 16, 32)
NCOL = 4
PEAK_HBM_GBS = 8000.0
WORKLOAD = "u4_16x16x16x32"
# counter summary of the operator kernel at this workload (HBM bytes per launch)
TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", f"k_dslash_fp32_{WORKLOAD}_counters.txt")


def workload(traj_length, nstep, nstep_gauge):
    return dict(dims=DIMS, PBC=-1, iseed=1234, Nroot=2, Norder=15, lam=2.0, kappa_u1=0.0,
                bmass=0.4, fmass=0.0, G=0.05, niter=5000, error_per_site=1e-5,
                traj_length=traj_length, nstep=nstep, nstep_gauge=nstep_gauge)


def dslash_bytes(V, nbytes_complex):
    """algorithmic bytes of one fermion_op: 16 source + 16 destination + 5
    link matrices of NCOL^2 complex numbers per site"""
    return 37 * NCOL * NCOL * nbytes_complex * V


def measured_traffic():
    """HBM bytes per launch of k_dslash<float> from the committed counter
    summary (FETCH_SIZE and WRITE_SIZE scaled by the copy-kernel factors)"""
    try:
        with open(TRAFFIC_SUMMARY) as f:
            m = re.search(r"^traffic_bytes_per_launch:\s*([0-9.eE+]+)", f.read(), re.M)
        return float(m.group(1)) if m else None
    except OSError:
        return None


def finite_or_none(x):
    x = float(x)
    return x if x == x and abs(x) < 1e300 else None


def run_form(lat, form, steps, start_links, prof_every):
    """`steps` trajectories from start_links with freshly seeded generators"""
    lat.set_links(start_links)
    lat.reseed(1234)
    lat.prof_enable(prof_every)
    t0 = time.perf_counter()
    stats = []
    for k in range(steps):
        stats.append(lat.update(form))
        print(f"bench: form {form} trajectory {k}: dH {stats[-1][2]:.4g} accept "
              f"{int(stats[-1][3])} iters {int(stats[-1][4])} "
              f"t {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    t1 = time.perf_counter()
    prec = form & 1
    avg_ms, launches, sampled = lat.prof_read(prec)
    lat.prof_enable(0)
    return stats, t1 - t0, (avg_ms, launches, sampled), lat.plaquette()


def cpu_baseline(fp64_stats, nforce_solves):
    """The C oracle (a restatement of the reference, single core) timed on a
    sub-volume of the same workload: per-site costs of one 15-shift CG
    iteration and one fermion-force pole on U(4) 8^3 x 16, scaled to
    16^3 x 32 and to the operation counts of the plain FP64 GPU trajectory."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import numpy as np

    from oracle_lib import Oracle, random_links, random_tf
    sub = (8, 8, 8, 16)
    kw = workload(1.0, 10, 10)
    kw["dims"] = sub
    O = Oracle(NCOL, **kw)
    O.set_links(random_links(O.link_shape(), 1, 0.1))
    src = random_tf(O.tf_shape(), 2)
    shifts = O.rhmc()[16:31]
    O.congrad_multi(src, shifts, 1, 1e-30)          # first touch of the buffers
    t = time.perf_counter()
    O.congrad_multi(src, shifts, 3, 1e-30)
    t_cg1 = (time.perf_counter() - t) / 3
    _, _, ps = O.congrad_multi(src, shifts[:1], 1, 1e-30)
    t = time.perf_counter()
    O.fermion_force(0.01, ps[:1], O.rhmc()[1:2])
    t_pole = time.perf_counter() - t
    scale = np.prod(DIMS) / np.prod(sub)
    iters = np.mean([s[4] for s in fp64_stats])
    t_traj = scale * (iters * t_cg1 + nforce_solves * 15 * t_pole)
    acc = np.mean([s[3] for s in fp64_stats])
    return {"value": 3600.0 / t_traj * acc, "unit": "accepted traj/hour", "cores": 1,
            "kind": "port",
            "sample": (f"C oracle, 1 core, U(4) 8^3x16 sub-volume: {t_cg1:.2f} s per 15-shift "
                       f"CG iteration, {t_pole:.2f} s per force pole; scaled x{scale:.0f} to "
                       f"16^3x32 and to the FP64 GPU trajectory's {iters:.0f} CG iterations "
                       f"and {nforce_solves}x15 force poles; acceptance {acc:.2f} from the GPU "
                       f"run (est. {t_traj:.0f} s per trajectory)")}


def kernels_only(lat):
    """short operator workload for counter runs"""
    import numpy as np
    rng = np.random.default_rng(5)
    shape = lat.link_shape()
    lat.set_links(np.eye(NCOL)[None, None] + 0.1 * (rng.standard_normal(shape)
                                                      + 1j * rng.standard_normal(shape)))
    h = lat.tf_alloc(2)
    lat.tf_upload(h, (rng.standard_normal(lat.tf_shape())
                      + 1j * rng.standard_normal(lat.tf_shape())) / np.sqrt(2))
    for _ in range(5):
        lat.stream_copy(h, h + 1)
    out = {}
    for name, form in (("fp64", 0), ("fp32", 1)):
        out[name + "_ms"] = lat.time_fermion_op(h, h + 1, form=form, reps=10)
    print(json.dumps({"kernels": out, "copy_bytes_read": 16 * NCOL * NCOL * 16 * lat.V,
                      "copy_bytes_written": 16 * NCOL * NCOL * 16 * lat.V}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1, help="timed FP32-form trajectories")
    ap.add_argument("--warmup", type=int, default=0, help="untimed FP32-form trajectories")
    ap.add_argument("--fp64-steps", type=int, default=1, help="timed plain FP64 trajectories")
    ap.add_argument("--traj-length", type=float, default=0.25)
    ap.add_argument("--nstep", type=int, default=5)
    ap.add_argument("--nstep-gauge", type=int, default=10)
    ap.add_argument("--prof-every", type=int, default=8,
                    help="time every n-th operator kernel launch with HIP events")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--kernels", action="store_true", help="operator launches only")
    ap.add_argument("--sg-steps", type=int, default=300,
                    help="Adam steps of the surrogate training (0: no surrogate form)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU decomposition is a design, not an implementation",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import susy_amd
    lat = susy_amd.Lattice(NCOL, **workload(args.traj_length, args.nstep, args.nstep_gauge))
    if args.kernels:
        kernels_only(lat)
        return
    susy_amd.load(NCOL).susy_set_verbose(1)   # one progress line per solve (stderr)
    V = lat.V
    start = lat.get_links()
    # warm-up trajectories move the start away from the cold configuration
    for _ in range(args.warmup):
        lat.update(susy_amd.FORM_FP32)
    if args.warmup:
        start = lat.get_links()

    # isolated operator timing (HIP events, 20 launches each)
    h = lat.tf_alloc(2)
    rng = np.random.default_rng(5)
    lat.tf_upload(h, (rng.standard_normal(lat.tf_shape())
                      + 1j * rng.standard_normal(lat.tf_shape())) / np.sqrt(2))
    iso = {nm: lat.time_fermion_op(h, h + 1, form=f, reps=20)
           for nm, f in (("fp64", 0), ("fp32", 1))}

    st32, t32, prof32, plaq32 = run_form(lat, susy_amd.FORM_FP32, args.steps, start,
                                         args.prof_every)
    if args.sg_steps > 0:
        # the plain run records the (links, fermion force) pairs
        lat.sg_init(seed=1, scale=1.0)
        lat.sg_record(4 * (args.nstep + 2) * max(args.fp64_steps, 1))
    st64, t64, prof64, plaq64 = run_form(lat, susy_amd.FORM_FP64, args.fp64_steps, start,
                                         args.prof_every)
    sg = None
    if args.sg_steps > 0:
        lat.sg_record(0)
        npairs = lat.sg_pairs()
        ntest = 2 if npairs > 4 else 0
        held = list(range(npairs - ntest, npairs))
        before = float(np.mean([lat.sg_pair_loss(i) for i in held])) if held else None
        t = time.perf_counter()
        hist = lat.sg_train(0, npairs - ntest, steps=args.sg_steps, batch=2, lr=1e-3, seed=3)
        t_train = time.perf_counter() - t
        after = float(np.mean([lat.sg_pair_loss(i) for i in held])) if held else None
        stsg, tsg, profsg, plaqsg = run_form(lat, susy_amd.FORM_FP32 | susy_amd.FORM_SURROGATE,
                                             args.steps, start, args.prof_every)
        sg = (stsg, tsg, profsg, plaqsg, {
            "pairs": npairs, "held_out": ntest, "train_steps": args.sg_steps,
            "train_s": t_train, "train_loss_first_last": [float(np.mean(hist[:10])),
                                                          float(np.mean(hist[-10:]))],
            "held_out_force_err2_untrained": before, "held_out_force_err2_trained": after})

    def summary(stats, secs, prof, plaq, nbytes):
        acc = sum(s[3] for s in stats)
        return {"traj_per_hour": 3600.0 * len(stats) / secs,
                "accepted_traj_per_hour": 3600.0 * acc / secs,
                "acceptance": acc / len(stats),
                "dH": [float(s[2]) for s in stats],
                # null when a cold-start transient (dH << 0) overflows exp(-dH)
                "exp_minus_dH": finite_or_none(np.mean([np.exp(min(-s[2], 700.0))
                                                        for s in stats])),
                "cg_iters": [int(s[4]) for s in stats],
                "s_per_traj": secs / len(stats),
                "plaquette_end": list(plaq),
                "dslash_ms": prof[0],
                "dslash_GBs": dslash_bytes(V, nbytes) / (prof[0] * 1e-3) / 1e9 if prof[0] else 0}

    f32 = summary(st32, t32, prof32, plaq32, 8)
    f64 = summary(st64, t64, prof64, plaq64, 16)
    forms = {"fp32": f32, "fp64": f64,
             "speedup_fp32_vs_fp64": f64["s_per_traj"] / f32["s_per_traj"],
             "dH_abs_diff_first_traj": abs(st32[0][2] - st64[0][2])}
    if sg is not None:
        fsg = summary(sg[0], sg[1], sg[2], sg[3], 8)
        fsg["training"] = sg[4]
        forms["fp32_surrogate"] = fsg
        forms["speedup_fp32_surrogate_vs_fp64"] = f64["s_per_traj"] / fsg["s_per_traj"]
    nforce_solves = args.nstep * 2 + 1
    bytes32 = dslash_bytes(V, 8)
    avg_ms = prof32[0]
    achieved = bytes32 / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    out = {
        "metric": "accepted traj/hour, U(4) 16^3x32: FP32+surrogate vs plain FP64; Dslash GB/s",
        "value": f32["accepted_traj_per_hour"],
        "unit": "accepted traj/hour",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * t32 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (cold start, generators seeded 1234)",
        "config": {"workload": "4d_Q16 RHMC U(4) 16^3x32, lambda 2.0, bmass 0.4, G 0.05, "
                               "kappa_u1 0, Nroot 2, Norder 15; FP32 form from a cold start",
                   "lattice": "16x16x16x32", "traj_length": args.traj_length,
                   "nstep": args.nstep, "nstep_gauge": args.nstep_gauge,
                   "form": "FP32 operator and CG with FP64 reliable updates; FP64 forces, "
                           "actions and dH",
                   "surrogate": "value is the FP32 form; the FP32+surrogate form (trained "
                                "in this run on the FP64 trajectory's force pairs) is "
                                "reported under forms.fp32_surrogate"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": achieved / PEAK_HBM_GBS,
                     "traffic": measured_traffic(),
                     "kernel": "k_dslash<float, false, 0>: D = fermion_op(sign +1), FP32",
                     "launches": prof32[1], "sampled": prof32[2], "avg_ms": avg_ms,
                     "algorithmic_bytes_per_launch": bytes32},
        "forms": forms,
        "dslash_isolated": {"fp64_ms": iso["fp64"],
                            "fp64_GBs": dslash_bytes(V, 16) / iso["fp64"] / 1e6,
                            "fp32_ms": iso["fp32"], "fp32_GBs": bytes32 / iso["fp32"] / 1e6},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(st64, 2 * nforce_solves)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
