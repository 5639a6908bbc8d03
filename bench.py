"""Benchmark of the RHMC hot path on one MI355X.

Workload: U(4) on a 16^3 x 32 lattice, lambda = 2.0, bmass = 0.4, G = 0.05,
kappa_u1 = 0, fmass = 0, Nroot = 2, Norder = 15, antiperiodic temporal
fermion BC, CG stopping at |r|^2 <= 1e-10 |src|^2 (error_per_site 1e-5),
MD step 0.05 on the fermion level (the step of the unit-length, nstep-20
trajectories of the 12^3 x 24 configuration) with 10 gauge steps inside each
half step.  Trajectories are 0.25 long (nstep 5): at this volume a
multi-shift solve takes 4000-6000 CG iterations (6-9 s in FP32), so a unit
trajectory (84 solves) takes 10-15 minutes and cannot end inside a
benchmark run; a 0.25 trajectory (24 solves) can.  Every per-hour figure is
counted in unit-length equivalents (a trajectory of length 0.25 counts 0.25),
i.e. it is MD time per hour, which does not depend on the trajectory length
chosen.

Start: a thermalised configuration.  A 16^3 x 32 lattice needs many
trajectories from a cold start before its actions and solver costs settle,
far more than a benchmark run can spend, so the start is an 8^3 x 16
configuration with the same couplings thermalised by this library's own RHMC
(committed as susy_amd/configs/therm_u4_8x8x8x16.milc with its history in the
.info file beside it; made by `python -m susy_amd.start_config`) tiled
2 x 2 x 2 x 2 onto 16^3 x 32.
Momenta and pseudofermions come from the reference's generators seeded with
1234.  Every form starts from the same links with the same seed, so their
first trajectories have identical momenta and pseudofermions.

One step = one multi-shift solve of one root and its fermion-force kick (the
first of each half outer step also runs the gauge level before it): the
unit of the hot path, congrad_multi + fermion_force (update_o.c:80-113).  A
trajectory is its start (heatbath: 2 x Nroot solves; first kick), nstep x
2 x Nroot = 20 steps, and its end (action, Metropolis test), so `--steps 20`
times exactly one trajectory's work: the timed steps run on from the warm-up
steps of the first trajectory into the next, and the end of one trajectory
and the start of the next fall inside the timed region.

Forms:
  * FP32 + surrogate + pole split (headline `value`): FP32 operator and
    multi-shift CG with FP64 reliable updates, FP64 forces, actions and dH,
    the learned surrogate force on the inner (gauge) MD level, and the
    rational action's large-shift poles (9..14, shifts >= 0.06) on a middle
    MD level (two Omelyan steps per half outer step, each kick a multi-shift
    solve over those six shifts only: ~80 CG iterations) so that the outer
    level, which keeps the small-shift poles, takes 2 steps per trajectory
    where the plain forms take 5; `--warmup` untimed steps, then `--steps`
    timed steps, trajectories completed and Metropolis-tested as they end.
    `--split-first 0` runs the FP32 + surrogate form without the split at
    the plain forms' nstep.
  * FP32 (no surrogate) and plain FP64, nstep 5: the start of one trajectory
    and `--fp32-steps` / `--fp64-steps` timed steps (default 2: both roots'
    solves of the first half step) from the same links and seed; their
    trajectory time is start + 2 Nroot nstep x the mean step time.
The surrogate's parameters come from `python -m susy_amd.surrogate_train`:
the library's own forward and backward kernels trained them (Adam) on
(links, outer-level fermion force) pairs of the same form's RHMC on the
8^3 x 16 lattice of the same couplings (susy_amd/configs/
surrogate_u4_split_theta.npy with the split, whose outer level keeps only the
small-shift poles; surrogate_u4_theta.npy without) -- the network is local,
so its parameters carry over to 16^3 x 32.  The run reports their force
error on the pairs its own warm-up steps record at this size (pairs they
were not trained on).

value = accepted trajectories per hour of the headline form, in unit-length
equivalents = 3600 x acceptance x traj_length x (timed steps / steps per
trajectory) / timed seconds.  The acceptance is the mean Metropolis
probability min(1, exp(-dH)) of the trajectories that ended in the run (its
expectation is the acceptance rate, with far less variance than the 0/1
outcomes of the one or two trajectories a run of this length ends; the
outcomes and dH are reported beside it).  The plain FP64 and FP32 forms
integrate the same Hamiltonian with the same integrator (their dH on
identical inputs agree to ~1e-4), so they share an acceptance rate; the
split's 2 outer steps were chosen for an energy error like the plain
integrator's at 5 (16^3 x 32, 6 trajectories each: <|dH|> 0.45 against
0.48), so all forms are compared by trajectories per hour, and the ratios to
the plain FP64 form are reported.
The CPU baseline prices a trajectory of the reference (plain algorithm) at
the FP64 leg's CG iterations.

`--kernels` runs only a short operator workload (5 field copies for the
memory-counter calibration, then 10 launches of the operator in each
precision) for counter runs under rocprofv3.

Prints one JSON line (progress lines go to stderr).  Only one GPU is
supported.
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
SEED = 1234
# surrogate parameters trained by susy_amd/surrogate_train.py
SG_THETA = os.path.join(ROOT, "susy_amd", "configs", "surrogate_u4_theta.npy")
# ... and on the split form's outer forces (the headline form with the split)
SG_THETA_SPLIT = os.path.join(ROOT, "susy_amd", "configs", "surrogate_u4_split_theta.npy")
# headline form's pole split: poles SPLIT_FIRST..14 (shifts >= 0.06) on a
# middle level of SPLIT_NMID Omelyan steps per half outer step, SPLIT_NSTEP
# outer steps per trajectory of the same length
SPLIT_FIRST, SPLIT_NMID, SPLIT_NSTEP = 9, 2, 2


def workload(traj_length, nstep, nstep_gauge):
    return dict(dims=DIMS, PBC=-1, iseed=SEED, Nroot=2, Norder=15, lam=2.0, kappa_u1=0.0,
                bmass=0.4, fmass=0.0, G=0.05, niter=20000, error_per_site=1e-5,
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


def say(msg):
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def finite_or_none(x):
    x = float(x)
    return x if x == x and abs(x) < 1e300 else None


def sync_time():
    return time.perf_counter()


def partial_form(lat, susy_amd, form, name, start, nsteps, prof_every):
    """start of one trajectory and nsteps timed outer steps"""
    lat.set_links(start)
    lat.reseed(SEED)
    lat.prof_enable(prof_every)
    t0 = time.perf_counter()
    lat.traj_begin(form)
    t_begin = time.perf_counter() - t0
    it_begin = lat.traj_iters()[0]
    say(f"{name}: start of trajectory {t_begin:.2f} s")
    steps = []
    for k in range(nsteps):
        t = time.perf_counter()
        lat.traj_substep()
        steps.append(time.perf_counter() - t)
        say(f"{name}: step {k + 1} {steps[-1]:.2f} s")
    it_steps = lat.traj_iters()[0] - it_begin
    prof = lat.prof_read(form & 1)
    lat.prof_enable(0)
    return {"s_begin": t_begin, "s_steps": steps, "dslash_ms": prof[0],
            "cg_iters_begin": it_begin, "cg_iters_per_step": it_steps / max(nsteps, 1)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20,
                    help="timed steps (one root's solve + kick) of the FP32+surrogate form")
    ap.add_argument("--warmup", type=int, default=5, help="untimed outer steps before them")
    ap.add_argument("--fp32-steps", type=int, default=2,
                    help="timed steps, FP32 form (2: both roots' solves of a half step)")
    ap.add_argument("--fp64-steps", type=int, default=2,
                    help="timed steps, plain FP64 form (2: both roots' solves of a half step)")
    ap.add_argument("--traj-length", type=float, default=0.25)
    ap.add_argument("--nstep", type=int, default=5)
    ap.add_argument("--nstep-gauge", type=int, default=10)
    ap.add_argument("--prof-every", type=int, default=8,
                    help="time every n-th operator kernel launch with HIP events")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--kernels", action="store_true", help="operator launches only")
    ap.add_argument("--split-first", type=int, default=SPLIT_FIRST,
                    help="headline form: first pole of the middle MD level (0: no split)")
    ap.add_argument("--split-nmid", type=int, default=SPLIT_NMID,
                    help="headline form: middle Omelyan steps per half outer step")
    ap.add_argument("--split-nstep", type=int, default=SPLIT_NSTEP,
                    help="headline form: outer steps per trajectory with the split")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU decomposition is a design, not an implementation",
              file=sys.stderr)
        sys.exit(2)
    if args.steps < 1 or args.warmup < 0:
        sys.exit("bench.py: --steps must be positive and --warmup non-negative")

    import numpy as np

    import susy_amd
    t_job = time.perf_counter()
    lat = susy_amd.Lattice(NCOL, **workload(args.traj_length, args.nstep, args.nstep_gauge))
    if args.kernels:
        kernels_only(lat)
        return
    susy_amd.load(NCOL).susy_set_verbose(1)   # one progress line per solve (stderr)
    V, nstep = lat.V, args.nstep
    split = args.split_first > 0
    h_nstep = args.split_nstep if split else nstep   # the headline form's outer steps
    per_traj = 2 * 2 * nstep          # steps per trajectory: 2 halves x Nroot roots x nstep
    h_per_traj = 2 * 2 * h_nstep
    ell = args.traj_length

    # ---- start configuration: thermalised 8^3 x 16 links tiled onto 16^3 x 32
    from susy_amd import start_config
    lat.set_links(start_config.tiled_start(susy_amd, DIMS))
    start = lat.get_links()
    plaq_start = lat.plaquette()
    say(f"start: tiled thermalised configuration, plaquette {plaq_start[0]:.6f}")

    # ---- isolated operator timing (HIP events, 20 launches each)
    h = lat.tf_alloc(2)
    rng = np.random.default_rng(5)
    lat.tf_upload(h, (rng.standard_normal(lat.tf_shape())
                      + 1j * rng.standard_normal(lat.tf_shape())) / np.sqrt(2))
    iso = {nm: lat.time_fermion_op(h, h + 1, form=f, reps=20)
           for nm, f in (("fp64", 0), ("fp32", 1))}
    say(f"isolated D: FP64 {iso['fp64']:.3f} ms, FP32 {iso['fp32']:.3f} ms")

    # ---- FP32 and FP64 forms: start + a few steps each
    f32 = partial_form(lat, susy_amd, susy_amd.FORM_FP32, "fp32", start, args.fp32_steps,
                       args.prof_every)
    f64 = partial_form(lat, susy_amd, susy_amd.FORM_FP64, "fp64", start, args.fp64_steps,
                       args.prof_every)

    # ---- headline: FP32 + surrogate (+ pole split), warmup steps then timed
    # steps; with the split, a context of its own outer step count.  The
    # surrogate's parameters were trained offline on the 8^3 x 16 chain of the
    # same couplings and form (the outer level's forces); the warm-up steps
    # record this lattice's (links, outer force) pairs, on which its force
    # error is reported.
    form = susy_amd.FORM_FP32 | susy_amd.FORM_SURROGATE
    sg_theta = SG_THETA_SPLIT if split else SG_THETA
    if split:
        lat.close()
        lat = susy_amd.Lattice(NCOL, **workload(args.traj_length, h_nstep, args.nstep_gauge))
        lat.set_split(args.split_first, args.split_nmid)
        form |= susy_amd.FORM_SPLIT
    lat.sg_init(seed=1, scale=1.0)
    lat.sg_set_params(np.load(sg_theta))
    lat.set_links(start)
    lat.reseed(SEED)
    trajs = []
    done = [0]   # steps of the open trajectory
    h_name = "fp32+surrogate+split" if split else "fp32+surrogate"

    def one_step():
        if done[0] == 0:
            lat.traj_begin(form)
        lat.traj_substep()
        done[0] += 1
        if done[0] == h_per_traj:
            st = lat.traj_end()
            plaq = lat.plaquette()
            trajs.append({"dH": float(st[2]), "accept": int(st[3]), "cg_iters": int(st[4]),
                          "middle_cg_iters": int(st[23]), "plaquette": list(plaq)})
            say(f"{h_name}: trajectory {len(trajs)} dH {st[2]:.5g} accept {int(st[3])} "
                f"iters {int(st[4])} (middle {int(st[23])}) plaq {plaq[0]:.6f}")
            done[0] = 0

    lat.sg_record(max(args.warmup, 1))
    for k in range(args.warmup):
        one_step()
        say(f"{h_name}: warm-up step {k + 1}")
    lat.sg_record(0)
    lat.prof_enable(args.prof_every)
    ntraj0 = len(trajs)
    t0 = sync_time()
    for k in range(args.steps):
        one_step()
        say(f"{h_name}: timed step {k + 1} t {sync_time() - t0:.1f} s")
    t_timed = sync_time() - t0
    prof = lat.prof_read(1)
    lat.prof_enable(0)
    err_pairs = [lat.sg_pair_loss(i) for i in range(lat.sg_pairs())]
    say(f"surrogate: relative force error^2 on the {len(err_pairs)} pairs of the warm-up "
        f"steps {float(np.mean(err_pairs)) if err_pairs else float('nan'):.4f}")
    ended = trajs[ntraj0:]
    boundaries = len(ended)

    # ---- figures
    traj_per_hour_sg = 3600.0 * ell * (args.steps / h_per_traj) / t_timed
    # acceptance: mean Metropolis probability of the ended trajectories
    acc = float(np.mean([min(1.0, np.exp(min(-tr["dH"], 700.0))) for tr in trajs])) \
        if trajs else None
    value = traj_per_hour_sg * acc if acc is not None else 0.0

    def per_hour(f):
        s_traj = f["s_begin"] + per_traj * float(np.mean(f["s_steps"]))
        return s_traj, 3600.0 * ell / s_traj

    s32, tph32 = per_hour(f32)
    s64, tph64 = per_hour(f64)
    bytes32 = dslash_bytes(V, 8)
    avg_ms = prof[0]
    achieved = bytes32 / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    h_key = "fp32_surrogate_split" if split else "fp32_surrogate"
    forms = {
        h_key: {
            "traj_per_hour": traj_per_hour_sg, "accepted_traj_per_hour": value,
            "acceptance": acc,
            "acceptance_estimator": "mean of min(1, exp(-dH)) over the ended trajectories",
            "accepted_outcomes": [tr["accept"] for tr in trajs],
            "trajectories_ended": len(trajs),
            "trajectories": trajs,
            "exp_minus_dH": finite_or_none(np.mean([np.exp(min(-tr["dH"], 700.0))
                                                    for tr in trajs])) if trajs else None,
            "s_per_traj": t_timed * h_per_traj / args.steps,
            "nstep": h_nstep,
            "split": {"first_pole": args.split_first, "nstep_mid": args.split_nmid} if split else None,
            "dslash_ms": avg_ms,
            "surrogate": {"parameters": os.path.relpath(sg_theta, ROOT),
                          "trained": "offline, susy_amd/surrogate_train.py on the 8^3x16 "
                                     "chain of the same couplings and form",
                          "force_err2_on_warmup_pairs": [float(x) for x in err_pairs]}},
        "fp32": {"traj_per_hour": tph32, "s_per_traj": s32, "s_begin": f32["s_begin"],
                 "s_steps": f32["s_steps"], "dslash_ms": f32["dslash_ms"], "nstep": nstep},
        "fp64": {"traj_per_hour": tph64, "s_per_traj": s64, "s_begin": f64["s_begin"],
                 "s_steps": f64["s_steps"], "dslash_ms": f64["dslash_ms"], "nstep": nstep,
                 "cg_iters_begin": f64["cg_iters_begin"],
                 "cg_iters_per_step": f64["cg_iters_per_step"]},
    }
    forms[f"traj_per_hour_ratio_{h_key}_vs_fp64"] = traj_per_hour_sg / tph64
    forms["traj_per_hour_ratio_fp32_vs_fp64"] = tph32 / tph64
    out = {
        "metric": "accepted traj/hour, U(4) 16^3x32: FP32+surrogate vs plain FP64; Dslash GB/s",
        "value": value,
        "unit": "accepted traj/hour",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * t_timed / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (thermalised 8^3x16 links of this library's RHMC "
                                "tiled to 16^3x32; generators seeded 1234)",
        "config": {"workload": "4d_Q16 RHMC U(4) 16^3x32, lambda 2.0, bmass 0.4, G 0.05, "
                               "kappa_u1 0, Nroot 2, Norder 15; MD step "
                               f"{ell / h_nstep:g}, trajectory length {ell}, nstep {h_nstep}, "
                               f"nstep_gauge {args.nstep_gauge}; {h_name} form"
                               + (f" (poles {args.split_first}-14 on a middle level of "
                                  f"{args.split_nmid} step(s); the plain FP32 and FP64 forms "
                                  f"at nstep {nstep})" if split else "")
                               + "; trajectories counted in unit-length equivalents",
                   "lattice": "16x16x16x32", "traj_length": args.traj_length,
                   "nstep": h_nstep, "nstep_gauge": args.nstep_gauge,
                   "step": "one root's multi-shift solve + fermion-force kick (with the "
                           + ("middle and gauge levels" if split else "gauge level")
                           + " before the first of each half outer step); "
                           f"{h_per_traj} steps = one trajectory, whose start and end fall "
                           f"inside the timed steps ({boundaries} in this run)",
                   "per_hour_unit": "unit-length equivalents (MD time units) per hour",
                   "start": "thermalised U(4) 8^3x16 configuration (same couplings) tiled "
                            f"2x2x2x2 (plaquette {plaq_start[0]:.5f})",
                   "form": "FP32 operator and CG with FP64 reliable updates; FP64 forces, "
                           "actions and dH; learned surrogate force on the inner MD level"
                           + ("; rational-action poles with shifts >= "
                              f"{lat.rhmc()[16 + args.split_first]:.3g} on a middle MD level"
                              if split else "")},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": achieved / PEAK_HBM_GBS,
                     "traffic": measured_traffic(),
                     "kernel": "k_dslash<float, false, 0>: D = fermion_op(sign +1), FP32",
                     "launches": prof[1], "sampled": prof[2], "avg_ms": avg_ms,
                     "algorithmic_bytes_per_launch": bytes32},
        "forms": forms,
        "dslash_isolated": {"fp64_ms": iso["fp64"],
                            "fp64_GBs": dslash_bytes(V, 16) / iso["fp64"] / 1e6,
                            "fp32_ms": iso["fp32"], "fp32_GBs": bytes32 / iso["fp32"] / 1e6},
    }
    if not args.no_cpu and trajs:
        # the reference runs the plain algorithm: a trajectory is priced at the
        # FP64 leg's CG iterations (start + per_traj mean steps) and the plain
        # integrator's force poles
        it64 = int(round(f64["cg_iters_begin"] + per_traj * f64["cg_iters_per_step"]))
        out["cpu_baseline"] = cpu_baseline(it64, (2 * nstep + 1) * 2 * 15, acc, ell)
    out["job_s"] = time.perf_counter() - t_job
    print(json.dumps(out), flush=True)


REF_BENCH = os.path.join(ROOT, "oracle", "_ref", "susy_refbench_U4")
MPIEXEC = "/opt/conda/bin/mpiexec"


def cpu_baseline(iters_per_traj, force_poles, acceptance, ell):
    """The reference itself (its U(4) MPI build, oracle/_ref/susy_refbench_U4)
    on the host's cores, timed on the full 16^3 x 32 lattice: one fermion_op,
    3 iterations of the 15-shift congrad_multi and a 2-pole fermion_force;
    a trajectory is priced at the CG iteration count and force poles of the
    plain algorithm (which the reference runs), from the FP64 leg.  Without that build (or without MPI), the C oracle on one core on
    an 8^3 x 16 sub-volume, scaled by volume."""
    import subprocess
    ranks = min(16, os.cpu_count() or 1)
    if os.path.exists(REF_BENCH) and os.path.exists(MPIEXEC):
        cmd = [MPIEXEC, "-n", str(ranks), REF_BENCH, *map(str, DIMS), "1", "3", "2"]
        say("cpu baseline: " + " ".join(cmd[2:]))
        try:
            p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                               timeout=240, cwd="/tmp")
            line = [ln for ln in p.stdout.splitlines() if ln.startswith('{"ranks"')]
            if p.returncode == 0 and line:
                r = json.loads(line[-1])
                t_traj = iters_per_traj * r["cg_iteration_s"] + force_poles * r["force_pole_s"]
                return {"value": 3600.0 * ell / t_traj * acceptance,
                        "unit": "accepted traj/hour (unit-length equivalents)",
                        "cores": r["ranks"], "kind": "reference",
                        "sample": (f"reference U(4) MPI build on {r['ranks']} ranks, 16^3x32: "
                                   f"fermion_op {r['fermion_op_s']:.3f} s, 15-shift CG "
                                   f"iteration {r['cg_iteration_s']:.3f} s (3 timed), force "
                                   f"pole {r['force_pole_s']:.3f} s (2 timed); priced at "
                                   f"{iters_per_traj} CG iterations and {force_poles} force "
                                   f"poles per trajectory (the plain algorithm's counts: the "
                                   f"FP64 leg's start + 20 mean steps), "
                                   f"acceptance {acceptance:.2f} from the GPU run: est. "
                                   f"{t_traj:.0f} s per trajectory"),
                        "fermion_op_GBs": dslash_bytes(DIMS[0] * DIMS[1] * DIMS[2] * DIMS[3], 16) / r["fermion_op_s"] / 1e9}
            say("cpu baseline: reference run failed, using the oracle\n" + p.stdout[-2000:])
        except (OSError, subprocess.SubprocessError, ValueError, KeyError) as e:
            say(f"cpu baseline: reference run failed ({e}), using the oracle")
    return cpu_baseline_port(iters_per_traj, force_poles, acceptance, ell)


def cpu_baseline_port(iters_per_traj, force_poles, acceptance, ell):
    """The C oracle (a restatement of the reference, single core) timed on a
    sub-volume of the same workload: per-site costs of one 15-shift CG
    iteration and one fermion-force pole on U(4) 8^3 x 16, scaled to
    16^3 x 32 and to one trajectory's CG iterations and force poles."""
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
    scale = float(np.prod(DIMS) / np.prod(sub))
    t_traj = scale * (iters_per_traj * t_cg1 + force_poles * t_pole)
    return {"value": 3600.0 * ell / t_traj * acceptance,
            "unit": "accepted traj/hour (unit-length equivalents)", "cores": 1,
            "kind": "port",
            "sample": (f"C oracle, 1 core, U(4) 8^3x16 sub-volume: {t_cg1:.2f} s per 15-shift "
                       f"CG iteration, {t_pole:.2f} s per force pole; scaled x{scale:.0f} to "
                       f"16^3x32 and to {iters_per_traj} CG iterations and {force_poles} "
                       f"force poles per trajectory, acceptance {acceptance:.2f} from the GPU "
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


if __name__ == "__main__":
    main()
