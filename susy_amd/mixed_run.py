"""The mixed-precision workload: U(4) 12^3 x 24, lambda 1, kappa_u1 0,
bmass 0.5, fmass 0, G 0.05, Nroot 2, Norder 15, trajectory length 1,
nstep 20, nstep_gauge 10, antiperiodic temporal fermion BCs, seed 7.

  1. thermalisation: `--therm` trajectories from a cold start (plain FP64 by
     default; `--therm-form fp32` runs the FP32 form, which samples the same
     distribution);
  2. on the thermalised links, with the FP32 form's reliable-update period
     set to `--reliable` (50): one pseudofermion source (heatbath, FP64),
     the 15-shift solve (errormin = error_per_site^2, as in a trajectory) in
     both forms, and the fermion force of each solution set;
  3. one trajectory in each form from the same links and generator seed
     (identical momenta and pseudofermions).

Prints one line per step and a final JSON line with the relative errors of
the FP32 form's solutions and force against the FP64 form's, both forms' dH,
plaquettes and iteration counts, and seconds per trajectory.

    python -m susy_amd.mixed_run [--therm 20] [--therm-form fp64] [--reliable 50]
        [--save PATH] [--load PATH] [--therm-only] [--dims 12,12,12,24]
        [--traj-length 1] [--nstep 20]
(--load PATH --therm K continues a thermalisation saved with --save.)
"""
import argparse
import json
import time

import numpy as np

import susy_amd

PARAMS = dict(dims=(12, 12, 12, 24), PBC=-1, iseed=7, Nroot=2, Norder=15, lam=1.0,
              kappa_u1=0.0, bmass=0.5, fmass=0.0, G=0.05, niter=5000, error_per_site=1e-5,
              traj_length=1.0, nstep=20, nstep_gauge=10)
FORMS = {"fp64": susy_amd.FORM_FP64, "fp32": susy_amd.FORM_FP32}


def say(*a):
    print(*a, flush=True)


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a - b)) / np.linalg.norm(np.ravel(b)))


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--therm", type=int, default=20)
    ap.add_argument("--therm-form", choices=sorted(FORMS), default="fp64")
    ap.add_argument("--reliable", type=int, default=50)
    ap.add_argument("--save", default=None, help="write the thermalised links (MILC file)")
    ap.add_argument("--load", default=None, help="start from this MILC file (then --therm more)")
    ap.add_argument("--therm-only", action="store_true", help="stop after thermalising (and --save)")
    ap.add_argument("--dims", default="12,12,12,24", help="lattice (nx,ny,nz,nt)")
    ap.add_argument("--traj-length", type=float, default=1.0)
    ap.add_argument("--nstep", type=int, default=20)
    a = ap.parse_args(argv)

    params = dict(PARAMS, dims=tuple(int(x) for x in a.dims.split(",")),
                  traj_length=a.traj_length, nstep=a.nstep)
    susy_amd.set_cg_options(4, reliable=a.reliable)
    lat = susy_amd.Lattice(4, **params)
    therm = []
    if a.load:
        lat.load_config(a.load)
        say("loaded", a.load, "plaquette", lat.plaquette()[0])
    for k in range(a.therm):
        t0 = time.perf_counter()
        st = lat.update(FORMS[a.therm_form])
        therm.append({"dH": float(st[2]), "accept": bool(st[3]), "iters": int(st[4]),
                      "plaq": lat.plaquette()[0], "s": time.perf_counter() - t0})
        say(f"therm {k} dH {st[2]:.4f} accept {int(st[3])} iters {int(st[4])} "
            f"plaq {therm[-1]['plaq']:.6f} {therm[-1]['s']:.1f} s")
    if a.save:
        lat.save_config(a.save)
    if a.therm_only:
        susy_amd.set_cg_options(4)
        return
    U0 = lat.get_links()
    rh = lat.rhmc()
    amps, shifts = rh[1:16], rh[16:31]
    errormin = params["error_per_site"] ** 2
    h = lat.tf_alloc(1 + 15 + 15)
    lat.reseed(11)
    lat.grsource(h, h + 1, susy_amd.FORM_FP64)
    solve = {}
    for k, (name, form) in enumerate(FORMS.items()):
        t0 = time.perf_counter()
        it, size_r = lat.congrad_multi(h, h + 1 + 15 * k, shifts, params["niter"], errormin,
                                       form=form)
        solve[name] = {"iters": it, "size_r": size_r, "s": time.perf_counter() - t0}
        say(f"solve {name} iters {it} size_r {size_r:.3e} {solve[name]['s']:.2f} s")
    sol_err = max(rel(lat.tf_download(h + 16 + j), lat.tf_download(h + 1 + j))
                  for j in range(15))
    zero = np.zeros(lat.link_shape(), dtype=complex)
    forces = []
    for k in range(2):
        lat.set_mom(zero)
        lat.fermion_force(1.0, h + 1 + 15 * k, amps)
        forces.append(lat.get_mom())
    force_err = rel(forces[1], forces[0])
    say(f"solution error {sol_err:.3e} force error {force_err:.3e}")
    traj = {}
    for name, form in FORMS.items():
        lat.set_links(U0)
        lat.reseed(13)
        t0 = time.perf_counter()
        st = lat.update(form)
        traj[name] = {"dH": float(st[2]), "accept": bool(st[3]), "iters": int(st[4]),
                      "plaquette": list(lat.plaquette()), "s": time.perf_counter() - t0}
        say(f"trajectory {name} dH {st[2]:.6f} accept {int(st[3])} iters {int(st[4])} "
            f"{traj[name]['s']:.1f} s")
    d64, d32 = traj["fp64"]["dH"], traj["fp32"]["dH"]
    susy_amd.set_cg_options(4)
    res = {"workload": f"mixed precision U(4) {a.dims}, lambda 1, bmass 0.5, G 0.05, Nroot 2, "
                       f"traj length {a.traj_length}, nstep {a.nstep}",
           "thermalisation": {"trajectories": len(therm), "form": a.therm_form,
                              "acceptance": float(np.mean([t["accept"] for t in therm]))
                              if therm else None,
                              "plaquette_last": therm[-1]["plaq"] if therm else None,
                              "s_per_traj": float(np.mean([t["s"] for t in therm]))
                              if therm else None},
           "reliable_period": a.reliable, "solve": solve,
           "solution_rel_err": sol_err, "force_rel_err": force_err,
           "trajectory": traj, "dH_abs_err": abs(d32 - d64),
           "plaquette_rel_err": abs(traj["fp32"]["plaquette"][0] /
                                    traj["fp64"]["plaquette"][0] - 1)}
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
