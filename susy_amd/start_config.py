"""Thermalised start configurations for the U(4) 16^3 x 32 workload.

A 16^3 x 32 lattice needs many trajectories from a cold start before its
actions, forces and solver costs settle, far more than a benchmark run can
spend.  The start is therefore made once on an 8^3 x 16 lattice with the same
couplings (lambda 2, bmass 0.4, G 0.05, kappa_u1 0, Nroot 2, Norder 15) by
this library's own RHMC, saved as a MILC gauge file
(configs/therm_u4_8x8x8x16.milc), and tiled 2 x 2 x 2 x 2 onto 16^3 x 32:
every site then sees the local environment of an equilibrated configuration,
and only modes longer than the small lattice start out of equilibrium.

The committed file was made in two stages, FP32 form throughout:

    python -m susy_amd.start_config --therm 1 --seed 2024 --save A.milc
    python -m susy_amd.start_config --load A.milc --seed 31 --therm 46 \
        --traj-length 0.25 --nstep 20 --save therm_u4_8x8x8x16.milc

The first unit trajectory from the cold links (nstep 20, dH -60.6) lands in a
rough configuration (plaquette 5.61) from which unit trajectories at nstep 20
have dH of +6 to +9 (a step-size error: it falls as 1/nstep^2, 6.66 / 1.64 /
0.40 at nstep 20 / 40 / 80 on identical momenta), so the second stage runs
at MD step 0.0125: 46 trajectories of length 0.25, all accepted, the
plaquette falling 5.47 -> 4.69 and flat (4.67-4.70) over the last 20, where
|dH| ~ 0.015.
"""
import argparse
import json
import os
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
START_FILE = os.path.join(HERE, "configs", "therm_u4_8x8x8x16.milc")
SMALL_DIMS = (8, 8, 8, 16)
COUPLINGS = dict(PBC=-1, Nroot=2, Norder=15, lam=2.0, kappa_u1=0.0, bmass=0.4, fmass=0.0,
                 G=0.05, niter=5000, error_per_site=1e-5, traj_length=1.0, nstep=20,
                 nstep_gauge=10)


def site_index(dims):
    """site index (even sites first, lexicographic within each parity) of
    every lexicographic position x + nx (y + ny (z + nz t))"""
    nx, ny, nz, nt = dims
    t, z, y, x = np.meshgrid(np.arange(nt), np.arange(nz), np.arange(ny), np.arange(nx),
                             indexing="ij")
    lex = (x + nx * (y + ny * (z + nz * t))).ravel()
    odd = ((x + y + z + t) & 1).ravel()
    V = nx * ny * nz * nt
    return np.where(odd == 1, (lex + V) // 2, lex // 2)


def tile(links, small, factors):
    """links of the lattice `small` (AoS, site order) repeated `factors`
    times in (x, y, z, t): AoS links of the larger lattice in its site order"""
    big = tuple(s * f for s, f in zip(small, factors))
    nx, ny, nz, nt = big
    t, z, y, x = np.meshgrid(np.arange(nt), np.arange(nz), np.arange(ny), np.arange(nx),
                             indexing="ij")
    sx, sy, sz, st = small
    small_lex = ((x % sx) + sx * ((y % sy) + sy * ((z % sz) + sz * (t % st)))).ravel()
    src = site_index(small)[small_lex]          # small-lattice site of each big lex position
    dst = site_index(big)                        # big-lattice site of each big lex position
    out = np.empty((len(dst),) + links.shape[1:], dtype=links.dtype)
    out[dst] = links[src]
    return out


def tiled_start(susy_amd, big_dims, path=START_FILE):
    """AoS links of `big_dims` tiled from the committed small configuration"""
    factors = tuple(b // s for b, s in zip(big_dims, SMALL_DIMS))
    if any(b % s for b, s in zip(big_dims, SMALL_DIMS)):
        raise ValueError(f"{big_dims} is not a multiple of {SMALL_DIMS}")
    small = susy_amd.Lattice(4, dims=SMALL_DIMS, iseed=1, **COUPLINGS)
    small.load_config(path)
    links = small.get_links()
    small.close()
    return tile(links, SMALL_DIMS, factors)


def main(argv=None):
    import susy_amd
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--therm", type=int, default=40)
    ap.add_argument("--dims", default=",".join(map(str, SMALL_DIMS)))
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--save", default=None)
    ap.add_argument("--max-seconds", type=float, default=None,
                    help="start no further trajectory after this many seconds")
    ap.add_argument("--load", default=None, help="continue from this MILC file")
    ap.add_argument("--traj-length", type=float, default=COUPLINGS["traj_length"])
    ap.add_argument("--nstep", type=int, default=COUPLINGS["nstep"])
    a = ap.parse_args(argv)
    dims = tuple(int(v) for v in a.dims.split(","))
    kw = dict(COUPLINGS, traj_length=a.traj_length, nstep=a.nstep)
    lat = susy_amd.Lattice(4, dims=dims, iseed=a.seed, **kw)
    if a.load:
        lat.load_config(a.load)
    rows = []
    t_start = time.perf_counter()
    for k in range(a.therm):
        if a.max_seconds is not None and time.perf_counter() - t_start > a.max_seconds:
            break
        t0 = time.perf_counter()
        st = lat.update(susy_amd.FORM_FP32)
        rows.append({"dH": float(st[2]), "accept": int(st[3]), "iters": int(st[4]),
                     "plaq": lat.plaquette()[0], "s": time.perf_counter() - t0})
        print(f"therm {k}: dH {st[2]:.5g} accept {int(st[3])} iters {int(st[4])} "
              f"plaq {rows[-1]['plaq']:.6f} {rows[-1]['s']:.1f} s", flush=True)
        if a.save:
            lat.save_config(a.save)
    print(json.dumps({"dims": dims, "seed": a.seed, "traj_length": a.traj_length,
                      "nstep": a.nstep, "trajectories": rows}), flush=True)


if __name__ == "__main__":
    main()
