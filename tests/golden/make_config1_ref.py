"""Write tests/golden/config1_ref.json: the kernel-level configuration --
U(3), 8^3 x 16, links I + 0.1 complex Gaussian (tests/oracle_lib.random_links,
seed 11), a unit-norm complex Gaussian Twist_Fermion source (random_tf, seed
12), the 15 shifts of the reference's Norder = 15 rational approximation of
x^(-1/4) on [1e-7, 1000] (setup_rhmc.c), solved by the C oracle's
congrad_multi (FP64) to relative residual 1e-8 (errormin 1e-16).  Per shift:
|psi_j|^2 and <src, psi_j>; the iteration count and final |r|^2.

    python tests/golden/make_config1_ref.py      (a few minutes)
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from oracle_lib import Oracle, random_links, random_tf  # noqa: E402

DIMS = (8, 8, 8, 16)
LINK_SEED, SRC_SEED, SCALE = 11, 12, 0.1
ERRORMIN = 1e-16


def inputs(shape_links, shape_tf):
    links = random_links(shape_links, LINK_SEED, SCALE)
    src = random_tf(shape_tf, SRC_SEED)
    return links, src / np.sqrt(np.vdot(src, src).real)


def main():
    O = Oracle(3, dims=DIMS)
    links, src = inputs(O.link_shape(), O.tf_shape())
    O.set_links(links)
    shifts = O.rhmc()[16:31]
    it, size_r, psi = O.congrad_multi(src, shifts, 5000, ERRORMIN)
    out = {"dims": DIMS, "link_seed": LINK_SEED, "src_seed": SRC_SEED, "scale": SCALE,
           "errormin": ERRORMIN, "shifts": [float(s) for s in shifts], "iters": int(it),
           "size_r": float(size_r),
           "psi_norm2": [float(np.vdot(p, p).real) for p in psi],
           "src_dot_psi": [[float(np.vdot(src, p).real), float(np.vdot(src, p).imag)]
                           for p in psi]}
    with open(os.path.join(HERE, "config1_ref.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("iterations", it, "size_r", size_r)


if __name__ == "__main__":
    main()
