"""Write tests/golden/config0_ref.json: ten plain FP64 RHMC trajectories of the
small CPU-runnable configuration -- U(2), 4^4, periodic spatial and
antiperiodic temporal fermion BCs, lambda 1.0, kappa_u1 0, bmass 0.5, fmass 0,
G 0.05, Nroot 1, Norder 15, trajectory length 1 with nstep 10 and
nstep_gauge 10, CG tolerance 1e-5, cold start, seed 1234 -- run by the C
oracle (oracle/susy_oracle.c; its trajectories reproduce the reference's
hmc.U*.ref regression runs, tests/test_oracle.py).  Per trajectory: the
update statistics (start/end action and its parts, dS, accept, CG
iterations, force monitors), the plaquettes, the bosonic action per site and
the Polyakov loop after it (Polyakov loop by the numpy checker
tests/measure_ref.py).

    python tests/golden/make_config0_ref.py      (about 90 s)
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import measure_ref  # noqa: E402
from oracle_lib import Oracle  # noqa: E402

KW = dict(dims=(4, 4, 4, 4), PBC=-1, iseed=1234, Nroot=1, Norder=15, lam=1.0, kappa_u1=0.0,
          bmass=0.5, fmass=0.0, G=0.05, niter=5000, error_per_site=1e-5, traj_length=1.0,
          nstep=10, nstep_gauge=10)
NTRAJ = 10


def trajectory_record(O, st):
    ss, stp = O.plaquette()
    geo = measure_ref.Geometry(O.dims)
    re_im, mod = measure_ref.wilson_line(geo, O.get_links(), 3)
    return {"stats": [float(x) for x in st[:23]], "plaq": [ss, stp],
            "sb_per_site": float(O.actions()[0] / O.V),
            "ploop": [float(re_im.real), float(re_im.imag), float(mod)]}


def main():
    O = Oracle(2, **KW)
    trajs = []
    for k in range(NTRAJ):
        trajs.append(trajectory_record(O, O.update()))
        print(k, trajs[-1]["stats"][:5], flush=True)
    with open(os.path.join(HERE, "config0_ref.json"), "w") as f:
        json.dump({"params": KW, "trajectories": trajs}, f, indent=1)


if __name__ == "__main__":
    main()
