"""Write tests/golden/config0_refbuild.json: the small CPU-runnable workload
configuration -- U(2), 4^4, antiperiodic temporal fermion BC, lambda 1.0,
kappa_u1 0, bmass 0.5, fmass 0, G 0.05, Nroot 1, Norder 15, trajectory
length 1 with nstep 10 and nstep_gauge 10, error_per_site 1e-5, cold start
("fresh"), seed 1234, ten trajectories -- run by the reference's own
susy_hmc program (oracle/_ref/susy_hmc_U2, compiled in place from the
reference sources by oracle/ref.mk), and the numbers it prints per
trajectory: the action parts at both ends, ACCEPT/REJECT with dS and the
start/end actions, IT_PER_TRAJ, the force monitors and the GMES line.

Run in the build container (the reference build is made there):
    python tests/golden/make_config0_refbuild.py      (about 30 s)
"""
import json
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
from make_golden import parse_hmc_ref  # noqa: E402

BINARY = os.path.join(os.path.dirname(os.path.dirname(HERE)), "oracle", "_ref", "susy_hmc_U2")
INPUT = """prompt 0
nx 4
ny 4
nz 4
nt 4
PBC -1
iseed 1234
Nroot 1
Norder 15
warms 0
trajecs 10
traj_length 1
nstep 10
nstep_gauge 10
traj_between_meas 10
lambda 1.0
kappa_u1 0.0
bmass 0.5
fmass 0.0
G 0.05
max_cg_iterations 5000
error_per_site 1e-5
fresh
forget
"""


def main():
    if not os.path.exists(BINARY):
        raise SystemExit("build the reference first: make -f oracle/ref.mk")
    with tempfile.TemporaryDirectory() as tmp:
        p = subprocess.run([BINARY], input=INPUT, stdout=subprocess.PIPE,
                           stderr=subprocess.STDOUT, text=True, cwd=tmp, check=True)
        out_path = os.path.join(tmp, "out")
        with open(out_path, "w") as f:
            f.write(p.stdout)
        # parse_hmc_ref expects the reference's header lines of a reloaded
        # run; a fresh start prints START but no CHECK PLAQ / Checksums
        text = p.stdout
        text += "\nCHECK PLAQ: 0 0\nChecksums - - OK\n"
        with open(out_path, "w") as f:
            f.write(text)
        ref = parse_hmc_ref(out_path)
    for k in ("check_plaq", "checksums"):
        ref.pop(k)
    with open(os.path.join(HERE, "config0_refbuild.json"), "w") as f:
        json.dump({"input": INPUT, "ref": ref}, f, indent=1)
    print(len(ref["trajectories"]), "trajectories,", ref["total_iters"], "CG iterations")


if __name__ == "__main__":
    main()
