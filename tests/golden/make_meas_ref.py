"""Write tests/golden/meas_ref.json: the measurements the reference prints
after every trajectory of its regression runs
4d_Q16/testsuite/scalar/hmc.U{2,3,4}.ref (control.c:85-120) -- the FLINK line
(Tr(U U^dag)/N per direction, average, width; printed %.6g) and the LINES
line (Wilson lines in the five link directions, Re and Im; %.6g).  The GMES
line is already in hmc_ref.json.

Run in the build container (it reads /root/reference):
    python tests/golden/make_meas_ref.py
"""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference/4d_Q16/testsuite/scalar"


def parse(path):
    trajs, cur = [], None
    for line in open(path):
        f = line.split()
        if not f:
            continue
        if f[0] in ("ACCEPT:", "REJECT:"):
            cur = {}
            trajs.append(cur)
        elif cur is not None and f[0] == "FLINK" and "flink" not in cur:
            cur["flink"] = [float(x) for x in f[1:]]
        elif cur is not None and f[0] == "LINES" and "lines" not in cur:
            cur["lines"] = [float(x) for x in f[1:]]
    return trajs


def main():
    data = {}
    for n in (2, 3, 4):
        data[f"U{n}"] = {"source": f"4d_Q16/testsuite/scalar/hmc.U{n}.ref",
                         "trajectories": parse(os.path.join(REF, f"hmc.U{n}.ref"))}
    with open(os.path.join(HERE, "meas_ref.json"), "w") as f:
        json.dump(data, f, indent=1)


if __name__ == "__main__":
    main()
