"""Write tests/golden/hmc_meas_ref.json: the U(5) and U(7) regression runs of
the reference, 4d_Q16/testsuite/scalar/hmc_meas.U{5,7}.ref (input
testsuite/in.U{5,7}.hmc_meas: 1x1x1x4, fresh unit links, nstep 15,
nstep_gauge 4, lambda 1, kappa_u1 1, bmass 1, G 0.1).

Only the trajectories printed before the first measurement block are kept
(the first two): that block (control.c, traj_between_meas 2) draws random
sources for the Ward-identity measurement, so the generator streams of the
third trajectory depend on code outside the RHMC path.  Per trajectory: the
action lines, ACCEPT/REJECT with dS and S, IT_PER_TRAJ, the force monitors,
GMES, FLINK, LINES, LINES_POLAR, DET, WIDTHS, UUBAR_EIG, POLAR_EIG.

Run in the build container (it reads /root/reference):
    python tests/golden/make_hmc_meas_ref.py
"""
import json
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference/4d_Q16/testsuite"
MEAS = ("FLINK", "LINES", "LINES_POLAR", "DET", "WIDTHS")
EIG = ("UUBAR_EIG", "POLAR_EIG")


def parse_input(path):
    vals = {}
    for line in open(path):
        parts = line.split()
        if len(parts) >= 2:
            vals[parts[0]] = parts[1]
    return vals


def parse(path):
    text = open(path).read().split("Doing ")[0]   # up to the first measurement block
    out = {"trajectories": []}
    m = re.search(r"^START (\S+) (\S+) (\S+) (\S+)", text, re.M)
    out["start"] = [float(x) for x in m.groups()]
    m = re.search(r"CHECK PLAQ: (\S+) (\S+)", text)
    out["check_plaq"] = [float(m.group(1)), float(m.group(2))]
    lines = text.split("\n")
    acc_rows = [i for i, l in enumerate(lines) if re.match(r"(ACCEPT|REJECT):", l)]
    for k, i in enumerate(acc_rows):
        # the two action lines print right before ACCEPT/REJECT; the
        # measurements follow it, up to the next trajectory's action lines
        stop = acc_rows[k + 1] - 2 if k + 1 < len(acc_rows) else len(lines)
        b = "\n".join(lines[i - 2:stop])
        act = re.findall(r"^action: gauge (\S+) bmass (\S+) det (\S+) fermion0 (\S+) mom (\S+) "
                         r"sum (\S+)", b, re.M)
        acc = re.search(r"^(ACCEPT|REJECT): delta S = (\S+) start S = (\S+) end S = (\S+)", b,
                        re.M)
        tr = {"start_action": [float(x) for x in act[0]],
              "end_action": [float(x) for x in act[1]],
              "accept": acc.group(1) == "ACCEPT", "delta_S": float(acc.group(2)),
              "start_S": float(acc.group(3)), "end_S": float(acc.group(4)),
              "iters": int(re.search(r"^IT_PER_TRAJ (\d+)", b, re.M).group(1)),
              "gauge_force": [float(x) for x in
                              re.search(r"^MONITOR_FORCE_GAUGE\s+(\S+) (\S+)", b, re.M).groups()],
              "fermion_force": [float(x) for x in
                                re.search(r"^MONITOR_FORCE_FERMION0 (\S+) (\S+)", b,
                                          re.M).groups()],
              "gmes": [float(x) for x in
                       re.search(r"^GMES (\S+) (\S+) (\d+) (\S+) (\S+) (\S+) (\S+)", b,
                                 re.M).groups()]}
        for name in MEAS:
            m = re.search(rf"^{name}\s+(.*)$", b, re.M)
            tr[name.lower()] = [float(x) for x in m.group(1).split()]
        for name in EIG:
            tr[name.lower()] = [[float(x) for x in row.split()[1:]]
                                for row in re.findall(rf"^{name} (.*)$", b, re.M)]
        out["trajectories"].append(tr)
    return out


def main():
    data = {}
    for n in (5, 7):
        data[f"U{n}"] = {"source": f"4d_Q16/testsuite/scalar/hmc_meas.U{n}.ref",
                         "input": parse_input(os.path.join(REF, f"in.U{n}.hmc_meas")),
                         "ref": parse(os.path.join(REF, "scalar", f"hmc_meas.U{n}.ref"))}
        assert len(data[f"U{n}"]["ref"]["trajectories"]) == 2, n
    path = os.path.join(HERE, "hmc_meas_ref.json")
    with open(path, "w") as f:
        json.dump(data, f, indent=1)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
