"""Write tests/golden/meas_ref.json: the measurements the reference prints
  * after every trajectory of its regression runs
    4d_Q16/testsuite/scalar/hmc.U{2,3,4}.ref (control.c:85-138): FLINK,
    LINES, LINES_POLAR, DET, WIDTHS, UUBAR_EIG, POLAR_EIG (all printed %.6g;
    the GMES line is already in hmc_ref.json);
  * for the loaded regression configuration config.U{2,3,4}.4444 in
    4d_Q16/testsuite/scalar/meas.U{2,3,4}.ref (control_meas.c:70-112): FLINK,
    FLINK_DET, LINES, LINES_POLAR -- the lines printed before that driver
    smears the links.

Run in the build container (it reads /root/reference):
    python tests/golden/make_meas_ref.py
"""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference/4d_Q16/testsuite/scalar"
TRAJ_LINES = ("FLINK", "LINES", "LINES_POLAR", "DET", "WIDTHS")
EIG_LINES = ("UUBAR_EIG", "POLAR_EIG")
CONFIG_LINES = ("FLINK", "FLINK_DET", "LINES", "LINES_POLAR")


def parse_hmc(path):
    trajs, cur = [], None
    for line in open(path):
        f = line.split()
        if not f:
            continue
        if f[0] in ("ACCEPT:", "REJECT:"):
            cur = {}
            trajs.append(cur)
        elif cur is None:
            continue
        elif f[0] in TRAJ_LINES and f[0].lower() not in cur:
            cur[f[0].lower()] = [float(x) for x in f[1:]]
        elif f[0] in EIG_LINES:
            cur.setdefault(f[0].lower(), []).append([float(x) for x in f[2:]])
    return trajs


def parse_meas(path):
    out = {}
    for line in open(path):
        f = line.split()
        if not f:
            continue
        if f[0].startswith("Doing"):      # the driver smears from here on
            break
        if f[0] in CONFIG_LINES and f[0].lower() not in out:
            out[f[0].lower()] = [float(x) for x in f[1:]]
    return out


def main():
    data = {}
    for n in (2, 3, 4):
        data[f"U{n}"] = {"source": f"4d_Q16/testsuite/scalar/hmc.U{n}.ref",
                         "trajectories": parse_hmc(os.path.join(REF, f"hmc.U{n}.ref")),
                         "config_source": f"4d_Q16/testsuite/scalar/meas.U{n}.ref",
                         "config": parse_meas(os.path.join(REF, f"meas.U{n}.ref"))}
    with open(os.path.join(HERE, "meas_ref.json"), "w") as f:
        json.dump(data, f, indent=1)


if __name__ == "__main__":
    main()
