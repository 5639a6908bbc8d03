"""Write the golden fixtures under tests/golden/ from the reference tree.

Run in the build container (it reads /root/reference, which the GPU box does
not have):  python tests/golden/make_golden.py

Outputs (data only):
  configs/config.U{2,3,4}.{2222,4444}   the reference's regression gauge files
                                        (4d_Q16/testsuite/), copied unchanged
  hmc_ref.json      numbers printed by testsuite/scalar/hmc.U{2,3,4}.ref and the
                    parameters of testsuite/in.U{2,3,4}.hmc
  ref_U2.npz        outputs of the reference build (oracle/_ref) on
                    config.U2.4444 with seeded inputs: gather tables (also on
                    an irregular 2x4x6x8 lattice), boundary-sign tables,
                    fermion_op(+1/-1), DSq, a 15-shift solve, gauge and fermion
                    force increments, generator streams.
"""
import json
import os
import re
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from oracle_lib import Ref, have_ref, random_tf  # noqa: E402

REF = "/root/reference/4d_Q16/testsuite"
SITES = np.r_[0:64, 128:192]     # of the 256 sites of a 4^4 lattice


def copy_configs():
    os.makedirs(os.path.join(HERE, "configs"), exist_ok=True)
    for n in (2, 3, 4):
        for d in ("2222", "4444"):
            for suffix in ("", ".info"):
                src = os.path.join(REF, f"config.U{n}.{d}{suffix}")
                shutil.copy(src, os.path.join(HERE, "configs", os.path.basename(src)))


def parse_input(path):
    vals = {}
    for line in open(path):
        parts = line.split()
        if len(parts) >= 2:
            vals[parts[0]] = parts[1]
    return vals


def parse_hmc_ref(path):
    text = open(path).read()
    out = {"trajectories": []}
    m = re.search(r"CHECK PLAQ: (\S+) (\S+)", text)
    out["check_plaq"] = [float(m.group(1)), float(m.group(2))]
    m = re.search(r"Checksums (\S+) (\S+) OK", text)
    out["checksums"] = [m.group(1), m.group(2)]
    m = re.search(r"^START (\S+) (\S+) (\S+) (\S+)", text, re.M)
    out["start"] = [float(x) for x in m.groups()]
    act = re.findall(r"^action: gauge (\S+) bmass (\S+) det (\S+) fermion0 (\S+) mom (\S+) "
                     r"sum (\S+)", text, re.M)
    acc = re.findall(r"^(ACCEPT|REJECT): delta S = (\S+) start S = (\S+) end S = (\S+)",
                     text, re.M)
    its = re.findall(r"^IT_PER_TRAJ (\d+)", text, re.M)
    mg = re.findall(r"^MONITOR_FORCE_GAUGE\s+(\S+) (\S+)", text, re.M)
    mf = re.findall(r"^MONITOR_FORCE_FERMION0 (\S+) (\S+)", text, re.M)
    gm = re.findall(r"^GMES (\S+) (\S+) (\d+) (\S+) (\S+) (\S+) (\S+)", text, re.M)
    for k in range(len(acc)):
        out["trajectories"].append({
            "start_action": [float(x) for x in act[2 * k]],
            "end_action": [float(x) for x in act[2 * k + 1]],
            "accept": acc[k][0] == "ACCEPT",
            "delta_S": float(acc[k][1]), "start_S": float(acc[k][2]),
            "end_S": float(acc[k][3]), "iters": int(its[k]),
            "gauge_force": [float(x) for x in mg[k]],
            "fermion_force": [float(x) for x in mf[k]],
            "gmes": [float(x) for x in gm[k]],
        })
    m = re.search(r"total_iters = (\d+)", text)
    out["total_iters"] = int(m.group(1))
    return out


def hmc_refs():
    res = {}
    for n in (2, 3, 4):
        res[f"U{n}"] = {"input": parse_input(os.path.join(REF, f"in.U{n}.hmc")),
                        "ref": parse_hmc_ref(os.path.join(REF, "scalar", f"hmc.U{n}.ref"))}
    with open(os.path.join(HERE, "hmc_ref.json"), "w") as f:
        json.dump(res, f, indent=1)


def ref_outputs():
    if not have_ref():
        raise SystemExit("build oracle/_ref first: make -f oracle/ref.mk")
    out = {}
    R = Ref(dims=(2, 4, 6, 8), iseed=5, PBC=-1)
    out["gather_2468"] = np.stack([R.gather(g) for g in range(74)])
    out["coords_2468"] = R.coords()
    kw = dict(iseed=41, kappa_u1=1.0, bmass=1.0, G=0.1)
    R = Ref(**kw)
    R.load_config(os.path.join(HERE, "configs", "config.U2.4444"))
    out["gather_4444"] = np.stack([R.gather(g) for g in range(74)])
    out["coords_4444"] = R.coords()
    # boundary signs depend on t only: one representative site per time slice
    co = out["coords_4444"]
    reps = [int(np.where((co[:, 0] == 0) & (co[:, 1] == 0) & (co[:, 2] == 0) & (co[:, 3] == t))[0][0])
            for t in range(4)]
    out["bc_sites"] = np.array(reps)
    bcs = [R.bc(i) for i in reps]
    out["bc1"] = np.stack([b[0] for b in bcs])
    out["bc2"] = np.stack([b[1] for b in bcs])
    out["bc3"] = np.stack([b[2] for b in bcs])
    out["links"] = R.get_links()
    # inputs are regenerated from their seeds (oracle_lib.random_tf); large
    # outputs are kept on SITES (64 even + 64 odd sites) plus full-field norms
    src = random_tf(R.tf_shape(), 2024)
    for key, val in (("fop_plus", R.fermion_op(src, 1)), ("fop_minus", R.fermion_op(src, -1)),
                     ("dsq", R.dsq(src))):
        out[key] = val[SITES]
        out[key + "_norm"] = np.linalg.norm(val)
    pd, ui, dm, fm = R.fields()
    out["plaqdet"], out["uinv"], out["dmuumu"], out["fmunu"] = pd, ui, dm, fm
    shifts = np.array([3.1085594175442315e-08, 3.2994455960441383e-07,
                       1.9424842756552213e-06, 1.0453359626231250e-05,
                       5.5337819905761986e-05, 2.9204178440857227e-04,
                       1.5403300046437174e-03, 8.1233558140562465e-03,
                       4.2840454273820550e-02, 2.2594500626442715e-01,
                       1.1921171782283737e+00, 6.3026182343759860e+00,
                       3.3683411978650057e+01, 1.9083658214156412e+02,
                       1.5386784635765257e+03])
    it, sz, ps = R.congrad_multi(src, shifts, 5000, 1e-10)
    out["cg_iters"], out["cg_rsq"] = it, sz
    out["cg_psim_sites"] = ps[[0, 7, 14]][:, SITES]
    out["cg_psim_norms"] = np.linalg.norm(ps.reshape(len(shifts), -1), axis=1)
    mom = random_tf(R.link_shape(), 77)
    R.set_mom(mom)
    out["gforce_norm"] = R.gauge_force(0.1)
    out["gforce_mom"] = R.get_mom()
    R.set_mom(mom)
    out["fforce_norm"] = R.fermion_force(0.1, ps)
    out["fforce_mom"] = R.get_mom()
    out["rand_site0"] = R.rand_stream(0, 64)
    out["rand_site200"] = R.rand_stream(200, 64)
    out["rand_node"] = R.rand_stream(-1, 64)
    np.savez_compressed(os.path.join(HERE, "ref_U2.npz"), **out)


if __name__ == "__main__":
    copy_configs()
    hmc_refs()
    ref_outputs()
    print("golden fixtures written to", HERE)
