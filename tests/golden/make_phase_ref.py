"""Write tests/golden/phase_ref.json: what the reference's Pfaffian phase
measurement prints for the regression configurations config.U{2,3,4}.2222
(4d_Q16/testsuite/scalar/phase.U{2,3,4}.ref, input in.U{N}.phase):
  * the PFAFF line (phase.c:388: log |Pf|, phase in [0, 2 pi), |cos|, |sin|,
    printed %.8g);
  * the diagonal element printed after each pair of columns (phase.c:349,
    %.16g): all of them for U(2), the first 128 for U(3) and U(4).

Run in the build container (it reads /root/reference):
    python tests/golden/make_phase_ref.py
"""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference/4d_Q16/testsuite"
KEEP = {2: None, 3: 128, 4: 128}
PARAMS = ("nx", "ny", "nz", "nt", "PBC", "lambda", "kappa_u1", "bmass", "fmass", "G")


def parse(path):
    out = {"diag": [], "params": {}}
    for line in open(path):
        f = line.split()
        if not f:
            continue
        if f[0] in PARAMS and len(f) == 2 and f[0] not in out["params"]:
            out["params"][f[0]] = float(f[1])
        elif f[0] == "Columns":
            out["diag"].append([float(f[-2]), float(f[-1])])
            out["n"] = int(f[3].rstrip(":"))
        elif f[0] == "PFAFF":
            out["pfaff"] = [float(x) for x in f[1:]]
    return out


def main():
    data = {"_note": "PFAFF and per-pair diagonal elements printed by 4d_Q16/susy/phase.c "
                     "for config.U{N}.2222 (testsuite/scalar/phase.U{N}.ref)"}
    for n, keep in KEEP.items():
        d = parse(os.path.join(REF, "scalar", f"phase.U{n}.ref"))
        assert len(d["diag"]) == d["n"] // 2, (n, len(d["diag"]))
        if keep is not None:
            d["diag"] = d["diag"][:keep]
        d["source"] = f"4d_Q16/testsuite/scalar/phase.U{n}.ref"
        d["config"] = f"config.U{n}.2222"
        data[f"U{n}"] = d
    path = os.path.join(HERE, "phase_ref.json")
    with open(path, "w") as f:
        json.dump(data, f, indent=0)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
