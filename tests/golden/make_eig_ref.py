"""Write tests/golden/eig_ref.json: the D^dag D eigenvalues printed by the
reference's regression runs 4d_Q16/testsuite/scalar/eig.U{2,3,4}.ref (its
eigenvalue measurement on config.U{N}.4444) and the couplings of those runs.

Run in the build container (it reads /root/reference):
    python tests/golden/make_eig_ref.py
"""
import json
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference/4d_Q16/testsuite/scalar"


def parse(path):
    out = {"low": [], "high": [], "params": {}}
    for line in open(path):
        f = line.split()
        if not f:
            continue
        if f[0] == "EIGENVALUE" and len(f) >= 3:
            out["low"].append(float(f[2]))
        elif f[0] == "BIGEIGVAL" and len(f) >= 3:
            out["high"].append(float(f[2]))
        elif f[0] in ("nx", "ny", "nz", "nt", "PBC", "Nvec") and len(f) == 2:
            out["params"][f[0]] = int(f[1])
        elif f[0] in ("lambda", "kappa_u1", "bmass", "fmass", "G", "eig_tol") and len(f) == 2:
            out["params"][f[0]] = float(f[1])
        elif f[0] == "reload_serial":
            out["params"]["config"] = os.path.basename(f[1])
    return out


def main():
    data = {}
    for n in (2, 3, 4):
        path = os.path.join(REF, f"eig.U{n}.ref")
        data[f"U{n}"] = parse(path)
        data[f"U{n}"]["source"] = f"4d_Q16/testsuite/scalar/eig.U{n}.ref"
    with open(os.path.join(HERE, "eig_ref.json"), "w") as f:
        json.dump(data, f, indent=1)


if __name__ == "__main__":
    main()
