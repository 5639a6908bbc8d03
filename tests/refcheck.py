"""Comparison of trajectory statistics with the numbers printed in the
reference's regression outputs (testsuite/scalar/hmc.U*.ref)."""
import numpy as np


def hmc_kwargs(inp):
    """Lattice / checker keyword arguments from a testsuite in.U*.hmc file"""
    return dict(dims=(int(inp["nx"]), int(inp["ny"]), int(inp["nz"]), int(inp["nt"])),
                PBC=int(inp["PBC"]), iseed=int(inp["iseed"]), Nroot=int(inp["Nroot"]),
                Norder=int(inp["Norder"]), lam=float(inp["lambda"]),
                kappa_u1=float(inp["kappa_u1"]), bmass=float(inp["bmass"]),
                fmass=float(inp["fmass"]), G=float(inp["G"]),
                niter=int(inp["max_cg_iterations"]),
                error_per_site=float(inp["error_per_site"]),
                traj_length=float(inp["traj_length"]), nstep=int(inp["nstep"]),
                nstep_gauge=int(inp["nstep_gauge"]))


def printed_ok(value, printed, sig, slack=1.0):
    """value agrees with a number printed with %.<sig>g to its last digit"""
    if printed == 0:
        return abs(value) < 1e-300
    ulp = 10.0 ** (np.floor(np.log10(abs(printed))) - (sig - 1))
    return abs(value - printed) <= (0.5 + slack) * ulp


def check_traj(st, tr, iters_exact=True):
    """st: stats vector of one trajectory; tr: the matching .ref record"""
    names = ["gauge", "bmass", "det", "fermion", "mom", "sum"]
    for k, nm in enumerate(names):
        assert printed_ok(st[11 + k], tr["start_action"][k], 8), nm
        assert printed_ok(st[17 + k], tr["end_action"][k], 8), nm
    assert printed_ok(st[0], tr["start_S"], 12) and printed_ok(st[1], tr["end_S"], 12)
    assert printed_ok(st[2], tr["delta_S"], 4)
    if iters_exact:
        assert int(st[4]) == tr["iters"]
    assert bool(st[3]) == tr["accept"]
    assert printed_ok(st[5], tr["gauge_force"][0], 4)
    assert printed_ok(st[6], tr["gauge_force"][1], 4)
    assert printed_ok(st[7], tr["fermion_force"][0], 4)
    assert printed_ok(st[8], tr["fermion_force"][1], 4)


