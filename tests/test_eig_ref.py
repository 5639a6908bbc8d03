"""The D^dag D spectrum printed by the reference's eigenvalue runs
(testsuite/scalar/eig.U{2,3,4}.ref, fixture tests/golden/eig_ref.json) as an
independent pin of the operator: Lanczos (tests/lanczos.py) on the oracle's
DSq on the CPU, and the library's own eigenvalue tool (susy_eig_dsq) on the
GPU.  Distinct eigenvalues are compared (each is printed twice: the spectrum
comes in degenerate pairs); values to their 8 printed digits."""
import json
import os

import numpy as np
import pytest

from lanczos import distinct, lanczos_extremes
from oracle_lib import GOLDEN, Oracle, config_path
from refcheck import printed_ok

REF = json.load(open(os.path.join(GOLDEN, "eig_ref.json")))


def kwargs(ncol):
    p = REF[f"U{ncol}"]["params"]
    assert p["config"] == f"config.U{ncol}.4444" and p["fmass"] == 0
    return dict(dims=(p["nx"], p["ny"], p["nz"], p["nt"]), PBC=p["PBC"], iseed=41,
                lam=p["lambda"], kappa_u1=p["kappa_u1"], bmass=p["bmass"], G=p["G"])


def check(ncol, low, high, nlow, nhigh):
    ref = REF[f"U{ncol}"]
    want_lo, want_hi = distinct(ref["low"]), distinct(ref["high"])[::-1]
    got_lo, got_hi = distinct(low), distinct(high)[::-1]
    for k in range(nlow):
        assert printed_ok(got_lo[k], want_lo[k], 8), ("low", k, got_lo[k], want_lo[k])
    for k in range(nhigh):
        assert printed_ok(got_hi[k], want_hi[k], 8), ("high", k, got_hi[k], want_hi[k])


def test_fixture_pairs():
    for n in (2, 3, 4):
        r = REF[f"U{n}"]
        assert len(r["low"]) == 12 and len(distinct(r["low"])) == 6
        assert len(r["high"]) == 12 and len(distinct(r["high"])) == 6


def test_oracle_spectrum_u2():
    O = Oracle(2, **kwargs(2))
    assert O.load_config(config_path(2))[0] == 0
    lo, hi, it = lanczos_extremes(O.dsq, O.tf_shape(), 6, 6, maxiter=500, tol=1e-10)
    check(2, lo, hi, 3, 3)


@pytest.mark.gpu
@pytest.mark.parametrize("ncol", [2, 3, 4])
def test_library_spectrum(ncol):
    import susy_amd
    print("eig", ncol, flush=True)
    lat = susy_amd.Lattice(ncol, **kwargs(ncol))
    assert lat.load_config(config_path(ncol)) == 0
    lo, hi, it = lat.eig_dsq(12, 12, maxiter=900, tol=1e-10)
    print("lanczos iterations", it, flush=True)
    check(ncol, lo, hi, 4, 4)
