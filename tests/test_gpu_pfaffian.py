"""The Pfaffian phase measurement (4d_Q16/susy/phase.c) through the library:
on config.U{2,3,4}.2222 with the parameters of in.U{N}.phase, the fermion
matrix built from the FP64 operator on the device against the checker's
matrix from the C oracle, its leading pivots against the diagonal elements
phase.c printed, and log |Pf|, the phase, |cos| and |sin| against the
printed PFAFF line (testsuite/scalar/phase.U{2,3,4}.ref)."""
import json
import os

import numpy as np
import pytest

import pfaffian_ref as P
import susy_amd
from oracle_lib import GOLDEN, Oracle, config_path
from refcheck import printed_ok

pytestmark = pytest.mark.gpu

REF = json.load(open(os.path.join(GOLDEN, "phase_ref.json")))
PHASE_KW = dict(dims=(2, 2, 2, 2), PBC=-1, lam=1.0, kappa_u1=1.0, bmass=1.0, fmass=0.0, G=0.1)


def say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("ncol", [2, 3, 4])
def test_pfaffian_of_regression_config(ncol):
    ref = REF[f"U{ncol}"]
    say("pfaffian U", ncol, "n", ref["n"])
    lat = susy_amd.Lattice(ncol, **PHASE_KW)
    assert lat.load_config(config_path(ncol, "2222")) == 0
    M = lat.fermion_matrix()
    assert M.shape == (ref["n"], ref["n"])
    # the checker's matrix from the C oracle (leading 512 columns for U(4))
    O = Oracle(ncol, **PHASE_KW)
    assert O.load_config(config_path(ncol, "2222"))[0] == 0
    g = {k: O.gather(k) for k in (22, 30, 34)}
    cols = None if ncol < 4 else 512
    W = P.fermion_matrix(lambda s: O.fermion_op(s, 1), O.lambdas(), g, O.V, cols)
    err = np.abs(M[:, :W.shape[1]] - W).max() / np.abs(W).max()
    say("matrix vs checker", err)
    assert err < 1e-13
    want = np.array([complex(*x) for x in ref["diag"]])
    m = 2 * len(want)
    got = P.scheme_diag(M[:m, :m], len(want))
    say("pivots vs printed", np.max(np.abs(got - want) / np.abs(want)))
    assert np.max(np.abs(got - want) / np.abs(want)) < 1e-10
    del M, W
    lm, ph, asym = lat.pfaffian()
    out = [lm, ph, abs(np.cos(ph)), abs(np.sin(ph))]
    say("PFAFF", out, "printed", ref["pfaff"], "asym", asym)
    assert asym < 1e-14
    for k, (gv, w) in enumerate(zip(out, ref["pfaff"])):
        assert printed_ok(gv, w, 8), (k, gv, w)


def test_pfaffian_refuses_large_lattice():
    lat = susy_amd.Lattice(2, dims=(4, 4, 4, 4))
    with pytest.raises(RuntimeError, match="too large"):
        lat.pfaffian()
