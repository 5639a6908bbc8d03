"""CPU pins of the Pfaffian checker (tests/pfaffian_ref.py).

  * the two eliminations against what the mathematics fixes: the closed form
    (expansion along a row) for n <= 8, Pf^2 = det, Pf(B^T A B) = det B Pf A,
    a row-and-column interchange negating Pf; the reference's Q scheme
    against the pivoted LTL^T (equal for n/2 even, pi apart for n/2 odd);
  * the fermion matrix of the C oracle on config.U{2,3,4}.2222 against what
    4d_Q16/susy/phase.c printed for it (testsuite/scalar/phase.U*.ref, input
    in.U*.phase): antisymmetry, every per-pair diagonal element (all 512 for
    U(2), the first 128 for U(3), U(4)), and for U(2) the PFAFF line.
"""
import json
import os

import numpy as np
import pytest

import pfaffian_ref as P
from oracle_lib import GOLDEN, Oracle, config_path
from refcheck import printed_ok

REF = json.load(open(os.path.join(GOLDEN, "phase_ref.json")))
PHASE_KW = dict(dims=(2, 2, 2, 2), PBC=-1, lam=1.0, kappa_u1=1.0, bmass=1.0, fmass=0.0, G=0.1)


def pf(A):
    lm, ph = P.pfaffian_ltl(A)
    return np.exp(lm + 1j * ph)


def test_ltl_against_closed_forms():
    rng = np.random.default_rng(5)
    for n in (2, 4, 6, 8):
        for _ in range(3):
            A = P.random_antisym(n, rng)
            want = P.pfaffian_closed(A)
            assert abs(pf(A) - want) <= 1e-12 * abs(want), n
    a = 0.3 - 1.7j
    assert abs(pf(np.array([[0, a], [-a, 0]])) - a) < 1e-15
    for n in (20, 40):
        A = P.random_antisym(n, rng)
        lm, ph = P.pfaffian_ltl(A)
        sign, logdet = np.linalg.slogdet(A)
        assert abs(2 * lm - logdet) < 1e-10
        assert abs(np.exp(2j * ph) - sign) < 1e-10
        B = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
        assert abs(pf(B.T @ A @ B) / (np.linalg.det(B) * pf(A)) - 1) < 1e-9
        perm = np.arange(n)
        perm[[3, 11]] = perm[[11, 3]]
        assert abs(pf(A[np.ix_(perm, perm)]) / pf(A) + 1) < 1e-12


def test_reference_scheme_against_ltl():
    rng = np.random.default_rng(9)
    for n in (6, 8, 12):
        A = P.random_antisym(n, rng)
        lm, ph, diag = P.phase_scheme(A)
        lm2, ph2 = P.pfaffian_ltl(A)
        assert abs(lm - lm2) < 1e-11
        offset = np.pi * (n // 2)   # Pf(Q^T M Q) = (-1)^(n/2)
        assert abs(np.exp(1j * ph) - np.exp(1j * (ph2 + offset))) < 1e-11, n
        assert np.allclose(diag, P.scheme_diag(A, n // 2), rtol=1e-12, atol=0)


def _matrix(ncol, ncols=None):
    O = Oracle(ncol, **PHASE_KW)
    assert O.load_config(config_path(ncol, "2222"))[0] == 0
    g = {k: O.gather(k) for k in (22, 30, 34)}
    return P.fermion_matrix(lambda s: O.fermion_op(s, 1), O.lambdas(), g, O.V, ncols)


def test_u2_matrix_against_phase_ref():
    ref = REF["U2"]
    M = _matrix(2)
    assert M.shape == (ref["n"], ref["n"])
    assert np.abs(M + M.T).max() <= 1e-14 * np.abs(M).max()
    want = np.array([complex(*x) for x in ref["diag"]])
    got = P.scheme_diag(M, len(want))
    assert np.max(np.abs(got - want) / np.abs(want)) < 1e-10
    lm, ph = P.pfaffian_ltl(M)
    ph = P.fold(ph)
    out = [lm, ph, abs(np.cos(ph)), abs(np.sin(ph))]
    for k, (g, w) in enumerate(zip(out, ref["pfaff"])):
        assert printed_ok(g, w, 8), (k, g, w)


@pytest.mark.parametrize("ncol", [3, 4])
def test_leading_pivots_against_phase_ref(ncol):
    ref = REF[f"U{ncol}"]
    want = np.array([complex(*x) for x in ref["diag"]])
    m = 2 * len(want)
    M = _matrix(ncol, m)
    assert M.shape == (ref["n"], m)
    lead = M[:m, :m]
    assert np.abs(lead + lead.T).max() <= 1e-14 * np.abs(lead).max()
    got = P.scheme_diag(lead, len(want))
    assert np.max(np.abs(got - want) / np.abs(want)) < 1e-10
