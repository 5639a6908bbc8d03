"""Lanczos with full reorthogonalisation for the extreme eigenvalues of a
Hermitian positive operator (test infrastructure: applied to the oracle's DSq
on the CPU).  Each distinct eigenvalue appears once among the Ritz values."""
import numpy as np


def tridiag_eigvals(alpha, beta):
    """eigenvalues of the symmetric tridiagonal matrix (alpha; beta[1:])"""
    k = len(alpha)
    T = np.diag(alpha) + np.diag(beta[1:k], 1) + np.diag(beta[1:k], -1)
    return np.linalg.eigvalsh(T)


def lanczos_extremes(apply, shape, nlow, nhigh, maxiter=600, tol=1e-10, seed=1, check=20):
    """(nlow smallest, nhigh largest Ritz values, iterations): stops when all
    of them moved by less than tol (relative) over `check` iterations"""
    rng = np.random.default_rng(seed)
    n = int(np.prod(shape))
    V = np.zeros((maxiter + 1, n), dtype=np.complex128)
    v = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    V[0] = v / np.linalg.norm(v)
    alpha, beta = [], [0.0]
    last = None
    for j in range(maxiter):
        w = apply(V[j].reshape(shape)).ravel()
        a = np.vdot(V[j], w).real
        alpha.append(a)
        w = w - a * V[j] - (beta[j] * V[j - 1] if j > 0 else 0)
        for _ in range(2):   # classical Gram-Schmidt, twice
            w = w - V[:j + 1].T @ (V[:j + 1].conj() @ w)
        b = np.linalg.norm(w)
        beta.append(b)
        V[j + 1] = w / b
        if (j + 1) % check == 0:
            ev = tridiag_eigvals(np.array(alpha), np.array(beta))
            cur = np.r_[ev[:nlow], ev[-nhigh:][::-1]]
            if last is not None and np.all(np.abs(cur - last) <= tol * np.abs(cur)):
                return cur[:nlow], cur[nlow:], j + 1
            last = cur
    ev = tridiag_eigvals(np.array(alpha), np.array(beta))
    return ev[:nlow], ev[-nhigh:][::-1], maxiter


def distinct(values, rtol=1e-6):
    """sorted values with near-equal neighbours merged"""
    out = []
    for x in sorted(values):
        if not out or abs(x - out[-1]) > rtol * abs(x):
            out.append(x)
    return np.array(out)
