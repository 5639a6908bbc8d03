"""Checker (numpy) for the Pfaffian phase measurement of 4d_Q16/susy/phase.c.
Test infrastructure only.

  * basis (phase.c:36-120 in, 168-213 out): per site (the reference's site
    order), per generator Lambda_j, 16 slots naming Twist_Fermion components;
    the (2,3), (1,3), (1,2) plaquette slots of site x are carried by the
    Twist_Fermion of x + e0 + e1, x + e0 + e2, x + e0 + e3 (the reference
    gathers them through paths 23/31/35 in and 22/30/34 out).  Entry
    M_ab = Tr((D e_b)_slot(a) Lambda_ja), the bilinear product of phase.c:12.
  * phase_scheme: the reference's own elimination (phase.c:228-387, after
    hep-lat/0305002): upper-triangular Q with Q^T M Q block-diagonal,
    Pf M = 1 / det Q for n/2 even, printed as -log|det Q| and the phase
    -sum arg diag folded into [0, 2 pi).
  * pfaffian_ltl: Parlett-Reid with pivoting (M. Wimmer, ACM TOMS 38
    (2012) 30), the form the library computes.

Pins (tests/test_pfaffian_ref.py): Pf^2 = det and the 2x2 / 4x4 closed forms
on random antisymmetric matrices, Pf(B^T A B) = det B Pf A, both algorithms
agreeing, and the whole measurement against the printed PFAFF line of
testsuite/scalar/phase.U2.ref on config.U2.2222.
"""
import numpy as np

# Twist_Fermion components in the oracle's AoS order: 0 site, 1..5 links
# 0..4, 6..15 plaquettes 0..9 with (a,b) -> 0:01 1:02 2:03 3:04 4:12 5:13
# 6:14 7:23 8:24 9:34
_LINK = {a: 1 + a for a in range(5)}
_PLAQ = {(0, 1): 6, (0, 2): 7, (0, 3): 8, (0, 4): 9, (1, 2): 10, (1, 3): 11,
         (1, 4): 12, (2, 3): 13, (2, 4): 14, (3, 4): 15}
# the 16 slots of phase.c:50-100 in order: (component, carrier gather)
SLOTS = [(0, None), (_LINK[4], None),
         (_LINK[0], None), (_PLAQ[0, 4], None),
         (_LINK[1], None), (_PLAQ[1, 4], None),
         (_LINK[2], None), (_PLAQ[2, 4], None),
         (_LINK[3], None), (_PLAQ[3, 4], None),
         (_PLAQ[0, 1], None), (_PLAQ[2, 3], 22),
         (_PLAQ[0, 2], None), (_PLAQ[1, 3], 30),
         (_PLAQ[0, 3], None), (_PLAQ[1, 2], 34)]


def fermion_matrix(apply_d, lambdas, gathers, V, ncols=None):
    """n x n matrix M_ab (n = 16 N^2 V) of the linear map apply_d on (V, 16,
    N, N) Twist_Fermion arrays; gathers[g] = site index of x + offset g.
    With ncols, only the first ncols columns (n x ncols)."""
    nn = lambdas.shape[0]
    n = 16 * nn * V
    carrier = np.empty((16, V), dtype=np.int64)
    comp = np.empty(16, dtype=np.int64)
    for q, (k, g) in enumerate(SLOTS):
        comp[q] = k
        carrier[q] = np.arange(V) if g is None else gathers[g]
    N = lambdas.shape[1]
    ncols = n if ncols is None else ncols
    M = np.zeros((n, ncols), dtype=np.complex128)
    lam_t = np.transpose(lambdas, (0, 2, 1))   # Lambda_j^T, for Tr(X Lambda_j)
    for b in range(ncols):
        sb, jb, qb = b // (16 * nn), (b // 16) % nn, b % 16
        src = np.zeros((V, 16, N, N), dtype=np.complex128)
        src[carrier[qb, sb], comp[qb]] = lambdas[jb]
        res = apply_d(src)
        R = res[carrier.T, comp[None, :]]        # (V, 16, N, N): slot values per site
        # Tr(R Lambda_j) = sum_ik R_ik Lambda_j,ki
        col = np.einsum("sqik,jik->sjq", R, lam_t)
        M[:, b] = col.reshape(n)
    return M


def pfaffian_ltl(A):
    """(log |Pf A|, arg Pf A) of a complex antisymmetric matrix (even n)"""
    A = np.array(A, dtype=np.complex128)
    n = A.shape[0]
    lm, ph = 0.0, 0.0
    for k in range(0, n - 1, 2):
        kp = k + 1 + int(np.argmax(np.abs(A[k + 1:, k])))
        if kp != k + 1:
            A[[k + 1, kp], k:] = A[[kp, k + 1], k:]
            A[k:, [k + 1, kp]] = A[k:, [kp, k + 1]]
            ph += np.pi
        piv = A[k, k + 1]
        if piv == 0:
            return -np.inf, 0.0
        lm += np.log(abs(piv))
        ph += np.angle(piv)
        if k + 2 < n:
            tau = A[k, k + 2:] / piv
            col = A[k + 2:, k + 1].copy()
            A[k + 2:, k + 2:] += np.outer(tau, col) - np.outer(col, tau)
    return lm, ph


def fold(ph):
    ph = np.fmod(ph, 2 * np.pi)
    return ph + 2 * np.pi if ph < 0 else ph


def phase_scheme(M):
    """The reference's elimination on an explicit matrix (matvec = M @ q):
    returns what it prints, (-log|det Q|, phase in [0, 2 pi)), and the
    per-pair diagonal elements diag[i + 1].  O(n^4): small matrices only."""
    n = M.shape[0]
    Q = np.eye(n, dtype=np.complex128)
    diag = np.zeros(n, dtype=np.complex128)
    for i in range(0, n - 1, 2):
        tc = Q[:, i + 1] @ (M @ Q[:, i])
        Q[:, i + 1] /= tc
        for k in range(i + 2, n):
            mq = M @ Q[:, k]
            Q[:, k] -= Q[:, i] * (Q[:, i + 1] @ mq)
            Q[:, k] += Q[:, i + 1] * (Q[:, i] @ mq)
        diag[i + 1] = Q[i + 1, i + 1]
    phase = sum(np.angle(diag[i]) for i in range(1, n, 2))
    log_mag = sum(np.log(abs(diag[i]) ** 2) / 2 for i in range(1, n, 2))
    return -log_mag, fold(-phase), diag[1::2]


def scheme_diag(M, npairs):
    """The diagonal elements the reference prints after each pair of columns
    (phase.c:349): its elimination is the LTL^T one without pivoting, with
    diag[i + 1] = 1 / <q_{i+1}|D|q_i> = -1 / pivot_i.  Pair p depends on the
    leading (2p + 2) x (2p + 2) block only."""
    A = np.array(M[:2 * npairs, :2 * npairs], dtype=np.complex128)
    n = A.shape[0]
    out = np.zeros(npairs, dtype=np.complex128)
    for p, k in enumerate(range(0, n - 1, 2)):
        piv = A[k, k + 1]
        out[p] = -1.0 / piv
        if k + 2 < n:
            tau = A[k, k + 2:] / piv
            col = A[k + 2:, k + 1].copy()
            A[k + 2:, k + 2:] += np.outer(tau, col) - np.outer(col, tau)
    return out


def pfaffian_closed(A):
    """Pf by expansion along the first row (tiny n)"""
    n = A.shape[0]
    if n == 0:
        return 1.0 + 0j
    out = 0j
    for j in range(1, n):
        keep = [i for i in range(n) if i not in (0, j)]
        out += (-1) ** (j + 1) * A[0, j] * pfaffian_closed(A[np.ix_(keep, keep)])
    return out


def random_antisym(n, rng):
    X = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    return X - X.T
