"""numpy checker of the per-trajectory measurements (test infrastructure only).

Each function restates one measurement of the reference's driver
(control.c:85-138) on links U of shape (V, 5, N, N) in the reference's
even-first site order, with the Hermitian eigenproblems delegated to
numpy.linalg (the reference calls LAPACK zheev for the same thing):
  link_trace     FLINK / FLINK_DET   link_trace.c:7-55
  wilson_line    LINES / LINES_POLAR ploop.c:8-126 (+ polar, unit.c:53-110)
  measure_det    DET                 determinant.c:93-119
  widths         WIDTHS              widths.c:8-69
  scalar_eig     UUBAR_EIG / POLAR_EIG  scalar_eig.c:14-123 (+ matrix_log,
                                     unit.c:10-44)
tests/test_measure_ref.py pins these against the values the reference prints
for its regression configurations and trajectories.
"""
import numpy as np

from surrogate_ref import EV, Geometry

NUMLINK = 5


def dag(m):
    return np.conj(np.swapaxes(m, -1, -2))


def link_trace(U):
    """(linktr[5], ave, width, link_det[5], det_ave, det_width)"""
    V, _, N, _ = U.shape
    uu = U @ dag(U)
    tr = np.trace(uu, axis1=-2, axis2=-1).real / N          # (V, 5)
    tl = uu - tr[..., None, None] * np.eye(N)
    det = np.linalg.det(tl).real
    linktr = tr.mean(axis=0)
    ave = linktr.mean()
    width = np.sqrt((tr ** 2).sum() / (V * NUMLINK) - ave * ave)
    link_det = det.mean(axis=0)
    dave = link_det.mean()
    dwidth = np.sqrt((det ** 2).sum() / (V * NUMLINK) - dave * dave)
    return linktr, ave, width, link_det, dave, dwidth


def polar_unitary(m):
    """u = (m m^dag)^(-1/2) m, the unitary factor of m = P u (unit.c:53-110)"""
    lam, w = np.linalg.eigh(m @ dag(m))
    inv = (w * (1.0 / np.sqrt(lam))[..., None, :]) @ dag(w)
    return inv @ m


def wilson_line(geo, U, a, project=False):
    """average trace of the ordered product U_a(x) U_a(x+a) ... over the
    hyperplane where coordinate a (t for the fifth link) is 0, and the
    average modulus of the traces"""
    nx, ny, nz, nt = geo.dims
    lens = (nx, ny, nz, nt, nt)
    axis = a if a < 4 else 3
    L = lens[a]
    links = polar_unitary(U[:, a]) if project else U[:, a]
    step = geo.shift(EV[a])
    start = np.nonzero(geo.coords[:, axis] == 0)[0]
    x = start.copy()
    p = links[x]
    for _ in range(1, L):
        x = step[x]
        p = p @ links[x]
    tr = np.trace(p, axis1=-2, axis2=-1)
    norm = L / geo.V
    return tr.sum() * norm, np.abs(tr).sum() * norm


def plaqdet(geo, U):
    """plaqdet_ab(x) = det U_a(x)^* det U_b(x) det U_b(x+a)^* det U_a(x+b)"""
    d = np.linalg.det(U)                                     # (V, 5)
    pd = np.zeros((NUMLINK, NUMLINK, geo.V), dtype=complex)
    for a in range(NUMLINK):
        xa = geo.shift(EV[a])
        for b in range(NUMLINK):
            if a != b:
                xb = geo.shift(EV[b])
                pd[a, b] = np.conj(d[:, a]) * d[:, b] * np.conj(d[xa, b]) * d[xb, a]
    return pd


def measure_det(pd):
    """DET: <plaqdet>, <Re^2>, <Im^2>, <|plaqdet - 1|^2> over the 10 a > b"""
    vals = np.concatenate([pd[a, b] for a in range(1, NUMLINK) for b in range(a)])
    n = len(vals)
    tot = vals.sum() / n
    return (tot.real, tot.imag, (vals.real ** 2).sum() / n, (vals.imag ** 2).sum() / n,
            (np.abs(vals - 1) ** 2).sum() / n)


def widths(geo, U):
    """WIDTHS: widths of Re Tr P_ab and of Re, Im det P_ab over the 10 a < b,
    P_ab(x) = U_b(x) U_a(x+b) U_b(x+a)^dag U_a(x)^dag"""
    tr, dt = [], []
    for a in range(NUMLINK):
        xa = geo.shift(EV[a])
        for b in range(a + 1, NUMLINK):
            xb = geo.shift(EV[b])
            P = U[:, b] @ U[xb, a] @ dag(U[xa, b]) @ dag(U[:, a])
            tr.append(np.trace(P, axis1=-2, axis2=-1).real)
            dt.append(np.linalg.det(P))
    tr, dt = np.concatenate(tr), np.concatenate(dt)
    n = len(tr)

    def w(v):
        return np.sqrt((v ** 2).sum() / n - (v.sum() / n) ** 2)

    return w(tr), w(dt.real), w(dt.imag)


def scalar_eig(U, project=False):
    """(ave, width, min, max) per eigenvalue index (ascending) of the
    traceless part of U U^dag (project=False) or of log P, P = (U U^dag)^(1/2)
    (project=True), over all links"""
    N = U.shape[-1]
    uu = U @ dag(U)
    if project:
        lam = 0.5 * np.log(np.linalg.eigvalsh(uu))
    else:
        tr = np.trace(uu, axis1=-2, axis2=-1).real / N
        lam = np.linalg.eigvalsh(uu - tr[..., None, None] * np.eye(N))
    lam = lam.reshape(-1, N)
    ave = lam.mean(axis=0)
    width = np.sqrt((lam ** 2).mean(axis=0) - ave * ave)
    return ave, width, lam.min(axis=0), lam.max(axis=0)


def report(dims, U):
    """every measurement above, keyed by the reference's output line name,
    flattened in its print order"""
    geo = Geometry(dims)
    lt = link_trace(U)
    out = {"FLINK": list(lt[0]) + [lt[1], lt[2]],
           "FLINK_DET": list(lt[3]) + [lt[4], lt[5]]}
    for name, proj in (("LINES", False), ("LINES_POLAR", True)):
        row = []
        for a in range(NUMLINK):
            s, _ = wilson_line(geo, U, a, proj)
            row += [s.real, s.imag]
        out[name] = row
    out["DET"] = list(measure_det(plaqdet(geo, U)))
    out["WIDTHS"] = list(widths(geo, U))
    for name, proj in (("UUBAR_EIG", False), ("POLAR_EIG", True)):
        ave, wd, mn, mx = scalar_eig(U, proj)
        out[name] = [[ave[j], wd[j], mn[j], mx[j]] for j in range(U.shape[-1])]
    return out
