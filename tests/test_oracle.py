"""CPU checks that pin the C oracle (oracle/susy_oracle.c) to the reference:
golden fixtures written by the reference build (tests/golden/ref_U2.npz), the
reference's own regression outputs (testsuite/scalar/hmc.U*.ref, in
tests/golden/hmc_ref.json), and identities the mathematics fixes."""
import os
import re

import numpy as np
import pytest

from oracle_lib import Oracle, Ref, config_path, have_ref, random_links, random_tf, rel_err
from refcheck import check_traj, hmc_kwargs

REF_ROOT = "/root/reference/4d_Q16"
HMC_KW = dict(iseed=41, kappa_u1=1.0, bmass=1.0, G=0.1)


def hmc_oracle(ncol, ref):
    return Oracle(ncol, **hmc_kwargs(ref[f"U{ncol}"]["input"]))


def test_gather_tables_match_reference(golden_ref):
    O = Oracle(2, dims=(2, 4, 6, 8), iseed=5)
    assert np.array_equal(O.coords(), golden_ref["coords_2468"])
    for g in range(74):
        assert np.array_equal(O.gather(g), golden_ref["gather_2468"][g]), g
    O = Oracle(2, **HMC_KW)
    assert np.array_equal(O.coords(), golden_ref["coords_4444"])
    for g in range(74):
        assert np.array_equal(O.gather(g), golden_ref["gather_4444"][g]), g


def test_site_order_closed_form():
    """even sites first, lexicographic (x fastest) within each parity"""
    O = Oracle(3, dims=(4, 2, 6, 2))
    co = O.coords()
    lex = co[:, 0] + 4 * (co[:, 1] + 2 * (co[:, 2] + 6 * co[:, 3]))
    par = co.sum(1) % 2
    V = len(co)
    assert (par[: V // 2] == 0).all() and (par[V // 2:] == 1).all()
    assert (np.diff(lex[: V // 2]) > 0).all() and (np.diff(lex[V // 2:]) > 0).all()
    # gather 8+2a is x+e_a, 9+2a is x-e_a, e_4 = (-1,-1,-1,-1)
    dims = np.array([4, 2, 6, 2])
    where = {tuple(c): i for i, c in enumerate(co)}
    ev = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1], [-1, -1, -1, -1]])
    for a in range(5):
        fwd, bwd = O.gather(8 + 2 * a), O.gather(9 + 2 * a)
        for i in range(V):
            assert fwd[i] == where[tuple((co[i] + ev[a]) % dims)]
            assert bwd[i] == where[tuple((co[i] - ev[a]) % dims)]


def test_boundary_signs_match_reference(golden_ref):
    """bc[a](x) = PBC when the path from x leaves [0, nt-1] in t; the fixture
    holds the reference's per-site tables for one site per time slice."""
    tau = [0, 0, 0, 1, -1]
    for k, t in enumerate(range(4)):
        bc1, bc2, bc3 = golden_ref["bc1"][k], golden_ref["bc2"][k], golden_ref["bc3"][k]
        sign = lambda dt: -1.0 if (t + dt < 0 or t + dt > 3) else 1.0  # noqa: E731
        for a in range(5):
            assert bc1[a] == sign(tau[a]) and bc1[9 - a] == sign(-tau[a])
            for b in range(5):
                assert bc2[a][b] == sign(tau[a] + tau[b])
                assert bc2[9 - a][9 - b] == sign(-tau[a] - tau[b])
                for c in range(5):
                    assert bc3[a][b][c] == sign(tau[a] + tau[b] + tau[c])
                    assert bc3[9 - a][9 - b][9 - c] == sign(-tau[a] - tau[b] - tau[c])


def _oracle_on_config():
    O = Oracle(2, **HMC_KW)
    r, sums = O.load_config(config_path(2))
    assert r == 0 and sums[0] == sums[2] and sums[1] == sums[3]
    return O


def test_oracle_operator_matches_reference_fixture(golden_ref):
    O = _oracle_on_config()
    assert rel_err(O.get_links(), golden_ref["links"]) == 0
    src = random_tf(O.tf_shape(), 2024)
    sites = np.r_[0:64, 128:192]
    for key, val in (("fop_plus", O.fermion_op(src, 1)), ("fop_minus", O.fermion_op(src, -1)),
                     ("dsq", O.dsq(src))):
        assert rel_err(val[sites], golden_ref[key]) < 1e-13, key
        assert abs(np.linalg.norm(val) / golden_ref[key + "_norm"] - 1) < 1e-13, key
    pd, ui, dm, fm = O.fields()
    mask = ~np.eye(5, dtype=bool)
    assert rel_err(pd[mask], golden_ref["plaqdet"][mask]) < 1e-13
    assert rel_err(ui, golden_ref["uinv"]) < 1e-13
    assert rel_err(dm, golden_ref["dmuumu"]) < 1e-13
    assert rel_err(fm, golden_ref["fmunu"]) < 1e-13


def test_oracle_cg_and_forces_match_reference_fixture(golden_ref):
    O = _oracle_on_config()
    src = random_tf(O.tf_shape(), 2024)
    shifts = O.rhmc()[16:31]
    it, sz, ps = O.congrad_multi(src, shifts, 5000, 1e-10)
    assert it == golden_ref["cg_iters"]
    assert abs(sz / golden_ref["cg_rsq"] - 1) < 1e-8
    sites = np.r_[0:64, 128:192]
    assert rel_err(ps[[0, 7, 14]][:, sites], golden_ref["cg_psim_sites"]) < 1e-10
    norms = np.linalg.norm(ps.reshape(15, -1), axis=1)
    assert np.allclose(norms, golden_ref["cg_psim_norms"], rtol=1e-10)
    mom = random_tf(O.link_shape(), 77)
    O.set_mom(mom)
    nm = O.gauge_force(0.1)
    assert abs(nm / golden_ref["gforce_norm"] - 1) < 1e-12
    assert rel_err(O.get_mom() - mom, golden_ref["gforce_mom"] - mom) < 1e-12
    O.set_mom(mom)
    nm = O.fermion_force(0.1, ps, O.rhmc()[1:16])
    assert abs(nm / golden_ref["fforce_norm"] - 1) < 1e-9
    assert rel_err(O.get_mom() - mom, golden_ref["fforce_mom"] - mom) < 1e-9


def test_generator_streams_match_reference(golden_ref):
    O = Oracle(2, **HMC_KW)
    assert np.array_equal(O.rand_stream(0, 64), golden_ref["rand_site0"])
    assert np.array_equal(O.rand_stream(200, 64), golden_ref["rand_site200"])
    assert np.array_equal(O.rand_stream(-1, 64), golden_ref["rand_node"])


@pytest.mark.parametrize("ncol", [2, 3, 4])
def test_config_checksums_and_plaquette(ncol, hmc_ref):
    O = Oracle(ncol, **HMC_KW)
    r, sums = O.load_config(config_path(ncol))
    assert r == 0
    ref = hmc_ref[f"U{ncol}"]["ref"]
    assert [f"{sums[0]:x}", f"{sums[1]:x}"] == ref["checksums"]
    ss, st = O.plaquette()
    assert abs(ss / ref["check_plaq"][0] - 1) < 1e-15
    assert abs(st / ref["check_plaq"][1] - 1) < 1e-15
    g = O.actions()[0]
    assert abs(g / O.V / ref["start"][3] - 1) < 1e-7   # START line S_B/site (%.8g)


@pytest.mark.parametrize("ncol", [2, 3, 4])
def test_adjointness(ncol):
    """<a, D b> = <D^dag a, b> on random links, both BCs"""
    for pbc in (1, -1):
        O = Oracle(ncol, dims=(2, 2, 4, 4), PBC=pbc, G=0.3)
        O.set_links(random_links(O.link_shape(), 11 + ncol, 0.3))
        a, b = random_tf(O.tf_shape(), 1), random_tf(O.tf_shape(), 2)
        lhs = np.vdot(a, O.fermion_op(b, 1))
        rhs = np.vdot(O.fermion_op(a, -1), b)
        assert abs(lhs - rhs) / abs(lhs) < 1e-13


def test_gauge_covariance():
    """<a, D b> is invariant under U_a(x) -> g(x) U_a(x) g(x+a)^dag when the
    source components transform as eta(x) -> g eta g^dag (at x),
    psi_b(x) -> g(x) psi_b g(x+b)^dag, chi_de(x) -> g(x+d+e) chi_de g(x)^dag,
    and the output components as site: g(x) . g(x)^dag, link b:
    g(x+b) . g(x)^dag, plaquette ab: g(x) . g(x+a+b)^dag (the idea of
    random_gauge_trans.c:8-40).  The plaquette-determinant terms are
    invariant as well (plaqdet is gauge invariant), so G > 0 here."""
    dims = np.array([2, 2, 2, 4])
    O = Oracle(3, dims=tuple(dims), G=0.2, PBC=-1)
    links = random_links(O.link_shape(), 3, 0.4)
    O.set_links(links)
    a, b = random_tf(O.tf_shape(), 4), random_tf(O.tf_shape(), 5)
    ref = np.vdot(a, O.fermion_op(b, 1))
    rng = np.random.default_rng(9)
    z = rng.standard_normal((O.V, 3, 3)) + 1j * rng.standard_normal((O.V, 3, 3))
    g = np.linalg.qr(z)[0]
    gd = np.conj(np.transpose(g, (0, 2, 1)))
    co = O.coords()
    where = {tuple(c): i for i, c in enumerate(co)}
    ev = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1], [-1, -1, -1, -1]])

    def at(d):   # site index of x + d for every x
        return np.array([where[tuple((c + d) % dims)] for c in co])

    plaq = [(i, j) for i in range(5) for j in range(i + 1, 5)]
    newl = links.copy()
    for k in range(5):
        newl[:, k] = g @ links[:, k] @ gd[at(ev[k])]
    ra, rb = a.copy(), b.copy()
    ra[:, 0] = g @ a[:, 0] @ gd
    rb[:, 0] = g @ b[:, 0] @ gd
    for k in range(5):
        ra[:, 1 + k] = g[at(ev[k])] @ a[:, 1 + k] @ gd
        rb[:, 1 + k] = g @ b[:, 1 + k] @ gd[at(ev[k])]
    for p, (i, j) in enumerate(plaq):
        ra[:, 6 + p] = g @ a[:, 6 + p] @ gd[at(ev[i] + ev[j])]
        rb[:, 6 + p] = g[at(ev[i] + ev[j])] @ b[:, 6 + p] @ gd
    O.set_links(newl)
    val = np.vdot(ra, O.fermion_op(rb, 1))
    assert abs(val - ref) / abs(ref) < 1e-12


@pytest.mark.parametrize("ncol", [2, 3, 4])
def test_det_inverse_identities(ncol):
    rng = np.random.default_rng(ncol)
    O = Oracle(ncol, dims=(2, 2, 2, 2))
    for _ in range(20):
        m = rng.standard_normal((ncol, ncol)) + 1j * rng.standard_normal((ncol, ncol))
        det, inv = O.det_inv(m)
        assert abs(det - np.linalg.det(m)) / abs(det) < 1e-13
        assert np.allclose(m @ inv, np.eye(ncol), atol=1e-12)


def test_generators_normalisation():
    for n in (2, 3, 4):
        lam = Oracle(n, dims=(2, 2, 2, 2)).lambdas()
        gram = np.einsum("aij,bji->ab", lam, lam)
        assert np.allclose(gram, -np.eye(n * n), atol=1e-14)
        # completeness: sum_a L^a_kl L^a_ij = -delta_kj delta_il
        comp = np.einsum("akl,aij->klij", lam, lam)
        want = -np.einsum("kj,il->klij", np.eye(n), np.eye(n))
        assert np.allclose(comp, want, atol=1e-14)


def test_rational_approximation_accuracy():
    """sum_j amp/(x+shift) + ampdeg approximates x^(-1/4) and x^(1/8) within
    the quoted Remez errors on [1e-7, 1000] (setup_rhmc.c:879-881)"""
    O = Oracle(2, dims=(2, 2, 2, 2))
    r = O.rhmc()
    x = np.logspace(-7, 3, 2001)
    f4 = r[0] + sum(r[1 + j] / (x + r[16 + j]) for j in range(15))
    f8 = r[31] + sum(r[32 + j] / (x + r[47 + j]) for j in range(15))
    assert np.max(np.abs(f4 / x ** -0.25 - 1)) < 2.0e-5
    assert np.max(np.abs(f8 / x ** 0.125 - 1)) < 1.1e-5


@pytest.mark.skipif(not os.path.isdir(REF_ROOT), reason="reference tree absent")
def test_rhmc_tables_match_reference_source():
    src = open(os.path.join(REF_ROOT, "susy", "setup_rhmc.c")).read()
    body = src[src.index("void setup_rhmc15()"):src.index("void setup_rhmc()")]
    nums_ref = [float(v) for v in re.findall(r"=\s*(-?[0-9.]+e[-+][0-9]+);", body)]
    # reference order per Nroot: ampdeg4, amp4, shift4, ampdeg8, amp8, shift8;
    # the oracle reads include/susy_rhmc15.h, the table the library uses
    O = Oracle(2, dims=(2, 2, 2, 2), Nroot=1)
    assert np.array_equal(O.rhmc(), np.array(nums_ref[:62]))
    O = Oracle(2, dims=(2, 2, 2, 2), Nroot=2)
    assert np.array_equal(O.rhmc(), np.array(nums_ref[62:124]))


def test_oracle_trajectory_reproduces_hmc_U2_ref(hmc_ref):
    O = hmc_oracle(2, hmc_ref)
    assert O.load_config(config_path(2))[0] == 0
    for tr in hmc_ref["U2"]["ref"]["trajectories"][:2]:
        check_traj(O.update(), tr)


@pytest.mark.slow
@pytest.mark.parametrize("ncol", [3, 4])
def test_oracle_trajectory_reproduces_hmc_ref(ncol, hmc_ref):
    O = hmc_oracle(ncol, hmc_ref)
    assert O.load_config(config_path(ncol))[0] == 0
    check_traj(O.update(), hmc_ref[f"U{ncol}"]["ref"]["trajectories"][0])


@pytest.mark.skipif(not have_ref(), reason="oracle/_ref not built")
def test_oracle_matches_live_reference_on_random_links():
    """Live comparison with the reference build on an irregular lattice with
    periodic BCs and random links (not covered by the fixtures)."""
    kw = dict(dims=(2, 4, 6, 2), PBC=1, iseed=3, kappa_u1=0.7, bmass=0.3, G=0.2)
    R = Ref(**kw)
    links = random_links(R.link_shape(), 8, 0.3)
    R.set_links(links)
    src = random_tf(R.tf_shape(), 9)
    mom = random_tf(R.link_shape(), 10)
    want = {"p": R.fermion_op(src, 1), "m": R.fermion_op(src, -1)}
    R.set_mom(mom)
    gn = R.gauge_force(0.05)
    gm = R.get_mom()
    O = Oracle(2, **kw)
    O.set_links(links)
    assert rel_err(O.fermion_op(src, 1), want["p"]) < 1e-13
    assert rel_err(O.fermion_op(src, -1), want["m"]) < 1e-13
    O.set_mom(mom)
    assert abs(O.gauge_force(0.05) / gn - 1) < 1e-12
    assert rel_err(O.get_mom() - mom, gm - mom) < 1e-12
