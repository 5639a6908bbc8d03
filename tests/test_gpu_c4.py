"""GPU parity at the size the benchmark runs: U(4) on 16^3 x 32 (131072
sites; an FP64 Twist_Fermion field is 537 MB) with the benchmark's couplings
(lambda 2, bmass 0.4, G 0.05, kappa_u1 0, Nroot 2) on seeded random links
I + 0.1 (complex Gaussian).

  * all 74 gather tables against the closed form (site index of x + d, the
    offsets d decoded from the oracle's tables on 4^4, which the reference
    build pins bit-exactly);
  * FP64 fermion_op (both signs) and DSq against the C oracle within 1e-12
    relative, FP32 D within 1e-5 of the oracle's FP64 output;
  * one 15-shift solve (the heatbath's shift4 table) in each form, every
    shifted system's true FP64 residual |src - (DSq + s_j) psi_j|^2 below
    4 x the stopping bound;
  * the fermion force of three poles against the oracle within 1e-10.
The oracle is single-threaded C: the whole module takes a few minutes of
host time on the GPU box."""
import sys

import numpy as np
import pytest

import susy_amd
from oracle_lib import Oracle, random_links, random_tf, rel_err
from susy_amd.start_config import site_index

pytestmark = pytest.mark.gpu

DIMS = (16, 16, 16, 32)
KW = dict(dims=DIMS, PBC=-1, iseed=1234, Nroot=2, Norder=15, lam=2.0, kappa_u1=0.0,
          bmass=0.4, fmass=0.0, G=0.05, niter=5000, error_per_site=1e-5, traj_length=1.0,
          nstep=20, nstep_gauge=10)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def c4():
    say("c4: set-up")
    O = Oracle(4, **KW)
    lat = susy_amd.Lattice(4, **KW)
    lk = random_links(lat.link_shape(), 2024, 0.1)
    O.set_links(lk)
    lat.set_links(lk)
    yield O, lat
    lat.close()


def gather_offsets():
    """offset d of every gather id, decoded from the oracle's tables on 4^4"""
    O = Oracle(2, dims=(4, 4, 4, 4))
    xyzt = O.coords()
    i0 = int(np.flatnonzero((xyzt == 1).all(axis=1))[0])
    offs = []
    for g in range(susy_amd.NGATHER):
        d = xyzt[O.gather(g)[i0]] - 1
        assert np.abs(d).max() <= 1
        offs.append(d)
    return offs


def test_c4_gathers_closed_form(c4):
    O, lat = c4
    say("c4: gathers")
    offs = gather_offsets()
    nx, ny, nz, nt = DIMS
    idx = site_index(DIMS)                      # site index of each lexicographic position
    t, z, y, x = np.meshgrid(np.arange(nt), np.arange(nz), np.arange(ny), np.arange(nx),
                             indexing="ij")
    coords = [a.ravel() for a in (x, y, z, t)]
    for g, d in enumerate(offs):
        nb = [(c + int(dd)) % n for c, dd, n in zip(coords, d, DIMS)]
        lex_nb = nb[0] + nx * (nb[1] + ny * (nb[2] + nz * nb[3]))
        want = np.empty(lat.V, dtype=np.int64)
        want[idx] = idx[lex_nb]
        assert np.array_equal(lat.gather(g), want), g


def test_c4_operator(c4):
    O, lat = c4
    src = random_tf(lat.tf_shape(), 77)
    h = lat.tf_alloc(2)
    lat.tf_upload(h, src)
    for sign in (1, -1):
        say("c4: fermion_op", sign)
        want = O.fermion_op(src, sign)
        lat.fermion_op(h, h + 1, sign)
        e64 = rel_err(lat.tf_download(h + 1), want)
        lat.fermion_op(h, h + 1, sign, form=susy_amd.FORM_FP32)
        e32 = rel_err(lat.tf_download(h + 1), want)
        say("c4: fermion_op", sign, "FP64", e64, "FP32", e32)
        assert e64 < 1e-12, (sign, e64)
        assert e32 < 1e-5, (sign, e32)
        del want
    say("c4: DSq")
    want = O.dsq(src)
    lat.dsq(h, h + 1)
    e = rel_err(lat.tf_download(h + 1), want)
    say("c4: DSq", e)
    assert e < 1e-12


@pytest.mark.parametrize("form", [susy_amd.FORM_FP64, susy_amd.FORM_FP32])
def test_c4_solve_true_residuals(c4, form):
    O, lat = c4
    src = random_tf(lat.tf_shape(), 78)
    shifts = lat.rhmc()[16:31]
    err = KW["error_per_site"] ** 2
    h = lat.tf_alloc(1 + 15 + 1)
    lat.tf_upload(h, src)
    it, sz = lat.congrad_multi(h, h + 1, shifts, 5000, err, form=form)
    say("c4: solve form", form, "iterations", it)
    assert 0 < it < 5000
    src_n2 = np.vdot(src, src).real
    for j in range(15):
        lat.dsq(h + 1 + j, h + 16)
        r = src - (lat.tf_download(h + 16) + shifts[j] * lat.tf_download(h + 1 + j))
        rr = np.vdot(r, r).real
        assert rr <= 4 * err * src_n2, (form, j, rr / src_n2)


def test_c4_fermion_force(c4):
    O, lat = c4
    say("c4: fermion force")
    sol = random_tf((3,) + lat.tf_shape(), 79)
    amps = lat.rhmc()[1:4]
    mom = random_tf(lat.link_shape(), 80)
    O.set_mom(mom)
    lat.set_mom(mom)
    h = lat.tf_alloc(3)
    for j in range(3):
        lat.tf_upload(h + j, sol[j])
    n = lat.fermion_force(0.03, h, amps)
    n_o = O.fermion_force(0.03, sol, amps)
    e = rel_err(lat.get_mom() - mom, O.get_mom() - mom)
    say("c4: fermion force", e, n / n_o - 1)
    assert abs(n / n_o - 1) < 1e-10
    assert e < 1e-10
