"""GPU parity of the learned surrogate (features, action, force, force-matching
gradient) against the numpy checker tests/surrogate_ref.py, with the weights
the kernels use (bfloat16-rounded W1, W2, w3).  Activations are FP32 on the
GPU, so values agree to FP32 accumulation error: 1e-5 relative for features,
action and force, 1e-4 for the parameter gradient.  Then the integrator:
the surrogate form still integrates the exact Hamiltonian (dH falls with
the step size like the plain form's), and training on recorded pairs lowers
the force-matching loss."""
import sys

import numpy as np
import pytest

import susy_amd
import surrogate_ref as sr
from oracle_lib import random_links

pytestmark = pytest.mark.gpu

KW = dict(PBC=-1, iseed=41, kappa_u1=0.0, bmass=0.5, G=0.05, lam=1.0)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a - b)) / np.linalg.norm(np.ravel(b)))


def setup(ncol, dims=(2, 4, 2, 4), seed=3, eps=0.3, scale=2.0):
    lat = susy_amd.Lattice(ncol, dims=dims, **KW)
    U = random_links(lat.link_shape(), seed, eps)
    lat.set_links(U)
    lat.sg_init(seed=5, scale=scale)
    th = sr.unpack(lat.sg_get_params())
    th["b1"] = np.random.default_rng(6).standard_normal(sr.NCH) * 0.2
    th["b2"] = np.random.default_rng(7).standard_normal(sr.NCH) * 0.2
    lat.sg_set_params(sr.pack(th))
    return lat, U, sr.weights_used(sr.unpack(lat.sg_get_params())), sr.Geometry(dims)


@pytest.mark.parametrize("ncol", [2, 3, 4])
def test_features_action_force(ncol):
    say("surrogate eval", ncol)
    lat, U, th, geo = setup(ncol)
    phi = sr.features(geo, U)
    assert rel(lat.sg_features(), phi) < 1e-6
    S_ref, F_ref = sr.action_and_force(geo, th, U)
    S, F = lat.sg_eval()
    assert abs(S / S_ref - 1) < 1e-5
    assert rel(F, F_ref) < 1e-5


@pytest.mark.parametrize("ncol,dims", [(2, (2, 4, 2, 4)), (4, (2, 4, 2, 4)),
                                       (2, (6, 6, 6, 6))])
def test_force_matching_gradient(ncol, dims):
    """(6^4: 1296 sites, two site chunks of the weight-gradient reduction,
    the last one ragged)"""
    say("surrogate grad", ncol, dims)
    lat, U, th, geo = setup(ncol, dims=dims)
    Ft = random_links(lat.link_shape(), 9, 0.5) - np.eye(ncol)
    loss, g = lat.sg_loss_grad(U, Ft)
    loss_ref, g_ref = sr.force_loss_grad(geo, th, U, Ft)
    assert abs(loss / loss_ref - 1) < 1e-5
    assert rel(g, sr.pack(g_ref).astype(np.float64)) < 1e-4


def test_surrogate_form_integrates_the_exact_hamiltonian():
    """inner level S_g + S_theta, outer S_F - S_theta: the splitting leaves the
    Hamiltonian unchanged, so |dH| must fall ~ eps^2 for any parameters"""
    say("surrogate dH")
    dh = {}
    for nstep in (4, 8):
        lat = susy_amd.Lattice(2, dims=(4, 4, 4, 4), traj_length=0.5, nstep=nstep, nstep_gauge=4,
                               **KW)
        lat.set_links(random_links(lat.link_shape(), 21, 0.2))
        lat.sg_init(seed=2, scale=3.0)
        lat.reseed(77)
        st = lat.update(susy_amd.FORM_FP64 | susy_amd.FORM_SURROGATE)
        dh[nstep] = abs(st[2])
        say("nstep", nstep, "dH", st[2])
        lat.close()
    assert dh[8] < 0.5 * dh[4]


def test_training_lowers_the_loss():
    say("surrogate training")
    lat = susy_amd.Lattice(2, dims=(4, 4, 4, 4), traj_length=0.5, nstep=4, nstep_gauge=4, **KW)
    lat.set_links(random_links(lat.link_shape(), 23, 0.2))
    lat.sg_init(seed=4, scale=1.0)
    lat.sg_record(40)
    lat.reseed(5)
    for _ in range(3):
        lat.update(susy_amd.FORM_FP64)
    n = lat.sg_pairs()
    assert n == 3 * (2 * 4 + 1)
    before = np.mean([lat.sg_pair_loss(i) for i in range(n - 4, n)])
    hist = lat.sg_train(0, n - 4, steps=150, batch=2, lr=3e-3, seed=1)
    after = np.mean([lat.sg_pair_loss(i) for i in range(n - 4, n)])
    say("loss before", before, "after", after, "train", hist[0], hist[-1])
    assert np.all(np.isfinite(hist))
    assert np.mean(hist[-20:]) < np.mean(hist[:10])
    assert after < before
