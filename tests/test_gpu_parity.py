"""GPU parity of the HIP library against the C oracle on the same inputs,
and end-to-end against the reference's regression outputs.

Tolerances: FP64 operator and fields 1e-12 relative;
CG solutions 1e-10 relative with identical iteration counts; forces 1e-10;
FP32 operator 1e-5 relative to FP64."""
import sys

import numpy as np
import pytest

import susy_amd
from oracle_lib import Oracle, config_path, random_links, random_tf, rel_err

pytestmark = pytest.mark.gpu

HMC_KW = dict(iseed=41, kappa_u1=1.0, bmass=1.0, G=0.1)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def pair(ncol, cfg=True, dims=(4, 4, 4, 4), **kw):
    args = dict(HMC_KW, dims=dims)
    args.update(kw)
    O = Oracle(ncol, **args)
    lat = susy_amd.Lattice(ncol, **args)
    if cfg:
        assert O.load_config(config_path(ncol))[0] == 0
        assert lat.load_config(config_path(ncol)) == 0
    return O, lat


@pytest.mark.parametrize("ncol", [2, 3, 4])
def test_gather_tables(ncol):
    O, lat = pair(ncol, cfg=False, dims=(2, 4, 6, 8))
    for g in range(susy_amd.NGATHER):
        assert np.array_equal(lat.gather(g), O.gather(g)), g


@pytest.mark.parametrize("ncol", [2, 3, 4])
def test_links_and_fields(ncol):
    say("fields", ncol)
    O, lat = pair(ncol)
    assert rel_err(lat.get_links(), O.get_links()) == 0
    pd, ui, dm, fm = lat.fields()
    opd, oui, odm, ofm = O.fields()
    mask = ~np.eye(5, dtype=bool)
    assert rel_err(pd[mask], opd[mask]) < 1e-12
    assert rel_err(ui, oui) < 1e-12
    assert rel_err(dm, odm) < 1e-12
    assert rel_err(fm, ofm) < 1e-12
    ss, st = lat.plaquette()
    oss, ost = O.plaquette()
    assert abs(ss / oss - 1) < 1e-13 and abs(st / ost - 1) < 1e-13


@pytest.mark.parametrize("ncol", [2, 3, 4])
@pytest.mark.parametrize("links", ["config", "random"])
def test_fermion_op(ncol, links):
    say("fermion_op", ncol, links)
    if links == "config":
        O, lat = pair(ncol)
    else:
        O, lat = pair(ncol, cfg=False, dims=(2, 4, 2, 6), PBC=1, G=0.3)
        lk = random_links(O.link_shape(), 31 + ncol, 0.3)
        O.set_links(lk)
        lat.set_links(lk)
    src = random_tf(O.tf_shape(), 7)
    hs, hd = lat.tf_alloc(2), None
    hd = hs + 1
    lat.tf_upload(hs, src)
    for sign in (1, -1):
        lat.fermion_op(hs, hd, sign)
        want = O.fermion_op(src, sign)
        got = lat.tf_download(hd)
        assert rel_err(got, want) < 1e-12, (sign, rel_err(got, want))
        lat.fermion_op(hs, hd, sign, form=susy_amd.FORM_FP32)
        got32 = lat.tf_download(hd)
        assert rel_err(got32, want) < 1e-5, (sign, rel_err(got32, want))
    lat.dsq(hs, hd)
    assert rel_err(lat.tf_download(hd), O.dsq(src)) < 1e-12


@pytest.mark.parametrize("ncol", [2, 3, 4])
@pytest.mark.parametrize("window", [1, 16, 32])
def test_congrad_multi(ncol, window):
    say("cg", ncol, window)
    susy_amd.set_cg_options(ncol, window=window)
    O, lat = pair(ncol)
    src = random_tf(O.tf_shape(), 3)
    shifts = O.rhmc()[16:31]
    it_o, sz_o, ps_o = O.congrad_multi(src, shifts, 5000, 1e-10)
    hs = lat.tf_alloc(1 + 15)
    lat.tf_upload(hs, src)
    it, sz = lat.congrad_multi(hs, hs + 1, shifts, 5000, 1e-10)
    assert it == it_o
    assert abs(sz / sz_o - 1) < 1e-6
    for j in range(15):
        assert rel_err(lat.tf_download(hs + 1 + j), ps_o[j]) < 1e-10, j
    susy_amd.set_cg_options(ncol)


@pytest.mark.parametrize("ncol", [2, 3, 4])
def test_congrad_multi_fp32_form(ncol):
    """FP32 solve with FP64 reliable updates reaches the FP64 tolerance:
    every shifted system's true FP64 residual is below the stopping bound."""
    say("cg32", ncol)
    O, lat = pair(ncol)
    src = random_tf(O.tf_shape(), 3)
    shifts = O.rhmc()[16:31]
    hs = lat.tf_alloc(1 + 15 + 1)
    lat.tf_upload(hs, src)
    err = 1e-10
    it, sz = lat.congrad_multi(hs, hs + 1, shifts, 5000, err, form=susy_amd.FORM_FP32)
    it_o = O.congrad_multi(src, shifts, 5000, err)[0]
    assert it <= 1.2 * it_o + 20
    tmp = hs + 16
    src_n2 = np.vdot(src, src).real
    for j in range(15):
        lat.dsq(hs + 1 + j, tmp)
        psi = lat.tf_download(hs + 1 + j)
        r = src - (lat.tf_download(tmp) + shifts[j] * psi)
        assert np.vdot(r, r).real <= 4 * err * src_n2, j


@pytest.mark.parametrize("offset", [-1, 0, 1])
def test_congrad_multi_fp32_reliable_at_convergence(offset):
    """the solver enqueues iterations ahead of its convergence test: a reliable
    update falling on the converged iteration (or just after it) must leave
    the solutions intact"""
    say("cg32 reliable", offset)
    O, lat = pair(2)
    src = random_tf(O.tf_shape(), 3)
    shifts = O.rhmc()[16:31]
    err = 1e-10
    it_o = O.congrad_multi(src, shifts, 5000, err)[0]
    hs = lat.tf_alloc(1 + 15 + 1)
    lat.tf_upload(hs, src)
    for window in (32, 16, 1):
        susy_amd.set_cg_options(2, window=window, reliable=it_o + offset)
        it, sz = lat.congrad_multi(hs, hs + 1, shifts, 5000, err, form=susy_amd.FORM_FP32)
        assert it <= 1.2 * it_o + 20, (window, it, it_o)
        src_n2 = np.vdot(src, src).real
        for j in (0, 7, 14):
            lat.dsq(hs + 1 + j, hs + 16)
            psi = lat.tf_download(hs + 1 + j)
            r = src - (lat.tf_download(hs + 16) + shifts[j] * psi)
            assert np.vdot(r, r).real <= 4 * err * src_n2, (window, j)
    susy_amd.set_cg_options(2)


@pytest.mark.parametrize("ncol", [2, 3, 4])
def test_forces(ncol):
    say("forces", ncol)
    O, lat = pair(ncol)
    mom = random_tf(O.link_shape(), 5)
    O.set_mom(mom)
    lat.set_mom(mom)
    n_o = O.gauge_force(0.07)
    n = lat.gauge_force(0.07)
    assert abs(n / n_o - 1) < 1e-12
    assert rel_err(lat.get_mom() - mom, O.get_mom() - mom) < 1e-12
    sol = random_tf((15,) + O.tf_shape(), 8)
    amps = O.rhmc()[1:16]
    O.set_mom(mom)
    lat.set_mom(mom)
    n_o = O.fermion_force(0.05, sol, amps)
    hs = lat.tf_alloc(15)
    for j in range(15):
        lat.tf_upload(hs + j, sol[j])
    n = lat.fermion_force(0.05, hs, amps)
    assert abs(n / n_o - 1) < 1e-10
    assert rel_err(lat.get_mom() - mom, O.get_mom() - mom) < 1e-10


@pytest.mark.parametrize("ncol", [2, 3, 4])
def test_update_u_and_actions(ncol):
    say("update_u", ncol)
    O, lat = pair(ncol)
    mom = random_tf(O.link_shape(), 6) * 0.1
    O.set_mom(mom)
    lat.set_mom(mom)
    O.update_u(0.03)
    lat.update_u(0.03)
    assert rel_err(lat.get_links(), O.get_links()) < 1e-14
    a, ao = lat.actions(), O.actions()
    for k in range(4):
        assert abs(a[k] / ao[k] - 1) < 1e-12, k


@pytest.mark.parametrize("ncol", [2, 4])
def test_larger_volume(ncol):
    """8^3 x 16: many site chunks per XCD and larger buffer offsets"""
    say("large volume", ncol)
    O, lat = pair(ncol, cfg=False, dims=(8, 8, 8, 16), PBC=-1, G=0.05, bmass=0.4, lam=2.0,
                  kappa_u1=0.3)
    lk = random_links(O.link_shape(), 41, 0.2)
    O.set_links(lk)
    lat.set_links(lk)
    src = random_tf(O.tf_shape(), 42)
    h = lat.tf_alloc(5)
    lat.tf_upload(h, src)
    for sign in (1, -1):
        want = O.fermion_op(src, sign)
        lat.fermion_op(h, h + 1, sign)
        assert rel_err(lat.tf_download(h + 1), want) < 1e-12
        lat.fermion_op(h, h + 1, sign, form=susy_amd.FORM_FP32)
        got32 = lat.tf_download(h + 1)
        assert rel_err(got32, want) < 1e-5
        # the block order is a schedule only: every order is bit-identical
        # (8^3 x 16 admits order 2: 2 chunks per t-slice and parity)
        lat.fermion_op(h, h + 1, sign)
        got64 = lat.tf_download(h + 1)
        for order in (0, 1, 2, 1 + (2 << 8), 2 + (0 << 8) + 256):
            lat.set_site_order(order)
            lat.fermion_op(h, h + 1, sign, form=susy_amd.FORM_FP32)
            assert np.array_equal(lat.tf_download(h + 1), got32), order
            lat.fermion_op(h, h + 1, sign)
            assert np.array_equal(lat.tf_download(h + 1), got64), order
        lat.set_site_order(susy_amd.DEFAULT_SITE_ORDER)
    mom = random_tf(O.link_shape(), 43)
    O.set_mom(mom)
    lat.set_mom(mom)
    assert abs(lat.gauge_force(0.02) / O.gauge_force(0.02) - 1) < 1e-12
    assert rel_err(lat.get_mom() - mom, O.get_mom() - mom) < 1e-12
    sol = random_tf((3,) + O.tf_shape(), 44)
    amps = O.rhmc()[1:4]
    O.set_mom(mom)
    lat.set_mom(mom)
    for j in range(3):
        lat.tf_upload(h + 2 + j, sol[j])
    n = lat.fermion_force(0.03, h + 2, amps)
    assert abs(n / O.fermion_force(0.03, sol, amps) - 1) < 1e-10
    assert rel_err(lat.get_mom() - mom, O.get_mom() - mom) < 1e-10
    a, ao = lat.actions(), O.actions()
    for k in range(4):
        assert abs(a[k] / ao[k] - 1) < 1e-11, k
