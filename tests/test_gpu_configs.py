"""The workload configurations of the benchmark description, as parity cases
on the GPU:
  * U(2) 4^4 plain FP64 RHMC (lambda 1, bmass 0.5, G 0.05, Nroot 1, length 1,
    nstep 10 / 10, cold start, seed 1234): ten trajectories against the
    numbers the reference's own susy_hmc printed for the same run
    (tests/golden/config0_refbuild.json, every printed digit) and against the
    C oracle's run (tests/golden/config0_ref.json, 1e-11);
  * U(3) 8^3 x 16 kernel level: gather tables and fermion_op against the
    oracle, the 15-shift x^(-1/4) multi-shift CG to relative residual 1e-8
    against the oracle's solve (tests/golden/config1_ref.json), and the FP32
    form of the same solve;
  * U(4) 12^3 x 24 mixed precision: errors of the FP32 form's solutions,
    forces and dH against the FP64 form on the same links, pseudofermion
    source and momenta (short trajectory; the thermalised start of the
    description is replaced by one FP64 trajectory from a cold start).
"""
import json
import os

import numpy as np
import pytest

import measure_ref
import susy_amd
from oracle_lib import GOLDEN, Oracle, random_links, random_tf, rel_err
from refcheck import check_traj, printed_ok

pytestmark = pytest.mark.gpu

CFG0 = json.load(open(os.path.join(GOLDEN, "config0_ref.json")))
CFG1 = json.load(open(os.path.join(GOLDEN, "config1_ref.json")))
CFG0_REF = json.load(open(os.path.join(GOLDEN, "config0_refbuild.json")))


def say(*a):
    print(*a, flush=True)


def test_config0_trajectories_match_oracle():
    kw = dict(CFG0["params"])
    kw["dims"] = tuple(kw["dims"])
    lat = susy_amd.Lattice(2, **kw)
    geo = measure_ref.Geometry(kw["dims"])
    for k, ref in enumerate(CFG0["trajectories"]):
        st = lat.update(susy_amd.FORM_FP64)
        say("config0 traj", k, st[:5])
        tr = CFG0_REF["ref"]["trajectories"][k]
        check_traj(st, tr)
        re_, im_, _ = lat.wilson_line(3)
        ss_, st_ = lat.plaquette()
        for got, want_ in zip((re_, im_, ss_, st_), (tr["gmes"][0], tr["gmes"][1],
                                                    tr["gmes"][3], tr["gmes"][4])):
            assert printed_ok(got, want_, 8), ("gmes", k, got, want_)
        want = ref["stats"]
        assert int(st[4]) == int(want[4]), ("iterations", k)
        assert bool(st[3]) == bool(want[3]), ("accept", k)
        for q in (0, 1) + tuple(range(11, 23)):
            assert abs(st[q] - want[q]) <= 1e-11 * abs(want[q]) + 1e-9, (k, q, st[q], want[q])
        assert abs(st[2] - want[2]) < 1e-8, (k, st[2], want[2])
        for q in range(5, 11):
            assert abs(st[q] - want[q]) <= 1e-9 * abs(want[q]) + 1e-12, (k, q)
        ss, stp = lat.plaquette()
        assert abs(ss - ref["plaq"][0]) < 1e-11 and abs(stp - ref["plaq"][1]) < 1e-11
        assert abs(lat.actions()[0] / lat.V - ref["sb_per_site"]) < 1e-11
        re, im, mod = lat.wilson_line(3)
        assert np.allclose([re, im, mod], ref["ploop"], rtol=0, atol=1e-11), k
    # the FP32 form from the same cold start and seed: same first trajectory
    # up to the solver tolerance
    lat32 = susy_amd.Lattice(2, **kw)
    st = lat32.update(susy_amd.FORM_FP32)
    want = CFG0["trajectories"][0]["stats"]
    say("config0 fp32", st[:5], "fp64", want[:5])
    assert bool(st[3]) == bool(want[3])
    assert abs(st[2] - want[2]) < 1e-3
    assert int(st[4]) <= 1.2 * want[4] + 20


def test_config1_kernel_level():
    c = CFG1
    dims = tuple(c["dims"])
    O = Oracle(3, dims=dims)
    lat = susy_amd.Lattice(3, dims=dims)
    for g in (0, 8, 9, 17, 73):
        assert np.array_equal(lat.gather(g), O.gather(g)), g
    links = random_links(O.link_shape(), c["link_seed"], c["scale"])
    src = random_tf(O.tf_shape(), c["src_seed"])
    src = src / np.sqrt(np.vdot(src, src).real)
    O.set_links(links)
    lat.set_links(links)
    h = lat.tf_alloc(2 + 15 + 15 + 1)
    lat.tf_upload(h, src)
    for sign in (1, -1):
        want = O.fermion_op(src, sign)
        lat.fermion_op(h, h + 1, sign)
        assert rel_err(lat.tf_download(h + 1), want) < 1e-12, sign
        lat.fermion_op(h, h + 1, sign, form=susy_amd.FORM_FP32)
        assert rel_err(lat.tf_download(h + 1), want) < 1e-5, sign
    shifts = np.array(c["shifts"])
    it, size_r = lat.congrad_multi(h, h + 2, shifts, 5000, c["errormin"])
    say("config1 cg64", it, c["iters"])
    assert it == c["iters"]
    # the final |r|^2 (1e-16 of the source's) is the recurrence's: the
    # summation order of the reductions shows at that depth
    assert abs(size_r / c["size_r"] - 1) < 1e-5
    dev = 0.0
    for j in range(15):
        psi = lat.tf_download(h + 2 + j)
        dev = max(dev, abs(np.vdot(psi, psi).real / c["psi_norm2"][j] - 1))
        want = complex(*c["src_dot_psi"][j])
        dev = max(dev, abs(np.vdot(src, psi) - want) / abs(want))
    say("config1 solutions: max relative deviation", dev)
    assert dev < 1e-11
    it32, _ = lat.congrad_multi(h, h + 17, shifts, 5000, c["errormin"], form=susy_amd.FORM_FP32)
    say("config1 cg32", it32)
    assert it32 <= 1.2 * c["iters"] + 20
    # the base system's residual is FP64 (reliable updates, final FP64
    # check); the shifted systems carry FP32 accumulation, |r_j|^2 <= 1e-12
    tmp = h + 32
    for j in (0, 7, 14):
        lat.dsq(h + 17 + j, tmp)
        psi = lat.tf_download(h + 17 + j)
        r = src - (lat.tf_download(tmp) + shifts[j] * psi)
        r2 = np.vdot(r, r).real
        say("config1 cg32 shift", j, "true |r|^2", r2)
        assert r2 <= (4 * c["errormin"] if j == 0 else 1e-12), j


def test_config2_mixed_precision_errors():
    kw = dict(dims=(12, 12, 12, 24), PBC=-1, iseed=7, Nroot=2, Norder=15, lam=1.0,
              kappa_u1=0.0, bmass=0.5, fmass=0.0, G=0.05, niter=5000, error_per_site=1e-5,
              traj_length=0.1, nstep=2, nstep_gauge=10)
    lat = susy_amd.Lattice(4, **kw)
    lat.update(susy_amd.FORM_FP64)
    U0 = lat.get_links()
    rh = Oracle(4, dims=(2, 2, 2, 2), Nroot=2).rhmc()
    assert np.array_equal(lat.rhmc(), rh)
    amps, shifts = rh[1:16], rh[16:31]
    h = lat.tf_alloc(1 + 15 + 15 + 15)
    lat.reseed(11)
    lat.grsource(h, h + 1, susy_amd.FORM_FP64)
    errormin = 1e-10
    it64, _ = lat.congrad_multi(h, h + 1, shifts, 5000, errormin)
    it32, _ = lat.congrad_multi(h, h + 16, shifts, 5000, errormin, form=susy_amd.FORM_FP32)
    sol_err = max(rel_err(lat.tf_download(h + 16 + j), lat.tf_download(h + 1 + j))
                  for j in range(15))
    zero = np.zeros(lat.link_shape(), dtype=complex)
    forces = []
    for hs in (h + 1, h + 16):
        lat.set_mom(zero)
        lat.fermion_force(1.0, hs, amps)
        forces.append(lat.get_mom())
    force_err = rel_err(forces[1], forces[0])
    stats = []
    for form in (susy_amd.FORM_FP64, susy_amd.FORM_FP32):
        lat.set_links(U0)
        lat.reseed(13)
        stats.append((lat.update(form), lat.plaquette()))
    (s64, p64), (s32, p32) = stats
    dh_err = abs(s32[2] - s64[2])
    say("config2 iterations", it64, it32, "solution", sol_err, "force", force_err,
        "dH", s64[2], s32[2], "plaquette", p64, p32)
    assert it32 <= 1.2 * it64 + 20
    assert sol_err < 1e-4
    assert force_err < 1e-4
    assert dh_err < 1e-3 + 1e-3 * abs(s64[2])
    assert abs(p32[0] / p64[0] - 1) < 1e-8 and abs(p32[1] / p64[1] - 1) < 1e-8
