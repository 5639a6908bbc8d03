"""The pure-gauge form (the reference's -DPUREGAUGE build, update_o.c:74-177,
action.c:314: no pseudofermions, solves or fermion forces; the action has no
fermion part) against the same integrator driven step by step through the C
oracle's pieces: momenta from the site generators (grsource.c:44-64), the
Omelyan gauge level of update_gauge_step (update_o.c:42-61) nested in
update_step's outer loop (update_o.c:80-113) with every fermion kick
removed, the bosonic + momentum actions at both ends, and the Metropolis test
with one draw of the node generator (update_o.c:230-243).

U(2) 4^4 from config.U2.4444, lambda 1, bmass 0.5, kappa_u1 0.3 (the det
action and force on), G 0.05, trajectory length 1, nstep 4, nstep_gauge 3,
three trajectories: start S, end S and dS within 1e-10 relative, the same
accept/reject, and the links after each trajectory within 1e-12."""
import math

import numpy as np
import pytest

import susy_amd
from oracle_lib import Oracle, config_path, rel_err

pytestmark = pytest.mark.gpu

KW = dict(dims=(4, 4, 4, 4), PBC=-1, iseed=5, Nroot=1, Norder=15, lam=1.0, kappa_u1=0.3,
          bmass=0.5, fmass=0.0, G=0.05, niter=5000, error_per_site=1e-5, traj_length=1.0,
          nstep=4, nstep_gauge=3)
LAMBDA, TWO_LAMBDA, LAMBDA_MID = 0.193, 0.386, 0.614


def bosonic_action(O):
    g, bm, dt, mom = O.actions()
    tot = g + (bm if KW["bmass"] > 1e-8 else 0.0) + (dt if KW["kappa_u1"] > 1e-8 else 0.0)
    return tot + mom


def oracle_puregauge_trajectory(O):
    O._f("ranmom")()
    start = bosonic_action(O)
    old = O.get_links()
    f_eps = KW["traj_length"] / KW["nstep"]
    g_eps = f_eps / (2.0 * KW["nstep_gauge"])

    def gauge_step(eps):
        n = KW["nstep_gauge"]
        O.gauge_force(eps * LAMBDA)
        for i in range(1, n + 1):
            O.update_u(0.5 * eps)
            O.gauge_force(eps * LAMBDA_MID)
            O.update_u(0.5 * eps)
            O.gauge_force(eps * (TWO_LAMBDA if i < n else LAMBDA))

    for _ in range(KW["nstep"]):
        gauge_step(g_eps)
        gauge_step(g_eps)
    end = bosonic_action(O)
    dS = end - start
    xr = O.rand_stream(-1, 1)[0]
    accept = not (math.exp(-dS) < xr)
    if not accept:
        O.set_links(old)
    return start, end, dS, accept


def test_puregauge_trajectories_match_oracle_driver():
    O = Oracle(2, **KW)
    assert O.load_config(config_path(2))[0] == 0
    lat = susy_amd.Lattice(2, **KW)
    assert lat.load_config(config_path(2)) == 0
    for k in range(3):
        st = lat.update(susy_amd.FORM_PUREGAUGE)
        start, end, dS, accept = oracle_puregauge_trajectory(O)
        print("puregauge", k, st[:4], (start, end, dS, accept), flush=True)
        assert abs(st[0] / start - 1) < 1e-10 and abs(st[1] / end - 1) < 1e-10, k
        assert abs(st[2] - dS) <= 1e-10 * abs(start), k
        assert bool(st[3]) == accept, k
        assert int(st[4]) == 0 and st[7] == 0 and st[17 + 3] == 0, k   # no solves, no fermions
        assert rel_err(lat.get_links(), O.get_links()) < 1e-12, k


def test_puregauge_refuses_the_surrogate():
    lat = susy_amd.Lattice(2, **KW)
    with pytest.raises(RuntimeError):
        lat.update(susy_amd.FORM_PUREGAUGE | susy_amd.FORM_SURROGATE)
