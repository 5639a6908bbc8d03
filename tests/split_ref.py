"""The pole-split integrator (SUSY_FORM_SPLIT) restated step by step with the
C oracle's pieces, for the tests.

Trajectory frame as in update_o.c:127-282 (momenta, pseudofermions and their
solves, start action, Metropolis test with one draw of the node generator,
links restored on a rejection); the outer level is update_step's Omelyan loop
(update_o.c:80-113) with the fermion force of every pole j taken at amp4_j
for j < first and amp4_j - w_j above; in place of each gauge-level call
(update_o.c:42-61) a middle level of nmid Omelyan steps: kicks by the force
of the poles j >= first with weights w_j, from a multi-shift solve over those
shifts only, and the whole gauge level twice per step."""
import math

import numpy as np

from oracle_lib import _dp

LAMBDA, TWO_LAMBDA, LAMBDA_MID = 0.193, 0.386, 0.614


def total_action(O, srcs, sols, KW):
    g, bm, dt, mom = O.actions()
    tot = g + (bm if KW["bmass"] > 1e-8 else 0.0) + (dt if KW["kappa_u1"] > 1e-8 else 0.0)
    return tot + sum(O.fermion_action(s, p) for s, p in zip(srcs, sols)) + mom


def oracle_split_trajectory(O, KW, first, nmid, w):
    rh = O.rhmc()
    amp4, shift4 = rh[1:16], rh[16:31]
    amp_out = amp4.copy()
    amp_out[first:] -= w
    err, niter, Nr = KW["error_per_site"] ** 2, KW["niter"], KW["Nroot"]
    count = {"all": 0, "mid": 0}
    O._f("ranmom")()
    srcs = []
    for _ in range(Nr):
        g = np.zeros(O.tf_shape(), dtype=np.complex128)
        count["all"] += O._f("grsource")(_dp(g))
        srcs.append(g)
    sols = [None] * Nr

    def solve_all():
        for n in range(Nr):
            it, _, sols[n] = O.congrad_multi(srcs[n], shift4, niter, err)
            count["all"] += it

    def outer_kick(eps):
        for n in range(Nr):
            O.fermion_force(eps, sols[n], amp_out)

    def mid_kick(eps):
        for n in range(Nr):
            it, _, ps = O.congrad_multi(srcs[n], shift4[first:], niter, err)
            count["all"] += it
            count["mid"] += it
            O.fermion_force(eps, ps, w)

    def gauge_level(eps):
        n = KW["nstep_gauge"]
        O.gauge_force(eps * LAMBDA)
        for i in range(1, n + 1):
            O.update_u(0.5 * eps)
            O.gauge_force(eps * LAMBDA_MID)
            O.update_u(0.5 * eps)
            O.gauge_force(eps * (TWO_LAMBDA if i < n else LAMBDA))

    def mid_level(length):
        m = length / nmid
        g = m / (2.0 * KW["nstep_gauge"])
        mid_kick(m * LAMBDA)
        for i in range(1, nmid + 1):
            gauge_level(g)
            mid_kick(m * LAMBDA_MID)
            gauge_level(g)
            mid_kick(m * (TWO_LAMBDA if i < nmid else LAMBDA))

    solve_all()
    start = total_action(O, srcs, sols, KW)
    old = O.get_links()
    f_eps = KW["traj_length"] / KW["nstep"]
    outer_kick(f_eps * LAMBDA)
    for s in range(1, KW["nstep"] + 1):
        mid_level(0.5 * f_eps)
        solve_all()
        outer_kick(f_eps * LAMBDA_MID)
        mid_level(0.5 * f_eps)
        solve_all()
        outer_kick(f_eps * (TWO_LAMBDA if s < KW["nstep"] else LAMBDA))
    end = total_action(O, srcs, sols, KW)
    dS = end - start
    accept = not (math.exp(-dS) < O.rand_stream(-1, 1)[0])
    if not accept:
        O.set_links(old)
    return start, end, dS, accept, count["all"], count["mid"]
