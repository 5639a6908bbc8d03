"""CPU pin of the oracle-driven pole-split integrator (tests/split_ref.py)
that the GPU split form is compared with: the split changes only the
integrator, not the Hamiltonian (the outer level carries S_F - S_mid for any
middle weights), so the energy error of a trajectory must fall as the square
of the step (second-order Omelyan levels) and go to zero.
A pole counted on both levels, a wrong sign or a missing kick leaves an O(1)
error that does not shrink (the double-counted variant gives dS = 492 and
493 at the two step sizes).

U(2) 4^4 from config.U2.4444, lambda 1, bmass 0.5, kappa_u1 0.3, G 0.05,
Nroot 1, trajectory length 1, nstep_gauge 2, poles 10..14 on one middle step,
the same momenta and pseudofermions (fresh generators) at nstep 8 and 16."""
import pytest

from oracle_lib import Oracle, config_path
from split_ref import oracle_split_trajectory

KW = dict(dims=(4, 4, 4, 4), PBC=-1, iseed=5, Nroot=1, Norder=15, lam=1.0, kappa_u1=0.3,
          bmass=0.5, fmass=0.0, G=0.05, niter=5000, error_per_site=1e-5, traj_length=1.0,
          nstep_gauge=2)


@pytest.mark.parametrize("wscale", [1.0, 0.7])
def test_split_energy_error_falls_as_step_squared(wscale):
    """any middle weights leave the Hamiltonian unchanged (the outer level
    carries S_F - S_mid(w)); checked at the rational amplitudes and at 0.7
    of them, the weighting the GPU test uses"""
    dS = {}
    for ns in (8, 16):
        kw = dict(KW, nstep=ns)
        O = Oracle(2, **kw)
        assert O.load_config(config_path(2))[0] == 0
        r = oracle_split_trajectory(O, kw, 10, 1, wscale * O.rhmc()[11:16])
        dS[ns] = r[2]
        assert r[5] > 0 and r[4] > r[5]     # middle solves ran, and the outer ones
    ratio = dS[8] / dS[16]
    assert 3.0 < ratio < 5.0, dS
    assert abs(dS[16]) < 0.05, dS
