"""CPU pins of the measurement checker (tests/measure_ref.py) against the
numbers the reference prints: the loaded regression configurations
(meas.U{2,3,4}.ref, before that driver smears) and the first trajectory of
hmc.U2.ref, whose links the oracle reproduces (its own pin in
test_oracle.py)."""
import json
import os

import numpy as np
import pytest

import measure_ref as M
from oracle_lib import Oracle, config_path
from refcheck import hmc_kwargs, printed_ok

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MEAS = json.load(open(os.path.join(GOLDEN, "meas_ref.json")))
DIMS = (4, 4, 4, 4)


def check_rows(got, want, name):
    assert len(got) == len(want), name
    for k, (g, w) in enumerate(zip(got, want)):
        assert printed_ok(g, w, 6), (name, k, g, w)


@pytest.mark.parametrize("ncol", [2, 3, 4])
def test_measurements_of_regression_configs(ncol):
    O = Oracle(ncol, dims=DIMS, iseed=41)
    assert O.load_config(config_path(ncol))[0] == 0
    rep = M.report(DIMS, O.get_links())
    want = MEAS[f"U{ncol}"]["config"]
    for name in ("FLINK", "FLINK_DET", "LINES", "LINES_POLAR"):
        check_rows(rep[name], want[name.lower()], name)


def test_measurements_after_hmc_U2_trajectory(hmc_ref):
    O = Oracle(2, **hmc_kwargs(hmc_ref["U2"]["input"]))
    assert O.load_config(config_path(2))[0] == 0
    O.update()
    U = O.get_links()
    rep = M.report(DIMS, U)
    want = MEAS["U2"]["trajectories"][0]
    for name in ("FLINK", "LINES", "LINES_POLAR", "DET", "WIDTHS"):
        check_rows(rep[name], want[name.lower()], name)
    for name in ("UUBAR_EIG", "POLAR_EIG"):
        for j in range(2):
            check_rows(rep[name][j], want[name.lower()][j], (name, j))
    # the checker's plaquette determinants are the oracle's own
    pd, _, _, _ = O.fields()
    assert np.allclose(M.plaqdet(M.Geometry(DIMS), U), pd, rtol=0, atol=1e-13)


def test_measurement_identities():
    """closed forms: unit links give unit Wilson lines (trace N), zero
    scalar eigenvalues and plaquette determinants 1; a scaled unitary link
    c W has polar factor W and log-eigenvalues log c"""
    N, dims = 3, (2, 2, 2, 4)
    geo = M.Geometry(dims)
    U = np.broadcast_to(np.eye(N), (geo.V, 5, N, N)).astype(complex)
    for a in range(5):
        for proj in (False, True):
            s, mod = M.wilson_line(geo, U, a, proj)
            assert abs(s - N) < 1e-14 and abs(mod - N) < 1e-14
    assert np.allclose(M.measure_det(M.plaqdet(geo, U)), (1, 0, 1, 0, 0))
    ave, wd, mn, mx = M.scalar_eig(U)
    assert np.allclose([ave, wd, mn, mx], 0)
    rng = np.random.default_rng(3)
    Z = rng.standard_normal((geo.V, 5, N, N)) + 1j * rng.standard_normal((geo.V, 5, N, N))
    W, _ = np.linalg.qr(Z)
    c = np.exp(rng.uniform(-0.3, 0.3, (geo.V, 5)))
    cw = c[..., None, None] * W
    assert np.allclose(M.polar_unitary(cw), W, atol=1e-13)
    ave, wd, mn, mx = M.scalar_eig(cw, project=True)
    lc = np.log(c).ravel()
    for j in range(N):
        assert abs(ave[j] - lc.mean()) < 1e-13 and abs(mn[j] - lc.min()) < 1e-13
    # UUBAR eigenvalues are traceless: their averages sum to zero
    ave, _, _, _ = M.scalar_eig(np.eye(N) + 0.3 * Z)
    assert abs(ave.sum()) < 1e-12
