"""End-to-end RHMC trajectories on the GPU against the reference's own
regression outputs (testsuite/scalar/hmc.U{2,3,4}.ref): per-term actions to
the 8 printed digits, start/end S to 12 digits, accept/reject, force monitors
and the exact CG iteration count of every trajectory; then the measurements
printed after each trajectory (control.c:85-138): GMES (Polyakov loop,
plaquettes, bosonic action per site, mean |Polyakov|) to 8 digits; FLINK,
the Wilson lines in all five directions (LINES, and LINES_POLAR of the
polar-projected links), DET, WIDTHS, UUBAR_EIG and POLAR_EIG to their 6
printed digits.  Also the loaded regression configurations
(meas.U{2,3,4}.ref) and random links at 8^3 x 16 against the numpy checker
tests/measure_ref.py."""
import json
import os
import sys

import numpy as np
import pytest

import measure_ref
import susy_amd
from oracle_lib import GOLDEN, config_path, random_links
from refcheck import check_traj, hmc_kwargs, printed_ok

MEAS = json.load(open(os.path.join(GOLDEN, "meas_ref.json")))

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ncol", [2, 3, 4])
def test_trajectories_reproduce_hmc_ref(ncol, hmc_ref):
    print("hmc", ncol, flush=True)
    ref = hmc_ref[f"U{ncol}"]
    lat = susy_amd.Lattice(ncol, **hmc_kwargs(ref["input"]))
    assert lat.load_config(config_path(ncol)) == 0
    ss, st = lat.plaquette()
    assert abs(ss / ref["ref"]["check_plaq"][0] - 1) < 1e-14
    for k, tr in enumerate(ref["ref"]["trajectories"]):
        stats = lat.update(susy_amd.FORM_FP64)
        print("traj", k, stats[:5], flush=True)
        sys.stdout.flush()
        check_traj(stats, tr)
        # GMES Re P, Im P, CG iterations, ss and st plaquettes, S_B / V, <|P|>
        gm = tr["gmes"]
        re, im, mod = lat.wilson_line(3)
        ss, st = lat.plaquette()
        sb = lat.actions()[0] / lat.V
        for got, want in zip((re, im, ss, st, sb, mod), (gm[0], gm[1], gm[3], gm[4], gm[5],
                                                         gm[6])):
            assert printed_ok(got, want, 8), ("gmes", k, got, want)
        assert int(stats[4]) == int(gm[2])
        ms = MEAS[f"U{ncol}"]["trajectories"][k]
        lt, ave, width = lat.link_trace()[:3]
        for got, want in zip(list(lt) + [ave, width], ms["flink"]):
            assert printed_ok(got, want, 6), ("flink", k, got, want)
        check_lines(lat, ms, k)
        det5, w3 = lat.plaq_moments()
        check_rows(det5, ms["det"], ("det", k))
        check_rows(w3, ms["widths"], ("widths", k))
        for name, proj in (("uubar_eig", False), ("polar_eig", True)):
            ave, wd, mn, mx = lat.scalar_eig(proj)
            for j in range(ncol):
                check_rows((ave[j], wd[j], mn[j], mx[j]), ms[name][j], (name, k, j))


def check_rows(got, want, what):
    assert len(got) == len(want), what
    for q, (g, w) in enumerate(zip(got, want)):
        assert printed_ok(g, w, 6), (what, q, g, w)


def check_lines(lat, ms, k):
    for name, proj in (("lines", False), ("lines_polar", True)):
        row = []
        for a in range(5):
            re, im, _ = lat.wilson_line(a, proj)
            row += [re, im]
        check_rows(row, ms[name], (name, k))


@pytest.mark.parametrize("ncol", [2, 3, 4])
def test_measurements_of_regression_configs(ncol):
    """FLINK, FLINK_DET, LINES, LINES_POLAR of config.U<N>.4444 as printed
    in meas.U<N>.ref"""
    print("meas config", ncol, flush=True)
    lat = susy_amd.Lattice(ncol, dims=(4, 4, 4, 4))
    assert lat.load_config(config_path(ncol)) == 0
    ms = MEAS[f"U{ncol}"]["config"]
    lt, ave, width, ld, dave, dwidth = lat.link_trace()
    check_rows(list(lt) + [ave, width], ms["flink"], "flink")
    check_rows(list(ld) + [dave, dwidth], ms["flink_det"], "flink_det")
    check_lines(lat, ms, "config")


@pytest.mark.parametrize("ncol", [2, 3, 4])
def test_measurements_match_checker_on_random_links(ncol):
    """8^3 x 16, links I + 0.3 complex Gaussian: every measurement against
    the numpy checker to 1e-12 (several blocks of partial sums, min/max
    across blocks)"""
    print("meas random", ncol, flush=True)
    dims = (8, 8, 8, 16)
    lat = susy_amd.Lattice(ncol, dims=dims)
    U = random_links(lat.link_shape(), 17 + ncol, 0.3)
    lat.set_links(U)
    want = measure_ref.report(dims, U)

    def close(got, w):
        assert np.allclose(got, w, rtol=1e-12, atol=1e-12), (got, w)

    lt, ave, width, ld, dave, dwidth = lat.link_trace()
    close(list(lt) + [ave, width], want["FLINK"])
    close(list(ld) + [dave, dwidth], want["FLINK_DET"])
    for name, proj in (("LINES", False), ("LINES_POLAR", True)):
        row = []
        for a in range(5):
            re, im, _ = lat.wilson_line(a, proj)
            row += [re, im]
        close(row, want[name])
    det5, w3 = lat.plaq_moments()
    close(det5, want["DET"])
    close(w3, want["WIDTHS"])
    for name, proj in (("UUBAR_EIG", False), ("POLAR_EIG", True)):
        ave, wd, mn, mx = lat.scalar_eig(proj)
        close(np.array([ave, wd, mn, mx]).T, want[name])


def test_config_writer_round_trip(tmp_path):
    """a loaded reference configuration written back is byte-identical, and
    the written file loads with valid checksums"""
    path = config_path(2)
    raw = open(path, "rb").read()
    lat = susy_amd.Lattice(2, dims=(4, 4, 4, 4))
    assert lat.load_config(path) == 0
    out = tmp_path / "saved.U2.4444"
    lat.save_config(str(out), stamp=raw[20:84])
    assert open(out, "rb").read() == raw
    lat2 = susy_amd.Lattice(2, dims=(4, 4, 4, 4))
    assert lat2.load_config(str(out)) == 0
    assert (lat2.get_links() == lat.get_links()).all()
