"""The drop-in boundary end to end: the reference's own susy_hmc driver (its
objects compiled in place from the reference sources by oracle/ref.mk) with
its definitions of fermion_op, DSq, congrad_multi, fermion_force and
gauge_force localised and libsusy_hip_U2.so linked in their place
(oracle/_ref/susy_hmc_U2_hip; binding through oracle/hip_bind.c), run on the
input of the reference's regression test testsuite/in.U2.hmc.  Every
trajectory's printed action parts, ACCEPT/REJECT line, IT_PER_TRAJ, force
monitors and GMES line must agree with testsuite/scalar/hmc.U2.ref
(tests/golden/hmc_ref.json) to every printed digit.

The CPU test checks the link step: the five symbols of the reference's
objects are local in the binary and the library provides them."""
import os
import re
import subprocess

import pytest

from oracle_lib import ORACLE_DIR, config_path
from refcheck import printed_ok

BINARY = os.path.join(ORACLE_DIR, "_ref", "susy_hmc_U2_hip")
DROPIN = ("fermion_op", "DSq", "congrad_multi", "fermion_force", "gauge_force")
NUM = r"([-+0-9.eE]+)"


def input_text(inp, config):
    keys = ["prompt", "nx", "ny", "nz", "nt", "PBC", "iseed", "Nroot", "Norder", "warms",
            "trajecs", "traj_length", "nstep", "nstep_gauge", "traj_between_meas", "lambda",
            "kappa_u1", "bmass", "fmass", "G", "max_cg_iterations", "error_per_site"]
    return "".join(f"{k} {inp[k]}\n" for k in keys) + f"reload_serial {config}\nforget\n"


@pytest.mark.skipif(not os.path.exists(BINARY), reason="drop-in driver not built")
def test_dropin_driver_links_the_library():
    out = subprocess.run(["nm", BINARY], stdout=subprocess.PIPE, text=True, check=True).stdout
    for sym in DROPIN:
        kinds = {ln.split()[-2] for ln in out.splitlines() if ln.split()[-1:] == [sym]}
        # the reference's copy is local ("t"); calls are resolved from the library ("U")
        assert kinds == {"t", "U"}, (sym, kinds)
    ldd = subprocess.run(["ldd", BINARY], stdout=subprocess.PIPE, text=True).stdout
    assert "libsusy_hip_U2.so" in ldd


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(BINARY), reason="drop-in driver not built")
def test_reference_driver_on_the_library_reproduces_hmc_ref(hmc_ref, tmp_path):
    ref = hmc_ref["U2"]
    (tmp_path / "in.U2.hmc").write_text(input_text(ref["input"], config_path(2)))
    print("dropin driver: running", BINARY, flush=True)
    with open(tmp_path / "in.U2.hmc") as f:
        p = subprocess.run([BINARY], stdin=f, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                           text=True, timeout=300, cwd=str(tmp_path))
    out = p.stdout
    print(out[-3000:], flush=True)
    assert p.returncode == 0
    assert "hip_bind: fermion_op, DSq, congrad_multi, fermion_force and gauge_force" in out
    actions = re.findall(r"^action: gauge " + NUM + " bmass " + NUM + " det " + NUM +
                         " fermion0 " + NUM + " mom " + NUM + " sum " + NUM, out, re.M)
    decisions = re.findall(r"^(ACCEPT|REJECT): delta S = " + NUM + " start S = " + NUM +
                           " end S = " + NUM, out, re.M)
    iters = re.findall(r"^IT_PER_TRAJ (\d+)", out, re.M)
    gforce = re.findall(r"^MONITOR_FORCE_GAUGE\s+" + NUM + " " + NUM, out, re.M)
    fforce = re.findall(r"^MONITOR_FORCE_FERMION0\s+" + NUM + " " + NUM, out, re.M)
    gmes = re.findall(r"^GMES " + " ".join([NUM] * 7), out, re.M)
    trajs = ref["ref"]["trajectories"]
    assert len(decisions) == len(trajs) and len(actions) == 2 * len(trajs)
    for k, tr in enumerate(trajs):
        for got, want in zip(actions[2 * k], tr["start_action"]):
            assert printed_ok(float(got), want, 8), ("start action", k, got, want)
        for got, want in zip(actions[2 * k + 1], tr["end_action"]):
            assert printed_ok(float(got), want, 8), ("end action", k, got, want)
        word, ds, s0, s1 = decisions[k]
        assert (word == "ACCEPT") == tr["accept"], k
        assert printed_ok(float(ds), tr["delta_S"], 4)
        assert printed_ok(float(s0), tr["start_S"], 12) and printed_ok(float(s1), tr["end_S"], 12)
        assert int(iters[k]) == tr["iters"], k
        assert printed_ok(float(gforce[k][0]), tr["gauge_force"][0], 4)
        assert printed_ok(float(fforce[k][0]), tr["fermion_force"][0], 4)
        for got, want in zip(gmes[k], tr["gmes"]):
            assert printed_ok(float(got), want, 8), ("gmes", k, got, want)
