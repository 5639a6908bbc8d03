"""The workload drivers susy_amd.surrogate_run and susy_amd.mixed_run run end
to end at small sizes and report what their full-size runs report: the
surrogate run's held-out force loss falls with training and every form's
<exp(-dH)> is finite; the mixed-precision run's FP32 solutions, force and dH
agree with the FP64 ones to the bounds the configuration test uses."""
import contextlib
import io
import json
import math

import pytest

from susy_amd import mixed_run, surrogate_run

pytestmark = pytest.mark.gpu


def run(main, argv):
    buf = io.StringIO()
    with contextlib.redirect_stdout(buf):
        main(argv)
    lines = [l for l in buf.getvalue().splitlines() if l.startswith("{")]
    print(lines[-1][:300], flush=True)
    return json.loads(lines[-1])


def test_surrogate_run_small():
    print("surrogate_run", flush=True)
    res = run(surrogate_run.main, ["--train-traj", "3", "--pairs", "30", "--held", "6",
                                   "--adam", "40", "--traj", "2", "--nsteps", "10",
                                   "--sg-nsteps", "10"])
    before, after = res["training"]["held_out_loss"]
    assert after < before
    assert set(res["forms"]) == {"fp64@nstep10", "fp32@nstep10", "fp32_surrogate@nstep10"}
    for f in res["forms"].values():
        assert math.isfinite(f["exp_minus_dH"]) and f["iters_per_traj"] > 0


def test_mixed_run_small():
    print("mixed_run", flush=True)
    res = run(mixed_run.main, ["--therm", "1", "--dims", "4,4,4,8", "--traj-length", "0.1",
                               "--nstep", "2"])
    assert res["solution_rel_err"] < 1e-4 and res["force_rel_err"] < 1e-4
    assert res["dH_abs_err"] < 1e-3 + 1e-3 * abs(res["trajectory"]["fp64"]["dH"])
    assert res["plaquette_rel_err"] < 1e-8
    s = res["solve"]
    assert s["fp32"]["iters"] <= 1.2 * s["fp64"]["iters"] + 20
