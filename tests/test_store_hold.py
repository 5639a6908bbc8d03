"""The operator and CG kernels never overwrite a vector-memory store's data
registers before the store has read them (susy_amd/store_check.py): such
overwrites made repeated FP64 DSq / multi-shift CG runs differ in 64-byte
groups of the output.  Compiles the sources to device assembly; no GPU."""
import os
import shutil

import pytest

from susy_amd import build, store_check


@pytest.mark.skipif(shutil.which(build.HIPCC) is None and not os.path.exists(build.HIPCC),
                    reason="hipcc not available")
@pytest.mark.parametrize("ncol", [2, 4])
def test_operator_and_cg_hold_store_data(ncol):
    found = store_check.check(ncol=ncol)
    assert not found, found


def test_checker_finds_the_overwrite():
    lines = ["_Zk:  ; @k",
             "  buffer_store_dwordx4 v[2:5], v1, s[24:27], s33 offen",
             "  v_mul_f64 v[4:5], v[34:35], v[16:17]",
             "  s_endpgm"]
    assert store_check.overwrites(lines) == [("_Zk", 1, 2)]
    held = lines[:2] + ["  s_waitcnt expcnt(0)"] + lines[2:]
    assert store_check.overwrites(held) == []
    glob = ["_Zg:  ; @g",
            "  global_store_dwordx2 v[6:7], v[0:1], off",
            "  v_add_f64 v[0:1], v[0:1], 1.0",
            "  s_endpgm"]
    assert store_check.overwrites(glob) == [("_Zg", 1, 2)]
