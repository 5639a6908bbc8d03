"""No kernel of the library overwrites a vector-memory store's data registers
before the store has read them (tests/store_check.py): such overwrites made
repeated FP64 DSq / multi-shift CG runs differ in 64-byte groups of the
output.  Compiles every source to device assembly for every gauge group the
library is built for; no GPU."""
import os
import shutil

import pytest

import store_check
from susy_amd import build


@pytest.mark.skipif(shutil.which(build.HIPCC) is None and not os.path.exists(build.HIPCC),
                    reason="hipcc not available")
@pytest.mark.parametrize("ncol", build.NCOLS)
def test_every_kernel_holds_store_data(ncol):
    assert store_check.CHECKED == tuple(build.SOURCES)
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
    # a branch does not end the store's hazard window
    branch = ["_Zb:  ; @b",
              "  global_store_dwordx2 v0, v[0:1], s[2:3] offset:424",
              "  s_cbranch_scc1 .LBB2_6",
              ".LBB2_6:",
              "  v_mov_b32_e32 v1, 0",
              "  s_endpgm"]
    assert store_check.overwrites(branch) == [("_Zb", 1, 4)]
