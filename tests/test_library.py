"""CPU checks of the C-ABI libraries: every symbol declared in
include/susy_hip.h is exported by libsusy_hip_U{2,3,4}.so (no compute calls:
this container has no GPU)."""
import ctypes
import os
import re

import pytest

import susy_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "susy_hip.h")).read()
    text = re.sub(r"//.*", "", text)
    names = re.findall(r"^\s*(?:const\s+)?(?:void|int|double|char)\s*\*?\s*(\w+)\s*\(",
                       text, re.M)
    return sorted(set(names))


def test_header_lists_the_drop_in_signatures():
    names = declared_symbols()
    for sym in ("fermion_op", "DSq", "congrad_multi", "fermion_force", "gauge_force"):
        assert sym in names
    assert set(names) <= set(susy_amd.EXPORTS)


@pytest.mark.parametrize("ncol", [2, 3, 4, 5])
def test_library_exports_every_declared_symbol(ncol):
    path = susy_amd.lib_path(ncol)
    assert os.path.exists(path), "run __graft_entry__.build() first"
    lib = ctypes.CDLL(path)
    for name in declared_symbols() + ["susy_set_cg_options", "susy_ranmom",
                                      "susy_grsource_dev"]:
        assert hasattr(lib, name), name
    assert lib.susy_ncol() == ncol
    assert lib.susy_validate_tables() == 0


def test_loader_rejects_a_missing_library(tmp_path, monkeypatch):
    monkeypatch.setattr(susy_amd, "LIBDIR", str(tmp_path))
    monkeypatch.setattr(susy_amd, "_libs", {})
    with pytest.raises(RuntimeError):
        susy_amd.load(3)
