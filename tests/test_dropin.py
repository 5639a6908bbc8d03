"""The drop-in boundary: the reference's own entry points fermion_op, DSq,
congrad_multi, fermion_force and gauge_force (susy_includes.h:73-81), called
the way the reference driver calls them, with its state bound through
susy_bind_*: a site array with the reference's kind of layout (links and
momenta embedded in a larger per-site struct), and pointers to the RHMC
globals Norder / shift / amp4 / total_iters."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import susy_amd
from oracle_lib import Oracle, config_path, random_tf, rel_err

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
KW = dict(iseed=41, kappa_u1=1.0, bmass=1.0, G=0.1)

C_DRIVER = r"""
#include <stdio.h>
#include "susy_hip.h"
/* a reference-style site struct: coordinates, links and momenta among other
   per-site data (susy/lattice.h:29-59) */
typedef struct { short x, y, z, t; char parity; int index;
                 susy_matrix link[5]; susy_matrix old_link[5];
                 susy_matrix mom[5], f_U[5]; double bc[10]; } site;
int main(void) {
  int (*fop)(Twist_Fermion *, Twist_Fermion *, int) = (void *)fermion_op;
  printf("%p %p %p %p %p %zu\n", (void *)fop, (void *)DSq, (void *)congrad_multi,
         (void *)fermion_force, (void *)gauge_force, sizeof(site));
  return susy_ncol() == NCOL ? 0 : 1;
}
"""


def site_dtype(ncol):
    """numpy mirror of the C struct above (natural alignment)"""
    mat = (np.complex128, (ncol, ncol))
    return np.dtype([("xyzt", np.int16, 4), ("parity", np.int8), ("index", np.int32),
                     ("link", mat, (5,)), ("old_link", mat, (5,)), ("mom", mat, (5,)),
                     ("f_U", mat, (5,)), ("bc", np.float64, 10)], align=True)


@pytest.mark.parametrize("ncol", [2, 3, 4, 5])
def test_header_compiles_and_links_as_c(ncol, tmp_path):
    src = tmp_path / "driver.c"
    src.write_text(C_DRIVER)
    exe = tmp_path / "driver"
    libdir = os.path.dirname(susy_amd.lib_path(ncol))
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", f"-DNCOL={ncol}",
                    f"-I{ROOT}/include", str(src), f"-L{libdir}", f"-lsusy_hip_U{ncol}",
                    f"-Wl,-rpath,{libdir}", "-o", str(exe)], check=True)


@pytest.mark.gpu
@pytest.mark.parametrize("ncol", [2, 4])
def test_drop_in_symbols(ncol):
    print("dropin", ncol, flush=True)
    L = susy_amd.load(ncol)
    O = Oracle(ncol, **KW)
    assert O.load_config(config_path(ncol))[0] == 0
    V = O.V
    sites = np.zeros(V, dtype=site_dtype(ncol))
    sites["link"] = O.get_links()
    mom = random_tf(O.link_shape(), 12)
    sites["mom"] = mom
    dt = site_dtype(ncol)
    assert L.susy_bind_lattice(4, 4, 4, 4, -1) == 0
    assert L.susy_bind_couplings(1.0, KW["kappa_u1"], KW["bmass"], 0.0, KW["G"]) == 0
    assert L.susy_bind_sites(sites.ctypes.data, dt.itemsize, dt.fields["link"][1],
                             dt.fields["mom"][1]) == 0
    rh = O.rhmc()
    Norder = ctypes.c_int(15)
    shift = (ctypes.c_double * 15)(*rh[16:31])
    amp4 = (ctypes.c_double * 15)(*rh[1:16])
    shift_p = ctypes.cast(shift, ctypes.POINTER(ctypes.c_double))
    amp_p = ctypes.cast(amp4, ctypes.POINTER(ctypes.c_double))
    total = ctypes.c_int(0)
    assert L.susy_bind_rhmc(ctypes.byref(Norder), ctypes.byref(shift_p), ctypes.byref(amp_p),
                            ctypes.byref(total)) == 0
    assert L.susy_links_changed() == 0
    src = random_tf(O.tf_shape(), 13)
    dest = np.zeros_like(src)
    for sign in (1, -1):
        L.fermion_op(src.ctypes.data, dest.ctypes.data, sign)
        assert rel_err(dest, O.fermion_op(src, sign)) < 1e-12
    L.DSq(src.ctypes.data, dest.ctypes.data)
    assert rel_err(dest, O.dsq(src)) < 1e-12
    # congrad_multi with psim as an array of Twist_Fermion pointers
    psim = np.zeros((15,) + O.tf_shape(), dtype=np.complex128)
    ptrs = (ctypes.c_void_p * 15)(*[psim[j].ctypes.data for j in range(15)])
    size_r = ctypes.c_double()
    it = L.congrad_multi(src.ctypes.data, ptrs, 5000, 1e-10, ctypes.byref(size_r))
    it_o, sz_o, ps_o = O.congrad_multi(src, rh[16:31], 5000, 1e-10)
    assert it == it_o and total.value == it
    assert rel_err(psim, ps_o) < 1e-10
    # forces update mom inside the bound site array
    O.set_mom(mom)
    n = L.gauge_force(0.05)
    assert abs(n / O.gauge_force(0.05) - 1) < 1e-12
    assert rel_err(sites["mom"] - mom, O.get_mom() - mom) < 1e-12
    sites["mom"] = mom
    O.set_mom(mom)
    n = L.fermion_force(0.02, None, ptrs)
    assert abs(n / O.fermion_force(0.02, psim, rh[1:16]) - 1) < 1e-10
    assert rel_err(sites["mom"] - mom, O.get_mom() - mom) < 1e-10
    # links changed by the caller: the next call sees them
    sites["link"] = sites["link"] * 1.01
    O.set_links(sites["link"].copy())
    assert L.susy_links_changed() == 0
    L.fermion_op(src.ctypes.data, dest.ctypes.data, 1)
    assert rel_err(dest, O.fermion_op(src, 1)) < 1e-12
