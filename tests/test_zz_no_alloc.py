"""The allocation-free operator entry point susy_fermion_op_noalloc against
susy_fermion_op_dev (same kernels, scratch in the library's fixed device
array instead of the context's buffer): bitwise equal output for both signs,
on two calls in a row with the second on a smaller lattice, through the
C ABI with device buffers from the HIP runtime the library itself uses
(hipMalloc / hipMemcpy through ctypes); lattices beyond the array's size are
refused with an error."""
import ctypes

import numpy as np
import pytest

import susy_amd
from oracle_lib import random_links, random_tf

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def hip():
    """the HIP runtime already loaded by the library (same soname)"""
    h = ctypes.CDLL("libamdhip64.so.7", mode=ctypes.RTLD_GLOBAL)
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipFree.argtypes = [ctypes.c_void_p]
    return h


class DevBuf:
    def __init__(self, nbytes):
        self.h, self.n, self.p = hip(), nbytes, ctypes.c_void_p()
        assert self.h.hipMalloc(ctypes.byref(self.p), nbytes) == 0

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes == self.n
        assert self.h.hipMemcpy(self.p, a.ctypes.data, self.n, H2D) == 0

    def download(self, dtype):
        a = np.empty(self.n // np.dtype(dtype).itemsize, dtype=dtype)
        assert self.h.hipMemcpy(a.ctypes.data, self.p, self.n, D2H) == 0
        return a

    def free(self):
        self.h.hipFree(self.p)


def to_soa(tf):
    """(V, 16, N, N) AoS -> flat SoA: element (i,j) of component k at site s
    at index (k N^2 + i N + j) V + s"""
    V = tf.shape[0]
    return np.ascontiguousarray(tf.reshape(V, -1).T).ravel()


def from_soa(a, V, n):
    return np.ascontiguousarray(a.reshape(-1, V).T).reshape(V, 16, n, n)


def noalloc_vs_dev(ncol, dims, seed):
    lat = susy_amd.Lattice(ncol, dims=dims, G=0.1, kappa_u1=0.5, bmass=0.4)
    lat.set_links(random_links(lat.link_shape(), seed, 0.2))
    src = random_tf(lat.tf_shape(), seed + 1)
    h = lat.tf_alloc(2)
    lat.tf_upload(h, src)
    soa = to_soa(src)
    d_src, d_dst = DevBuf(soa.nbytes), DevBuf(soa.nbytes)
    d_src.upload(soa)
    d_dst.upload(np.zeros_like(soa))
    for sign in (1, -1):
        lat.fermion_op(h, h + 1, sign)
        want = lat.tf_download(h + 1)
        r = lat.L.susy_fermion_op_noalloc(lat.h, d_src.p, d_dst.p, sign)
        assert r == 0, lat.L.susy_last_error().decode()
        got = from_soa(d_dst.download(np.complex128), lat.V, ncol)
        print("noalloc", ncol, dims, sign, float(np.abs(got - want).max()), flush=True)
        assert np.array_equal(got, want), (ncol, dims, sign)
    d_src.free()
    d_dst.free()
    lat.close()


@pytest.mark.parametrize("ncol", [2, 4])
def test_noalloc_matches_resident_operator(ncol):
    noalloc_vs_dev(ncol, (8, 8, 8, 16), 3)       # first call
    noalloc_vs_dev(ncol, (4, 4, 4, 4), 5)        # second, smaller lattice


def test_noalloc_refuses_lattices_beyond_its_scratch():
    lat = susy_amd.Lattice(2, dims=(16, 16, 16, 48))     # 196608 sites > 174762
    buf = DevBuf(16 * 4 * 16 * lat.V)
    r = lat.L.susy_fermion_op_noalloc(lat.h, buf.p, buf.p, 1)
    assert r == -1                                        # SUSY_E_ARG
    assert "scratch" in lat.L.susy_last_error().decode()
    buf.free()
    lat.close()
