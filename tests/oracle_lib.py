"""ctypes access to the CPU checkers (test infrastructure only):
  * oracle/liboracle_U<N>.so  -- the C restatement of the reference algorithm
  * oracle/_ref/libsusyref_U2.so -- the reference itself, built in place from
    /root/reference by oracle/ref.mk (U(2) only: the reference header fixes
    NCOL), with oracle/ref_harness.c.
Both take fields in the reference's AoS layout as complex128 numpy arrays.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
REF_LIB = os.path.join(ORACLE_DIR, "_ref", "libsusyref_U2.so")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def oracle_path(ncol):
    return os.path.join(ORACLE_DIR, f"liboracle_U{ncol}.so")


def have_ref():
    return os.path.exists(REF_LIB)


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


class _Checker:
    prefix = ""

    def __init__(self, path, ncol, dims=(4, 4, 4, 4), PBC=-1, iseed=41, Nroot=1, Norder=15,
                 lam=1.0, kappa_u1=0.0, bmass=0.5, fmass=0.0, G=0.05, niter=5000,
                 error_per_site=1e-5, traj_length=1.0, nstep=10, nstep_gauge=10):
        # the checkers keep their lattice in globals, like the reference:
        # creating an instance re-initialises the library for all holders
        self.L = ctypes.CDLL(path, mode=ctypes.RTLD_LOCAL)
        self.ncol = ncol
        self.dims = dims
        self.V = int(np.prod(dims))
        self.Norder = Norder
        f = self._f("init")
        f.argtypes = [ctypes.c_int] * 8 + [ctypes.c_double] * 5 + [
            ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int]
        r = f(*dims, PBC, iseed, Nroot, Norder, lam, kappa_u1, bmass, fmass, G, niter,
              error_per_site, traj_length, nstep, nstep_gauge)
        self._init_ret = r
        for name, res in [("gauge_force", ctypes.c_double), ("fermion_force", ctypes.c_double),
                          ("fermion_action", ctypes.c_double)]:
            self._f(name).restype = res

    def _f(self, name):
        return getattr(self.L, self.prefix + name)

    def tf_shape(self):
        return (self.V, 16, self.ncol, self.ncol)

    def link_shape(self):
        return (self.V, 5, self.ncol, self.ncol)

    def gather(self, g):
        out = np.zeros(self.V, dtype=np.int32)
        self._f("get_gather")(g, _ip(out))
        return out

    def coords(self):
        out = np.zeros((self.V, 4), dtype=np.int32)
        self._f("get_coords")(_ip(out))
        return out

    def set_links(self, links):
        a = np.ascontiguousarray(links, dtype=np.complex128)
        self._f("set_links")(_dp(a))

    def get_links(self):
        a = np.zeros(self.link_shape(), dtype=np.complex128)
        self._f("get_links")(_dp(a))
        return a

    def set_mom(self, mom):
        a = np.ascontiguousarray(mom, dtype=np.complex128)
        self._f("set_mom")(_dp(a))

    def get_mom(self):
        a = np.zeros(self.link_shape(), dtype=np.complex128)
        self._f("get_mom")(_dp(a))
        return a

    def fermion_op(self, src, sign):
        s = np.ascontiguousarray(src, dtype=np.complex128)
        d = np.zeros(self.tf_shape(), dtype=np.complex128)
        self._f("fermion_op")(_dp(s), _dp(d), sign)
        return d

    def dsq(self, src):
        s = np.ascontiguousarray(src, dtype=np.complex128)
        d = np.zeros(self.tf_shape(), dtype=np.complex128)
        self._f("DSq")(_dp(s), _dp(d))
        return d

    def congrad_multi(self, src, shifts, maxcg, errormin):
        s = np.ascontiguousarray(src, dtype=np.complex128)
        sh = np.ascontiguousarray(shifts, dtype=np.float64)
        ps = np.zeros((len(sh),) + self.tf_shape(), dtype=np.complex128)
        sz = ctypes.c_double()
        f = self._f("congrad_multi")
        f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                      ctypes.c_int, ctypes.c_double, ctypes.c_void_p]
        it = f(s.ctypes.data, ps.ctypes.data, len(sh), sh.ctypes.data, maxcg, errormin,
               ctypes.addressof(sz))
        return it, sz.value, ps

    def gauge_force(self, eps):
        f = self._f("gauge_force")
        f.argtypes = [ctypes.c_double]
        return f(eps)

    def update_u(self, eps):
        f = self._f("update_u")
        f.argtypes = [ctypes.c_double]
        f(eps)

    def plaquette(self):
        a, b = ctypes.c_double(), ctypes.c_double()
        self._f("plaquette")(ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def actions(self):
        out = np.zeros(4)
        self._f("actions")(_dp(out))
        return out

    def fermion_action(self, src, sol):
        s = np.ascontiguousarray(src, dtype=np.complex128)
        p = np.ascontiguousarray(sol, dtype=np.complex128)
        f = self._f("fermion_action")
        f.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        return f(s.ctypes.data, p.ctypes.data)

    def rand_stream(self, site, n):
        out = np.zeros(n)
        self._f("rand_stream")(site, n, _dp(out))
        return out

    def fields(self):
        V, n = self.V, self.ncol
        pd = np.zeros((5, 5, V), dtype=np.complex128)
        ui = np.zeros((5, V, n, n), dtype=np.complex128)
        dm = np.zeros((V, n, n), dtype=np.complex128)
        fm = np.zeros((10, V, n, n), dtype=np.complex128)
        self._f("get_fields")(_dp(pd), _dp(ui), _dp(dm), _dp(fm))
        return pd, ui.transpose(1, 0, 2, 3).copy(), dm, fm


class Oracle(_Checker):
    """The C restatement (oracle/susy_oracle.c)."""
    prefix = "orc_"

    def __init__(self, ncol, **kw):
        super().__init__(oracle_path(ncol), ncol, **kw)
        if self._init_ret != 0:
            raise RuntimeError(f"orc_init failed: {self._init_ret}")
        self._f("set_quiet")(1)

    def load_config(self, path):
        sums = (ctypes.c_uint * 4)()
        r = self._f("load_config")(path.encode(), sums)
        return r, list(sums)

    def fermion_force(self, eps, sol, amps):
        s = np.ascontiguousarray(sol, dtype=np.complex128)
        a = np.ascontiguousarray(amps, dtype=np.float64)
        f = self._f("fermion_force")
        f.argtypes = [ctypes.c_double, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
        return f(eps, s.ctypes.data, len(a), a.ctypes.data)

    def update(self):
        st = np.zeros(32)
        self._f("update")(_dp(st))
        return st

    def set_quiet(self, q):
        self._f("set_quiet")(int(q))

    def rhmc(self):
        out = np.zeros(62)
        self._f("rhmc")(_dp(out))
        return out

    def lambdas(self):
        out = np.zeros((self.ncol ** 2, self.ncol, self.ncol), dtype=np.complex128)
        self._f("get_lambda")(_dp(out))
        return out

    def det_inv(self, m):
        a = np.ascontiguousarray(m, dtype=np.complex128)
        det = np.zeros(1, dtype=np.complex128)
        inv = np.zeros_like(a)
        self._f("det_inv")(_dp(a), _dp(det), _dp(inv))
        return det[0], inv


class Ref(_Checker):
    """The reference's own code (oracle/_ref), U(2) only."""
    prefix = "ref_"

    def __init__(self, **kw):
        if not have_ref():
            raise RuntimeError("oracle/_ref not built")
        super().__init__(REF_LIB, 2, **kw)

    def load_config(self, path):
        self._f("load_config")(path.encode())

    def fermion_force(self, eps, sol):
        s = np.ascontiguousarray(sol, dtype=np.complex128)
        f = self._f("fermion_force")
        f.argtypes = [ctypes.c_double, ctypes.c_void_p]
        return f(eps, s.ctypes.data)

    def bc(self, site):
        out = np.zeros(1110)
        self._f("get_bc")(site, _dp(out))
        return out[:10], out[10:110].reshape(10, 10), out[110:].reshape(10, 10, 10)


def random_tf(shape, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) / np.sqrt(2)


def random_links(shape, seed, scale=0.1):
    """Identity + scale * complex Gaussian per link (the kernel-level test
    configuration of the benchmark workloads)."""
    n = shape[-1]
    return np.eye(n)[None, None] + scale * random_tf(shape, seed)


def config_path(ncol, dims="4444"):
    return os.path.join(GOLDEN, "configs", f"config.U{ncol}.{dims}")


def rel_err(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))
