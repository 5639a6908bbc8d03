"""MI355X-native 4d N=4 SYM RHMC hot path (fermion_op / DSq, congrad_multi,
fermion and gauge forces) behind the reference's C interface.

The compute lives in the HIP libraries susy_amd/lib/libsusy_hip_U<N>.so;
this package only loads them (ctypes) and exposes the resident API with
numpy arrays in the reference's AoS layout.  There is no CPU fallback: if the
library for a gauge group is missing, loading raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(HERE, "lib")

FORM_FP64 = 0
FORM_FP32 = 1
FORM_SURROGATE = 4      # flag: learned surrogate force on the inner MD level
FORM_PUREGAUGE = 8      # flag: fermions off (the reference's -DPUREGAUGE build)
FORM_SPLIT = 16         # flag: large-shift poles on a middle MD level (set_split)
# operator-kernel block order a new context starts with (susy_set_site_order:
# FP64 order in the low byte, FP32 order in the next)
DEFAULT_SITE_ORDER = 2 + (2 << 8)
NGATHER = 74


class SusyParams(ctypes.Structure):
    """Mirror of susy_params in include/susy_hip.h (input-file parameters of
    the reference, setup.c:16-119 and 307-556)."""
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("nt", ctypes.c_int), ("PBC", ctypes.c_int), ("iseed", ctypes.c_int),
                ("Nroot", ctypes.c_int), ("Norder", ctypes.c_int),
                ("lambda_", ctypes.c_double), ("kappa_u1", ctypes.c_double),
                ("bmass", ctypes.c_double), ("fmass", ctypes.c_double),
                ("G", ctypes.c_double), ("niter", ctypes.c_int),
                ("error_per_site", ctypes.c_double), ("traj_length", ctypes.c_double),
                ("nstep", ctypes.c_int), ("nstep_gauge", ctypes.c_int)]


# every symbol declared in include/susy_hip.h
EXPORTS = [
    "fermion_op", "DSq", "congrad_multi", "fermion_force", "gauge_force",
    "susy_bind_lattice", "susy_bind_sites", "susy_bind_couplings", "susy_bind_rhmc",
    "susy_links_changed", "susy_ncol", "susy_last_error", "susy_create", "susy_destroy",
    "susy_volume", "susy_rhmc_coeffs", "susy_get_gather", "susy_set_links", "susy_get_links", "susy_set_mom",
    "susy_get_mom", "susy_load_config", "susy_cold_start", "susy_get_fields",
    "susy_tf_alloc", "susy_tf_upload", "susy_tf_download", "susy_fermion_op_dev",
    "susy_dsq_dev", "susy_congrad_multi_dev", "susy_fermion_force_dev",
    "susy_gauge_force_dev", "susy_update_u_dev", "susy_actions_dev",
    "susy_fermion_action_dev", "susy_plaquette_dev", "susy_update", "susy_time_fermion_op",
    "susy_set_cg_options", "susy_ranmom", "susy_grsource_dev", "susy_validate_tables",
    "susy_prof_enable", "susy_prof_read", "susy_reseed",
    "susy_set_verbose", "susy_stream_copy_dev", "susy_set_site_order",
    "susy_sg_nparams", "susy_sg_init", "susy_sg_set_params", "susy_sg_get_params",
    "susy_sg_eval", "susy_sg_features", "susy_sg_loss_grad", "susy_sg_record", "susy_sg_pairs",
    "susy_sg_pair_loss", "susy_sg_train", "susy_eig_dsq", "susy_save_config",
    "susy_link_trace_dev", "susy_wilson_line_dev", "susy_scalar_eig_dev",
    "susy_plaq_moments_dev", "susy_pfaffian_dev", "susy_fermion_matrix_dev",
    "susy_traj_begin", "susy_traj_step", "susy_traj_substep", "susy_traj_end", "susy_fermion_op_noalloc",
    "susy_set_split", "susy_traj_iters",
]

_libs = {}


def lib_path(ncol):
    return os.path.join(LIBDIR, f"libsusy_hip_U{ncol}.so")


def load(ncol):
    """Load libsusy_hip_U<ncol>.so (raises if it was not built)."""
    if ncol in _libs:
        return _libs[ncol]
    path = lib_path(ncol)
    if not os.path.exists(path):
        raise RuntimeError(f"{path} not built: run __graft_entry__.build()")
    L = ctypes.CDLL(path)
    P = ctypes.POINTER
    d, i, vp = ctypes.c_double, ctypes.c_int, ctypes.c_void_p
    dp, ip = P(d), P(i)
    sig = {
        "susy_ncol": (i, []), "susy_validate_tables": (i, []), "susy_last_error": (ctypes.c_char_p, []),
        "susy_create": (i, [P(SusyParams), P(vp)]), "susy_destroy": (None, [vp]),
        "susy_volume": (i, [vp]), "susy_rhmc_coeffs": (i, [vp, dp]),
        "susy_get_gather": (i, [vp, i, ip]),
        "susy_set_links": (i, [vp, dp]), "susy_get_links": (i, [vp, dp]),
        "susy_set_mom": (i, [vp, dp]), "susy_get_mom": (i, [vp, dp]),
        "susy_load_config": (i, [vp, ctypes.c_char_p]), "susy_cold_start": (i, [vp]),
        "susy_get_fields": (i, [vp, dp, dp, dp, dp]),
        "susy_tf_alloc": (i, [vp, i]), "susy_tf_upload": (i, [vp, i, dp]),
        "susy_tf_download": (i, [vp, i, dp]),
        "susy_fermion_op_dev": (i, [vp, i, i, i, i]), "susy_dsq_dev": (i, [vp, i, i, i]),
        "susy_congrad_multi_dev": (i, [vp, i, i, i, dp, i, d, dp, i]),
        "susy_fermion_force_dev": (i, [vp, d, i, i, dp, dp]),
        "susy_gauge_force_dev": (i, [vp, d, dp]), "susy_update_u_dev": (i, [vp, d]),
        "susy_actions_dev": (i, [vp, dp]), "susy_fermion_action_dev": (i, [vp, i, i, dp]),
        "susy_plaquette_dev": (i, [vp, dp, dp]), "susy_update": (i, [vp, i, dp]),
        "susy_traj_begin": (i, [vp, i]), "susy_traj_step": (i, [vp]),
        "susy_traj_substep": (i, [vp]),
        "susy_traj_end": (i, [vp, dp]),
        "susy_fermion_op_noalloc": (i, [vp, vp, vp, i]),
        "susy_set_split": (i, [vp, i, i, dp]),
        "susy_traj_iters": (i, [vp, ip, ip]),
        "susy_time_fermion_op": (i, [vp, i, i, i, i, dp]),
        "susy_prof_enable": (i, [vp, i]), "susy_reseed": (i, [vp, i]), "susy_set_verbose": (i, [i]),
        "susy_set_site_order": (i, [vp, i]),
        "susy_stream_copy_dev": (i, [vp, i, i]),
        "susy_prof_read": (i, [vp, i, dp, P(ctypes.c_long), P(ctypes.c_long)]),
        "susy_set_cg_options": (i, [i, i]), "susy_ranmom": (i, [vp]),
        "susy_grsource_dev": (i, [vp, i, i, i, ip]),
        "susy_bind_lattice": (i, [i, i, i, i, i]),
        "susy_bind_sites": (i, [vp, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_size_t]),
        "susy_bind_couplings": (i, [d, d, d, d, d]),
        "susy_bind_rhmc": (i, [ip, P(dp), P(dp), ip]), "susy_links_changed": (i, []),
        "fermion_op": (None, [vp, vp, i]), "DSq": (None, [vp, vp]),
        "congrad_multi": (i, [vp, P(vp), i, d, dp]),
        "fermion_force": (d, [d, vp, P(vp)]), "gauge_force": (d, [d]),
        "susy_sg_nparams": (i, []), "susy_sg_init": (i, [vp, i, d]),
        "susy_sg_set_params": (i, [vp, P(ctypes.c_float)]),
        "susy_sg_get_params": (i, [vp, P(ctypes.c_float)]),
        "susy_sg_eval": (i, [vp, dp, dp]), "susy_sg_features": (i, [vp, P(ctypes.c_float)]),
        "susy_sg_loss_grad": (i, [vp, dp, dp, dp, dp]), "susy_sg_record": (i, [vp, i]),
        "susy_sg_pairs": (i, [vp]), "susy_sg_pair_loss": (i, [vp, i, dp]),
        "susy_sg_train": (i, [vp, i, i, i, i, d, i, dp]),
        "susy_eig_dsq": (i, [vp, i, i, i, d, i, dp, dp, ip]),
        "susy_save_config": (i, [vp, ctypes.c_char_p, ctypes.c_char_p]),
        "susy_link_trace_dev": (i, [vp, dp, dp, dp, dp, dp, dp]),
        "susy_wilson_line_dev": (i, [vp, i, i, dp, dp, dp]),
        "susy_scalar_eig_dev": (i, [vp, i, dp, dp, dp, dp]),
        "susy_plaq_moments_dev": (i, [vp, dp, dp]),
        "susy_pfaffian_dev": (i, [vp, dp, dp, dp]),
        "susy_fermion_matrix_dev": (i, [vp, vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    if L.susy_ncol() != ncol:
        raise RuntimeError(f"{path} reports NCOL={L.susy_ncol()}")
    _libs[ncol] = L
    return L


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Lattice:
    """Device-resident lattice: links, momenta, derived fields and
    Twist_Fermion vectors in HBM (one susy_ctx)."""

    def __init__(self, ncol, dims=(4, 4, 4, 4), PBC=-1, iseed=41, Nroot=1, Norder=15,
                 lam=1.0, kappa_u1=0.0, bmass=0.5, fmass=0.0, G=0.05, niter=5000,
                 error_per_site=1e-5, traj_length=1.0, nstep=10, nstep_gauge=10):
        self.ncol = ncol
        self.L = load(ncol)
        p = SusyParams(*dims, PBC, iseed, Nroot, Norder, lam, kappa_u1, bmass, fmass, G,
                       niter, error_per_site, traj_length, nstep, nstep_gauge)
        self.params = p
        h = ctypes.c_void_p()
        self._check(self.L.susy_create(ctypes.byref(p), ctypes.byref(h)))
        self.h = h
        self.V = self.L.susy_volume(h)
        self.nn = ncol * ncol

    def _check(self, r):
        if r < 0:
            raise RuntimeError(f"susy_hip error {r}: {self.L.susy_last_error().decode()}")
        return r

    def close(self):
        if getattr(self, "h", None):
            self.L.susy_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- shapes (reference AoS layout, complex128)
    def tf_shape(self):
        return (self.V, 16, self.ncol, self.ncol)

    def link_shape(self):
        return (self.V, 5, self.ncol, self.ncol)

    def gather(self, g):
        out = np.zeros(self.V, dtype=np.int32)
        self._check(self.L.susy_get_gather(self.h, g, out.ctypes.data_as(
            ctypes.POINTER(ctypes.c_int))))
        return out

    def set_links(self, links):
        a = np.ascontiguousarray(links, dtype=np.complex128)
        self._check(self.L.susy_set_links(self.h, _dp(a)))

    def get_links(self):
        a = np.zeros(self.link_shape(), dtype=np.complex128)
        self._check(self.L.susy_get_links(self.h, _dp(a)))
        return a

    def set_mom(self, mom):
        a = np.ascontiguousarray(mom, dtype=np.complex128)
        self._check(self.L.susy_set_mom(self.h, _dp(a)))

    def rhmc(self):
        """(62,) ampdeg4, amp4[15], shift4[15], ampdeg8, amp8[15], shift8[15]"""
        out = np.zeros(62)
        self._check(self.L.susy_rhmc_coeffs(self.h, _dp(out)))
        return out

    def get_mom(self):
        a = np.zeros(self.link_shape(), dtype=np.complex128)
        self._check(self.L.susy_get_mom(self.h, _dp(a)))
        return a

    def load_config(self, path, allow_bad_checksum=False):
        """read a MILC gauge file; a checksum mismatch raises unless allowed
        (then the return value is 1)"""
        r = self._check(self.L.susy_load_config(self.h, os.fsencode(path)))
        if r != 0 and not allow_bad_checksum:
            raise RuntimeError(f"{path}: gauge-file checksums do not match")
        return r

    def fields(self):
        V, n = self.V, self.ncol
        pd = np.zeros((5, 5, V), dtype=np.complex128)
        ui = np.zeros((V, 5, n, n), dtype=np.complex128)
        dm = np.zeros((V, 1, n, n), dtype=np.complex128)
        fm = np.zeros((10, V, 1, n, n), dtype=np.complex128)
        self._check(self.L.susy_get_fields(self.h, _dp(pd), _dp(ui), _dp(dm), _dp(fm)))
        return pd, ui, dm[:, 0], fm[:, :, 0]

    def tf_alloc(self, count=1):
        return self._check(self.L.susy_tf_alloc(self.h, count))

    def tf_upload(self, h, tf):
        a = np.ascontiguousarray(tf, dtype=np.complex128)
        assert a.shape == self.tf_shape()
        self._check(self.L.susy_tf_upload(self.h, h, _dp(a)))

    def tf_download(self, h):
        a = np.zeros(self.tf_shape(), dtype=np.complex128)
        self._check(self.L.susy_tf_download(self.h, h, _dp(a)))
        return a

    def fermion_op(self, hs, hd, sign, form=FORM_FP64):
        self._check(self.L.susy_fermion_op_dev(self.h, hs, hd, sign, form))

    def stream_copy(self, hs, hd):
        self._check(self.L.susy_stream_copy_dev(self.h, hs, hd))

    def dsq(self, hs, hd, form=FORM_FP64):
        self._check(self.L.susy_dsq_dev(self.h, hs, hd, form))

    def congrad_multi(self, hs, hpsim, shifts, maxcg, errormin, form=FORM_FP64):
        sh = np.ascontiguousarray(shifts, dtype=np.float64)
        sz = ctypes.c_double()
        it = self._check(self.L.susy_congrad_multi_dev(
            self.h, hs, hpsim, len(sh), _dp(sh), maxcg, errormin, ctypes.byref(sz), form))
        return it, sz.value

    def fermion_force(self, eps, hsol, amps):
        a = np.ascontiguousarray(amps, dtype=np.float64)
        nm = ctypes.c_double()
        self._check(self.L.susy_fermion_force_dev(self.h, eps, hsol, len(a), _dp(a),
                                                  ctypes.byref(nm)))
        return nm.value

    def gauge_force(self, eps):
        nm = ctypes.c_double()
        self._check(self.L.susy_gauge_force_dev(self.h, eps, ctypes.byref(nm)))
        return nm.value

    def update_u(self, eps):
        self._check(self.L.susy_update_u_dev(self.h, eps))

    def actions(self):
        out = np.zeros(4)
        self._check(self.L.susy_actions_dev(self.h, _dp(out)))
        return out

    def fermion_action(self, hs, hsol):
        out = ctypes.c_double()
        self._check(self.L.susy_fermion_action_dev(self.h, hs, hsol, ctypes.byref(out)))
        return out.value

    def plaquette(self):
        ss, st = ctypes.c_double(), ctypes.c_double()
        self._check(self.L.susy_plaquette_dev(self.h, ctypes.byref(ss), ctypes.byref(st)))
        return ss.value, st.value

    def update(self, form=FORM_FP64):
        """One RHMC trajectory; returns the stats vector (see susy_update)."""
        st = np.zeros(32)
        self._check(self.L.susy_update(self.h, form, _dp(st)))
        return st

    def set_split(self, first_pole, nstep_mid=1, weights=None):
        """poles j >= first_pole on a middle MD level of nstep_mid Omelyan
        steps per half outer step (FORM_SPLIT); weights: the middle level's
        amplitudes for those poles (None: the rational approximation's)"""
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
        self._check(self.L.susy_set_split(self.h, int(first_pole), int(nstep_mid),
                                          None if w is None else _dp(w)))

    def traj_begin(self, form=FORM_FP64):
        """start a trajectory (heatbath, start action, first fermion kick)"""
        self._check(self.L.susy_traj_begin(self.h, form))

    def traj_step(self):
        """one outer MD step of the open trajectory; returns the steps done"""
        return self._check(self.L.susy_traj_step(self.h))

    def traj_substep(self):
        """one solve + fermion kick of one root; returns the sub-steps done"""
        return self._check(self.L.susy_traj_substep(self.h))

    def traj_iters(self):
        """CG iterations of the current trajectory so far: (all, middle level)"""
        a, b = ctypes.c_int(), ctypes.c_int()
        self._check(self.L.susy_traj_iters(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def traj_end(self):
        """Metropolis test of the open trajectory; returns the stats vector"""
        st = np.zeros(32)
        self._check(self.L.susy_traj_end(self.h, _dp(st)))
        return st

    def time_fermion_op(self, hs, hd, form=FORM_FP64, reps=20):
        ms = ctypes.c_double()
        self._check(self.L.susy_time_fermion_op(self.h, hs, hd, form, reps, ctypes.byref(ms)))
        return ms.value

    def prof_enable(self, every):
        self._check(self.L.susy_prof_enable(self.h, every))

    def set_site_order(self, paired):
        """block order of the operator kernel (tuning only)"""
        self._check(self.L.susy_set_site_order(self.h, int(paired)))

    def prof_read(self, prec):
        """(average ms of sampled operator-kernel launches, launches, sampled)"""
        ms, n, k = ctypes.c_double(), ctypes.c_long(), ctypes.c_long()
        self._check(self.L.susy_prof_read(self.h, prec, ctypes.byref(ms), ctypes.byref(n),
                                          ctypes.byref(k)))
        return ms.value, n.value, k.value

    def save_config(self, path, stamp=None):
        """MILC serial gauge file; stamp: 64 bytes or None (current time)"""
        st = None if stamp is None else bytes(stamp).ljust(64, b"\0")[:64]
        self._check(self.L.susy_save_config(self.h, os.fsencode(path), st))

    def link_trace(self):
        """(Tr(UU^dag)/N per direction, average, width,
            det of the traceless part per direction, average, width)"""
        lt, ld = np.zeros(5), np.zeros(5)
        a, w, da, dw = (ctypes.c_double() for _ in range(4))
        self._check(self.L.susy_link_trace_dev(self.h, _dp(lt), ctypes.byref(a), ctypes.byref(w),
                                               _dp(ld), ctypes.byref(da), ctypes.byref(dw)))
        return lt, a.value, w.value, ld, da.value, dw.value

    def wilson_line(self, a, project=False):
        """(Re, Im, average modulus) of the Wilson line in link direction a
        (project: of the polar-projected unitary links)"""
        re, im, mod = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        self._check(self.L.susy_wilson_line_dev(self.h, a, int(project), ctypes.byref(re),
                                                ctypes.byref(im), ctypes.byref(mod)))
        return re.value, im.value, mod.value

    def scalar_eig(self, project=False):
        """(average, width, minimum, maximum) per ascending eigenvalue index of
        the traceless part of U U^dag, or of log (U U^dag)^(1/2) (project)"""
        out = [np.zeros(self.ncol) for _ in range(4)]
        self._check(self.L.susy_scalar_eig_dev(self.h, int(project), *(_dp(o) for o in out)))
        return tuple(out)

    def plaq_moments(self):
        """(DET: Re, Im <plaqdet>, <Re^2>, <Im^2>, <|plaqdet - 1|^2>;
            WIDTHS: widths of Re Tr P, Re det P, Im det P)"""
        d5, w3 = np.zeros(5), np.zeros(3)
        self._check(self.L.susy_plaq_moments_dev(self.h, _dp(d5), _dp(w3)))
        return d5, w3

    def pfaffian(self):
        """(log |Pf M|, arg Pf M in [0, 2 pi), max|M + M^T| / max|M|) of the
        fermion operator in the reference's phase-measurement basis"""
        lm, ph, asym = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        self._check(self.L.susy_pfaffian_dev(self.h, ctypes.byref(lm), ctypes.byref(ph),
                                             ctypes.byref(asym)))
        return lm.value, ph.value, asym.value

    def fermion_matrix(self):
        """n x n complex matrix of the fermion operator in the same basis"""
        n = 16 * self.ncol ** 2 * self.V
        m = np.zeros((n, n), dtype=np.complex128)
        self._check(self.L.susy_fermion_matrix_dev(self.h, m.ctypes.data))
        return m

    def eig_dsq(self, nlow, nhigh, maxiter=600, tol=1e-10, seed=1):
        """(nlow smallest, nhigh largest eigenvalues of DSq, Lanczos iterations)"""
        lo, hi = np.zeros(max(nlow, 1)), np.zeros(max(nhigh, 1))
        it = ctypes.c_int()
        self._check(self.L.susy_eig_dsq(self.h, nlow, nhigh, maxiter, tol, seed, _dp(lo),
                                        _dp(hi), ctypes.byref(it)))
        return lo[:nlow], hi[:nhigh], it.value

    # ---- learned surrogate
    def sg_init(self, seed=1, scale=1.0):
        self._check(self.L.susy_sg_init(self.h, seed, scale))

    def sg_set_params(self, theta):
        t = np.ascontiguousarray(theta, dtype=np.float32)
        self._check(self.L.susy_sg_set_params(self.h, t.ctypes.data_as(
            ctypes.POINTER(ctypes.c_float))))

    def sg_get_params(self):
        t = np.zeros(self.L.susy_sg_nparams(), dtype=np.float32)
        self._check(self.L.susy_sg_get_params(self.h, t.ctypes.data_as(
            ctypes.POINTER(ctypes.c_float))))
        return t

    def sg_eval(self, force=True):
        """(S_theta, F_theta in AoS (V, 5, N, N) or None) at the current links"""
        S = ctypes.c_double()
        F = np.zeros(self.link_shape(), dtype=np.complex128) if force else None
        self._check(self.L.susy_sg_eval(self.h, ctypes.byref(S),
                                        _dp(F) if force else None))
        return S.value, F

    def sg_features(self):
        """features as (V, 65)"""
        phi = np.zeros((65, self.V), dtype=np.float32)
        self._check(self.L.susy_sg_features(self.h, phi.ctypes.data_as(
            ctypes.POINTER(ctypes.c_float))))
        return phi.T.copy()

    def sg_loss_grad(self, U, F):
        u = np.ascontiguousarray(U, dtype=np.complex128)
        f = np.ascontiguousarray(F, dtype=np.complex128)
        loss = ctypes.c_double()
        g = np.zeros(self.L.susy_sg_nparams())
        self._check(self.L.susy_sg_loss_grad(self.h, _dp(u), _dp(f), ctypes.byref(loss), _dp(g)))
        return loss.value, g

    def sg_record(self, capacity):
        self._check(self.L.susy_sg_record(self.h, capacity))

    def sg_pairs(self):
        return self._check(self.L.susy_sg_pairs(self.h))

    def sg_pair_loss(self, i):
        loss = ctypes.c_double()
        self._check(self.L.susy_sg_pair_loss(self.h, i, ctypes.byref(loss)))
        return loss.value

    def sg_train(self, first, count, steps, batch=1, lr=1e-3, seed=1):
        hist = np.zeros(max(steps, 1))
        self._check(self.L.susy_sg_train(self.h, first, count, steps, batch, lr, seed,
                                         _dp(hist)))
        return hist[:steps]

    def reseed(self, seed):
        self._check(self.L.susy_reseed(self.h, seed))

    def ranmom(self):
        self._check(self.L.susy_ranmom(self.h))

    def grsource(self, hsrc, hpsim, form=FORM_FP64):
        it = ctypes.c_int()
        self._check(self.L.susy_grsource_dev(self.h, hsrc, hpsim, form, ctypes.byref(it)))
        return it.value


def set_cg_options(ncol, window=0, reliable=128):
    """multi-shift CG window (0: the default, 32) and FP32-form
    reliable-update period"""
    r = load(ncol).susy_set_cg_options(window, reliable)
    if r < 0:
        raise RuntimeError(load(ncol).susy_last_error().decode())
