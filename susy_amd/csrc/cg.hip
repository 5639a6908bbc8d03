// Multi-shift conjugate gradient (Jegerlehner) for (D^dag D + sigma_j) psi_j = src
// (reference: 4d_Q16/susy/congrad_multi.c:26-221), zero initial guess.
//
// The scalar recurrences (beta, zeta, alpha, convergence) are the reference's,
// in FP64 on the host, so iteration counts follow the reference exactly.
// The vector work is restructured for HBM bandwidth:
//   * D^dag D p0, + sigma_0 p0 and the <p0, A p0> partial sums are one
//     operator pair with a fused epilogue;
//   * r += beta_0 A p0 and |r|^2 are one pass that writes the new r into
//     the next slot of a ring of the last m residuals (r has no buffer of
//     its own);
//   * p0 = r + alpha_0 p0 is one pass;
//   * no solution psi_j (the base system included) and no shifted direction
//     p_j (j >= 1) is touched every iteration.  Within a window they stay
//     linear combinations of their window-start values and the residuals
//     of the ring:
//       p_j   = A_j p_j(k0)        + sum_l C_jl r_l
//       psi_j = psi_j(k0) + E_j p_j(k0) + sum_l D_jl r_l
//     with scalar coefficients advanced on the host by the same recurrences
//     (zeta_0 = 1 for the base system, whose p0(k0) is kept by writing p0
//     into the second of two buffers during the window); one fused flush
//     applies them to all active systems every m iterations.
// Precision forms:
//   FP64: all vectors FP64 (the parity form).
//   FP32: p0, r, A p0, the residual window and p_j in FP32 with the FP32
//     operator; psi_j accumulate in FP64.  Every `reliable` iterations, and
//     before declaring convergence, the true residual src - (D^dag D +
//     sigma_0) psi_0 is recomputed with the FP64 operator and replaces r
//     (the shifted residuals stay zeta_j r by construction).
#include <math.h>
#include <stdio.h>
#include <string.h>

#include "susy_internal.h"

#define CB 256
#define MAXSH 16
#define MAXWIN 32

// ---------------------------------------------------------------------
template <typename C>
__device__ __forceinline__ double2 to64(C v) {
  return make_double2((double)v.x, (double)v.y);
}
template <typename C>
__device__ __forceinline__ C from64(double2 v);
template <>
__device__ __forceinline__ double2 from64<double2>(double2 v) { return v; }
template <>
__device__ __forceinline__ float2 from64<float2>(double2 v) {
  return make_float2((float)v.x, (float)v.y);
}

__device__ __forceinline__ void cg_block_sum(double v, double *partial) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  __shared__ double red[CB / 64];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0;
    for (int w = 0; w < CB / 64; w++) t += red[w];
    partial[blockIdx.x] = t;
  }
}

struct PsiPtrs {
  double2 *p[MAXSH];
};

// r = src, p0 = src, p_j = src (j >= 1), psi_j = 0; partial |src|^2
template <typename C>
__global__ void __launch_bounds__(CB)
k_cg_init(const double2 *__restrict__ src, C *__restrict__ r, C *__restrict__ p0,
          C *__restrict__ pj, int nsh, PsiPtrs psi, size_t n, double *__restrict__ partial) {
  double s = 0;
  for (size_t i = (size_t)blockIdx.x * CB + threadIdx.x; i < n; i += (size_t)gridDim.x * CB) {
    double2 v = src[i];
    C w = from64<C>(v);
    r[i] = w;
    p0[i] = w;
    for (int j = 1; j < nsh; j++) pj[(size_t)(j - 1) * n + i] = w;
    for (int j = 0; j < nsh; j++) psi.p[j][i] = make_double2(0, 0);
    s += v.x * v.x + v.y * v.y;
  }
  cg_block_sum(s, partial);
}

// rout = rin + beta0 * ap; partial |rout|^2.  rout is the next slot of the
// residual window and rin the previous one (the same slot when the window
// holds one residual: the update is elementwise, so in place is safe).
template <typename C>
__global__ void __launch_bounds__(CB)
k_cg_r(const C *rin, const C *__restrict__ ap, C *rout, double beta0, size_t n,
       double *__restrict__ partial) {
  double s = 0;
  for (size_t i = (size_t)blockIdx.x * CB + threadIdx.x; i < n; i += (size_t)gridDim.x * CB) {
    C a = ap[i], v = rin[i];
    v.x = fma(beta0, a.x, v.x);
    v.y = fma(beta0, a.y, v.y);
    rout[i] = v;
    s += (double)v.x * (double)v.x + (double)v.y * (double)v.y;
  }
  cg_block_sum(s, partial);
}

// pout = r + alpha0 pin (pout is pin, or the other p0 buffer at the first
// iteration of a window, which leaves the window-start p0 intact)
template <typename C>
__global__ void __launch_bounds__(CB)
k_cg_p0(const C *pin, const C *__restrict__ r, C *pout, double alpha0, size_t n) {
  size_t i = (size_t)blockIdx.x * CB + threadIdx.x;
  if (i >= n) return;
  C p = pin[i], rv = r[i];
  p.x = rv.x + alpha0 * p.x;
  p.y = rv.y + alpha0 * p.y;
  pout[i] = p;
}

// Window flush.  Slot k updates psi_j (j = sh[k]) and, for j >= 1, p_j:
//   p_j   <- A p_j(k0) + sum_l C_l r_l
//   psi_j <- psi_j + E p_j(k0) + sum_l D_l r_l
// coef layout per slot: [A, E, C_0..C_{MAXWIN-1}, D_0..D_{MAXWIN-1}].  For the
// base system (j = 0) p(k0) is the frozen window-start p0 and is not written.
struct FlushArgs {
  int nact;
  int sh[MAXSH];
};
struct PPtrs {
  void *p[MAXSH];
};
template <typename C, int M>
__global__ void __launch_bounds__(CB)
k_cg_flush(PPtrs pp, PsiPtrs psi, const C *__restrict__ win,
           const double *__restrict__ coef, FlushArgs fa, size_t n) {
  size_t i = (size_t)blockIdx.x * CB + threadIdx.x;
  if (i >= n) return;
  // all M slots are read; slots beyond this window's count carry zero weights
  double2 w[M];
#pragma unroll
  for (int l = 0; l < M; l++) w[l] = to64(win[(size_t)l * n + i]);
  const int stride = 2 + 2 * MAXWIN;
  for (int k = 0; k < fa.nact; k++) {
    const int j = fa.sh[k];
    const double *cf = coef + k * stride;
    C *pk = (C *)pp.p[k] + i;
    double2 p = to64(*pk);
    double2 x = psi.p[j][i];
    double A = cf[0], E = cf[1];
    double2 np = make_double2(A * p.x, A * p.y);
    x.x += E * p.x;
    x.y += E * p.y;
#pragma unroll
    for (int l = 0; l < M; l++) {
      double cl = cf[2 + l], dl = cf[2 + MAXWIN + l];
      np.x += cl * w[l].x;
      np.y += cl * w[l].y;
      x.x += dl * w[l].x;
      x.y += dl * w[l].y;
    }
    if (j > 0) *pk = from64<C>(np);
    psi.p[j][i] = x;
  }
}

// r = src - ap (FP64 true residual, stored in the CG precision); partial |r|^2
template <typename C>
__global__ void __launch_bounds__(CB)
k_true_res(const double2 *__restrict__ src, const double2 *__restrict__ ap,
           C *__restrict__ r, size_t n, double *__restrict__ partial) {
  double s = 0;
  for (size_t i = (size_t)blockIdx.x * CB + threadIdx.x; i < n; i += (size_t)gridDim.x * CB) {
    double2 a = src[i], b = ap[i];
    double2 v = make_double2(a.x - b.x, a.y - b.y);
    r[i] = from64<C>(v);
    s += v.x * v.x + v.y * v.y;
  }
  cg_block_sum(s, partial);
}

// ---------------------------------------------------------------------
static int nblk(size_t n) {
  size_t nb = (n + CB - 1) / CB;
  return (int)(nb < 2048 ? nb : 2048);
}

static int g_window = 32;     // residuals kept per window (<= MAXWIN)
static int g_reliable = 50;   // FP32 form: iterations between true residuals

extern "C" int susy_set_cg_options(int window, int reliable) {
  if ((window != 1 && window != 4 && window != 8 && window != 16 && window != 32) ||
      reliable < 1) {
    set_error("cg options: window in {1,4,8,16,32}, reliable >= 1");
    return SUSY_E_ARG;
  }
  g_window = window;
  g_reliable = reliable;
  return 0;
}

template <typename R>
static int cg_impl(susy_ctx *c, const double2 *src, double2 *const *psim, int nsh,
                   const double *shifts, int MaxCG, double errormin, double *size_r) {
  typedef typename CT<R>::t C;
  const bool fp32 = sizeof(R) == sizeof(float);
  const size_t n = tf_elems(c);
  const int m = g_window;
  // workspace in the CG precision: two p0 buffers, ap, tmp, p_j (nsh-1) and
  // the window of the last m residuals, which also holds the current r;
  // the FP32 form also uses two FP64 fields for the true residual
  size_t need = (size_t)(4 + (nsh - 1) + m) * n * sizeof(C);
  if (c->cg_ws_bytes < need) {
    if (c->cg_ws) HIPCHK(hipFree(c->cg_ws));
    c->cg_ws = nullptr;
    c->cg_ws_bytes = 0;
    HIPCHK(hipMalloc(&c->cg_ws, need));
    // stale window slots are read with zero weights: keep them finite
    HIPCHK(hipMemsetAsync(c->cg_ws, 0, need, c->s));
    c->cg_ws_bytes = need;
  }
  C *base = (C *)c->cg_ws;
  C *p0buf[2] = {base, base + n};
  C *ap = base + 2 * n, *tmp = base + 3 * n;
  C *pj = base + 4 * n;
  C *win = pj + (size_t)(nsh - 1) * n;
  auto slot = [&](int l) { return win + (size_t)l * n; };
  // r is the newest window slot; before the first iteration the last one
  C *r = slot(m - 1);
  // p0: pcur is the current direction, pfrozen its value at window start
  C *pcur = p0buf[0], *pfrozen = p0buf[0];
  double2 *ap64 = nullptr, *tmp64 = nullptr;
  if (fp32) {
    if (!c->cg_t64) HIPCHK(hipMalloc(&c->cg_t64, 2 * n * sizeof(double2)));
    ap64 = c->cg_t64;
    tmp64 = c->cg_t64 + n;
  }

  PsiPtrs psi;
  for (int j = 0; j < nsh; j++) psi.p[j] = psim[j];
  double sigma0 = shifts[0] + ((c->p.fmass > IMAG_TOL) ? c->p.fmass * c->p.fmass : 0.0);

  int nb = nblk(n);
  hipLaunchKernelGGL((k_cg_init<C>), dim3(nb), dim3(CB), 0, c->s, src, r, pcur, pj, nsh, psi,
                     n, c->d_part);
  HIPCHK(hipGetLastError());
  int e = reduce_partials(c, c->d_part, nb, 0);
  if (e) return e;
  double source_norm;
  if ((e = read_scalars(c, &source_norm, 1))) return e;
  double rsq = source_norm, rsqstop = errormin * source_norm;

  double zi[MAXSH], zim1[MAXSH], zip1[MAXSH], bi[MAXSH], bim1[MAXSH], al[MAXSH];
  int conv[MAXSH];
  for (int j = 0; j < nsh; j++) {
    zim1[j] = 1; zi[j] = 1; al[j] = 0; bim1[j] = 1; conv[j] = 0; zip1[j] = 1; bi[j] = 0;
  }
  // window coefficients, per system j (0 = base): A, E, C[m], D[m]
  const int stride = 2 + 2 * MAXWIN;
  if (!c->h_coef) {
    HIPCHK(hipHostMalloc((void **)&c->h_coef, sizeof(double) * stride * MAXSH));
    HIPCHK(hipMalloc((void **)&c->d_coef, sizeof(double) * stride * MAXSH));
  }
  double *h_coef = c->h_coef, *d_coef = c->d_coef;
  double A[MAXSH], E[MAXSH], Cw[MAXSH][MAXWIN], Dw[MAXSH][MAXWIN];
  int touched[MAXSH];
  int cnt = 0;
  auto reset_window = [&]() {
    for (int j = 0; j < nsh; j++) {
      A[j] = 1; E[j] = 0; touched[j] = 0;
      for (int l = 0; l < MAXWIN; l++) Cw[j][l] = Dw[j][l] = 0;
    }
    cnt = 0;
    pfrozen = pcur;
  };
  auto flush = [&]() -> int {
    FlushArgs fa;
    PPtrs pp;
    fa.nact = 0;
    for (int j = 0; j < nsh; j++) {
      if (!touched[j]) continue;
      double *cf = h_coef + fa.nact * stride;
      cf[0] = A[j]; cf[1] = E[j];
      for (int l = 0; l < MAXWIN; l++) { cf[2 + l] = Cw[j][l]; cf[2 + MAXWIN + l] = Dw[j][l]; }
      pp.p[fa.nact] = j == 0 ? (void *)pfrozen : (void *)(pj + (size_t)(j - 1) * n);
      fa.sh[fa.nact++] = j;
    }
    if (fa.nact > 0) {
      HIPCHK(hipMemcpyAsync(d_coef, h_coef, sizeof(double) * stride * fa.nact,
                            hipMemcpyHostToDevice, c->s));
      dim3 g((n + CB - 1) / CB), b(CB);
      switch (m) {
        case 1: hipLaunchKernelGGL((k_cg_flush<C, 1>), g, b, 0, c->s, pp, psi, win, d_coef, fa, n); break;
        case 4: hipLaunchKernelGGL((k_cg_flush<C, 4>), g, b, 0, c->s, pp, psi, win, d_coef, fa, n); break;
        case 8: hipLaunchKernelGGL((k_cg_flush<C, 8>), g, b, 0, c->s, pp, psi, win, d_coef, fa, n); break;
        case 16: hipLaunchKernelGGL((k_cg_flush<C, 16>), g, b, 0, c->s, pp, psi, win, d_coef, fa, n); break;
        default: hipLaunchKernelGGL((k_cg_flush<C, 32>), g, b, 0, c->s, pp, psi, win, d_coef, fa, n); break;
      }
      HIPCHK(hipGetLastError());
      // the pinned coefficient buffer is reused by the next flush
      HIPCHK(hipStreamSynchronize(c->s));
    }
    reset_window();
    return 0;
  };
  reset_window();

  int iteration = 0, since_reliable = 0;
  auto dsq = [&](const C *in, C *out, double *part, int *nbo) -> int {
    int r1, nbd = 0;
    if (fp32) {
      r1 = dslash32(c, (const float2 *)in, (float2 *)tmp, 1, nullptr, 0.f, nullptr, nullptr);
      if (!r1) r1 = dslash32(c, (const float2 *)tmp, (float2 *)out, -1, (const float2 *)in,
                             (float)sigma0, part, &nbd);
    } else {
      r1 = dslash64(c, (const double2 *)in, (double2 *)tmp, 1, nullptr, 0.0, nullptr, nullptr);
      if (!r1) r1 = dslash64(c, (const double2 *)tmp, (double2 *)out, -1, (const double2 *)in,
                             sigma0, part, &nbd);
    }
    *nbo = nbd;
    return r1;
  };
  // FP32 form: replace r by the FP64 true residual of the base system
  auto reliable_update = [&]() -> int {
    int r1, nbd = 0;
    if (cnt > 0 && (r1 = flush())) return r1;
    r1 = dslash64(c, psim[0], tmp64, 1, nullptr, 0.0, nullptr, nullptr);
    if (!r1) r1 = dslash64(c, tmp64, ap64, -1, psim[0], sigma0, c->d_part, &nbd);
    if (r1) return r1;
    hipLaunchKernelGGL((k_true_res<C>), dim3(nb), dim3(CB), 0, c->s, src, ap64, r, n, c->d_part);
    HIPCHK(hipGetLastError());
    if ((r1 = reduce_partials(c, c->d_part, nb, 0))) return r1;
    return read_scalars(c, &rsq, 1);
  };

  for (;;) {
    while (iteration < MaxCG && rsq > rsqstop) {
      int nbd;
      if ((e = dsq(pcur, ap, c->d_part, &nbd))) return e;
      if ((e = reduce_partials(c, c->d_part, nbd, 0))) return e;
      double cd;
      if ((e = read_scalars(c, &cd, 1))) return e;
      iteration++;
      c->total_iters++;
      bi[0] = -rsq / cd;
      zip1[0] = 1;
      for (int j = 1; j < nsh; j++) {
        if (conv[j]) continue;
        zip1[j] = zi[j] * zim1[j] * bim1[0];
        double c1 = bi[0] * al[0] * (zim1[j] - zi[j]);
        double c2 = zim1[j] * bim1[0] * (1 - (shifts[j] - shifts[0]) * bi[0]);
        zip1[j] /= c1 + c2;
        bi[j] = bi[0] * zip1[j] / zi[j];
      }
      C *rnew = slot(cnt);
      hipLaunchKernelGGL((k_cg_r<C>), dim3(nb), dim3(CB), 0, c->s, r, ap, rnew, bi[0], n,
                         c->d_part);
      HIPCHK(hipGetLastError());
      r = rnew;
      if ((e = reduce_partials(c, c->d_part, nb, 0))) return e;
      double rsqnew;
      if ((e = read_scalars(c, &rsqnew, 1))) return e;
      al[0] = rsqnew / rsq;
      for (int j = 1; j < nsh; j++)
        if (!conv[j]) al[j] = al[0] * zip1[j] * bi[j] / (zi[j] * bi[0]);
      C *pnext = pcur != pfrozen ? pcur : (pcur == p0buf[0] ? p0buf[1] : p0buf[0]);
      hipLaunchKernelGGL((k_cg_p0<C>), dim3((n + CB - 1) / CB), dim3(CB), 0, c->s, pcur, r,
                         pnext, al[0], n);
      HIPCHK(hipGetLastError());
      pcur = pnext;
      // advance the window coefficients: psi_j -= beta_j p_j, then
      // p_j = zeta_j r + alpha_j p_j (zeta_0 = 1)
      for (int j = 0; j < nsh; j++) {
        if (conv[j]) continue;
        touched[j] = 1;
        E[j] -= bi[j] * A[j];
        for (int l = 0; l < cnt; l++) Dw[j][l] -= bi[j] * Cw[j][l];
        A[j] *= al[j];
        for (int l = 0; l < cnt; l++) Cw[j][l] *= al[j];
        Cw[j][cnt] = zip1[j];
      }
      cnt++;
      rsq = rsqnew;
      for (int j = 1; j < nsh; j++)
        if (!conv[j] && rsq * zip1[j] * zip1[j] <= rsqstop) conv[j] = 1;
      for (int j = 0; j < nsh; j++)
        if (!conv[j]) { bim1[j] = bi[j]; zim1[j] = zi[j]; zi[j] = zip1[j]; }
      if (cnt == m && (e = flush())) return e;
      if (fp32 && ++since_reliable >= g_reliable && rsq > rsqstop) {
        since_reliable = 0;
        if ((e = reliable_update())) return e;
      }
    }
    if (cnt > 0 && (e = flush())) return e;
    if (!fp32 || iteration >= MaxCG) break;
    // FP32 form: confirm convergence on the FP64 true residual
    double iter_rsq = rsq;
    if ((e = reliable_update())) return e;
    since_reliable = 0;
    if (rsq <= rsqstop) break;
    (void)iter_rsq;
  }
  if (rsq > rsqstop) printf(" multi CONGRAD not converged\n rsq = %.4g\n", rsq);
  *size_r = rsq;
  return iteration;
}

int cg_multi(susy_ctx *c, const double2 *src, double2 *const *psim, int nshift,
             const double *shifts, int MaxCG, double errormin, double *size_r, int form) {
  if (nshift < 1 || nshift > MAXSH) {
    set_error("congrad_multi: nshift %d outside 1..%d", nshift, MAXSH);
    return SUSY_E_ARG;
  }
  if (form == SUSY_FORM_FP32)
    return cg_impl<float>(c, src, psim, nshift, shifts, MaxCG, errormin, size_r);
  return cg_impl<double>(c, src, psim, nshift, shifts, MaxCG, errormin, size_r);
}
