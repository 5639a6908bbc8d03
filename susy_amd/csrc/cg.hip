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

#include <vector>

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
    hold(t);
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
    vm_drain();
    C w = from64<C>(v);
    st_keep(r + i, w);
    st_keep(p0 + i, w);
    for (int j = 1; j < nsh; j++) st_keep(pj + (size_t)(j - 1) * n + i, w);
    for (int j = 0; j < nsh; j++) st_keep(psi.p[j] + i, make_double2(0, 0));
    s += v.x * v.x + v.y * v.y;
  }
  cg_block_sum(s, partial);
}

// Scalar state of one solve, in device memory.  The reference's recurrences
// (congrad_multi.c:60-180) run in single-thread kernels between the vector
// passes, so the host enqueues iterations without waiting for them; a
// solver that has converged turns every queued launch into a no-op.
struct CgState {
  int done, iteration;          // read back by the host (first 24 bytes)
  double rsq, rsqstop;
  int skip_rel, maxcg, nsh, cnt;
  double shifts[MAXSH];
  double zi[MAXSH], zim1[MAXSH], zip1[MAXSH], bi[MAXSH], bim1[MAXSH], al[MAXSH];
  int conv[MAXSH], touched[MAXSH];
  double A[MAXSH], E[MAXSH], Cw[MAXSH][MAXWIN], Dw[MAXSH][MAXWIN];
};
struct Shifts {
  double v[MAXSH];
};

// Arithmetic of the scalar kernels is not contracted into FMAs, so it rounds
// like the reference's host code.
// Initial state, one thread per window coefficient (MAXSH x MAXWIN threads):
// few stores per thread, so no store's data register is reused behind it.
#define STATE_THREADS (MAXSH * MAXWIN)
__global__ void __launch_bounds__(STATE_THREADS)
k_cgs_init(CgState *st, const double *__restrict__ source_norm, double errormin, int nsh,
           Shifts sh, int maxcg) {
#pragma clang fp contract(off)
  const int j = threadIdx.x / MAXWIN, l = threadIdx.x % MAXWIN;
  st->Cw[j][l] = 0;
  st->Dw[j][l] = 0;
  if (l == 0) {
    const double shj = j < nsh ? sh.v[j] : 0.0;
    st->shifts[j] = shj;
    st->zim1[j] = 1; st->zi[j] = 1; st->al[j] = 0; st->bim1[j] = 1; st->conv[j] = 0;
    st->zip1[j] = 1; st->bi[j] = 0;
    st->A[j] = 1; st->E[j] = 0; st->touched[j] = 0;
    hold(shj);
  }
  if (threadIdx.x == 0) {
    const double rsq = *source_norm;
    const double stop = errormin * rsq;
    st->rsq = rsq;
    st->rsqstop = stop;
    st->iteration = 0;
    st->maxcg = maxcg;
    st->nsh = nsh;
    st->cnt = 0;
    st->done = !(0 < maxcg && rsq > stop);
    st->skip_rel = 0;
    hold(rsq);
    hold(stop);
  }
}

// the while condition of the iteration loop, evaluated at the end of each
// iteration (after its p update, which the converged iteration still makes):
// everything queued after it -- flushes, reliable updates, further
// iterations -- then sees the solve finished
__global__ void k_cgs_stop(CgState *st) {
  if (!st->done && !(st->iteration < st->maxcg && st->rsq > st->rsqstop)) st->done = 1;
}

// beta and the shifted zeta / beta after A p0 and <p0, A p0> = *cd
__global__ void k_cgs_beta(CgState *st, const double *__restrict__ cd) {
#pragma clang fp contract(off)
  if (st->done) return;
  // every load before the first store (vm_drain: see its definition)
  const int it = st->iteration, nsh = st->nsh;
  const double rsq = st->rsq, cdv = *cd, al0 = st->al[0], bim10 = st->bim1[0],
               s0 = st->shifts[0];
  vm_drain();
  const double bi0 = -rsq / cdv;
  st_keep(&st->iteration, it + 1);
  st->bi[0] = bi0;
  hold(bi0);
  st_keep(&st->zip1[0], 1.0);
  for (int j = 1; j < nsh; j++) {
    const int cv = st->conv[j];
    const double zij = st->zi[j], zim1j = st->zim1[j], sj = st->shifts[j];
    vm_drain();
    if (cv) continue;
    double zp = zij * zim1j * bim10;
    const double c1 = bi0 * al0 * (zim1j - zij);
    const double c2 = zim1j * bim10 * (1 - (sj - s0) * bi0);
    zp /= c1 + c2;
    const double bj = bi0 * zp / zij;
    st->zip1[j] = zp;
    st->bi[j] = bj;
    hold(zp);
    hold(bj);
  }
}

// alpha, the window coefficients, convergence of the shifted systems and the
// advance of the recurrences after |r_new|^2 = *rsqnew.  One block of
// MAXSH x MAXWIN threads: thread (j, l) updates the window coefficients
// Cw[j][l], Dw[j][l]; thread l = 0 of shift j does that shift's scalars.
// Every element sees the same operations in the same order as a serial
// loop over j and l would make.
#define ALPHA_THREADS (MAXSH * MAXWIN)
__global__ void __launch_bounds__(ALPHA_THREADS)
k_cgs_alpha(CgState *st, const double *__restrict__ rsqnew) {
#pragma clang fp contract(off)
  if (st->done) return;
  __shared__ double al_s[MAXSH], bi_s[MAXSH];
  __shared__ int conv_s[MAXSH];
  const int j = threadIdx.x / MAXWIN, l = threadIdx.x % MAXWIN;
  // every load before the first store (vm_drain: see its definition)
  const int nsh = st->nsh, cnt = st->cnt;
  const double rn = *rsqnew, rsq = st->rsq, rsqstop = st->rsqstop;
  const bool lead = l == 0 && j < nsh;
  int cv = 0;
  double zij = 0, zp1 = 0, bij = 0, bi0 = 0, alj = 0, Ej = 0, Aj = 0, cw = 0, dw = 0;
  if (lead) {
    cv = st->conv[j];
    zij = st->zi[j];
    zp1 = st->zip1[j];
    bij = st->bi[j];
    bi0 = st->bi[0];
    alj = st->al[j];
    Ej = st->E[j];
    Aj = st->A[j];
  }
  if (j < nsh && l < cnt) {
    cw = st->Cw[j][l];
    dw = st->Dw[j][l];
  }
  vm_drain();
  if (lead) {
    const double a0 = rn / rsq;
    if (j == 0) alj = a0;
    else if (!cv) alj = a0 * zp1 * bij / (zij * bi0);
    al_s[j] = alj;
    bi_s[j] = bij;
    conv_s[j] = cv;
  }
  __syncthreads();
  if (j < nsh && !conv_s[j]) {
    // psi_j -= beta_j p_j, then p_j = zeta_j r + alpha_j p_j (zeta_0 = 1)
    if (l < cnt) {
      const double d2 = dw - bi_s[j] * cw, c2 = cw * al_s[j];
      st->Dw[j][l] = d2;
      st->Cw[j][l] = c2;
      hold(d2);
      hold(c2);
    }
    if (l == 0) {
      const double e2 = Ej - bij * Aj, a2 = Aj * alj;
      st->al[j] = alj;
      st->touched[j] = 1;
      st->E[j] = e2;
      st->A[j] = a2;
      st->Cw[j][cnt] = zp1;
      hold(e2);
      hold(a2);
    }
  }
  if (lead) {
    int c2 = cv;
    if (j > 0 && !cv && rn * zp1 * zp1 <= rsqstop) {
      c2 = 1;
      st->conv[j] = 1;
    }
    if (!c2) {
      st->bim1[j] = bij;
      st->zim1[j] = zij;
      st->zi[j] = zp1;
    }
  }
  if (threadIdx.x == 0) {
    st->cnt = cnt + 1;
    st->rsq = rn;
  }
}

// window state after a flush (MAXSH x MAXWIN threads)
__global__ void __launch_bounds__(STATE_THREADS)
k_cgs_reset(CgState *st, const int *__restrict__ skip) {
  const bool idle = (skip != nullptr && *skip) || st->cnt == 0;
  __syncthreads();   // every thread has read cnt before thread 0 clears it
  if (idle) return;
  const int j = threadIdx.x / MAXWIN, l = threadIdx.x % MAXWIN;
  st->Cw[j][l] = 0;
  st->Dw[j][l] = 0;
  if (l == 0) {
    st->A[j] = 1;
    st->E[j] = 0;
    st->touched[j] = 0;
  }
  if (threadIdx.x == 0) st->cnt = 0;
}

// the reliable update runs only while the solve is open and unconverged
__global__ void k_cgs_relcheck(CgState *st) {
  st->skip_rel = st->done || !(st->rsq > st->rsqstop);
}
__global__ void k_cgs_setrsq(CgState *st, const double *__restrict__ v,
                             const int *__restrict__ skip) {
  if (skip != nullptr && *skip) return;
  st->rsq = *v;
}
__global__ void k_cgs_reopen(CgState *st) { st->done = 0; }

// rout = rin + beta_0 * ap; partial |rout|^2.  rout is the next slot of the
// residual window and rin the previous one (the same slot when the window
// holds one residual: the update is elementwise, so in place is safe).
template <typename C>
__global__ void __launch_bounds__(CB)
k_cg_r(const C *rin, const C *__restrict__ ap, C *rout, const CgState *__restrict__ st,
       size_t n, double *__restrict__ partial) {
  if (st->done) return;
  const double beta0 = st->bi[0];
  double s = 0;
  for (size_t i = (size_t)blockIdx.x * CB + threadIdx.x; i < n; i += (size_t)gridDim.x * CB) {
    C a = ap[i], v = rin[i];
    vm_drain();   // loads of this element complete before its store
    v.x = fma(beta0, a.x, v.x);
    v.y = fma(beta0, a.y, v.y);
    st_keep(rout + i, v);
    s += (double)v.x * (double)v.x + (double)v.y * (double)v.y;
  }
  cg_block_sum(s, partial);
}

// pout = r + alpha_0 pin (pout is pin, or the other p0 buffer at the first
// iteration of a window, which leaves the window-start p0 intact), one thread
// per (site, Twist_Fermion component); the thread also stores the
// operator's pre-pass values of the new p for the next D:
// aux[0] = Tr p_0, aux[1+a] = Tr(U_a^-1 p_{1+a}) (aux = nullptr: no G terms)
template <typename C> struct ScalarOf;
template <> struct ScalarOf<float2> { typedef float t; };
template <> struct ScalarOf<double2> { typedef double t; };
template <typename C>
__global__ void __launch_bounds__(CB)
k_cg_p0_prep(const C *pin, const C *__restrict__ r, C *pout, const C *__restrict__ uinv,
             C *__restrict__ aux, const CgState *__restrict__ st, int V) {
  typedef typename ScalarOf<C>::t R;
  if (st->done) return;
  const int s = blockIdx.x * CB + threadIdx.x, k = blockIdx.y;
  if (s >= V) return;
  const double alpha0 = st->al[0];
  const bool pre = aux != nullptr && k <= NUMLINK;
  // every load before the first store (vm_drain: see its definition)
  C v[NN], rv[NN], u[NN];
#pragma unroll
  for (int e = 0; e < NN; e++) {
    const size_t i = ((size_t)k * NN + e) * V + s;
    v[e] = pin[i];
    rv[e] = r[i];
  }
  if (pre && k > 0) {
#pragma unroll
    for (int e = 0; e < NN; e++) u[e] = uinv[((size_t)(k - 1) * NN + e) * V + s];
  }
  vm_drain();
#pragma unroll
  for (int e = 0; e < NN; e++) {
    C p = v[e];
    p.x = rv[e].x + alpha0 * p.x;
    p.y = rv[e].y + alpha0 * p.y;
    pout[((size_t)k * NN + e) * V + s] = p;
    v[e] = p;
  }
#pragma unroll
  for (int e = 0; e < NN; e++) hold(v[e]);
  if (!pre) return;
  R tx = 0, ty = 0;
  if (k == 0) {
#pragma unroll
    for (int i = 0; i < NCOL; i++) {
      tx += v[i * NCOL + i].x;
      ty += v[i * NCOL + i].y;
    }
  } else {
#pragma unroll
    for (int i = 0; i < NCOL; i++)
#pragma unroll
      for (int j = 0; j < NCOL; j++) {
        const C uu = u[i * NCOL + j], q = v[j * NCOL + i];
        tx += uu.x * q.x - uu.y * q.y;
        ty += uu.x * q.y + uu.y * q.x;
      }
  }
  C t;
  t.x = tx;
  t.y = ty;
  st_keep(aux + (size_t)k * V + s, t);
}

// Window flush.  Every touched system j updates psi_j and, for j >= 1, p_j:
//   p_j   <- A_j p_j(k0) + sum_l C_jl r_l
//   psi_j <- psi_j + E_j p_j(k0) + sum_l D_jl r_l
// with the coefficients of the device state; pp.p[j] is p_j(k0) (for the
// base system the frozen window-start p0, which is not written).
struct PPtrs {
  void *p[MAXSH];
};
template <typename C, int M>
__global__ void __launch_bounds__(CB)
k_cg_flush(PPtrs pp, PsiPtrs psi, const C *__restrict__ win, const CgState *__restrict__ st,
           const int *__restrict__ skip, size_t n) {
  const int cnt = st->cnt;
  if ((skip != nullptr && *skip) || cnt == 0) return;
  size_t i = (size_t)blockIdx.x * CB + threadIdx.x;
  if (i >= n) return;
  // only this window's slots are read (the others may hold data of another
  // solve, in either precision)
  double2 w[M];
#pragma unroll
  for (int l = 0; l < M; l++) w[l] = l < cnt ? to64(win[(size_t)l * n + i]) : make_double2(0, 0);
  const int nsh = st->nsh;
  for (int j = 0; j < nsh; j++) {
    if (!st->touched[j]) continue;
    C *pk = (C *)pp.p[j] + i;
    double2 p = to64(*pk);
    double2 x = psi.p[j][i];
    const double A = st->A[j], E = st->E[j];
    vm_drain();   // this shift's loads complete before its stores
    double2 np = make_double2(A * p.x, A * p.y);
    x.x += E * p.x;
    x.y += E * p.y;
#pragma unroll
    for (int l = 0; l < M; l++) {
      const double cl = st->Cw[j][l], dl = st->Dw[j][l];
      np.x += cl * w[l].x;
      np.y += cl * w[l].y;
      x.x += dl * w[l].x;
      x.y += dl * w[l].y;
    }
    if (j > 0) st_keep(pk, from64<C>(np));
    st_keep(psi.p[j] + i, x);
  }
}

// r = src - ap (FP64 true residual, stored in the CG precision); partial |r|^2
template <typename C>
__global__ void __launch_bounds__(CB)
k_true_res(const double2 *__restrict__ src, const double2 *__restrict__ ap,
           C *__restrict__ r, size_t n, double *__restrict__ partial,
           const int *__restrict__ skip) {
  if (skip != nullptr && *skip) return;
  double s = 0;
  for (size_t i = (size_t)blockIdx.x * CB + threadIdx.x; i < n; i += (size_t)gridDim.x * CB) {
    double2 a = src[i], b = ap[i];
    vm_drain();
    double2 v = make_double2(a.x - b.x, a.y - b.y);
    st_keep(r + i, from64<C>(v));
    s += v.x * v.x + v.y * v.y;
  }
  cg_block_sum(s, partial);
}

// ---------------------------------------------------------------------
static int nblk(size_t n) {
  size_t nb = (n + CB - 1) / CB;
  return (int)(nb < 2048 ? nb : 2048);
}

// residuals kept per window (<= MAXWIN), per precision
static int g_window64 = 32, g_window32 = 32;
// FP32 form: iterations between true residuals; a multiple of the window, so
// the reliable update's flush finds the window already flushed
static int g_reliable = 128;

extern "C" int susy_set_cg_options(int window, int reliable) {
  if ((window != 0 && window != 1 && window != 4 && window != 8 && window != 16 &&
       window != 32) ||
      reliable < 1) {
    set_error("cg options: window in {0,1,4,8,16,32}, reliable >= 1");
    return SUSY_E_ARG;
  }
  g_window64 = window ? window : 32;
  g_window32 = window ? window : 32;
  g_reliable = reliable;
  return 0;
}

// iterations enqueued between two read-backs of the solver status
#define CG_BATCH 16

template <typename R>
static int cg_impl(susy_ctx *c, const double2 *src, double2 *const *psim, int nsh,
                   const double *shifts, int MaxCG, double errormin, double *size_r) {
  typedef typename CT<R>::t C;
  const bool fp32 = sizeof(R) == sizeof(float);
  const size_t n = tf_elems(c);
  const int m = fp32 ? g_window32 : g_window64;
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
  if (!c->cg_state) HIPCHK(hipMalloc(&c->cg_state, sizeof(CgState)));
  CgState *st = (CgState *)c->cg_state;
  const int *d_done = &st->done, *d_skiprel = &st->skip_rel;
  C *base = (C *)c->cg_ws;
  C *p0buf[2] = {base, base + n};
  C *ap = base + 2 * n, *tmp = base + 3 * n;
  C *pj = base + 4 * n;
  C *win = pj + (size_t)(nsh - 1) * n;
  auto slot = [&](int l) { return win + (size_t)l * n; };
  double2 *ap64 = nullptr, *tmp64 = nullptr;
  if (fp32) {
    if (!c->cg_t64) HIPCHK(hipMalloc(&c->cg_t64, 2 * n * sizeof(double2)));
    ap64 = c->cg_t64;
    tmp64 = c->cg_t64 + n;
  }
  PsiPtrs psi;
  for (int j = 0; j < nsh; j++) psi.p[j] = psim[j];
  double sigma0 = shifts[0] + ((c->p.fmass > IMAG_TOL) ? c->p.fmass * c->p.fmass : 0.0);
  double *d_norm = c->d_scal, *d_cd = c->d_scal + 1, *d_rn = c->d_scal + 2,
         *d_true = c->d_scal + 3;

  // Host bookkeeping of the buffers, a function of the iteration count: r is
  // the newest window slot (before the first iteration the last one), pcur
  // the current p0 and pfrozen its value at the window start.
  C *r = slot(m - 1), *pcur = p0buf[0], *pfrozen = p0buf[0];
  int host_cnt = 0, since_reliable = 0;
  // the p update also writes the operator pre-pass of the new p (G terms
  // only); the first iteration's p (the source) gets the ordinary pre-pass
  if (fp32 && !c->f_fresh) {
    const int e0 = fields_refresh32(c);
    if (e0) return e0;
  }
  C *aux_c = !c->doG ? nullptr : fp32 ? (C *)c->f_aux : (C *)c->d_aux;
  const C *uinv_c = fp32 ? (const C *)c->f_uinv : (const C *)c->d_uinv;
  bool p_prepped = false;

  const int nb = nblk(n);
  hipLaunchKernelGGL((k_cg_init<C>), dim3(nb), dim3(CB), 0, c->s, src, r, pcur, pj, nsh, psi,
                     n, c->d_part);
  HIPCHK(hipGetLastError());
  int e = reduce_partials(c, c->d_part, nb, 0);
  if (e) return e;
  Shifts shv;
  for (int j = 0; j < MAXSH; j++) shv.v[j] = j < nsh ? shifts[j] : 0.0;
  hipLaunchKernelGGL(k_cgs_init, dim3(1), dim3(STATE_THREADS), 0, c->s, st, d_norm, errormin, nsh,
                     shv, MaxCG);
  HIPCHK(hipGetLastError());

  auto one = [&](void (*k)(CgState *)) -> int {
    hipLaunchKernelGGL(k, dim3(1), dim3(1), 0, c->s, st);
    HIPCHK(hipGetLastError());
    return 0;
  };
  // flush the window (if the device's window holds any iteration) and reset it
  auto flush = [&](const int *skip) -> int {
    PPtrs pp;
    pp.p[0] = pfrozen;
    for (int j = 1; j < nsh; j++) pp.p[j] = pj + (size_t)(j - 1) * n;
    dim3 g((n + CB - 1) / CB), b(CB);
    switch (m) {
      case 1: hipLaunchKernelGGL((k_cg_flush<C, 1>), g, b, 0, c->s, pp, psi, win, st, skip, n); break;
      case 4: hipLaunchKernelGGL((k_cg_flush<C, 4>), g, b, 0, c->s, pp, psi, win, st, skip, n); break;
      case 8: hipLaunchKernelGGL((k_cg_flush<C, 8>), g, b, 0, c->s, pp, psi, win, st, skip, n); break;
      case 16: hipLaunchKernelGGL((k_cg_flush<C, 16>), g, b, 0, c->s, pp, psi, win, st, skip, n); break;
      default: hipLaunchKernelGGL((k_cg_flush<C, 32>), g, b, 0, c->s, pp, psi, win, st, skip, n); break;
    }
    hipLaunchKernelGGL(k_cgs_reset, dim3(1), dim3(STATE_THREADS), 0, c->s, st, skip);
    HIPCHK(hipGetLastError());
    host_cnt = 0;
    pfrozen = pcur;
    return 0;
  };
  // in_prepped: the operator's pre-pass values of `in` are already in the
  // context's aux buffer (written by k_cg_p0_prep)
  auto dsq = [&](const C *in, C *out, int *nbo, bool in_prepped) -> int {
    int r1, nbd = 0;
    if (fp32) {
      r1 = dsq_pair32(c, (const float2 *)in, (float2 *)tmp, (float2 *)out, (float)sigma0,
                      c->d_part, &nbd, d_done, !in_prepped);
    } else {
      r1 = dsq_pair64(c, (const double2 *)in, (double2 *)tmp, (double2 *)out, sigma0,
                      c->d_part, &nbd, d_done, !in_prepped);
    }
    *nbo = nbd;
    return r1;
  };
  // FP32 form: replace r by the FP64 true residual of the base system (skip:
  // the device flag that cancels it, or nullptr to run it unconditionally)
  auto reliable_update = [&](const int *skip) -> int {
    int r1, nbd = 0;
    if ((r1 = flush(skip))) return r1;
    r1 = dsq_pair64(c, psim[0], tmp64, ap64, sigma0, c->d_part, &nbd, skip);
    if (r1) return r1;
    hipLaunchKernelGGL((k_true_res<C>), dim3(nb), dim3(CB), 0, c->s, src, ap64, r, n, c->d_part,
                       skip);
    HIPCHK(hipGetLastError());
    if ((r1 = reduce_partials(c, c->d_part, nb, 3))) return r1;
    hipLaunchKernelGGL(k_cgs_setrsq, dim3(1), dim3(1), 0, c->s, st, d_true, skip);
    HIPCHK(hipGetLastError());
    return 0;
  };
  struct Status {
    int done, iteration;
    double rsq, rsqstop;
  } stat = {0, 0, 0.0, 0.0};
  auto read_status = [&]() -> int {
    HIPCHK(hipMemcpyAsync(c->h_scal, st, sizeof(Status), hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    memcpy(&stat, c->h_scal, sizeof(Status));
    return 0;
  };

  // buffers after each queued iteration (before its flush / reliable update)
  struct Pos {
    C *r, *p, *pfrozen;
  };
  Pos log[CG_BATCH];
  for (;;) {
    // batches of iterations until the device reports the loop finished
    for (;;) {
      const int it_before = stat.iteration;
      const Pos start = {r, pcur, pfrozen};
      for (int k = 0; k < CG_BATCH; k++) {
        int nbd;
        if ((e = dsq(pcur, ap, &nbd, p_prepped))) return e;
        if ((e = reduce_partials(c, c->d_part, nbd, 1))) return e;
        hipLaunchKernelGGL(k_cgs_beta, dim3(1), dim3(1), 0, c->s, st, d_cd);
        C *rnew = slot(host_cnt);
        hipLaunchKernelGGL((k_cg_r<C>), dim3(nb), dim3(CB), 0, c->s, r, ap, rnew, st, n,
                           c->d_part);
        HIPCHK(hipGetLastError());
        if ((e = reduce_partials(c, c->d_part, nb, 2))) return e;
        hipLaunchKernelGGL(k_cgs_alpha, dim3(1), dim3(ALPHA_THREADS), 0, c->s, st, d_rn);
        C *pnext = pcur != pfrozen ? pcur : (pcur == p0buf[0] ? p0buf[1] : p0buf[0]);
        hipLaunchKernelGGL((k_cg_p0_prep<C>), dim3((c->V + CB - 1) / CB, NTF), dim3(CB), 0, c->s,
                           pcur, rnew, pnext, uinv_c, aux_c, st, c->V);
        HIPCHK(hipGetLastError());
        p_prepped = aux_c != nullptr;
        r = rnew;
        pcur = pnext;
        host_cnt++;
        if ((e = one(k_cgs_stop))) return e;
        log[k] = {r, pcur, pfrozen};
        // window end: a solve that has finished keeps its open window for
        // the final flush below, made with the restored buffers
        if (host_cnt == m && (e = flush(d_done))) return e;
        if (fp32 && ++since_reliable >= g_reliable) {
          since_reliable = 0;
          if ((e = one(k_cgs_relcheck))) return e;
          if ((e = reliable_update(d_skiprel))) return e;
        }
      }
      if ((e = read_status())) return e;
      if (!stat.done) continue;
      // The loop ended inside this batch: the launches queued after it did
      // nothing (flushes and reliable updates included), so the buffers are
      // those of its last real iteration.
      const int real = stat.iteration - it_before;
      const Pos at = real > 0 ? log[real - 1] : start;
      r = at.r;
      pcur = at.p;
      pfrozen = at.pfrozen;
      break;
    }
    if ((e = flush(nullptr))) return e;
    if (!fp32 || stat.iteration >= MaxCG) break;
    // FP32 form: confirm convergence on the FP64 true residual
    if ((e = reliable_update(nullptr))) return e;
    since_reliable = 0;
    if ((e = read_status())) return e;
    // resume only as the loop condition would (a NaN residual ends the solve)
    if (!(stat.rsq > stat.rsqstop)) break;
    if ((e = one(k_cgs_reopen))) return e;
    stat.done = 0;
  }
  if ((e = read_status())) return e;
  c->total_iters += stat.iteration;
  if (!(stat.rsq <= stat.rsqstop))
    printf(" multi CONGRAD not converged\n rsq = %.4g\n", stat.rsq);
  *size_r = stat.rsq;
  return stat.iteration;
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
