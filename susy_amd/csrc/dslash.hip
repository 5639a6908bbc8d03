// Twisted fermion operator D and D^dagger on the A4* lattice
// (reference: 4d_Q16/susy/utilities.c:141-865), one fused gather-only kernel.
//
// Every thread owns one output matrix (one of the 16 Twist_Fermion components)
// of one site and sums all terms landing on it:
//   site  eta      : DbminusLtoS + detLtoS                 (10 matmuls)
//   link  psi_b    : Dminus + DbplusStoL + detStoL          (10 matmuls)
//   plaq  chi_ab   : Dplus + DbplusPtoP + DbminusPtoP       (16 matmuls)
// so no output is written twice and no atomics are needed.  D^dagger uses
// D^dagger v = -(D v^dagger)^dagger (utilities.c:780-788, 856-864): the
// adjoint is applied on the loads and the stores, in registers.
//
// NCOL <= 4 complex matrix products are VALU FMA chains: an output matrix
// accumulates in registers, the right operand of each product is held in
// registers while the left operand streams element by element.
//
// Block mapping is XCD-aware: consecutive hardware block ids are dealt to the
// 8 XCDs round-robin, so the 16 component-blocks of one site chunk are given
// ids that land on the same XCD and share its L2.
#include "susy_internal.h"

__constant__ PlaqTerms c_pt;

int dslash_init_tables(const PlaqTerms &pt) {
  HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(c_pt), &pt, sizeof(PlaqTerms)));
  return 0;
}

#define DS_BLOCK 128

template <typename C, typename R>
__device__ __forceinline__ void cfma(C &acc, R ar, R ai, C b) {
  acc.x = fma(ar, b.x, acc.x);
  acc.x = fma(-ai, b.y, acc.x);
  acc.y = fma(ar, b.y, acc.y);
  acc.y = fma(ai, b.x, acc.y);
}

// acc += s * op(A) * op(B); A, B point at element (0,0) of a matrix at the
// wanted site; elements are V apart.  TA/TB select the dagger.
template <typename R, bool TA, bool TB>
__device__ __forceinline__ void mmacc(typename CT<R>::t (&acc)[NCOL][NCOL], R s,
                                      const typename CT<R>::t *__restrict__ A,
                                      const typename CT<R>::t *__restrict__ B,
                                      int V) {
  typedef typename CT<R>::t C;
  C b[NCOL][NCOL];
#pragma unroll
  for (int i = 0; i < NCOL; i++)
#pragma unroll
    for (int j = 0; j < NCOL; j++) {
      C v = TB ? B[(j * NCOL + i) * V] : B[(i * NCOL + j) * V];
      if (TB) v.y = -v.y;
      b[i][j] = v;
    }
#pragma unroll
  for (int i = 0; i < NCOL; i++)
#pragma unroll
    for (int j = 0; j < NCOL; j++) {
      C a = TA ? A[(j * NCOL + i) * V] : A[(i * NCOL + j) * V];
      R ar = s * a.x, ai = TA ? -s * a.y : s * a.y;
#pragma unroll
      for (int k = 0; k < NCOL; k++) cfma(acc[i][k], ar, ai, b[j][k]);
    }
}

template <typename R>
__device__ __forceinline__ R bcsign(int t, int dt, int nt, R pbc) {
  int tt = t + dt;
  return (tt < 0 || tt > nt - 1) ? pbc : (R)1;
}

// Pre-pass for the plaquette-determinant terms (doG): per site
//   aux[0] = Tr eta,  aux[1+a] = Tr(U_a^-1 psi_a)
// with eta, psi taken adjointed for D^dagger (utilities.c:793-799, 706-712).
template <typename R, bool DAG>
__global__ void __launch_bounds__(256)
k_dslash_prep(const typename CT<R>::t *__restrict__ src,
              const typename CT<R>::t *__restrict__ Uinv,
              typename CT<R>::t *__restrict__ aux, int V) {
  typedef typename CT<R>::t C;
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= V) return;
  C tr = mkc<C>(0, 0);
#pragma unroll
  for (int i = 0; i < NCOL; i++) {
    C v = src[(i * NCOL + i) * V + s];
    tr.x += v.x;
    tr.y += DAG ? -v.y : v.y;
  }
  aux[s] = tr;
#pragma unroll
  for (int a = 0; a < NUMLINK; a++) {
    const C *psi = src + (size_t)(1 + a) * NN * V + s;
    const C *ui = Uinv + (size_t)a * NN * V + s;
    C t = mkc<C>(0, 0);
#pragma unroll
    for (int i = 0; i < NCOL; i++)
#pragma unroll
      for (int j = 0; j < NCOL; j++) {
        C u = ui[(i * NCOL + j) * V];
        C p = DAG ? psi[(i * NCOL + j) * V] : psi[(j * NCOL + i) * V];
        if (DAG) p.y = -p.y;
        t.x += u.x * p.x - u.y * p.y;
        t.y += u.x * p.y + u.y * p.x;
      }
    aux[(size_t)(1 + a) * V + s] = t;
  }
}

// EPI = 0: dst = result.  EPI = 1: dst = result + shift * p0, and the block's
// partial sum of Re Tr(p0^dag dst) goes to partial[block].
template <typename R, bool DAG, int EPI>
__global__ void __launch_bounds__(DS_BLOCK)
k_dslash(const typename CT<R>::t *__restrict__ src, typename CT<R>::t *__restrict__ dst,
         const typename CT<R>::t *__restrict__ U, const typename CT<R>::t *__restrict__ Uinv,
         const typename CT<R>::t *__restrict__ pdet, const typename CT<R>::t *__restrict__ aux,
         const int *__restrict__ nbr, const int *__restrict__ tco, int V, int nt,
         R pbc, int doG, R G, int nchunks, const typename CT<R>::t *__restrict__ p0,
         R shift, double *__restrict__ partial) {
  typedef typename CT<R>::t C;
  const int L = blockIdx.x;
  const int slot = L >> 3;
  const int comp = slot % NTF;
  const int chunk = (slot / NTF) * 8 + (L & 7);
  const int s = chunk * DS_BLOCK + threadIdx.x;
  const bool live = chunk < nchunks && s < V;

  double dot = 0.0;
  if (live) {
    const size_t M = (size_t)NN * V;
    const C *S_site = src + s * 0;
    C acc[NCOL][NCOL];
#pragma unroll
    for (int i = 0; i < NCOL; i++)
#pragma unroll
      for (int k = 0; k < NCOL; k++) acc[i][k] = mkc<C>(0, 0);
    const int t = tco[s];
    (void)S_site;

    if (comp == 0) {
      // ---- site output: 1/2 sum_a [psi_a(x) Ubar_a(x) - bc(-a) Ubar_a(x-a) psi_a(x-a)]
#pragma unroll 1
      for (int a = 0; a < NUMLINK; a++) {
        int xm = nbr[g_bwd(a) * V + s];
        mmacc<R, DAG, true>(acc, (R)0.5, src + (1 + a) * M + s, U + a * M + s, V);
        mmacc<R, true, DAG>(acc, (R)-0.5 * bcsign<R>(t, -link_tau(a), nt, pbc),
                            U + a * M + xm, src + (1 + a) * M + xm, V);
      }
      if (doG) {
        // detLtoS: -(i G C2 sqrt(N)/2) sum_{a!=b} plaqdet_ab [T_b(x) + bc(b) T_a(x+b)]
        //          * Lambda_{DIMF-1},  Lambda_{DIMF-1} = i I / sqrt(N)
        C sum = mkc<C>(0, 0);
#pragma unroll 1
        for (int a = 0; a < NUMLINK; a++) {
#pragma unroll 1
          for (int b = 0; b < NUMLINK; b++) {
            if (a == b) continue;
            int xb = nbr[g_fwd(b) * V + s];
            C tb = aux[(1 + b) * V + s], ta = aux[(1 + a) * V + xb];
            R sg = bcsign<R>(t, link_tau(b), nt, pbc);
            C t2 = mkc<C>(tb.x + sg * ta.x, tb.y + sg * ta.y);
            C pd = pdet[(a * NUMLINK + b) * V + s];
            sum.x += pd.x * t2.x - pd.y * t2.y;
            sum.y += pd.x * t2.y + pd.y * t2.x;
          }
        }
        // trd = i lg sum; acc -= trd * (i/sqrt N) I = (lg/sqrt N) sum I
        R f = (R)(0.5 * C2) * G;  // lg / sqrt(N) = G C2 / 2
#pragma unroll
        for (int i = 0; i < NCOL; i++) {
          acc[i][i].x += f * sum.x;
          acc[i][i].y += f * sum.y;
        }
      }
    } else if (comp <= NUMLINK) {
      // ---- link output b
      const int b = comp - 1;
      const int xb = nbr[g_fwd(b) * V + s];
#pragma unroll 1
      for (int a = 0; a < NUMLINK; a++) {
        if (a == b) continue;
        // Dminus: U_a(x+b) chi_ab(x) - bc(-a) chi_ab(x-a) U_a(x-a)
        const int xm = nbr[g_bwd(a) * V + s];
        const R sg = a < b ? (R)1 : (R)-1;   // chi_ab = sg * stored chi_{min,max}
        const C *chi = src + (6 + plaq_idx(a, b)) * M;
        mmacc<R, false, DAG>(acc, sg, U + a * M + xb, chi + s, V);
        mmacc<R, DAG, false>(acc, -sg * bcsign<R>(t, -link_tau(a), nt, pbc), chi + xm,
                             U + a * M + xm, V);
      }
      // DbplusStoL: 1/2 [bc(b) eta(x+b) Ubar_b(x) - Ubar_b(x) eta(x)]
      mmacc<R, DAG, true>(acc, (R)0.5 * bcsign<R>(t, link_tau(b), nt, pbc), src + xb,
                          U + b * M + s, V);
      mmacc<R, true, DAG>(acc, (R)-0.5, U + b * M + s, src + s, V);
      if (doG) {
        // detStoL: -G C2/2 U_b^-1(x) sum_{c!=b} [Tr eta(x) plaqdet_cb(x)
        //                                        + bc(-c) Tr eta(x-c) plaqdet_bc(x-c)]
        C sum = mkc<C>(0, 0);
        const C te = aux[s];
#pragma unroll 1
        for (int c = 0; c < NUMLINK; c++) {
          if (c == b) continue;
          const int xm = nbr[g_bwd(c) * V + s];
          C p1 = pdet[(c * NUMLINK + b) * V + s];
          C p2 = pdet[(b * NUMLINK + c) * V + xm];
          C tm = aux[xm];
          R sg = bcsign<R>(t, -link_tau(c), nt, pbc);
          sum.x += te.x * p1.x - te.y * p1.y + sg * (tm.x * p2.x - tm.y * p2.y);
          sum.y += te.x * p1.y + te.y * p1.x + sg * (tm.x * p2.y + tm.y * p2.x);
        }
        R f = (R)(-0.5 * C2) * G;
        R sr = f * sum.x, si = f * sum.y;
        const C *ui = Uinv + b * M + s;
#pragma unroll
        for (int i = 0; i < NCOL; i++)
#pragma unroll
          for (int k = 0; k < NCOL; k++) {
            C u = ui[(i * NCOL + k) * V];
            acc[i][k].x += sr * u.x - si * u.y;
            acc[i][k].y += sr * u.y + si * u.x;
          }
      }
    } else {
      // ---- plaquette output (a,b), a<b
      const int p = comp - 1 - NUMLINK;
      const int a = c_pt.a[p], b = c_pt.b[p];
      const int xa = nbr[g_fwd(a) * V + s], xb = nbr[g_fwd(b) * V + s];
      // Dplus: bc(a) U_a(x) psi_b(x+a) - psi_b(x) U_a(x+b)
      //        - bc(b) U_b(x) psi_a(x+b) + psi_a(x) U_b(x+a)
      mmacc<R, false, DAG>(acc, bcsign<R>(t, link_tau(a), nt, pbc), U + a * M + s,
                           src + (1 + b) * M + xa, V);
      mmacc<R, DAG, false>(acc, (R)-1, src + (1 + b) * M + s, U + a * M + xb, V);
      mmacc<R, false, DAG>(acc, -bcsign<R>(t, link_tau(b), nt, pbc), U + b * M + s,
                           src + (1 + a) * M + xb, V);
      mmacc<R, DAG, false>(acc, (R)1, src + (1 + a) * M + s, U + b * M + xa, V);
#pragma unroll 1
      for (int n = 0; n < 3; n++) {
        // DbplusPtoP
        const int c = c_pt.pp_c[p][n];
        const C *chi = src + (6 + c_pt.pp_de[p][n]) * M;
        const int x1 = nbr[c_pt.pp_g1[p][n] * V + s];
        const int x2 = nbr[c_pt.pp_g2[p][n] * V + s];
        const int xmc = nbr[g_bwd(c) * V + s];
        const R tr = (R)c_pt.pp_tr[p][n];
        mmacc<R, DAG, true>(acc, tr * bcsign<R>(t, c_pt.pp_dt3[p][n], nt, pbc), chi + x2,
                            U + c * M + x1, V);
        mmacc<R, true, DAG>(acc, -tr * bcsign<R>(t, c_pt.pp_dt2[p][n], nt, pbc),
                            U + c * M + xmc, chi + x1, V);
      }
#pragma unroll 1
      for (int n = 0; n < 3; n++) {
        // DbminusPtoP
        const int c = c_pt.pm_c[p][n];
        const C *chi = src + (6 + c_pt.pm_ab[p][n]) * M;
        const int x2 = nbr[c_pt.pm_g2[p][n] * V + s];
        const int x1 = nbr[c_pt.pm_g1[p][n] * V + s];
        const int xmc = nbr[g_bwd(c) * V + s];
        const R tr = (R)c_pt.pm_tr[p][n];
        mmacc<R, DAG, true>(acc, tr * bcsign<R>(t, c_pt.pm_dt2[p][n], nt, pbc), chi + x2,
                            U + c * M + x1, V);
        mmacc<R, true, DAG>(acc, -tr * bcsign<R>(t, c_pt.pm_dt3[p][n], nt, pbc),
                            U + c * M + xmc, chi + x1, V);
      }
    }

    // ---- store (D^dagger: negated adjoint)
    C *out = dst + comp * M + s;
    const C *pp = (EPI == 1) ? p0 + comp * M + s : nullptr;
#pragma unroll
    for (int i = 0; i < NCOL; i++)
#pragma unroll
      for (int k = 0; k < NCOL; k++) {
        C v = DAG ? mkc<C>(-acc[k][i].x, acc[k][i].y) : acc[i][k];
        if (EPI == 1) {
          C q = pp[(i * NCOL + k) * V];
          v.x = fma(shift, q.x, v.x);
          v.y = fma(shift, q.y, v.y);
          dot += (double)q.x * (double)v.x + (double)q.y * (double)v.y;
        }
        out[(i * NCOL + k) * V] = v;
      }
  }
  if (EPI == 1) {
    // deterministic block reduction: wavefront butterfly, then across waves
    for (int off = 32; off > 0; off >>= 1) dot += __shfl_xor(dot, off, 64);
    __shared__ double red[DS_BLOCK / 64];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = dot;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0;
      for (int w = 0; w < DS_BLOCK / 64; w++) t += red[w];
      partial[L] = t;
    }
  }
}

template <typename R>
static int dslash_launch(susy_ctx *c, const typename CT<R>::t *src,
                         typename CT<R>::t *dst, int sign, const typename CT<R>::t *p0,
                         R shift, double *partial, int *nblocks,
                         const typename CT<R>::t *U, const typename CT<R>::t *Uinv,
                         const typename CT<R>::t *pdet, typename CT<R>::t *aux) {
  if (sign != 1 && sign != -1) {
    set_error("fermion_op: incorrect sign %d", sign);
    return SUSY_E_ARG;
  }
  const int V = c->V;
  if (c->doG) {
    int nb = (V + 255) / 256;
    if (sign == 1)
      hipLaunchKernelGGL((k_dslash_prep<R, false>), dim3(nb), dim3(256), 0, c->s, src, Uinv, aux, V);
    else
      hipLaunchKernelGGL((k_dslash_prep<R, true>), dim3(nb), dim3(256), 0, c->s, src, Uinv, aux, V);
    HIPCHK(hipGetLastError());
  }
  int nchunks = (V + DS_BLOCK - 1) / DS_BLOCK;
  int padded = (nchunks + 7) / 8 * 8;
  int nb = padded * NTF;
  if (nblocks) *nblocks = nb;
  R pbc = (R)c->p.PBC, G = (R)c->p.G;
#define DS_ARGS src, dst, U, Uinv, pdet, aux, c->d_nbr, c->d_t, V, c->p.nt, pbc, c->doG, G, \
                nchunks, p0, shift, partial
  if (p0 == nullptr) {
    if (sign == 1)
      hipLaunchKernelGGL((k_dslash<R, false, 0>), dim3(nb), dim3(DS_BLOCK), 0, c->s, DS_ARGS);
    else
      hipLaunchKernelGGL((k_dslash<R, true, 0>), dim3(nb), dim3(DS_BLOCK), 0, c->s, DS_ARGS);
  } else {
    if (sign == 1)
      hipLaunchKernelGGL((k_dslash<R, false, 1>), dim3(nb), dim3(DS_BLOCK), 0, c->s, DS_ARGS);
    else
      hipLaunchKernelGGL((k_dslash<R, true, 1>), dim3(nb), dim3(DS_BLOCK), 0, c->s, DS_ARGS);
  }
#undef DS_ARGS
  HIPCHK(hipGetLastError());
  return 0;
}

int dslash64(susy_ctx *c, const double2 *src, double2 *dst, int sign,
             const double2 *p0, double shift, double *partial, int *nblocks) {
  return dslash_launch<double>(c, src, dst, sign, p0, shift, partial, nblocks,
                               c->d_links, c->d_uinv, c->d_pdet, c->d_aux);
}

int dslash32(susy_ctx *c, const float2 *src, float2 *dst, int sign,
             const float2 *p0, float shift, double *partial, int *nblocks) {
  if (!c->f_fresh) {
    int r = fields_refresh32(c);
    if (r) return r;
  }
  return dslash_launch<float>(c, src, dst, sign, p0, shift, partial, nblocks,
                              c->f_links, c->f_uinv, c->f_pdet, c->f_aux);
}
