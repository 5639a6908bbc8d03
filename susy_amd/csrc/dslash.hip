// Twisted fermion operator D and D^dagger on the A4* lattice
// (reference: 4d_Q16/susy/utilities.c:141-865), one fused gather-only kernel.
//
// Every thread owns one output matrix (one of the 16 Twist_Fermion components)
// of one site and sums all terms landing on it:
//   site  eta      : DbminusLtoS + detLtoS                 (10 matmuls)
//   link  psi_b    : Dminus + DbplusStoL + detStoL          (10 matmuls)
//   plaq  chi_ab   : Dplus + DbplusPtoP + DbminusPtoP       (10 matmuls)
// so no output is written twice and no atomics are needed.  D^dagger uses
// D^dagger v = -(D v^dagger)^dagger (utilities.c:780-788, 856-864): the
// adjoint is applied on the loads and the stores, in registers.
//
// NCOL <= 4 complex matrix products are VALU FMA chains: an output matrix
// accumulates in registers, the right operand of each product is held in
// registers while the left operand streams element by element.
//
// Memory access goes through raw buffer loads: each field has a 128-bit buffer
// descriptor in SGPRs, a lane supplies only the 32-bit byte offset of its
// site, and the matrix-element offset is a uniform SGPR offset computed on
// the scalar unit.  This keeps address arithmetic out of the VGPR budget, and
// the descriptor's record count bounds every access.
//
// Block mapping is XCD-aware: consecutive hardware block ids are dealt to the
// 8 XCDs round-robin.  The ids are remapped so that each XCD walks one
// contiguous slab of site chunks, all 16 component-blocks of a chunk back to
// back: neighbouring chunks (the z and t neighbours of a site) then share the
// same L2.
#include "susy_internal.h"

#include <mutex>

__constant__ PlaqTerms c_pt;

int dslash_init_tables(const PlaqTerms &pt) {
  HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(c_pt), &pt, sizeof(PlaqTerms)));
  return 0;
}

// minimum wavefronts per SIMD the register allocator must allow.  From
// U(5) on an FP64 output matrix and its operands exceed 256 registers; at one
// wave per SIMD the allocator keeps the excess in AGPRs instead of spilling
// it to scratch memory.
#ifndef DS_WAVES
#define DS_WAVES (NCOL >= 5 ? 1 : 2)
#endif
#ifndef DS_WAVES64
#define DS_WAVES64 DS_WAVES
#endif
#ifndef DS_WAVES32
#define DS_WAVES32 DS_WAVES
#endif
// dynamic LDS bytes requested per block: caps the blocks resident per CU
// (160 KB of LDS per CU), and with it the sites in flight per XCD, whose
// neighbourhood must fit the XCD's L2.  0 leaves occupancy to registers.
#ifndef DS_OCC_LDS64
#define DS_OCC_LDS64 0
#endif
#ifndef DS_OCC_LDS32
#define DS_OCC_LDS32 0
#endif
// unroll factor of the term loops (a wave can issue the next product's
// loads before the current product's FMAs when registers allow)
#ifndef DS_UNROLL
#define DS_UNROLL 1
#endif
#define DS_STR(x) #x
#define DS_UNROLL_LOOP(n) _Pragma(DS_STR(unroll n))

template <typename C, typename R>
__device__ __forceinline__ void cfma(C &acc, R ar, R ai, C b) {
  acc.x = fma(ar, b.x, acc.x);
  acc.x = fma(-ai, b.y, acc.x);
  acc.y = fma(ar, b.y, acc.y);
  acc.y = fma(ai, b.x, acc.y);
}

typedef __amdgpu_buffer_rsrc_t Rsrc;

__device__ __forceinline__ Rsrc make_rsrc(const void *p, uint32_t bytes) {
  // dword3 0x00020000: raw (untyped) access on the gfx9 family
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes,
                                           0x00020000);
}
template <typename C> __device__ __forceinline__ C bload(Rsrc r, uint32_t voff, uint32_t soff);
template <>
__device__ __forceinline__ double2 bload<double2>(Rsrc r, uint32_t voff, uint32_t soff) {
  return __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
}
template <>
__device__ __forceinline__ float2 bload<float2>(Rsrc r, uint32_t voff, uint32_t soff) {
  return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
__device__ __forceinline__ int bload_i32(Rsrc r, uint32_t voff, uint32_t soff) {
  return (int)__builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0);
}
// cache policy of the output stores: non-temporal (2), so the written lines,
// which this launch never reads again, do not displace the neighbour
// gathers' source lines in L2 (FP32 CG iteration -0.8%, FP64 -1.8%)
#ifndef DS_STORE_POLICY
#define DS_STORE_POLICY 2
#endif
template <typename C> __device__ __forceinline__ void bstore(C v, Rsrc r, uint32_t voff, uint32_t soff);
// (callers keep the stored values live until the stores have read them:
// hold() after the last store, see st_keep)
template <>
__device__ __forceinline__ void bstore<double2>(double2 v, Rsrc r, uint32_t voff, uint32_t soff) {
  __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(__attribute__((ext_vector_type(4))) unsigned, v), r, voff, soff, DS_STORE_POLICY);
}
template <>
__device__ __forceinline__ void bstore<float2>(float2 v, Rsrc r, uint32_t voff, uint32_t soff) {
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(__attribute__((ext_vector_type(2))) unsigned, v), r, voff, soff, DS_STORE_POLICY);
}

// A matrix operand: field descriptor, lane byte offset of the site, uniform
// byte offset of element (0,0); element (i,j) is (i*NCOL+j) planes further.
struct Opnd {
  Rsrc r;
  uint32_t v, s;
};

// acc += s * op(A) * op(B).  TA/TB select the dagger.  B is streamed one row
// at a time (row j of op(B) meets column j of op(A)), so every operand
// element is loaded once and only the accumulator plus one row stay live.
template <typename R, bool TA, bool TB>
__device__ __forceinline__ void mmacc(typename CT<R>::t (&acc)[NCOL][NCOL], R s, Opnd A,
                                      Opnd B, uint32_t plane) {
  typedef typename CT<R>::t C;
#pragma unroll
  for (int j = 0; j < NCOL; j++) {
    C brow[NCOL];
#pragma unroll
    for (int k = 0; k < NCOL; k++) {
      C v = bload<C>(B.r, B.v, B.s + (uint32_t)(TB ? k * NCOL + j : j * NCOL + k) * plane);
      if (TB) v.y = -v.y;
      brow[k] = v;
    }
#pragma unroll
    for (int i = 0; i < NCOL; i++) {
      C a = bload<C>(A.r, A.v, A.s + (uint32_t)(TA ? j * NCOL + i : i * NCOL + j) * plane);
      R ar = s * a.x, ai = TA ? -s * a.y : s * a.y;
#pragma unroll
      for (int k = 0; k < NCOL; k++) cfma(acc[i][k], ar, ai, brow[k]);
    }
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
              typename CT<R>::t *__restrict__ aux, int V, const int *__restrict__ skip) {
  typedef typename CT<R>::t C;
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= V || (skip != nullptr && *skip)) return;
  C t[1 + NUMLINK];
  t[0] = mkc<C>(0, 0);
#pragma unroll
  for (int i = 0; i < NCOL; i++) {
    C v = src[(i * NCOL + i) * V + s];
    t[0].x += v.x;
    t[0].y += DAG ? -v.y : v.y;
  }
#pragma unroll
  for (int a = 0; a < NUMLINK; a++) {
    const C *psi = src + (size_t)(1 + a) * NN * V + s;
    const C *ui = Uinv + (size_t)a * NN * V + s;
    C ta = mkc<C>(0, 0);
#pragma unroll
    for (int i = 0; i < NCOL; i++)
#pragma unroll
      for (int j = 0; j < NCOL; j++) {
        C u = ui[(i * NCOL + j) * V];
        C p = DAG ? psi[(i * NCOL + j) * V] : psi[(j * NCOL + i) * V];
        if (DAG) p.y = -p.y;
        ta.x += u.x * p.x - u.y * p.y;
        ta.y += u.x * p.y + u.y * p.x;
      }
    t[1 + a] = ta;
  }
  vm_drain();
#pragma unroll
  for (int a = 0; a <= NUMLINK; a++) st_keep(aux + (size_t)a * V + s, t[a]);
}

template <typename R> struct DsSites;
template <> struct DsSites<double> {
  static constexpr int n = DS_SITES64, waves = DS_WAVES64, cpb = DS_CPB64;
};
template <> struct DsSites<float> {
  static constexpr int n = DS_SITES32, waves = DS_WAVES32, cpb = DS_CPB32;
};

// EPI = 0: dst = result.  EPI = 1: dst = result + shift * p0, and the block's
// partial sum of Re Tr(p0^dag dst) goes to partial[block].
// With aux_out (D, EPI = 0 only) the blocks of the site and link components
// also store the pre-pass values of D^dagger applied to this output:
// aux_out[0] = Tr(out_0)^*, aux_out[1+a] = Tr(U_a^-1 out_{1+a}^dag).
//
// A block covers DS_SITES sites of one output component.  With DS_SPLIT = 2
// two threads share each output matrix: the first computes the site terms
// a < 3, the first three Dminus terms, or Dplus and the first PtoP pair; the
// second the rest and the G terms.  Their partial sums meet in LDS and the
// first stores.  Fewer sites are then in flight per XCD at the same
// occupancy, so their neighbourhoods compete less for the XCD's L2.
template <typename R, bool DAG, int EPI>
__global__ void __launch_bounds__(DsSites<R>::n * DS_SPLIT * DsSites<R>::cpb)
    __attribute__((amdgpu_waves_per_eu(DsSites<R>::waves)))
k_dslash(const typename CT<R>::t *__restrict__ src, typename CT<R>::t *__restrict__ dst,
         const typename CT<R>::t *__restrict__ U, const typename CT<R>::t *__restrict__ Uinv,
         const typename CT<R>::t *__restrict__ pdet, const typename CT<R>::t *__restrict__ aux,
         typename CT<R>::t *__restrict__ aux_out,
         const int *__restrict__ nbr, const int *__restrict__ tco, int V, int nt,
         R pbc, int doG, R G, DsOrder ord,
         const typename CT<R>::t *__restrict__ p0,
         R shift, double *__restrict__ partial, const int *__restrict__ skip) {
  typedef typename CT<R>::t C;
  // a solver that has converged leaves its remaining queued launches idle
  // (the flag is one word read by every thread: the exit is block-uniform)
  if (skip != nullptr && *skip) return;
  const int L = blockIdx.x;
  constexpr int DS_SITES = DsSites<R>::n, CPB = DsSites<R>::cpb;
  constexpr int DS_THREADS = DS_SITES * DS_SPLIT * CPB;
  static_assert(CPB == 1 || DS_SPLIT == 1, "components per block need DS_SPLIT 1");
  static_assert(NTF % CPB == 0 && DS_SITES % 64 == 0, "block shape");
  const int lane = threadIdx.x % DS_SITES, w = CPB == 1 ? threadIdx.x / DS_SITES : 0;
  const int q = L >> 3;                       // position in this XCD's queue
  // a block holds CPB consecutive output components of one site chunk, one
  // group of DS_SITES threads (whole wavefronts) per component
  const int comp = CPB == 1 ? q % NTF : (q % (NTF / CPB)) * CPB + (int)threadIdx.x / DS_SITES;
  const int qq = q / (NTF / CPB);             // chunk position in the queue
  int s;
  bool live;
  if (ord.paired) {
    // even chunk k, odd chunk k, even chunk k+1, ... of this XCD's slab
    const int par = qq & 1, pos = qq >> 1;
    const int chunk = (L & 7) * ord.slab + (ord.paired == 2 ? (pos % ord.T) * ord.P + pos / ord.T : pos);
    const int i = chunk * DS_SITES + lane;
    s = par ? ord.half + i : i;
    live = i < (par ? V - ord.half : ord.half);
  } else {
    const int chunk = (L & 7) * ord.slab + qq;
    s = chunk * DS_SITES + lane;
    live = s < V;
  }

  const uint32_t cs = sizeof(C);
  const uint32_t pl = (uint32_t)V * cs;     // one element plane
  const uint32_t Mb = NN * pl;              // one matrix field
  const Rsrc rS = make_rsrc(src, NTF * Mb), rU = make_rsrc(U, NUMLINK * Mb);
  const Rsrc rN = make_rsrc(nbr, (uint32_t)NGATHER * V * 4u);
  auto nb = [&](int g) { return bload_i32(rN, (uint32_t)s * 4u, (uint32_t)g * V * 4u); };
  auto S = [&](int k, int x) { return Opnd{rS, (uint32_t)x * cs, (uint32_t)k * Mb}; };
  auto Ul = [&](int a, int x) { return Opnd{rU, (uint32_t)x * cs, (uint32_t)a * Mb}; };

  C acc[NCOL][NCOL];
#pragma unroll
  for (int i = 0; i < NCOL; i++)
#pragma unroll
    for (int k = 0; k < NCOL; k++) acc[i][k] = mkc<C>(0, 0);
  const bool first = DS_SPLIT == 1 || w == 0, last = DS_SPLIT == 1 || w == DS_SPLIT - 1;

  if (live) {
    const int t = tco[s];

    if (comp == 0) {
      // ---- site output: 1/2 sum_a [psi_a(x) Ubar_a(x) - bc(-a) Ubar_a(x-a) psi_a(x-a)]
      const int a0 = first ? 0 : 3, a1 = last ? NUMLINK : 3;
      DS_UNROLL_LOOP(DS_UNROLL)
      for (int a = a0; a < a1; a++) {
        const int xm = nb(g_bwd(a));
        mmacc<R, DAG, true>(acc, (R)0.5, S(1 + a, s), Ul(a, s), pl);
        mmacc<R, true, DAG>(acc, (R)-0.5 * bcsign<R>(t, -link_tau(a), nt, pbc), Ul(a, xm),
                            S(1 + a, xm), pl);
      }
      if (doG && last) {
        // detLtoS: -(i G C2 sqrt(N)/2) sum_{a!=b} plaqdet_ab [T_b(x) + bc(b) T_a(x+b)]
        //          * Lambda_{DIMF-1},  Lambda_{DIMF-1} = i I / sqrt(N)
        const Rsrc rA = make_rsrc(aux, 6 * pl), rP = make_rsrc(pdet, NUMLINK * NUMLINK * pl);
        C sum = mkc<C>(0, 0);
#pragma unroll 1
        for (int b = 0; b < NUMLINK; b++) {
          const int xb = nb(g_fwd(b));
          const C tb = bload<C>(rA, (uint32_t)s * cs, (1 + b) * pl);
          const R sg = bcsign<R>(t, link_tau(b), nt, pbc);
#pragma unroll 1
          for (int a = 0; a < NUMLINK; a++) {
            if (a == b) continue;
            const C ta = bload<C>(rA, (uint32_t)xb * cs, (1 + a) * pl);
            const C t2 = mkc<C>(tb.x + sg * ta.x, tb.y + sg * ta.y);
            const C pd = bload<C>(rP, (uint32_t)s * cs, (a * NUMLINK + b) * pl);
            sum.x += pd.x * t2.x - pd.y * t2.y;
            sum.y += pd.x * t2.y + pd.y * t2.x;
          }
        }
        // trd = i lg sum; acc -= trd * (i/sqrt N) I = (lg/sqrt N) sum I
        const R f = (R)(0.5 * C2) * G;  // lg / sqrt(N) = G C2 / 2
#pragma unroll
        for (int i = 0; i < NCOL; i++) {
          acc[i][i].x += f * sum.x;
          acc[i][i].y += f * sum.y;
        }
      }
    } else if (comp <= NUMLINK) {
      // ---- link output b
      const int b = comp - 1;
      const int xb = nb(g_fwd(b));
      DS_UNROLL_LOOP(DS_UNROLL)
      for (int a = 0; a < NUMLINK; a++) {
        if (a == b) continue;
        // the first thread takes the first three a != b, the other the last
        if (DS_SPLIT == 2 && ((a - (a > b)) < 3) != first) continue;
        // Dminus: U_a(x+b) chi_ab(x) - bc(-a) chi_ab(x-a) U_a(x-a)
        const int xm = nb(g_bwd(a));
        const R sg = a < b ? (R)1 : (R)-1;   // chi_ab = sg * stored chi_{min,max}
        const int pc = 6 + plaq_idx(a, b);
        mmacc<R, false, DAG>(acc, sg, Ul(a, xb), S(pc, s), pl);
        mmacc<R, DAG, false>(acc, -sg * bcsign<R>(t, -link_tau(a), nt, pbc), S(pc, xm),
                             Ul(a, xm), pl);
      }
      if (last) {
        // DbplusStoL: 1/2 [bc(b) eta(x+b) Ubar_b(x) - Ubar_b(x) eta(x)]
        mmacc<R, DAG, true>(acc, (R)0.5 * bcsign<R>(t, link_tau(b), nt, pbc), S(0, xb),
                            Ul(b, s), pl);
        mmacc<R, true, DAG>(acc, (R)-0.5, Ul(b, s), S(0, s), pl);
      }
      if (doG && last) {
        // detStoL: -G C2/2 U_b^-1(x) sum_{c!=b} [Tr eta(x) plaqdet_cb(x)
        //                                        + bc(-c) Tr eta(x-c) plaqdet_bc(x-c)]
        const Rsrc rA = make_rsrc(aux, 6 * pl), rP = make_rsrc(pdet, NUMLINK * NUMLINK * pl);
        const Rsrc rI = make_rsrc(Uinv, NUMLINK * Mb);
        C sum = mkc<C>(0, 0);
        const C te = bload<C>(rA, (uint32_t)s * cs, 0);
#pragma unroll 1
        for (int c = 0; c < NUMLINK; c++) {
          if (c == b) continue;
          const int xm = nb(g_bwd(c));
          const C p1 = bload<C>(rP, (uint32_t)s * cs, (c * NUMLINK + b) * pl);
          const C p2 = bload<C>(rP, (uint32_t)xm * cs, (b * NUMLINK + c) * pl);
          const C tm = bload<C>(rA, (uint32_t)xm * cs, 0);
          const R sg = bcsign<R>(t, -link_tau(c), nt, pbc);
          sum.x += te.x * p1.x - te.y * p1.y + sg * (tm.x * p2.x - tm.y * p2.y);
          sum.y += te.x * p1.y + te.y * p1.x + sg * (tm.x * p2.y + tm.y * p2.x);
        }
        const R f = (R)(-0.5 * C2) * G;
        const R sr = f * sum.x, si = f * sum.y;
#pragma unroll
        for (int i = 0; i < NCOL; i++)
#pragma unroll
          for (int k = 0; k < NCOL; k++) {
            const C u = bload<C>(rI, (uint32_t)s * cs, b * Mb + (i * NCOL + k) * pl);
            acc[i][k].x += sr * u.x - si * u.y;
            acc[i][k].y += sr * u.y + si * u.x;
          }
      }
    } else {
      // ---- plaquette output (a,b), a<b
      const int p = comp - 1 - NUMLINK;
      const int a = c_pt.a[p], b = c_pt.b[p];
      const int xa = nb(g_fwd(a)), xb = nb(g_fwd(b));
      // Dplus: bc(a) U_a(x) psi_b(x+a) - psi_b(x) U_a(x+b)
      //        - bc(b) U_b(x) psi_a(x+b) + psi_a(x) U_b(x+a)
      if (first) {
        mmacc<R, false, DAG>(acc, bcsign<R>(t, link_tau(a), nt, pbc), Ul(a, s), S(1 + b, xa),
                             pl);
        mmacc<R, DAG, false>(acc, (R)-1, S(1 + b, s), Ul(a, xb), pl);
        mmacc<R, false, DAG>(acc, -bcsign<R>(t, link_tau(b), nt, pbc), Ul(b, s), S(1 + a, xb),
                             pl);
        mmacc<R, DAG, false>(acc, (R)1, S(1 + a, s), Ul(b, xa), pl);
      }
      const int n0 = first ? 0 : 1, n1 = last ? 3 : 1;
      // DbplusPtoP + DbminusPtoP.  Each DbminusPtoP term landing on (a,b)
      // coincides with the DbplusPtoP term of the same c: with (d,e) the
      // other two links, x-d-e = x+a+b+c and x-d-e-c = x+a+b (the five e_a
      // sum to zero), perm(d,e,c,a,b) = perm(a,b,c,d,e), and the temporal
      // offsets agree for the same reason.  The DbplusPtoP terms are applied
      // at double weight (checked term by term when the tables are built).
      DS_UNROLL_LOOP(DS_UNROLL)
      for (int n = n0; n < n1; n++) {
        const int c = c_pt.pp_c[p][n];
        const int pc = 6 + c_pt.pp_de[p][n];
        const int x1 = nb(c_pt.pp_g1[p][n]), x2 = nb(c_pt.pp_g2[p][n]), xmc = nb(g_bwd(c));
        const R tr = (R)(2.0f * c_pt.pp_tr[p][n]);
        mmacc<R, DAG, true>(acc, tr * bcsign<R>(t, c_pt.pp_dt3[p][n], nt, pbc), S(pc, x2),
                            Ul(c, x1), pl);
        mmacc<R, true, DAG>(acc, -tr * bcsign<R>(t, c_pt.pp_dt2[p][n], nt, pbc), Ul(c, xmc),
                            S(pc, x1), pl);
      }
    }
  }

#if DS_SPLIT > 1
  // the second thread's partial sum meets the first's in LDS
  __shared__ C part[DS_SPLIT - 1][NN][DS_SITES];
  if (!first) {
#pragma unroll
    for (int i = 0; i < NCOL; i++)
#pragma unroll
      for (int k = 0; k < NCOL; k++) part[w - 1][i * NCOL + k][lane] = acc[i][k];
  }
  __syncthreads();
  if (first) {
#pragma unroll
    for (int u = 0; u < DS_SPLIT - 1; u++)
#pragma unroll
      for (int i = 0; i < NCOL; i++)
#pragma unroll
        for (int k = 0; k < NCOL; k++) {
          const C v = part[u][i * NCOL + k][lane];
          acc[i][k].x += v.x;
          acc[i][k].y += v.y;
        }
  }
#endif

  double dot = 0.0;
  if (live && first) {
    // ---- epilogue loads (p0 for the DSq epilogue, U^-1 for the D^dagger
    // pre-pass values), all drained before the first store
    const Rsrc rD = make_rsrc(dst, NTF * Mb);
    const bool pre = !DAG && EPI == 0 && aux_out != nullptr && comp <= NUMLINK;
    C q[NCOL][NCOL];
    if (EPI == 1) {
      const Rsrc rQ = make_rsrc(p0, NTF * Mb);
#pragma unroll
      for (int i = 0; i < NCOL; i++)
#pragma unroll
        for (int k = 0; k < NCOL; k++)
          q[i][k] = bload<C>(rQ, (uint32_t)s * cs, comp * Mb + (i * NCOL + k) * pl);
    } else if (pre && comp > 0) {
      const Rsrc rI = make_rsrc(Uinv, NUMLINK * Mb);
#pragma unroll
      for (int i = 0; i < NCOL; i++)
#pragma unroll
        for (int k = 0; k < NCOL; k++)
          q[i][k] = bload<C>(rI, (uint32_t)s * cs, (comp - 1) * Mb + (i * NCOL + k) * pl);
    }
    vm_drain();
    // ---- store (D^dagger: negated adjoint); the stored values stay live
    // until one wait after the last store reports their registers read
    C out[NCOL][NCOL];
#pragma unroll
    for (int i = 0; i < NCOL; i++)
#pragma unroll
      for (int k = 0; k < NCOL; k++) {
        C v = DAG ? mkc<C>(-acc[k][i].x, acc[k][i].y) : acc[i][k];
        const uint32_t so = comp * Mb + (i * NCOL + k) * pl;
        if (EPI == 1) {
          v.x = fma(shift, q[i][k].x, v.x);
          v.y = fma(shift, q[i][k].y, v.y);
          dot += (double)q[i][k].x * (double)v.x + (double)q[i][k].y * (double)v.y;
        }
        bstore<C>(v, rD, (uint32_t)s * cs, so);
        out[i][k] = v;
      }
#pragma unroll
    for (int i = 0; i < NCOL; i++)
#pragma unroll
      for (int k = 0; k < NCOL; k++) hold(out[i][k]);
    if (pre) {
      C tr = mkc<C>(0, 0);
      if (comp == 0) {
#pragma unroll
        for (int i = 0; i < NCOL; i++) {
          tr.x += acc[i][i].x;
          tr.y -= acc[i][i].y;
        }
      } else {
#pragma unroll
        for (int i = 0; i < NCOL; i++)
#pragma unroll
          for (int j = 0; j < NCOL; j++) {
            const C u = q[i][j], p = acc[i][j];
            tr.x += u.x * p.x + u.y * p.y;
            tr.y += u.y * p.x - u.x * p.y;
          }
      }
      st_keep(aux_out + (size_t)comp * V + s, tr);
    }
  }
  if (EPI == 1) {
    // deterministic block reduction: wavefront butterfly, then across waves
    for (int off = 32; off > 0; off >>= 1) dot += __shfl_xor(dot, off, 64);
    __shared__ double red[DS_THREADS / 64];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = dot;
    __syncthreads();
    if (threadIdx.x == 0) {
      double tsum = 0;
      for (int u = 0; u < DS_THREADS / 64; u++) tsum += red[u];
      partial[L] = tsum;
    }
  }
}

template <typename R>
static int dslash_launch(susy_ctx *c, const typename CT<R>::t *src,
                         typename CT<R>::t *dst, int sign, const typename CT<R>::t *p0,
                         R shift, double *partial, int *nblocks,
                         const typename CT<R>::t *U, const typename CT<R>::t *Uinv,
                         const typename CT<R>::t *pdet, typename CT<R>::t *aux,
                         const int *skip, typename CT<R>::t *aux_out = nullptr,
                         bool prep = true, bool sampled = true) {
  if (sign != 1 && sign != -1) {
    set_error("fermion_op: incorrect sign %d", sign);
    return SUSY_E_ARG;
  }
  const int V = c->V;
  if ((size_t)NTF * NN * V * sizeof(typename CT<R>::t) > 0x7fffffffu) {
    set_error("lattice too large for 32-bit buffer offsets (V=%d)", V);
    return SUSY_E_ARG;
  }
  if (sign != 1 || p0 != nullptr || !c->doG) aux_out = nullptr;
  if (c->doG && prep) {
    int nb = (V + 255) / 256;
    if (sign == 1)
      hipLaunchKernelGGL((k_dslash_prep<R, false>), dim3(nb), dim3(256), 0, c->s, src, Uinv, aux, V,
                         skip);
    else
      hipLaunchKernelGGL((k_dslash_prep<R, true>), dim3(nb), dim3(256), 0, c->s, src, Uinv, aux, V,
                         skip);
    HIPCHK(hipGetLastError());
  }
  DsOrder ord;
  constexpr int threads = DsSites<R>::n * DS_SPLIT * DsSites<R>::cpb;
  const int nb = ds_grid(V, c->p.nx * c->p.ny * c->p.nz,
                         sizeof(R) == sizeof(float) ? c->site_order32 : c->site_order,
                         DsSites<R>::n, &ord, DsSites<R>::cpb);
  if (nblocks) *nblocks = nb;
  R pbc = (R)c->p.PBC, G = (R)c->p.G;
#define DS_ARGS src, dst, U, Uinv, pdet, aux, aux_out, c->d_nbr, c->d_t, V, c->p.nt, pbc, \
                c->doG, G, ord, p0, shift, partial, skip
  // sampled timing covers the plain operator D (sign +1, no epilogue) only
  const int pi = sizeof(R) == sizeof(float) ? 1 : 0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (sampled && c->prof_every > 0 && sign == 1 && p0 == nullptr &&
      (c->prof_launch[pi]++ % c->prof_every) == 0) {
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, c->s));
  }
  const unsigned lds = sizeof(R) == sizeof(double) ? DS_OCC_LDS64 : DS_OCC_LDS32;
  if (p0 == nullptr) {
    if (sign == 1)
      hipLaunchKernelGGL((k_dslash<R, false, 0>), dim3(nb), dim3(threads), lds, c->s, DS_ARGS);
    else
      hipLaunchKernelGGL((k_dslash<R, true, 0>), dim3(nb), dim3(threads), lds, c->s, DS_ARGS);
  } else {
    if (sign == 1)
      hipLaunchKernelGGL((k_dslash<R, false, 1>), dim3(nb), dim3(threads), lds, c->s, DS_ARGS);
    else
      hipLaunchKernelGGL((k_dslash<R, true, 1>), dim3(nb), dim3(threads), lds, c->s, DS_ARGS);
  }
#undef DS_ARGS
  HIPCHK(hipGetLastError());
  if (e0) {
    HIPCHK(hipEventRecord(e1, c->s));
    c->prof_ev[pi].push_back(e0);
    c->prof_ev[pi].push_back(e1);
  }
  return 0;
}

// Sampled timing of the operator kernel on the context's stream: every
// prof_every-th launch of D (sign +1, no epilogue) is bracketed by events.
// prof_every = 0 switches sampling off.
extern "C" int susy_prof_enable(susy_ctx *c, int every) {
  if (!c || every < 0) return SUSY_E_ARG;
  HIPCHK(hipStreamSynchronize(c->s));
  for (int p = 0; p < 2; p++) {
    for (hipEvent_t e : c->prof_ev[p]) (void)hipEventDestroy(e);
    c->prof_ev[p].clear();
    c->prof_launch[p] = 0;
  }
  c->prof_every = every;
  return 0;
}
// Average duration (ms) of the sampled launches and the count of D launches,
// for FP64 (prec 0) or FP32 (prec 1) operator kernels.
extern "C" int susy_prof_read(susy_ctx *c, int prec, double *avg_ms, long *launches,
                              long *sampled) {
  if (!c || prec < 0 || prec > 1) return SUSY_E_ARG;
  HIPCHK(hipStreamSynchronize(c->s));
  double tot = 0;
  size_t n = c->prof_ev[prec].size() / 2;
  for (size_t k = 0; k < n; k++) {
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, c->prof_ev[prec][2 * k], c->prof_ev[prec][2 * k + 1]));
    tot += ms;
  }
  *avg_ms = n ? tot / n : 0.0;
  *launches = c->prof_launch[prec];
  *sampled = (long)n;
  return 0;
}

int dslash64(susy_ctx *c, const double2 *src, double2 *dst, int sign,
             const double2 *p0, double shift, double *partial, int *nblocks, const int *skip) {
  return dslash_launch<double>(c, src, dst, sign, p0, shift, partial, nblocks,
                               c->d_links, c->d_uinv, c->d_pdet, c->d_aux, skip);
}
int dslash32(susy_ctx *c, const float2 *src, float2 *dst, int sign,
             const float2 *p0, float shift, double *partial, int *nblocks, const int *skip) {
  if (!c->f_fresh) {
    int r = fields_refresh32(c);
    if (r) return r;
  }
  return dslash_launch<float>(c, src, dst, sign, p0, shift, partial, nblocks,
                              c->f_links, c->f_uinv, c->f_pdet, c->f_aux, skip);
}

// dst = D^dagger D src + shift src, tmp = D src, per-block partials of
// Re <src, dst> in partial (the DSq of utilities.c:874-885 with the CG's
// shift).  D stores the D^dagger pre-pass values of its output (aux2), so
// D^dagger runs without its pre-pass.  prep_src = false: the caller has
// already stored the pre-pass values of src in the context's aux buffer.
int dsq_pair64(susy_ctx *c, const double2 *src, double2 *tmp, double2 *dst, double shift,
               double *partial, int *nblocks, const int *skip, bool prep_src) {
  int r = dslash_launch<double>(c, src, tmp, 1, nullptr, 0.0, nullptr, nullptr, c->d_links,
                                c->d_uinv, c->d_pdet, c->d_aux, skip, c->d_aux2, prep_src);
  if (r) return r;
  return dslash_launch<double>(c, tmp, dst, -1, src, shift, partial, nblocks, c->d_links,
                               c->d_uinv, c->d_pdet, c->d_aux2, skip, nullptr, false);
}
int dsq_pair32(susy_ctx *c, const float2 *src, float2 *tmp, float2 *dst, float shift,
               double *partial, int *nblocks, const int *skip, bool prep_src) {
  if (!c->f_fresh) {
    int r = fields_refresh32(c);
    if (r) return r;
  }
  int r = dslash_launch<float>(c, src, tmp, 1, nullptr, 0.f, nullptr, nullptr, c->f_links,
                               c->f_uinv, c->f_pdet, c->f_aux, skip, c->f_aux2, prep_src);
  if (r) return r;
  return dslash_launch<float>(c, tmp, dst, -1, src, shift, partial, nblocks, c->f_links,
                              c->f_uinv, c->f_pdet, c->f_aux2, skip, nullptr, false);
}

// ---------------------------------------------------------------------
// Operator without allocation (susy_fermion_op_noalloc).  Its only scratch
// memory -- the G-term pre-pass values, 6 complex numbers per site -- lives
// in this fixed device array instead of the context's buffer.  Every device
// the library runs on holds one copy until the process ends.  The pre-pass
// writes every value the operator then reads, so nothing in it needs
// clearing between calls; calls that overlap on one device race on it.
#define NOALLOC_SCRATCH_BYTES (16u << 20)
#define NOALLOC_MAX_DEVICES 64
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];
// device address of `scratch` on each device (the name itself is only a
// host-side handle of the symbol)
static void *g_scratch_dev[NOALLOC_MAX_DEVICES];
static std::mutex g_scratch_mu;

// Called by susy_create with the context's device current: looks up that
// device's copy of the array once.
int noalloc_init_device(susy_ctx *c) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) {
    set_error("device %d beyond the %d the library tracks", dev, NOALLOC_MAX_DEVICES);
    return SUSY_E_STATE;
  }
  c->dev = dev;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
      set_error("hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
      return SUSY_E_HIP;
    }
    g_scratch_dev[dev] = p;
  }
  return 0;
}

extern "C" int susy_fermion_op_noalloc(susy_ctx *c, const double *src, double *dst, int sign) {
  if (!c) return SUSY_E_STATE;
  if (!src || !dst) {
    set_error("susy_fermion_op_noalloc: null field pointer");
    return SUSY_E_ARG;
  }
  if (sign != 1 && sign != -1) {
    set_error("susy_fermion_op_noalloc: incorrect sign %d", sign);
    return SUSY_E_ARG;
  }
  const size_t need = (size_t)6 * c->V * sizeof(double2);
  if (need > NOALLOC_SCRATCH_BYTES) {
    set_error("susy_fermion_op_noalloc: %d sites need %zu bytes of scratch; the fixed "
              "array holds %u (at most %u sites)", c->V, need, NOALLOC_SCRATCH_BYTES,
              (unsigned)(NOALLOC_SCRATCH_BYTES / (6 * sizeof(double2))));
    return SUSY_E_ARG;
  }
  void *aux = g_scratch_dev[c->dev];
  if (!aux) {
    set_error("susy_fermion_op_noalloc: scratch of device %d not set up", c->dev);
    return SUSY_E_STATE;
  }
  int r = dslash_launch<double>(c, (const double2 *)src, (double2 *)dst, sign, nullptr, 0.0,
                                nullptr, nullptr, c->d_links, c->d_uinv, c->d_pdet,
                                (double2 *)aux, nullptr, nullptr, true, false);
  if (r) return r;
  HIPCHK(hipStreamSynchronize(c->s));
  return 0;
}
