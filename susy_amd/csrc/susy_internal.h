// Internal declarations shared by the translation units of libsusy_hip_U<N>.
//
// Device layout (structure of arrays, site index fastest):
//   a field of K matrices is K * NCOL^2 planes of V complex numbers;
//   element (i, j) of matrix k at site s lives at [(k * NN + i * NCOL + j) * V + s].
//   Twist_Fermion fields: K = 16 (0 = site, 1..5 = links, 6..15 = plaquettes
//   in plaq_index order); links and momenta: K = 5.
// Lanes of a wavefront take consecutive sites, so every matrix-element load
// of a wavefront is one contiguous 64 x 16 B (FP64) or 64 x 8 B (FP32) run.
#ifndef SUSY_INTERNAL_H
#define SUSY_INTERNAL_H

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/susy_hip.h"

#define NN (NCOL * NCOL)
#define NUMLINK 5
#define NPLAQ 10
#define NTF 16
#define NGATHER 74
#define NTERMS 30
#define NQLINK 28
#define IMAG_TOL 1.0e-8
// operator kernel: a block covers DS_SITES64 (FP64) / DS_SITES32 (FP32)
// sites of one output component, with DS_SPLIT wavefronts sharing the
// products of each output matrix (measured best: 256 / 128 / 1)
#ifndef DS_SITES64
#define DS_SITES64 256
#endif
#ifndef DS_SITES32
#define DS_SITES32 128
#endif
#ifndef DS_SPLIT
#define DS_SPLIT 1
#endif
// output components per block of the operator kernel (each a whole number
// of wavefronts over the block's site chunk); 1 = one component per block
#ifndef DS_CPB64
#define DS_CPB64 1
#endif
#ifndef DS_CPB32
#define DS_CPB32 1
#endif
#define C2 1.0

// Block order of the fused operator kernel.  Hardware block ids are dealt to
// the 8 XCDs round-robin; each XCD walks one slab of `slab` site chunks.
//   paired = 0: a slab is a contiguous range of site indices;
//   paired = 1: a slab is the same chunk range in the even half [0, half) and
//     the odd half [half, V) of the even-first site order, even and odd chunks
//     alternating, so a site's nearest neighbours (opposite parity, same t
//     range) are fetched into the same L2;
//   paired = 2: as 1, with the slab's chunks visited t fastest (P chunks per
//     t-slice and parity, T t-slices per slab: chunk (pos % T) P + pos / T),
//     so t-neighbours are a chunk apart instead of a t-slice.  Needs the
//     slab to hold whole t-slices of whole chunks; otherwise it is order 1.
struct DsOrder {
  int paired, slab, half, P, T;
};
// grid size (blocks) of the operator kernel, and its order parameters;
// tslice = sites per t-slice (nx ny nz)
inline int ds_grid(int V, int tslice, int paired, int DS_SITES, DsOrder *o, int cpb = 1) {
  DsOrder d;
  d.paired = paired;
  d.half = V / 2;
  d.P = d.T = 1;
  if (paired) {
    const int chunks = (V - d.half + DS_SITES - 1) / DS_SITES;   // odd half is the larger
    d.slab = (chunks + 7) / 8;
    if (paired == 2) {
      const bool fit = V % 2 == 0 && tslice % (2 * DS_SITES) == 0 && d.half % DS_SITES == 0 &&
                       chunks % 8 == 0 && d.slab % (tslice / (2 * DS_SITES)) == 0;
      if (fit) {
        d.P = tslice / (2 * DS_SITES);
        d.T = d.slab / d.P;
      } else {
        d.paired = 1;
      }
    }
    if (o) *o = d;
    return 8 * d.slab * 2 * (NTF / cpb);
  }
  const int chunks = (V + DS_SITES - 1) / DS_SITES;
  d.slab = (chunks + 7) / 8;
  if (o) *o = d;
  return 8 * d.slab * (NTF / cpb);
}

// ---------------------------------------------------------------------
// Complex helpers for double2 / float2
// Completes every outstanding vector-memory access of the wave and keeps the
// compiler from moving memory accesses across it: the kernels issue all
// their loads, drain, then store.
__device__ __forceinline__ void vm_drain() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

// A vector-memory store reads its data VGPRs some cycles after it issues.
// On this hardware an instruction that overwrites those VGPRs right after
// the store (which the compiler schedules freely) can win the race, and the
// store then writes the new value -- seen as sporadic wrong 64-byte groups
// of the operator's FP64 output at U(3) 8^3x16 and up, never at 4^4.  These
// stores keep the data live until s_waitcnt expcnt(0) reports it read.
__device__ __forceinline__ void hold(double v) { asm volatile("s_waitcnt expcnt(0)" ::"v"(v) : "memory"); }
__device__ __forceinline__ void hold(float v) { asm volatile("s_waitcnt expcnt(0)" ::"v"(v) : "memory"); }
__device__ __forceinline__ void hold(double2 v) {
  asm volatile("s_waitcnt expcnt(0)" ::"v"(v.x), "v"(v.y) : "memory");
}
__device__ __forceinline__ void hold(float2 v) {
  asm volatile("s_waitcnt expcnt(0)" ::"v"(v.x), "v"(v.y) : "memory");
}
__device__ __forceinline__ void hold(int v) { asm volatile("s_waitcnt expcnt(0)" ::"v"(v) : "memory"); }
__device__ __forceinline__ void hold(unsigned v) { asm volatile("s_waitcnt expcnt(0)" ::"v"(v) : "memory"); }
__device__ __forceinline__ void hold(unsigned short v) {
  asm volatile("s_waitcnt expcnt(0)" ::"v"(v) : "memory");
}
// Store one value and keep its registers until the store has read them.
// Every global store of the library goes through this (or through a batch of
// stores followed by hold() of each stored value).
template <typename T> __device__ __forceinline__ void st_keep(T *p, T v) {
  *p = v;
  hold(v);
}
// lhs = rhs through st_keep, rhs converted to the element type of lhs
#define ST(lhs, rhs) st_keep(&(lhs), (std::remove_reference_t<decltype(lhs)>)(rhs))

template <typename R> struct CT;
template <> struct CT<double> { typedef double2 t; };
template <> struct CT<float> { typedef float2 t; };

template <typename C>
__host__ __device__ __forceinline__ C mkc(decltype(C().x) r, decltype(C().x) i) {
  C c; c.x = r; c.y = i; return c;
}

// ---------------------------------------------------------------------
// Term tables of the fermion operator and force, built on the host from the
// reference's lookup tables (setup_lambda.c:202-297) and gather numbering
// (setup_offset.c:207-485); copied to __constant__ memory by each TU.
struct PlaqTerms {
  int a[NPLAQ], b[NPLAQ];
  // DbplusPtoP terms whose output is plaquette p: dest_ab +=
  //   tr [bc3 chi_de(x+a+b+c) Ubar_c(x+a+b) - bc2 Ubar_c(x-c) chi_de(x+a+b)]
  int pp_c[NPLAQ][3], pp_de[NPLAQ][3], pp_g1[NPLAQ][3], pp_g2[NPLAQ][3];
  int pp_dt3[NPLAQ][3], pp_dt2[NPLAQ][3];
  float pp_tr[NPLAQ][3];
  // DbminusPtoP terms whose output (d,e) is plaquette p: dest_de +=
  //   tr [bc2' chi_ab(x-a-b) Ubar_c(x-a-b-c) - bc3' Ubar_c(x-c) chi_ab(x-a-b-c)]
  // Each coincides with the DbplusPtoP term of the same p and c; the kernel
  // applies DbplusPtoP at double weight and these entries serve only the
  // set-up check of that coincidence.
  int pm_c[NPLAQ][3], pm_ab[NPLAQ][3], pm_g2[NPLAQ][3], pm_g1[NPLAQ][3];
  int pm_dt2[NPLAQ][3], pm_dt3[NPLAQ][3];
  float pm_tr[NPLAQ][3];
};
// Q-closed force terms landing on link c (F1Q/F2Q, update_h.c:272-462):
// six (a,b,d,e) per c.
struct ForceTerms {
  int a[NUMLINK][6], b[NUMLINK][6], ab[NUMLINK][6], de[NUMLINK][6];
  int g_abc[NUMLINK][6];   // gather x + e_a + e_b + e_c
  int g_mab[NUMLINK][6];   // gather x - e_a - e_b
  int dt3[NUMLINK][6];     // tau_a + tau_b + tau_c
  int dt2m[NUMLINK][6];    // -tau_a - tau_b
  float perm1[NUMLINK][6]; // perm[d][e][c][a][b]
  float perm2[NUMLINK][6]; // perm[a][b][c][d][e]
};

// temporal offset of link a: links 3 and 4 step +1 / -1 in t
__host__ __device__ __forceinline__ int link_tau(int a) {
  return a == 3 ? 1 : (a == 4 ? -1 : 0);
}
// gather ids of x + e_a and x - e_a (setup_offset.c:175-177)
__host__ __device__ __forceinline__ int g_fwd(int a) { return 8 + 2 * a; }
__host__ __device__ __forceinline__ int g_bwd(int a) { return 9 + 2 * a; }
__host__ __device__ __forceinline__ int plaq_idx(int a, int b) {
  int lo = a < b ? a : b, hi = a < b ? b : a;
  return lo * (NUMLINK - 1) - lo * (lo + 1) / 2 + hi - 1;
}

// ---------------------------------------------------------------------
// Reference generator state (generic/ranstuff.c, include/random.h)
struct Prn {
  uint64_t r[7];
  uint64_t multiplier, addend, ic_state;
  float scale;
};

struct Geometry {
  int nx, ny, nz, nt, V;
  std::vector<int> x, y, z, t;      // coordinates per site index
  std::vector<int> gathers;         // NGATHER * V
  int q_offset[NQLINK][4];
};

// host geometry and tables (lattice.hip)
int site_index(const Geometry &g, int x, int y, int z, int t);
void build_geometry(Geometry &g, int nx, int ny, int nz, int nt);
int gather_id(const int q_offset[NQLINK][4], const int d[4]);
void build_plaq_terms(const int q_offset[NQLINK][4], PlaqTerms &pt);
void build_force_terms(const int q_offset[NQLINK][4], ForceTerms &ft);
void build_lambda(std::vector<double> &lam);   // DIMF matrices, AoS re/im
void prn_init(Prn *p, int seed, int index);
double prn_next(Prn *p);

// ---------------------------------------------------------------------
// Context
struct SgState;   // learned surrogate (surrogate.hip)
struct Traj;      // trajectory in progress (capi.hip)

struct susy_ctx {
  susy_params p;
  Geometry geo;
  int V = 0;
  double kappa = 0;
  int doG = 0;
  double ampdeg4 = 0, amp4[15], shift4[15], ampdeg8 = 0, amp8[15], shift8[15];
  std::vector<double> lambda_gen;   // Lambda generators, DIMF * NN * 2

  hipStream_t s = nullptr;
  int dev = 0;                      // HIP device the context was created on
  int *d_nbr = nullptr;             // NGATHER * V
  int *d_t = nullptr;               // t coordinate per site
  // gauge fields (FP64, SoA) and derived fields
  double2 *d_links = nullptr, *d_mom = nullptr, *d_old_links = nullptr,
          *d_old_mom = nullptr;
  double2 *d_uinv = nullptr;        // 5 matrices
  double2 *d_det = nullptr;         // 5 complex
  double2 *d_pdet = nullptr;        // 25 complex [a][b]
  double2 *d_zw = nullptr;          // 25 complex
  double2 *d_dmu = nullptr;         // 1 matrix
  double2 *d_fmunu = nullptr;       // 10 matrices
  double2 *d_force = nullptr;       // 5 matrices (force accumulator)
  // pole-split integrator (susy_set_split): first middle-level pole, middle
  // steps per half outer step, middle-level weights (entries >= split_first)
  int split_first = 0, split_nmid = 1;
  double split_w[16] = {0};
  // FP32 copies of the gauge-derived operands of the FP32 operator
  float2 *f_links = nullptr, *f_uinv = nullptr, *f_pdet = nullptr;
  int f_fresh = 0;                  // FP32 copies current?
  // operator scratch
  double2 *d_aux = nullptr;         // 12 complex per site (operator/force prep)
  float2 *f_aux = nullptr;
  double2 *d_aux2 = nullptr;        // 6 complex per site: D^dagger prep written by D
  float2 *f_aux2 = nullptr;
  double2 *d_tmp = nullptr;         // one TF (inner D of DSq)
  float2 *f_tmp = nullptr;
  double2 *d_boundary = nullptr;    // AoS staging for host transfers (one TF)
  // reductions
  double *d_part = nullptr;         // per-block partial sums
  int npart = 0;
  double *d_scal = nullptr;         // final reduced scalars
  double *h_scal = nullptr;         // pinned host copy of d_scal (64 doubles)
  // Twist_Fermion vectors
  std::vector<double2 *> tf;
  // CG workspace (allocated on demand, grown as needed)
  void *cg_ws = nullptr;
  size_t cg_ws_bytes = 0;
  double2 *cg_t64 = nullptr;        // two FP64 fields for true residuals
  void *cg_state = nullptr;         // scalar state of the multi-shift CG (device)
  float2 *f_in = nullptr, *f_out = nullptr;   // FP32 operator staging
  std::vector<double2 *> traj_pool;           // pseudofermions + solutions
  // RNG (host, bit-exact with the reference)
  std::vector<Prn> site_prn;
  Prn node_prn;
  int g_iset = 0;
  double g_gset = 0;
  int total_iters = 0;
  // sampled HIP-event timing of the operator kernel launches (per precision:
  // index 0 FP64, 1 FP32): every prof_every-th launch is bracketed by events
  int site_order = 2;          // DsOrder::paired of the operator kernel (FP64)
  int site_order32 = 2;        // the same for the FP32 operator
  int pf_handle = -1;          // two Twist_Fermion handles of the Pfaffian (susy_pfaffian_dev)
  SgState *sg = nullptr;       // learned surrogate state (susy_sg_init)
  Traj *traj = nullptr;        // trajectory between susy_traj_begin and susy_traj_end
  int prof_every = 0;
  long prof_launch[2] = {0, 0};
  std::vector<hipEvent_t> prof_ev[2];
};

void set_error(const char *fmt, ...);
#define HIPCHK(x)                                                   \
  do {                                                              \
    hipError_t _e = (x);                                            \
    if (_e != hipSuccess) {                                         \
      set_error("%s:%d %s: %s", __FILE__, __LINE__, #x,             \
                hipGetErrorString(_e));                             \
      return SUSY_E_HIP;                                            \
    }                                                               \
  } while (0)

// ---------------------------------------------------------------------
// Launchers (device code in dslash.hip, fields.hip, force.hip, blas.hip)
int dslash_init_tables(const PlaqTerms &pt);
// dst = D src (sign +1) or D^dag src (sign -1).  If p0 != nullptr the
// epilogue writes dst + shift*p0 and a per-block partial of Re<p0, dst>.
// skip: optional device flag; when it reads non-zero the launches do nothing
int dslash64(susy_ctx *c, const double2 *src, double2 *dst, int sign,
             const double2 *p0, double shift, double *partial, int *nblocks,
             const int *skip = nullptr);
int dslash32(susy_ctx *c, const float2 *src, float2 *dst, int sign,
             const float2 *p0, float shift, double *partial, int *nblocks,
             const int *skip = nullptr);
int dsq_pair64(susy_ctx *c, const double2 *src, double2 *tmp, double2 *dst, double shift,
               double *partial, int *nblocks, const int *skip = nullptr, bool prep_src = true);
int dsq_pair32(susy_ctx *c, const float2 *src, float2 *tmp, float2 *dst, float shift,
               double *partial, int *nblocks, const int *skip = nullptr, bool prep_src = true);

int noalloc_init_device(susy_ctx *c);   // (dslash.hip) fixed scratch of this device

int fields_recompute(susy_ctx *c);      // det, Uinv, plaqdet, ZW*, DmuUmu, Fmunu
int fields_refresh32(susy_ctx *c);      // FP32 copies of links, Uinv, plaqdet
int gauge_force_dev(susy_ctx *c, double eps, double *norm);
int update_u_dev(susy_ctx *c, double eps);
int actions_dev(susy_ctx *c, double *out4);
int plaquette_dev(susy_ctx *c, double *ss, double *st);

int force_init_tables(const ForceTerms &ft);
int fermion_force_dev(susy_ctx *c, double eps, double2 *const *sol, int nsol,
                      const double *amps, double *norm);

// BLAS-1 and reductions (blas.hip)
int reduce_partials(susy_ctx *c, const double *part, int n, int slot);
int read_scalars(susy_ctx *c, double *out, int n);
int tf_dot_re(susy_ctx *c, const double2 *a, const double2 *b, double *out);
int tf_axpby(susy_ctx *c, double2 *y, double a, const double2 *x, double b);
int tf_copy(susy_ctx *c, double2 *dst, const double2 *src);
int tf_stream_copy(susy_ctx *c, double2 *dst, const double2 *src);
int tf_zero(susy_ctx *c, double2 *dst);
int aos_to_soa(susy_ctx *c, const double *aos, double2 *soa, int nmat);
int soa_to_aos(susy_ctx *c, const double2 *soa, double *aos, int nmat);

// CG (cg.hip)
int cg_multi(susy_ctx *c, const double2 *src, double2 *const *psim, int nshift,
             const double *shifts, int MaxCG, double errormin, double *size_r,
             int form);

size_t tf_elems(const susy_ctx *c);   // NTF * NN * V complex

void traj_free(susy_ctx *c);           // (capi.hip)

// learned surrogate (surrogate.hip)
void sg_free(susy_ctx *c);
int sg_force_dev(susy_ctx *c, double eps);   // mom -= eps F_theta at the links
int sg_record_begin(susy_ctx *c);            // training pairs of one outer step
int sg_record_root(susy_ctx *c);
int sg_record_end(susy_ctx *c);
#endif  // SUSY_INTERNAL_H
