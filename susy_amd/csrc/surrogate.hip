// Learned surrogate action S_theta(U) for the inner level of the MD
// integrator (no counterpart in the reference; it is trained on and checked
// against the plain fermion force of update_h.c:928-978).
//
//   S_theta(U) = sum_x w3 . h2(x),
//   h2 = tanh(b2 + sum_t W2[t] h1(x + d_t)),  h1 = tanh(b1 + sum_t W1[t] phi(x + d_t))
// over the 11 taps d_t = 0, +-e_0, ..., +-e_4 of the A4* lattice, with 65
// gauge-invariant features per site: Re, Im Tr / N of the 10 plaquettes
// P_ab = U_a(x) U_b(x+a) U_a(x+b)^dag U_b(x)^dag (a < b) and of the 20
// rectangles R_ab = U_a(x) U_a(x+a) U_b(x+2a) U_a(x+a+b)^dag U_a(x+b)^dag
// U_b(x)^dag (a != b), and Tr(U_a U_a^dag) / N - 1.  W1, W2, w3 are used in
// bfloat16 with FP32 accumulation; loop products are FP64.
//
// The force is F = dS/dUbar in the reference's convention (mom -= eps F,
// U += eps mom): each loop contributes G (rest)^dag where U_a(x) appears and
// conj(G) (rest) where U_a(x)^dag appears, G = (g_re + i g_im) / (2N) from
// the feature gradients g = dS/dphi, "rest" being the loop's other factors
// in cyclic order.  The link owning thread sums its 32 occurrences.
//
// Force matching: loss = |F_theta - F*|^2 / |F*|^2.  With R = F_theta - F*,
// d loss / dtheta = d/dtheta [ d/deps S_theta(phi + eps v) ] / |F*|^2, where
// v = d/dt phi(U + t R) is the tangent of the features along R: one
// forward-mode pass through the network followed by reverse mode.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <map>
#include <random>
#include <vector>

#include "dev_mat.h"

#define SGB 256
#define SG_NT 11
#define SG_NC 16
#define SG_NF 65
#define SG_NLOOP 30
#define SG_NOCC 32
#define SG_W1 0
#define SG_B1 (SG_NT * SG_NC * SG_NF)
#define SG_W2 (SG_B1 + SG_NC)
#define SG_B2 (SG_W2 + SG_NT * SG_NC * SG_NC)
#define SG_W3 (SG_B2 + SG_NC)
#define SG_NP (SG_W3 + SG_NC)

static const int EVEC[NUMLINK][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1},
                                     {-1, -1, -1, -1}};

struct SgTables {
  int nfac[SG_NLOOP];
  int fdir[SG_NLOOP][6], fdag[SG_NLOOP][6], foff[SG_NLOOP][6];   // foff: base-offset table
  // occurrences of link direction a: loop k, position m, base site (diff
  // table id of -off_m) and the other factors in cyclic order m+1 .. m-1
  int ok[NUMLINK][SG_NOCC], om[NUMLINK][SG_NOCC], obase[NUMLINK][SG_NOCC];
  int odir[NUMLINK][SG_NOCC][5], odag[NUMLINK][SG_NOCC][5], odiff[NUMLINK][SG_NOCC][5];
};
__constant__ SgTables c_sg;

struct SgState {
  int nbase = 0, ndiff = 0;
  int *d_base = nullptr, *d_diff = nullptr;     // shift tables (id * V + site)
  float *theta = nullptr, *adam_m = nullptr, *adam_v = nullptr;
  unsigned short *wbf = nullptr;                // bfloat16 copies (same layout)
  double *gacc = nullptr;                       // gradient accumulator
  double *wpart = nullptr;                      // per-chunk weight-gradient partials
  float *phi = nullptr, *h1 = nullptr, *h2 = nullptr, *gz2 = nullptr, *gz1 = nullptr,
        *gphi = nullptr;
  float *v = nullptr, *z1d = nullptr, *h1d = nullptr, *h2d = nullptr, *az2d = nullptr,
        *az2 = nullptr, *az1d = nullptr, *az1 = nullptr;
  double2 *F = nullptr, *R = nullptr, *Frec = nullptr;
  std::vector<double2 *> pairU, pairF;          // recorded (links, fermion force)
  int capacity = 0, npairs = 0, recording = 0;
  long adam_t = 0;
  bool tables = false;
};

// ---------------------------------------------------------------------
// host: loops, occurrences and shift tables
struct Fac {
  int a, off[4], dag;
};
static std::vector<std::vector<Fac>> sg_loops() {
  std::vector<std::vector<Fac>> L;
  auto f = [](int a, const int *o, int d) {
    Fac r;
    r.a = a;
    r.dag = d;
    for (int k = 0; k < 4; k++) r.off[k] = o[k];
    return r;
  };
  for (int a = 0; a < NUMLINK; a++)
    for (int b = a + 1; b < NUMLINK; b++) {
      int z[4] = {0, 0, 0, 0};
      L.push_back({f(a, z, 0), f(b, EVEC[a], 0), f(a, EVEC[b], 1), f(b, z, 1)});
    }
  for (int a = 0; a < NUMLINK; a++)
    for (int b = 0; b < NUMLINK; b++) {
      if (a == b) continue;
      int z[4] = {0, 0, 0, 0}, a2[4], ab[4];
      for (int k = 0; k < 4; k++) {
        a2[k] = 2 * EVEC[a][k];
        ab[k] = EVEC[a][k] + EVEC[b][k];
      }
      L.push_back({f(a, z, 0), f(a, EVEC[a], 0), f(b, a2, 0), f(a, ab, 1), f(a, EVEC[b], 1),
                   f(b, z, 1)});
    }
  return L;
}

static std::vector<int> shift_table(const Geometry &g, const int *v) {
  std::vector<int> t(g.V);
  for (int i = 0; i < g.V; i++)
    t[i] = site_index(g, ((g.x[i] + v[0]) % g.nx + g.nx) % g.nx,
                      ((g.y[i] + v[1]) % g.ny + g.ny) % g.ny,
                      ((g.z[i] + v[2]) % g.nz + g.nz) % g.nz,
                      ((g.t[i] + v[3]) % g.nt + g.nt) % g.nt);
  return t;
}

static int sg_build_tables(susy_ctx *c) {
  SgState &S = *c->sg;
  const Geometry &g = c->geo;
  auto loops = sg_loops();
  if ((int)loops.size() != SG_NLOOP) {
    set_error("surrogate: %d loops, expected %d", (int)loops.size(), SG_NLOOP);
    return SUSY_E_STATE;
  }
  SgTables T;
  memset(&T, 0, sizeof(T));
  std::map<std::vector<int>, int> base_id, diff_id;
  auto id_of = [](std::map<std::vector<int>, int> &m, const int *v) {
    std::vector<int> key(v, v + 4);
    auto it = m.find(key);
    if (it != m.end()) return it->second;
    int n = (int)m.size();
    m[key] = n;
    return n;
  };
  int nocc[NUMLINK] = {0, 0, 0, 0, 0};
  for (int k = 0; k < SG_NLOOP; k++) {
    const auto &lp = loops[k];
    const int n = (int)lp.size();
    T.nfac[k] = n;
    for (int m = 0; m < n; m++) {
      T.fdir[k][m] = lp[m].a;
      T.fdag[k][m] = lp[m].dag;
      T.foff[k][m] = id_of(base_id, lp[m].off);
    }
    for (int m = 0; m < n; m++) {
      const int a = lp[m].a, o = nocc[a]++;
      if (o >= SG_NOCC) {
        set_error("surrogate: more than %d occurrences of link %d", SG_NOCC, a);
        return SUSY_E_STATE;
      }
      T.ok[a][o] = k;
      T.om[a][o] = m;
      int mb[4];
      for (int q = 0; q < 4; q++) mb[q] = -lp[m].off[q];
      T.obase[a][o] = id_of(diff_id, mb);
      for (int j = 1; j < n; j++) {
        const Fac &fj = lp[(m + j) % n];
        int d[4];
        for (int q = 0; q < 4; q++) d[q] = fj.off[q] - lp[m].off[q];
        T.odir[a][o][j - 1] = fj.a;
        T.odag[a][o][j - 1] = fj.dag;
        T.odiff[a][o][j - 1] = id_of(diff_id, d);
      }
    }
  }
  for (int a = 0; a < NUMLINK; a++)
    if (nocc[a] != SG_NOCC) {
      set_error("surrogate: link %d has %d loop occurrences, expected %d", a, nocc[a], SG_NOCC);
      return SUSY_E_STATE;
    }
  const int V = c->V;
  S.nbase = (int)base_id.size();
  S.ndiff = (int)diff_id.size();
  std::vector<int> hb((size_t)S.nbase * V), hd((size_t)S.ndiff * V);
  for (auto &kv : base_id) {
    auto t = shift_table(g, kv.first.data());
    memcpy(&hb[(size_t)kv.second * V], t.data(), sizeof(int) * V);
  }
  for (auto &kv : diff_id) {
    auto t = shift_table(g, kv.first.data());
    memcpy(&hd[(size_t)kv.second * V], t.data(), sizeof(int) * V);
  }
  HIPCHK(hipMalloc(&S.d_base, sizeof(int) * hb.size()));
  HIPCHK(hipMalloc(&S.d_diff, sizeof(int) * hd.size()));
  HIPCHK(hipMemcpy(S.d_base, hb.data(), sizeof(int) * hb.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(S.d_diff, hd.data(), sizeof(int) * hd.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(c_sg), &T, sizeof(T)));
  S.tables = true;
  return 0;
}

// ---------------------------------------------------------------------
// device helpers
__device__ __forceinline__ float bf2f(unsigned short b) {
  return __uint_as_float((unsigned)b << 16);
}
__device__ __forceinline__ unsigned short f2bf(float x) {
  unsigned u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
// site x + d_t: t = 0 self, 1 + 2a: x + e_a (gather 8 + 2a), 2 + 2a: x - e_a
__device__ __forceinline__ int tap_site(const int *__restrict__ nbr, int t, int s, int V) {
  return t == 0 ? s : nbr[(7 + t) * V + s];
}
__device__ __forceinline__ int opp_tap(int t) { return t == 0 ? 0 : ((t & 1) ? t + 1 : t - 1); }

__device__ __forceinline__ void load_op(Mat m, const double2 *__restrict__ U, int a, int site,
                                        int dg, int V) {
  const double2 *p = U + (size_t)a * NN * V + site;
#pragma unroll
  for (int i = 0; i < NCOL; i++)
#pragma unroll
    for (int j = 0; j < NCOL; j++) {
      if (dg) {
        double2 v = p[(j * NCOL + i) * V];
        m[i][j] = make_double2(v.x, -v.y);
      } else {
        m[i][j] = p[(i * NCOL + j) * V];
      }
    }
}
__device__ __forceinline__ void mmul_into(Mat acc, const Mat f) {   // acc = acc f
  Mat t;
  mzero(t);
  mmadd<false, false>(t, 1.0, acc, f);
#pragma unroll
  for (int i = 0; i < NCOL; i++)
#pragma unroll
    for (int j = 0; j < NCOL; j++) acc[i][j] = t[i][j];
}
__device__ __forceinline__ double2 mtrace(const Mat m) {
  double2 t = make_double2(0, 0);
#pragma unroll
  for (int i = 0; i < NCOL; i++) {
    t.x += m[i][i].x;
    t.y += m[i][i].y;
  }
  return t;
}

// ---------------------------------------------------------------------
// features phi[k * V + y]
__global__ void __launch_bounds__(SGB)
k_sg_features(const double2 *__restrict__ U, const int *__restrict__ base,
              float *__restrict__ phi, int V) {
  const int y = blockIdx.x * SGB + threadIdx.x;
  if (y >= V) return;
#pragma unroll 1
  for (int k = 0; k < SG_NLOOP; k++) {
    Mat acc, f;
    load_op(acc, U, c_sg.fdir[k][0], base[c_sg.foff[k][0] * V + y], c_sg.fdag[k][0], V);
#pragma unroll 1
    for (int m = 1; m < c_sg.nfac[k]; m++) {
      load_op(f, U, c_sg.fdir[k][m], base[c_sg.foff[k][m] * V + y], c_sg.fdag[k][m], V);
      mmul_into(acc, f);
    }
    double2 tr = mtrace(acc);
    ST(phi[(size_t)(2 * k) * V + y], (float)(tr.x / NCOL));
    ST(phi[(size_t)(2 * k + 1) * V + y], (float)(tr.y / NCOL));
  }
#pragma unroll 1
  for (int a = 0; a < NUMLINK; a++) {
    const double2 *p = U + (size_t)a * NN * V + y;
    double s = 0;
    for (int e = 0; e < NN; e++) {
      double2 v = p[(size_t)e * V];
      s += v.x * v.x + v.y * v.y;
    }
    ST(phi[(size_t)(60 + a) * V + y], (float)(s / NCOL - 1.0));
  }
}

// v = d/dt phi(U + t R)
__global__ void __launch_bounds__(SGB)
k_sg_jvp(const double2 *__restrict__ U, const double2 *__restrict__ R,
         const int *__restrict__ base, float *__restrict__ v, int V) {
  const int y = blockIdx.x * SGB + threadIdx.x;
  if (y >= V) return;
#pragma unroll 1
  for (int k = 0; k < SG_NLOOP; k++) {
    const int n = c_sg.nfac[k];
    double2 tot = make_double2(0, 0);
#pragma unroll 1
    for (int r = 0; r < n; r++) {
      Mat acc, f;
      load_op(acc, r == 0 ? R : U, c_sg.fdir[k][0], base[c_sg.foff[k][0] * V + y],
              c_sg.fdag[k][0], V);
#pragma unroll 1
      for (int m = 1; m < n; m++) {
        load_op(f, r == m ? R : U, c_sg.fdir[k][m], base[c_sg.foff[k][m] * V + y],
                c_sg.fdag[k][m], V);
        mmul_into(acc, f);
      }
      double2 tr = mtrace(acc);
      tot.x += tr.x;
      tot.y += tr.y;
    }
    ST(v[(size_t)(2 * k) * V + y], (float)(tot.x / NCOL));
    ST(v[(size_t)(2 * k + 1) * V + y], (float)(tot.y / NCOL));
  }
#pragma unroll 1
  for (int a = 0; a < NUMLINK; a++) {
    const double2 *p = U + (size_t)a * NN * V + y, *q = R + (size_t)a * NN * V + y;
    double s = 0;
    for (int e = 0; e < NN; e++) {
      double2 u = p[(size_t)e * V], w = q[(size_t)e * V];
      s += u.x * w.x + u.y * w.y;
    }
    ST(v[(size_t)(60 + a) * V + y], (float)(2.0 * s / NCOL));
  }
}

// out[o](s) = sum_t sum_i W[t][o][i] in[i](s + d_t)       (ADJ = false)
// out[i](s) = sum_t sum_o W[t][o][i] in[o](s - d_t)       (ADJ = true)
// W in bfloat16, row-major [t][o][i] with o < CO, i < CI.
template <int CI, int CO, bool ADJ>
__device__ __forceinline__ void conv_site(const unsigned short *__restrict__ W,
                                          const float *__restrict__ in,
                                          const int *__restrict__ nbr, int s, int V,
                                          float *out) {
  constexpr int NOUT = ADJ ? CI : CO, NIN = ADJ ? CO : CI;
#pragma unroll
  for (int q = 0; q < NOUT; q++) out[q] = 0.f;
#pragma unroll 1
  for (int t = 0; t < SG_NT; t++) {
    const int x = tap_site(nbr, ADJ ? opp_tap(t) : t, s, V);
    const unsigned short *Wt = W + t * CO * CI;
#pragma unroll 1
    for (int p = 0; p < NIN; p++) {
      const float xin = in[(size_t)p * V + x];
#pragma unroll
      for (int q = 0; q < NOUT; q++) {
        const float w = ADJ ? bf2f(Wt[p * CI + q]) : bf2f(Wt[q * CI + p]);
        out[q] = fmaf(w, xin, out[q]);
      }
    }
  }
}

// h1 = tanh(b1 + conv(W1, phi))
__global__ void __launch_bounds__(SGB)
k_sg_conv1(const unsigned short *__restrict__ wbf, const float *__restrict__ theta,
           const float *__restrict__ phi, const int *__restrict__ nbr, float *__restrict__ h1,
           int V) {
  const int s = blockIdx.x * SGB + threadIdx.x;
  if (s >= V) return;
  float z[SG_NC];
  conv_site<SG_NF, SG_NC, false>(wbf + SG_W1, phi, nbr, s, V, z);
#pragma unroll
  for (int o = 0; o < SG_NC; o++) ST(h1[(size_t)o * V + s], tanhf(z[o] + theta[SG_B1 + o]));
}

// h2 = tanh(b2 + conv(W2, h1)); gz2 = (1 - h2^2) w3; partial of sum w3 . h2
__global__ void __launch_bounds__(SGB)
k_sg_conv2(const unsigned short *__restrict__ wbf, const float *__restrict__ theta,
           const float *__restrict__ h1, const int *__restrict__ nbr, float *__restrict__ h2,
           float *__restrict__ gz2, double *__restrict__ partial, int V) {
  const int s = blockIdx.x * SGB + threadIdx.x;
  double act = 0;
  if (s < V) {
    float z[SG_NC];
    conv_site<SG_NC, SG_NC, false>(wbf + SG_W2, h1, nbr, s, V, z);
#pragma unroll
    for (int o = 0; o < SG_NC; o++) {
      const float h = tanhf(z[o] + theta[SG_B2 + o]), w = bf2f(wbf[SG_W3 + o]);
      ST(h2[(size_t)o * V + s], h);
      ST(gz2[(size_t)o * V + s], (1.f - h * h) * w);
      act += (double)w * (double)h;
    }
  }
  block_partial(act, partial, 0, 0);
}

// gz1 = (1 - h1^2) conv^T(W2, gz2)
__global__ void __launch_bounds__(SGB)
k_sg_back1(const unsigned short *__restrict__ wbf, const float *__restrict__ gz2,
           const float *__restrict__ h1, const int *__restrict__ nbr, float *__restrict__ gz1,
           int V) {
  const int s = blockIdx.x * SGB + threadIdx.x;
  if (s >= V) return;
  float g[SG_NC];
  conv_site<SG_NC, SG_NC, true>(wbf + SG_W2, gz2, nbr, s, V, g);
#pragma unroll
  for (int i = 0; i < SG_NC; i++) {
    const float h = h1[(size_t)i * V + s];
    ST(gz1[(size_t)i * V + s], (1.f - h * h) * g[i]);
  }
}

// gphi = conv^T(W1, gz1)
__global__ void __launch_bounds__(SGB)
k_sg_back0(const unsigned short *__restrict__ wbf, const float *__restrict__ gz1,
           const int *__restrict__ nbr, float *__restrict__ gphi, int V) {
  const int s = blockIdx.x * SGB + threadIdx.x;
  if (s >= V) return;
  float g[SG_NF];
  conv_site<SG_NF, SG_NC, true>(wbf + SG_W1, gz1, nbr, s, V, g);
#pragma unroll
  for (int i = 0; i < SG_NF; i++) ST(gphi[(size_t)i * V + s], g[i]);
}

// F = dS/dUbar on link (x, a)
__global__ void __launch_bounds__(SGB)
k_sg_force(const double2 *__restrict__ U, const int *__restrict__ diff,
           const float *__restrict__ gphi, double2 *__restrict__ F, int V) {
  const int x = blockIdx.x * SGB + threadIdx.x, a = blockIdx.y;
  if (x >= V) return;
  Mat acc;
  mzero(acc);
#pragma unroll 1
  for (int o = 0; o < SG_NOCC; o++) {
    const int k = c_sg.ok[a][o], n = c_sg.nfac[k];
    const int y = diff[c_sg.obase[a][o] * V + x];
    const double gr = gphi[(size_t)(2 * k) * V + y], gi = gphi[(size_t)(2 * k + 1) * V + y];
    const double sc = 0.5 / NCOL;
    Mat rest, f;
    load_op(rest, U, c_sg.odir[a][o][0], diff[c_sg.odiff[a][o][0] * V + x], c_sg.odag[a][o][0],
            V);
#pragma unroll 1
    for (int j = 1; j < n - 1; j++) {
      load_op(f, U, c_sg.odir[a][o][j], diff[c_sg.odiff[a][o][j] * V + x], c_sg.odag[a][o][j],
              V);
      mmul_into(rest, f);
    }
    if (c_sg.fdag[k][c_sg.om[a][o]]) {
      // conj(G) rest,  G = (gr + i gi) sc
#pragma unroll
      for (int i = 0; i < NCOL; i++)
#pragma unroll
        for (int j = 0; j < NCOL; j++) {
          const double2 r = rest[i][j];
          acc[i][j].x += sc * (gr * r.x + gi * r.y);
          acc[i][j].y += sc * (gr * r.y - gi * r.x);
        }
    } else {
      // G rest^dag
#pragma unroll
      for (int i = 0; i < NCOL; i++)
#pragma unroll
        for (int j = 0; j < NCOL; j++) {
          const double2 r = make_double2(rest[j][i].x, -rest[j][i].y);
          acc[i][j].x += sc * (gr * r.x - gi * r.y);
          acc[i][j].y += sc * (gr * r.y + gi * r.x);
        }
    }
  }
  const double gs = gphi[(size_t)(60 + a) * V + x] / (double)NCOL;
  const double2 *u = U + (size_t)a * NN * V + x;
  double2 *out = F + (size_t)a * NN * V + x;
#pragma unroll
  for (int e = 0; e < NN; e++) {
    const double2 w = u[(size_t)e * V];
    st_keep(out + (size_t)e * V, make_double2(acc[e / NCOL][e % NCOL].x + gs * w.x,
                                              acc[e / NCOL][e % NCOL].y + gs * w.y));
  }
}

// mom -= eps F
__global__ void k_sg_apply(const double2 *__restrict__ F, double2 *__restrict__ mom, double eps,
                           size_t n) {
  size_t i = (size_t)blockIdx.x * SGB + threadIdx.x;
  if (i < n) {
    double2 f = F[i], m = mom[i];
    m.x -= eps * f.x;
    m.y -= eps * f.y;
    ST(mom[i], m);
  }
}

// R = F - Ft; partials |R|^2 (slot 0) and |Ft|^2 (slot 1)
__global__ void __launch_bounds__(FB)
k_sg_resid(const double2 *__restrict__ F, const double2 *__restrict__ Ft,
           double2 *__restrict__ R, size_t n, int nblk, double *__restrict__ partial) {
  double r2 = 0, f2 = 0;
  for (size_t i = (size_t)blockIdx.x * FB + threadIdx.x; i < n; i += (size_t)gridDim.x * FB) {
    const double2 f = F[i], t = Ft[i];
    const double2 r = make_double2(f.x - t.x, f.y - t.y);
    ST(R[i], r);
    r2 += r.x * r.x + r.y * r.y;
    f2 += t.x * t.x + t.y * t.y;
  }
  block_partial(r2, partial, nblk, 0);
  block_partial(f2, partial, nblk, 1);
}

// tangent pass 1: z1d = conv(W1, v), h1d = (1 - h1^2) z1d
__global__ void __launch_bounds__(SGB)
k_sg_tan1(const unsigned short *__restrict__ wbf, const float *__restrict__ v,
          const float *__restrict__ h1, const int *__restrict__ nbr, float *__restrict__ z1d,
          float *__restrict__ h1d, int V) {
  const int s = blockIdx.x * SGB + threadIdx.x;
  if (s >= V) return;
  float z[SG_NC];
  conv_site<SG_NF, SG_NC, false>(wbf + SG_W1, v, nbr, s, V, z);
#pragma unroll
  for (int o = 0; o < SG_NC; o++) {
    const float h = h1[(size_t)o * V + s];
    ST(z1d[(size_t)o * V + s], z[o]);
    ST(h1d[(size_t)o * V + s], (1.f - h * h) * z[o]);
  }
}

// tangent pass 2: z2d = conv(W2, h1d); h2d = a2 z2d; adjoints
// az2d = a2 w3, az2 = -2 h2 a2 w3 z2d  (a2 = 1 - h2^2)
__global__ void __launch_bounds__(SGB)
k_sg_tan2(const unsigned short *__restrict__ wbf, const float *__restrict__ h1d,
          const float *__restrict__ h2, const int *__restrict__ nbr, float *__restrict__ h2d,
          float *__restrict__ az2d, float *__restrict__ az2, int V) {
  const int s = blockIdx.x * SGB + threadIdx.x;
  if (s >= V) return;
  float z[SG_NC];
  conv_site<SG_NC, SG_NC, false>(wbf + SG_W2, h1d, nbr, s, V, z);
#pragma unroll
  for (int o = 0; o < SG_NC; o++) {
    const float h = h2[(size_t)o * V + s], a2 = 1.f - h * h, w = bf2f(wbf[SG_W3 + o]);
    ST(h2d[(size_t)o * V + s], a2 * z[o]);
    ST(az2d[(size_t)o * V + s], a2 * w);
    ST(az2[(size_t)o * V + s], -2.f * h * a2 * w * z[o]);
  }
}

// tangent pass 3: ah1d = conv^T(W2, az2d); ah1 = conv^T(W2, az2) - 2 h1 ah1d z1d;
// az1d = a1 ah1d, az1 = a1 ah1  (a1 = 1 - h1^2)
__global__ void __launch_bounds__(SGB)
k_sg_tan3(const unsigned short *__restrict__ wbf, const float *__restrict__ az2d,
          const float *__restrict__ az2, const float *__restrict__ h1,
          const float *__restrict__ z1d, const int *__restrict__ nbr,
          float *__restrict__ az1d, float *__restrict__ az1, int V) {
  const int s = blockIdx.x * SGB + threadIdx.x;
  if (s >= V) return;
  float gd[SG_NC], gp[SG_NC];
  conv_site<SG_NC, SG_NC, true>(wbf + SG_W2, az2d, nbr, s, V, gd);
  conv_site<SG_NC, SG_NC, true>(wbf + SG_W2, az2, nbr, s, V, gp);
#pragma unroll
  for (int i = 0; i < SG_NC; i++) {
    const float h = h1[(size_t)i * V + s], a1 = 1.f - h * h;
    const float ah1 = gp[i] - 2.f * h * gd[i] * z1d[(size_t)i * V + s];
    ST(az1d[(size_t)i * V + s], a1 * gd[i]);
    ST(az1[(size_t)i * V + s], a1 * ah1);
  }
}

// parameter gradients of the tangent, summed over sites and added to gacc
struct SgGradIn {
  const float *phi, *v, *h1, *h1d, *h2d, *az2d, *az2, *az1d, *az1;
};
// Two passes.  k_sg_wpart: block (chunk of WG_CHUNK sites, tap t) stages
// WG_SUB sites at a time in LDS -- the adjoints at x and the layer inputs at
// x + d_t -- and accumulates the tap's 1040 + 256 weight entries (thread
// p, p + 256, ...) over its chunk; the tap-0 blocks also sum the 48 bias
// and w3 entries.  k_sg_wsum adds the chunk partials of each parameter in
// chunk order (deterministic).
#define WG_SUB 64
#define WG_CHUNK 1024
#define WG_P1 (SG_NC * SG_NF)
#define WG_PT (WG_P1 + SG_NC * SG_NC)
#define WG_NB (3 * SG_NC)
#define WG_SLOTS ((WG_PT + SGB - 1) / SGB)
__global__ void __launch_bounds__(SGB)
k_sg_wpart(SgGradIn in, const int *__restrict__ nbr, double *__restrict__ part,
           double *__restrict__ bpart, int V) {
  __shared__ float A1[2][WG_SUB][SG_NC], B1[2][WG_SUB][SG_NF];
  __shared__ float A2[2][WG_SUB][SG_NC], B2[2][WG_SUB][SG_NC];
  __shared__ int ys[WG_SUB];
  const int chunk = blockIdx.x, t = blockIdx.y, tid = threadIdx.x;
  const int x0 = chunk * WG_CHUNK, x1 = min(V, x0 + WG_CHUNK);
  double acc[WG_SLOTS];
#pragma unroll
  for (int k = 0; k < WG_SLOTS; k++) acc[k] = 0;
  for (int xs = x0; xs < x1; xs += WG_SUB) {
    const int n = min(WG_SUB, x1 - xs);
    if (tid < n) ys[tid] = tap_site(nbr, t, xs + tid, V);
    __syncthreads();
    for (int e = tid; e < WG_SUB * SG_NC; e += SGB) {
      const int x = e % WG_SUB, o = e / WG_SUB;
      const bool ok = x < n;
      const size_t ix = (size_t)o * V + xs + x;
      A1[0][x][o] = ok ? in.az1d[ix] : 0.f;
      A1[1][x][o] = ok ? in.az1[ix] : 0.f;
      A2[0][x][o] = ok ? in.az2d[ix] : 0.f;
      A2[1][x][o] = ok ? in.az2[ix] : 0.f;
      const size_t iy = (size_t)o * V + (ok ? ys[x] : 0);
      B2[0][x][o] = ok ? in.h1d[iy] : 0.f;
      B2[1][x][o] = ok ? in.h1[iy] : 0.f;
    }
    for (int e = tid; e < WG_SUB * SG_NF; e += SGB) {
      const int x = e % WG_SUB, i = e / WG_SUB;
      const bool ok = x < n;
      const size_t iy = (size_t)i * V + (ok ? ys[x] : 0);
      B1[0][x][i] = ok ? in.v[iy] : 0.f;
      B1[1][x][i] = ok ? in.phi[iy] : 0.f;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < WG_SLOTS; k++) {
      const int p = tid + k * SGB;
      if (p < WG_P1) {
        const int o = p / SG_NF, i = p % SG_NF;
        double s = 0;
        for (int x = 0; x < WG_SUB; x++)
          s += (double)A1[0][x][o] * B1[0][x][i] + (double)A1[1][x][o] * B1[1][x][i];
        acc[k] += s;
      } else if (p < WG_PT) {
        const int q = p - WG_P1, o = q / SG_NC, i = q % SG_NC;
        double s = 0;
        for (int x = 0; x < WG_SUB; x++)
          s += (double)A2[0][x][o] * B2[0][x][i] + (double)A2[1][x][o] * B2[1][x][i];
        acc[k] += s;
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int k = 0; k < WG_SLOTS; k++) {
    const int p = tid + k * SGB;
    if (p < WG_PT) ST(part[((size_t)chunk * SG_NT + t) * WG_PT + p], acc[k]);
  }
  if (t == 0 && tid < WG_NB) {
    const float *f = tid < SG_NC ? in.az1 : tid < 2 * SG_NC ? in.az2 : in.h2d;
    const int o = tid % SG_NC;
    double s = 0;
    for (int x = x0; x < x1; x++) s += f[(size_t)o * V + x];
    ST(bpart[(size_t)chunk * WG_NB + tid], s);
  }
}

__global__ void __launch_bounds__(SGB)
k_sg_wsum(const double *__restrict__ part, const double *__restrict__ bpart, int nchunk,
          double *__restrict__ gacc, double scale) {
  const int p = blockIdx.x * SGB + threadIdx.x;
  if (p >= SG_NP) return;
  double s = 0;
  if (p < SG_B1 || (p >= SG_W2 && p < SG_B2)) {
    const int t = p < SG_B1 ? p / WG_P1 : (p - SG_W2) / (SG_NC * SG_NC);
    const int loc = p < SG_B1 ? p % WG_P1 : WG_P1 + (p - SG_W2) % (SG_NC * SG_NC);
    for (int c = 0; c < nchunk; c++) s += part[((size_t)c * SG_NT + t) * WG_PT + loc];
  } else {
    const int b = p < SG_W2 ? p - SG_B1 : p < SG_W3 ? SG_NC + p - SG_B2 : 2 * SG_NC + p - SG_W3;
    for (int c = 0; c < nchunk; c++) s += bpart[(size_t)c * WG_NB + b];
  }
  gacc[p] += s * scale;
}

// Adam step on the FP32 master parameters; refreshes the bfloat16 copies and
// clears the gradient accumulator
__global__ void k_sg_adam(float *__restrict__ theta, float *__restrict__ m,
                          float *__restrict__ v, unsigned short *__restrict__ wbf,
                          double *__restrict__ gacc, double scale, float lr, float b1c,
                          float b2c) {
  const int p = blockIdx.x * SGB + threadIdx.x;
  if (p >= SG_NP) return;
  const float g = (float)(gacc[p] * scale);
  const float mm = 0.9f * m[p] + 0.1f * g, vv = 0.999f * v[p] + 0.001f * g * g;
  ST(m[p], mm);
  ST(v[p], vv);
  const float th = theta[p] - lr * (mm / b1c) / (sqrtf(vv / b2c) + 1e-8f);
  ST(theta[p], th);
  ST(wbf[p], f2bf(th));
  ST(gacc[p], 0.0);
}

__global__ void k_sg_round(const float *__restrict__ theta, unsigned short *__restrict__ wbf) {
  const int p = blockIdx.x * SGB + threadIdx.x;
  if (p < SG_NP) ST(wbf[p], f2bf(theta[p]));
}

__global__ void k_sg_addf(double2 *__restrict__ y, const double2 *__restrict__ x, size_t n) {
  size_t i = (size_t)blockIdx.x * SGB + threadIdx.x;
  if (i < n) {
    double2 v = y[i], w = x[i];
    v.x += w.x;
    v.y += w.y;
    st_keep(y + i, v);
  }
}

// ---------------------------------------------------------------------
// host
static int sg_alloc_f(float **p, size_t n) {
  if (hipMalloc(p, n * sizeof(float)) != hipSuccess) {
    set_error("surrogate: hipMalloc of %zu floats failed", n);
    return SUSY_E_NOMEM;
  }
  return 0;
}
#define SGRET(x)           \
  do {                     \
    int _r = (x);          \
    if (_r < 0) return _r; \
  } while (0)

void sg_free(susy_ctx *c) {
  if (!c->sg) return;
  SgState &S = *c->sg;
  void *ptrs[] = {S.d_base, S.d_diff, S.theta, S.adam_m, S.adam_v, S.wbf, S.gacc, S.wpart, S.phi, S.h1,
                  S.h2, S.gz2, S.gz1, S.gphi, S.v, S.z1d, S.h1d, S.h2d, S.az2d, S.az2, S.az1d,
                  S.az1, S.F, S.R, S.Frec};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  for (double2 *p : S.pairU) (void)hipFree(p);
  for (double2 *p : S.pairF) (void)hipFree(p);
  delete c->sg;
  c->sg = nullptr;
}

static int sg_ready(susy_ctx *c) {
  if (!c || !c->sg) {
    set_error("surrogate not initialised (susy_sg_init)");
    return SUSY_E_STATE;
  }
  return 0;
}

extern "C" int susy_sg_init(susy_ctx *c, int seed, double scale) {
  if (!c) return SUSY_E_STATE;
  sg_free(c);
  c->sg = new SgState();
  SgState &S = *c->sg;
  SGRET(sg_build_tables(c));
  const size_t V = c->V, mat5 = (size_t)NUMLINK * NN * V;
  SGRET(sg_alloc_f(&S.theta, SG_NP));
  SGRET(sg_alloc_f(&S.adam_m, SG_NP));
  SGRET(sg_alloc_f(&S.adam_v, SG_NP));
  HIPCHK(hipMalloc(&S.wbf, sizeof(unsigned short) * SG_NP));
  HIPCHK(hipMalloc(&S.gacc, sizeof(double) * SG_NP));
  {
    const size_t nchunk = (V + WG_CHUNK - 1) / WG_CHUNK;
    HIPCHK(hipMalloc(&S.wpart, sizeof(double) * nchunk * (SG_NT * WG_PT + WG_NB)));
  }
  float **planes65[] = {&S.phi, &S.gphi, &S.v};
  for (float **p : planes65) SGRET(sg_alloc_f(p, SG_NF * V));
  float **planes16[] = {&S.h1, &S.h2, &S.gz2, &S.gz1, &S.z1d, &S.h1d, &S.h2d, &S.az2d, &S.az2,
                        &S.az1d, &S.az1};
  for (float **p : planes16) SGRET(sg_alloc_f(p, SG_NC * V));
  HIPCHK(hipMalloc(&S.F, sizeof(double2) * mat5));
  HIPCHK(hipMalloc(&S.R, sizeof(double2) * mat5));
  HIPCHK(hipMalloc(&S.Frec, sizeof(double2) * mat5));
  // initial parameters: N(0, scale^2 / fan-in) weights, zero biases
  std::vector<float> th(SG_NP, 0.f);
  std::mt19937_64 gen((unsigned long long)seed);
  std::normal_distribution<double> nd(0.0, 1.0);
  for (int p = 0; p < SG_NP; p++) {
    double fan = p < SG_B1 ? SG_NT * SG_NF : (p >= SG_W2 && p < SG_B2) ? SG_NT * SG_NC
                 : p >= SG_W3 ? SG_NC : 0;
    th[p] = fan > 0 ? (float)(nd(gen) * scale / sqrt(fan)) : 0.f;
  }
  HIPCHK(hipMemcpy(S.theta, th.data(), sizeof(float) * SG_NP, hipMemcpyHostToDevice));
  HIPCHK(hipMemset(S.adam_m, 0, sizeof(float) * SG_NP));
  HIPCHK(hipMemset(S.adam_v, 0, sizeof(float) * SG_NP));
  HIPCHK(hipMemset(S.gacc, 0, sizeof(double) * SG_NP));
  hipLaunchKernelGGL(k_sg_round, dim3((SG_NP + SGB - 1) / SGB), dim3(SGB), 0, c->s, S.theta,
                     S.wbf);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->s));
  return 0;
}

extern "C" int susy_sg_nparams(void) { return SG_NP; }

extern "C" int susy_sg_set_params(susy_ctx *c, const float *theta) {
  SGRET(sg_ready(c));
  SgState &S = *c->sg;
  HIPCHK(hipMemcpy(S.theta, theta, sizeof(float) * SG_NP, hipMemcpyHostToDevice));
  HIPCHK(hipMemset(S.adam_m, 0, sizeof(float) * SG_NP));
  HIPCHK(hipMemset(S.adam_v, 0, sizeof(float) * SG_NP));
  S.adam_t = 0;
  hipLaunchKernelGGL(k_sg_round, dim3((SG_NP + SGB - 1) / SGB), dim3(SGB), 0, c->s, S.theta,
                     S.wbf);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->s));
  return 0;
}

extern "C" int susy_sg_get_params(susy_ctx *c, float *theta) {
  SGRET(sg_ready(c));
  HIPCHK(hipStreamSynchronize(c->s));
  HIPCHK(hipMemcpy(theta, c->sg->theta, sizeof(float) * SG_NP, hipMemcpyDeviceToHost));
  return 0;
}

// S_theta and F_theta (into S.F) for links U; the action is read back when
// `action` is non-null
static int sg_eval(susy_ctx *c, const double2 *U, double *action) {
  SgState &S = *c->sg;
  const int V = c->V, nb = (V + SGB - 1) / SGB;
  hipLaunchKernelGGL(k_sg_features, dim3(nb), dim3(SGB), 0, c->s, U, S.d_base, S.phi, V);
  hipLaunchKernelGGL(k_sg_conv1, dim3(nb), dim3(SGB), 0, c->s, S.wbf, S.theta, S.phi, c->d_nbr,
                     S.h1, V);
  hipLaunchKernelGGL(k_sg_conv2, dim3(nb), dim3(SGB), 0, c->s, S.wbf, S.theta, S.h1, c->d_nbr,
                     S.h2, S.gz2, c->d_part, V);
  HIPCHK(hipGetLastError());
  if (action) {
    SGRET(reduce_partials(c, c->d_part, nb, 0));
    SGRET(read_scalars(c, action, 1));
  }
  hipLaunchKernelGGL(k_sg_back1, dim3(nb), dim3(SGB), 0, c->s, S.wbf, S.gz2, S.h1, c->d_nbr,
                     S.gz1, V);
  hipLaunchKernelGGL(k_sg_back0, dim3(nb), dim3(SGB), 0, c->s, S.wbf, S.gz1, c->d_nbr, S.gphi,
                     V);
  hipLaunchKernelGGL(k_sg_force, dim3(nb, NUMLINK), dim3(SGB), 0, c->s, U, S.d_diff, S.gphi,
                     S.F, V);
  HIPCHK(hipGetLastError());
  return 0;
}

// mom -= eps F_theta at the context's links (the inner level of the
// surrogate integrator uses eps > 0, the outer correction eps < 0)
int sg_force_dev(susy_ctx *c, double eps) {
  SGRET(sg_ready(c));
  SGRET(sg_eval(c, c->d_links, nullptr));
  const size_t n = (size_t)NUMLINK * NN * c->V;
  hipLaunchKernelGGL(k_sg_apply, dim3((n + SGB - 1) / SGB), dim3(SGB), 0, c->s, c->sg->F,
                     c->d_mom, eps, n);
  HIPCHK(hipGetLastError());
  return 0;
}

extern "C" int susy_sg_eval(susy_ctx *c, double *action, double *F_aos) {
  SGRET(sg_ready(c));
  SGRET(sg_eval(c, c->d_links, action));
  if (F_aos) {
    double *h;
    const size_t n = (size_t)NUMLINK * NN * c->V;
    HIPCHK(hipMalloc(&h, sizeof(double2) * n));
    int r = soa_to_aos(c, c->sg->F, h, NUMLINK);
    if (!r) {
      HIPCHK(hipStreamSynchronize(c->s));
      HIPCHK(hipMemcpy(F_aos, h, sizeof(double2) * n, hipMemcpyDeviceToHost));
    }
    (void)hipFree(h);
    return r;
  }
  return 0;
}

// features of the context's links as planes [k][site] (test access)
extern "C" int susy_sg_features(susy_ctx *c, float *phi) {
  SGRET(sg_ready(c));
  SGRET(sg_eval(c, c->d_links, nullptr));
  HIPCHK(hipStreamSynchronize(c->s));
  HIPCHK(hipMemcpy(phi, c->sg->phi, sizeof(float) * SG_NF * c->V, hipMemcpyDeviceToHost));
  return 0;
}

// loss of one (links, target force) pair; with `grad`, its gradient (which
// carries the pair's 1 / |F*|^2) is added to the accumulator times `gscale`
static int sg_loss_grad(susy_ctx *c, const double2 *U, const double2 *Ft, double *loss,
                        bool grad, double gscale) {
  SgState &S = *c->sg;
  const int V = c->V, nb = (V + SGB - 1) / SGB;
  SGRET(sg_eval(c, U, nullptr));
  const size_t n = (size_t)NUMLINK * NN * V;
  const int rb = 256;
  hipLaunchKernelGGL(k_sg_resid, dim3(rb), dim3(FB), 0, c->s, S.F, Ft, S.R, n, rb, c->d_part);
  HIPCHK(hipGetLastError());
  SGRET(reduce_partials(c, c->d_part, rb, 0));
  SGRET(reduce_partials(c, c->d_part + rb, rb, 1));
  double sc[2];
  SGRET(read_scalars(c, sc, 2));
  if (sc[1] <= 0) {
    set_error("surrogate: target force is zero");
    return SUSY_E_ARG;
  }
  *loss = sc[0] / sc[1];
  if (!grad) return 0;
  hipLaunchKernelGGL(k_sg_jvp, dim3(nb), dim3(SGB), 0, c->s, U, S.R, S.d_base, S.v, V);
  hipLaunchKernelGGL(k_sg_tan1, dim3(nb), dim3(SGB), 0, c->s, S.wbf, S.v, S.h1, c->d_nbr, S.z1d,
                     S.h1d, V);
  hipLaunchKernelGGL(k_sg_tan2, dim3(nb), dim3(SGB), 0, c->s, S.wbf, S.h1d, S.h2, c->d_nbr,
                     S.h2d, S.az2d, S.az2, V);
  hipLaunchKernelGGL(k_sg_tan3, dim3(nb), dim3(SGB), 0, c->s, S.wbf, S.az2d, S.az2, S.h1, S.z1d,
                     c->d_nbr, S.az1d, S.az1, V);
  SgGradIn in = {S.phi, S.v, S.h1, S.h1d, S.h2d, S.az2d, S.az2, S.az1d, S.az1};
  const int nchunk = (V + WG_CHUNK - 1) / WG_CHUNK;
  hipLaunchKernelGGL(k_sg_wpart, dim3(nchunk, SG_NT), dim3(SGB), 0, c->s, in, c->d_nbr, S.wpart,
                     S.wpart + (size_t)nchunk * SG_NT * WG_PT, V);
  hipLaunchKernelGGL(k_sg_wsum, dim3((SG_NP + SGB - 1) / SGB), dim3(SGB), 0, c->s, S.wpart,
                     S.wpart + (size_t)nchunk * SG_NT * WG_PT, nchunk, S.gacc, gscale / sc[1]);
  HIPCHK(hipGetLastError());
  return 0;
}

// Test access: loss and gradient (over the parameters, as used in bfloat16)
// for links and a target force given in the reference's AoS layout.
extern "C" int susy_sg_loss_grad(susy_ctx *c, const double *U_aos, const double *F_aos,
                                 double *loss, double *grad) {
  SGRET(sg_ready(c));
  SgState &S = *c->sg;
  const size_t n = (size_t)NUMLINK * NN * c->V;
  double *h;
  double2 *Ud, *Fd;
  HIPCHK(hipMalloc(&h, sizeof(double2) * n));
  HIPCHK(hipMalloc(&Ud, sizeof(double2) * n));
  HIPCHK(hipMalloc(&Fd, sizeof(double2) * n));
  int r = 0;
  HIPCHK(hipMemcpy(h, U_aos, sizeof(double2) * n, hipMemcpyHostToDevice));
  if (!r) r = aos_to_soa(c, h, Ud, NUMLINK);
  HIPCHK(hipStreamSynchronize(c->s));
  HIPCHK(hipMemcpy(h, F_aos, sizeof(double2) * n, hipMemcpyHostToDevice));
  if (!r) r = aos_to_soa(c, h, Fd, NUMLINK);
  HIPCHK(hipMemsetAsync(S.gacc, 0, sizeof(double) * SG_NP, c->s));
  if (!r) r = sg_loss_grad(c, Ud, Fd, loss, grad != nullptr, 1.0);
  if (!r && grad) {
    HIPCHK(hipStreamSynchronize(c->s));
    HIPCHK(hipMemcpy(grad, S.gacc, sizeof(double) * SG_NP, hipMemcpyDeviceToHost));
    HIPCHK(hipMemsetAsync(S.gacc, 0, sizeof(double) * SG_NP, c->s));
  }
  HIPCHK(hipStreamSynchronize(c->s));
  (void)hipFree(h);
  (void)hipFree(Ud);
  (void)hipFree(Fd);
  return r;
}

// Recording of (links, fermion force) pairs during susy_update: capacity
// pairs are kept (later ones overwrite the oldest); 0 stops recording.
extern "C" int susy_sg_record(susy_ctx *c, int capacity) {
  SGRET(sg_ready(c));
  SgState &S = *c->sg;
  if (capacity < 0) return SUSY_E_ARG;
  const size_t n = (size_t)NUMLINK * NN * c->V;
  while ((int)S.pairU.size() < capacity) {
    double2 *u, *f;
    if (hipMalloc(&u, sizeof(double2) * n) != hipSuccess ||
        hipMalloc(&f, sizeof(double2) * n) != hipSuccess) {
      set_error("surrogate: no memory for %d recorded pairs", capacity);
      return SUSY_E_NOMEM;
    }
    S.pairU.push_back(u);
    S.pairF.push_back(f);
  }
  if (capacity > 0) S.capacity = capacity;
  S.recording = capacity > 0;
  return 0;
}
extern "C" int susy_sg_pairs(susy_ctx *c) {
  SGRET(sg_ready(c));
  return c->sg->npairs < c->sg->capacity ? c->sg->npairs : c->sg->capacity;
}

// integrator hooks: the fermion force of one outer step is the sum over
// roots of the assembled per-root force left in d_force
int sg_record_begin(susy_ctx *c) {
  if (!c->sg || !c->sg->recording) return 0;
  HIPCHK(hipMemsetAsync(c->sg->Frec, 0, sizeof(double2) * NUMLINK * NN * c->V, c->s));
  return 0;
}
int sg_record_root(susy_ctx *c) {
  if (!c->sg || !c->sg->recording) return 0;
  const size_t n = (size_t)NUMLINK * NN * c->V;
  hipLaunchKernelGGL(k_sg_addf, dim3((n + SGB - 1) / SGB), dim3(SGB), 0, c->s, c->sg->Frec,
                     c->d_force, n);
  HIPCHK(hipGetLastError());
  return 0;
}
int sg_record_end(susy_ctx *c) {
  if (!c->sg || !c->sg->recording) return 0;
  SgState &S = *c->sg;
  const size_t n = (size_t)NUMLINK * NN * c->V;
  const int slot = S.npairs % S.capacity;
  HIPCHK(hipMemcpyAsync(S.pairU[slot], c->d_links, sizeof(double2) * n,
                        hipMemcpyDeviceToDevice, c->s));
  HIPCHK(hipMemcpyAsync(S.pairF[slot], S.Frec, sizeof(double2) * n, hipMemcpyDeviceToDevice,
                        c->s));
  S.npairs++;
  return 0;
}

// Force-matching loss of recorded pair i with the current parameters.
extern "C" int susy_sg_pair_loss(susy_ctx *c, int i, double *loss) {
  SGRET(sg_ready(c));
  SgState &S = *c->sg;
  const int np = S.npairs < S.capacity ? S.npairs : S.capacity;
  if (i < 0 || i >= np) return SUSY_E_ARG;
  return sg_loss_grad(c, S.pairU[i], S.pairF[i], loss, false, 0.0);
}

// Adam on recorded pairs [first, first + count): `steps` steps of `batch`
// pairs drawn with a seeded generator; loss_hist[step] = mean batch loss.
extern "C" int susy_sg_train(susy_ctx *c, int first, int count, int steps, int batch,
                             double lr, int seed, double *loss_hist) {
  SGRET(sg_ready(c));
  SgState &S = *c->sg;
  const int np = S.npairs < S.capacity ? S.npairs : S.capacity;
  if (first < 0 || count < 1 || first + count > np || batch < 1 || steps < 0) {
    set_error("susy_sg_train: pairs [%d, %d) outside the %d recorded", first, first + count, np);
    return SUSY_E_ARG;
  }
  std::mt19937_64 gen((unsigned long long)seed);
  std::uniform_int_distribution<int> pick(first, first + count - 1);
  HIPCHK(hipMemsetAsync(S.gacc, 0, sizeof(double) * SG_NP, c->s));
  for (int st = 0; st < steps; st++) {
    double tot = 0;
    for (int b = 0; b < batch; b++) {
      const int i = pick(gen);
      double l;
      SGRET(sg_loss_grad(c, S.pairU[i], S.pairF[i], &l, true, 1.0 / batch));
      tot += l;
    }
    S.adam_t++;
    const float b1c = 1.f - powf(0.9f, (float)S.adam_t);
    const float b2c = 1.f - powf(0.999f, (float)S.adam_t);
    hipLaunchKernelGGL(k_sg_adam, dim3((SG_NP + SGB - 1) / SGB), dim3(SGB), 0, c->s, S.theta,
                       S.adam_m, S.adam_v, S.wbf, S.gacc, 1.0, (float)lr, b1c, b2c);
    HIPCHK(hipGetLastError());
    if (loss_hist) loss_hist[st] = tot / batch;
  }
  HIPCHK(hipStreamSynchronize(c->s));
  return 0;
}
