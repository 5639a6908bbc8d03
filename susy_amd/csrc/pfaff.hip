// Pfaffian of the fermion operator on small lattices (the reference's phase
// measurement, 4d_Q16/susy/phase.c, printed as PFAFF in
// testsuite/scalar/phase.U*.ref).
//
// M_ab = <e_a | D e_b> with the bilinear product sum_c a_c b_c, in the
// reference's basis (phase.c:32-213): per site and generator Lambda_j the 16
// components site, link 4, link 0, plaq(0,4), link 1, plaq(1,4), link 2,
// plaq(2,4), link 3, plaq(3,4), plaq(0,1), plaq(2,3), plaq(0,2), plaq(1,3),
// plaq(0,3), plaq(1,2), where the (2,3), (1,3) and (1,2) entries of site x
// sit on the Twist_Fermion of site x + e_0 + e_1, x + e_0 + e_2 and
// x + e_0 + e_3 (gathers 22, 30, 34); in this basis M is antisymmetric.
// Columns come from the FP64 operator on the device; the Pfaffian of the
// n x n matrix (n = 16 N^2 V) is computed on the host by the Parlett-Reid
// LTL^T elimination with pivoting (M. Wimmer, ACM TOMS 38 (2012) 30,
// "pfaffian_LTL"), the rank-2 updates spread over threads.  The result is
// independent of the elimination order, so it compares directly with the
// reference's (which eliminates by its own column-pair scheme).
#include <math.h>

#include <algorithm>
#include <complex>
#include <thread>
#include <vector>

#include "susy_internal.h"

typedef std::complex<double> cd;

namespace {

// basis slot k of a site: Twist_Fermion component and the gather that moves
// it to the site whose Twist_Fermion carries it (-1: the site itself)
const int kComp[16] = {0, 5, 1, 9, 2, 12, 3, 14, 4, 15, 6, 13, 7, 11, 8, 10};
const int kGather[16] = {-1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, 22, -1, 30, -1, 34};

// log |Pf(A)| and arg Pf(A) of the antisymmetric n x n matrix A (row-major,
// destroyed); returns false if a pivot vanishes (Pf = 0)
bool pfaffian_ltl(std::vector<cd> &A, int n, double *log_mag, double *arg) {
  double lm = 0, ph = 0;
  const int nthr = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<cd> tau(n), col(n);
  for (int k = 0; k + 1 < n; k += 2) {
    // largest |A[i][k]|, i > k, moved to row / column k + 1
    int kp = k + 1;
    double best = std::abs(A[(size_t)(k + 1) * n + k]);
    for (int i = k + 2; i < n; i++) {
      const double v = std::abs(A[(size_t)i * n + k]);
      if (v > best) {
        best = v;
        kp = i;
      }
    }
    if (kp != k + 1) {
      for (int j = k; j < n; j++) std::swap(A[(size_t)(k + 1) * n + j], A[(size_t)kp * n + j]);
      for (int i = k; i < n; i++) std::swap(A[(size_t)i * n + k + 1], A[(size_t)i * n + kp]);
      ph += M_PI;   // a row-and-column interchange flips the sign of Pf
    }
    const cd piv = A[(size_t)k * n + k + 1];
    if (best == 0.0 || piv == cd(0, 0)) return false;
    lm += std::log(std::abs(piv));
    ph += std::arg(piv);
    const int m0 = k + 2;
    if (m0 >= n) break;
    for (int j = m0; j < n; j++) {
      tau[j] = A[(size_t)k * n + j] / piv;
      col[j] = A[(size_t)j * n + k + 1];
    }
    // A[i][j] += tau_i A[j][k+1] - A[i][k+1] tau_j over i, j >= k + 2
    auto rows = [&](int r0, int r1) {
      for (int i = r0; i < r1; i++) {
        cd *row = &A[(size_t)i * n];
        const cd ti = tau[i], ci = col[i];
        double *rw = reinterpret_cast<double *>(row);
        const double *cj = reinterpret_cast<const double *>(col.data());
        const double *tj = reinterpret_cast<const double *>(tau.data());
        const double tr = ti.real(), tim = ti.imag(), cr = ci.real(), cim = ci.imag();
        for (int j = m0; j < n; j++) {
          const double c0 = cj[2 * j], c1 = cj[2 * j + 1], t0 = tj[2 * j], t1 = tj[2 * j + 1];
          rw[2 * j] += (tr * c0 - tim * c1) - (cr * t0 - cim * t1);
          rw[2 * j + 1] += (tr * c1 + tim * c0) - (cr * t1 + cim * t0);
        }
      }
    };
    const int m = n - m0;
    if (m < 256 || nthr == 1) {
      rows(m0, n);
    } else {
      std::vector<std::thread> th;
      for (int t = 0; t < nthr; t++) {
        const int r0 = m0 + (int)((long)m * t / nthr), r1 = m0 + (int)((long)m * (t + 1) / nthr);
        th.emplace_back(rows, r0, r1);
      }
      for (auto &x : th) x.join();
    }
  }
  *log_mag = lm;
  *arg = ph;
  return true;
}

}  // namespace

// M[a][b] (row-major n x n, re/im interleaved; n = 16 N^2 V) into m
static int fermion_matrix(susy_ctx *c, double *m) {
  const int V = c->V, n = 16 * NN * V;
  const Geometry &g = c->geo;
  // Twist_Fermion site carrying basis entry (site s, slot k); the read-out
  // uses the same sites (the reference gathers back with the opposite paths)
  auto carrier = [&](int s, int k) { return kGather[k] < 0 ? s : g.gathers[kGather[k] * V + s]; };
  const size_t tfn = (size_t)V * NTF * NN;   // complex numbers per field
  std::vector<double> host(2 * tfn), out(2 * tfn);
  const double *lam = c->lambda_gen.data();   // [j][r][s] re/im
  std::vector<size_t> slot(n);
  for (int a = 0; a < n; a++) {
    const int sa = a / (16 * NN), ka = a % 16;
    slot[a] = ((size_t)carrier(sa, ka) * NTF + kComp[ka]) * NN;
  }
  int h = c->pf_handle;
  if (h < 0) {
    h = susy_tf_alloc(c, 2);
    if (h < 0) return h;
    c->pf_handle = h;
  }
  for (int b = 0; b < n; b++) {
    // basis vector b = (site, generator jb, slot): Lambda_jb in the
    // component kComp[slot] of the Twist_Fermion at the carrier site
    const int jb = (b / 16) % NN;
    std::fill(host.begin(), host.end(), 0.0);
    for (int e = 0; e < NN; e++) {
      host[2 * (slot[b] + e)] = lam[2 * (jb * NN + e)];
      host[2 * (slot[b] + e) + 1] = lam[2 * (jb * NN + e) + 1];
    }
    int r = susy_tf_upload(c, h, (const Twist_Fermion *)host.data());
    if (!r) r = susy_fermion_op_dev(c, h, h + 1, 1, SUSY_FORM_FP64);
    if (!r) r = susy_tf_download(c, h + 1, (Twist_Fermion *)out.data());
    if (r) return r;
    // read-out: M[a][b] = Tr(res_slot(a) Lambda_ja) = sum_ik x_ik Lambda_ja,ki
    for (int a = 0; a < n; a++) {
      const int ja = (a / 16) % NN;
      const double *x = &out[2 * slot[a]];
      const double *l = &lam[2 * ja * NN];
      double re = 0, im = 0;
      for (int i = 0; i < NCOL; i++)
        for (int k = 0; k < NCOL; k++) {
          const double xr = x[2 * (i * NCOL + k)], xi = x[2 * (i * NCOL + k) + 1];
          const double lr = l[2 * (k * NCOL + i)], li = l[2 * (k * NCOL + i) + 1];
          re += xr * lr - xi * li;
          im += xr * li + xi * lr;
        }
      m[2 * ((size_t)a * n + b)] = re;
      m[2 * ((size_t)a * n + b) + 1] = im;
    }
  }
  return 0;
}

static int check_size(susy_ctx *c, const char *who) {
  const int n = 16 * NN * c->V;
  if (n > 8192) {
    set_error("%s: %d x %d matrix too large (at most 8192)", who, n, n);
    return SUSY_E_ARG;
  }
  return 0;
}

// The n x n matrix of the fermion operator in the basis above (phase.c's
// matvec, 4d_Q16/susy/phase.c:36-213), M[a][b] row-major with (re, im)
// interleaved into the caller's 2 n^2 doubles.  Refuses n > 8192.
extern "C" int susy_fermion_matrix_dev(susy_ctx *c, double *m) {
  if (!c) return SUSY_E_STATE;
  if (!m) {
    set_error("susy_fermion_matrix_dev: null output");
    return SUSY_E_ARG;
  }
  if (int r = check_size(c, "susy_fermion_matrix_dev")) return r;
  return fermion_matrix(c, m);
}

// log |Pf M| and its phase in [0, 2 pi) (phase.c:369-390 prints them, with
// |cos| and |sin| of the phase); asym = max |M + M^T| / max |M| as a check
// of the basis (may be NULL).  Refuses n > 8192 (the matrix is held on the
// host).
extern "C" int susy_pfaffian_dev(susy_ctx *c, double *log_mag, double *phase, double *asym) {
  if (!c) return SUSY_E_STATE;
  if (!log_mag || !phase) {
    set_error("susy_pfaffian_dev: null output");
    return SUSY_E_ARG;
  }
  if (int r = check_size(c, "susy_pfaffian_dev")) return r;
  const int n = 16 * NN * c->V;
  std::vector<cd> M((size_t)n * n);
  if (int r = fermion_matrix(c, (double *)M.data())) return r;
  if (asym) {
    double mx = 0, ma = 0;
    for (int a = 0; a < n; a++)
      for (int b = 0; b < n; b++) {
        mx = std::max(mx, std::abs(M[(size_t)a * n + b]));
        ma = std::max(ma, std::abs(M[(size_t)a * n + b] + M[(size_t)b * n + a]));
      }
    *asym = mx > 0 ? ma / mx : 0.0;
  }
  double lm, ph;
  if (!pfaffian_ltl(M, n, &lm, &ph)) {
    *log_mag = -INFINITY;
    *phase = 0.0;
    return 0;
  }
  const double two_pi = 2.0 * M_PI;
  ph = fmod(ph, two_pi);
  if (ph < 0) ph += two_pi;
  *log_mag = lm;
  *phase = ph;
  return 0;
}
