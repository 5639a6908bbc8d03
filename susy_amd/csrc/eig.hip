// Extreme eigenvalues of DSq = D^dag D + fmass^2 (the reference's eigenvalue measurement
// prints them for its regression configurations, testsuite/scalar/eig.U*.ref;
// the reference computes them with PRIMME, eig.c).  Here: Lanczos with full
// reorthogonalisation (classical Gram-Schmidt, twice) on the FP64 operator
// of the context, Lanczos vectors kept in HBM, the tridiagonal eigenvalues
// by Sturm-sequence bisection on the host.  Each distinct eigenvalue appears
// among the Ritz values; exact degeneracies are resolved only as far as
// rounding seeds the second direction.
#include <math.h>
#include <string.h>

#include <random>
#include <vector>

#include "susy_internal.h"

#define EB 256

// partial[i * nb + b] = sum over block b of conj(V_i) w   (grid: nb x k)
__global__ void __launch_bounds__(EB)
k_eig_dots(const double2 *__restrict__ V, const double2 *__restrict__ w, size_t n,
           double2 *__restrict__ partial) {
  const int i = blockIdx.y, nb = gridDim.x;
  const double2 *v = V + (size_t)i * n;
  double sr = 0, si = 0;
  for (size_t e = (size_t)blockIdx.x * EB + threadIdx.x; e < n; e += (size_t)nb * EB) {
    const double2 a = v[e], b = w[e];
    sr += a.x * b.x + a.y * b.y;
    si += a.x * b.y - a.y * b.x;
  }
  for (int off = 32; off > 0; off >>= 1) {
    sr += __shfl_xor(sr, off, 64);
    si += __shfl_xor(si, off, 64);
  }
  __shared__ double rr[EB / 64], ri[EB / 64];
  if ((threadIdx.x & 63) == 0) {
    rr[threadIdx.x >> 6] = sr;
    ri[threadIdx.x >> 6] = si;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double tr = 0, ti = 0;
    for (int q = 0; q < EB / 64; q++) {
      tr += rr[q];
      ti += ri[q];
    }
    partial[(size_t)i * nb + blockIdx.x] = make_double2(tr, ti);
  }
}

// c[i] = sum_b partial[i * nb + b], one thread per vector, fixed order
__global__ void k_eig_sum(const double2 *__restrict__ partial, int nb, int k,
                          double2 *__restrict__ c) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= k) return;
  double2 s = make_double2(0, 0);
  for (int b = 0; b < nb; b++) {
    s.x += partial[(size_t)i * nb + b].x;
    s.y += partial[(size_t)i * nb + b].y;
  }
  c[i] = s;
}

// w -= sum_i c[i] V_i
__global__ void __launch_bounds__(EB)
k_eig_sub(const double2 *__restrict__ V, const double2 *__restrict__ c, int k,
          double2 *__restrict__ w, size_t n) {
  const size_t e = (size_t)blockIdx.x * EB + threadIdx.x;
  if (e >= n) return;
  double2 x = w[e];
  for (int i = 0; i < k; i++) {
    const double2 ci = c[i], v = V[(size_t)i * n + e];
    x.x -= ci.x * v.x - ci.y * v.y;
    x.y -= ci.x * v.y + ci.y * v.x;
  }
  w[e] = x;
}

// y = s x
__global__ void k_eig_scale(double2 *__restrict__ y, const double2 *__restrict__ x, double s,
                            size_t n) {
  const size_t e = (size_t)blockIdx.x * EB + threadIdx.x;
  if (e < n) y[e] = make_double2(s * x[e].x, s * x[e].y);
}

// number of eigenvalues of the tridiagonal matrix (a; b[1..k-1]) below x
static int sturm_count(const std::vector<double> &a, const std::vector<double> &b, int k,
                       double x) {
  int cnt = 0;
  double q = a[0] - x;
  if (q < 0) cnt++;
  for (int i = 1; i < k; i++) {
    if (q == 0) q = 1e-300;
    q = a[i] - x - b[i] * b[i] / q;
    if (q < 0) cnt++;
  }
  return cnt;
}
// m-th smallest eigenvalue (0-based) of the k x k tridiagonal matrix
static double tridiag_eig(const std::vector<double> &a, const std::vector<double> &b, int k,
                          int m) {
  double lo = a[0], hi = a[0];
  for (int i = 0; i < k; i++) {
    const double r = (i > 0 ? fabs(b[i]) : 0) + (i + 1 < k ? fabs(b[i + 1]) : 0);
    lo = fmin(lo, a[i] - r);
    hi = fmax(hi, a[i] + r);
  }
  for (int it = 0; it < 200 && hi - lo > 1e-15 * fmax(fabs(lo), fabs(hi)); it++) {
    const double mid = 0.5 * (lo + hi);
    if (sturm_count(a, b, k, mid) > m) hi = mid;
    else lo = mid;
  }
  return 0.5 * (lo + hi);
}

extern "C" int susy_eig_dsq(susy_ctx *c, int nlow, int nhigh, int maxiter, double tol,
                            int seed, double *low, double *high, int *iters) {
  if (!c) return SUSY_E_STATE;
  if (nlow < 0 || nhigh < 0 || nlow + nhigh < 1 || maxiter < nlow + nhigh + 1 || tol <= 0) {
    set_error("susy_eig_dsq: need nlow + nhigh >= 1, maxiter > nlow + nhigh, tol > 0");
    return SUSY_E_ARG;
  }
  const size_t n = tf_elems(c);
  double2 *V = nullptr, *w = nullptr, *cbuf = nullptr, *part = nullptr;
  const int nb = (int)((n + EB - 1) / EB < 256 ? (n + EB - 1) / EB : 256);
  if (hipMalloc(&V, sizeof(double2) * n * (size_t)(maxiter + 1)) != hipSuccess) {
    set_error("susy_eig_dsq: no memory for %d Lanczos vectors", maxiter + 1);
    return SUSY_E_NOMEM;
  }
  int rc = 0;
  auto chk = [&](hipError_t e) {
    if (e != hipSuccess && !rc) {
      set_error("HIP error %s", hipGetErrorString(e));
      rc = SUSY_E_HIP;
    }
  };
  chk(hipMalloc(&w, sizeof(double2) * n));
  chk(hipMalloc(&cbuf, sizeof(double2) * (maxiter + 1)));
  chk(hipMalloc(&part, sizeof(double2) * (size_t)nb * (maxiter + 1)));
  std::vector<double> a, b(1, 0.0), last;
  int done_iters = maxiter;
  if (!rc) {
    // seeded Gaussian start vector, normalised
    std::vector<double2> h(n);
    std::mt19937_64 gen((unsigned long long)seed);
    std::normal_distribution<double> nd(0.0, 1.0);
    double nrm = 0;
    for (size_t e = 0; e < n; e++) {
      h[e] = make_double2(nd(gen), nd(gen));
      nrm += h[e].x * h[e].x + h[e].y * h[e].y;
    }
    const double s = 1.0 / sqrt(nrm);
    for (size_t e = 0; e < n; e++) h[e] = make_double2(s * h[e].x, s * h[e].y);
    chk(hipMemcpy(V, h.data(), sizeof(double2) * n, hipMemcpyHostToDevice));
  }
  const dim3 ge((unsigned)((n + EB - 1) / EB)), be(EB);
  const double f2 = c->p.fmass > IMAG_TOL ? c->p.fmass * c->p.fmass : 0.0;
  for (int j = 0; j < maxiter && !rc; j++) {
    double2 *vj = V + (size_t)j * n;
    int nbd = 0;
    // w = DSq v_j = (D^dag D + fmass^2) v_j, as susy_dsq_dev
    rc = dsq_pair64(c, vj, c->d_tmp, w, f2, c->d_part, &nbd);
    if (rc) break;
    double alpha;
    if ((rc = tf_dot_re(c, vj, w, &alpha))) break;
    a.push_back(alpha);
    if ((rc = tf_axpby(c, w, 1.0, vj, -alpha))) break;
    if (j > 0 && (rc = tf_axpby(c, w, 1.0, vj - n, -b[j]))) break;
    for (int pass = 0; pass < 2; pass++) {
      hipLaunchKernelGGL(k_eig_dots, dim3(nb, j + 1), be, 0, c->s, V, w, n, part);
      hipLaunchKernelGGL(k_eig_sum, dim3((j + 1 + 63) / 64), dim3(64), 0, c->s, part, nb, j + 1,
                         cbuf);
      hipLaunchKernelGGL(k_eig_sub, ge, be, 0, c->s, V, cbuf, j + 1, w, n);
      chk(hipGetLastError());
    }
    double nn;
    if ((rc = tf_dot_re(c, w, w, &nn))) break;
    const double beta = sqrt(nn);
    b.push_back(beta);
    hipLaunchKernelGGL(k_eig_scale, ge, be, 0, c->s, V + (size_t)(j + 1) * n, w,
                       beta > 0 ? 1.0 / beta : 0.0, n);
    chk(hipGetLastError());
    const int k = j + 1;
    if ((k % 20 == 0 && k >= nlow + nhigh) || k == maxiter || beta == 0) {
      std::vector<double> cur;
      for (int m = 0; m < nlow && m < k; m++) cur.push_back(tridiag_eig(a, b, k, m));
      for (int m = 0; m < nhigh && m < k; m++) cur.push_back(tridiag_eig(a, b, k, k - 1 - m));
      bool conv = last.size() == cur.size();
      for (size_t q = 0; conv && q < cur.size(); q++)
        conv = fabs(cur[q] - last[q]) <= tol * fabs(cur[q]);
      last = cur;
      if (conv || beta == 0 || k == maxiter) {
        done_iters = k;
        break;
      }
    }
  }
  if (!rc) {
    for (int m = 0; m < nlow; m++) low[m] = m < (int)last.size() ? last[m] : NAN;
    for (int m = 0; m < nhigh; m++) high[m] = nlow + m < (int)last.size() ? last[nlow + m] : NAN;
    if (iters) *iters = done_iters;
  }
  (void)hipStreamSynchronize(c->s);
  (void)hipFree(V);
  if (w) (void)hipFree(w);
  if (cbuf) (void)hipFree(cbuf);
  if (part) (void)hipFree(part);
  return rc;
}
