// Per-trajectory measurements printed by the reference's driver after every
// trajectory (control.c:85-138):
//   FLINK   Tr(U_a U_a^dag) / N per direction, their average and width, and
//           the determinant of the traceless part of U_a U_a^dag
//           (link_trace.c:7-55);
//   GMES    the Polyakov loop in the t direction (ploop.c:8-126);
//   LINES   the Wilson lines in all five link directions (ploop.c), and
//   LINES_POLAR  the same with every link replaced by the unitary factor of
//           its polar decomposition (ploop.c:16-24, unit.c:53-110);
//   DET     moments of the plaquette determinants (determinant.c:93-119);
//   WIDTHS  widths of Re Tr and of Re, Im det of the plaquettes (widths.c);
//   UUBAR_EIG / POLAR_EIG  eigenvalue statistics of the traceless part of
//           U U^dag and of log (U U^dag)^(1/2) (scalar_eig.c:14-123).
// The reference diagonalises with LAPACK zheev; here every link's Hermitian
// eigenproblem is solved in its thread by cyclic complex Jacobi rotations.
#include <math.h>

#include <vector>

#include "dev_mat.h"

#define MB 256
#define NMEAS 12

__device__ __forceinline__ void mcopy(Mat d, const Mat m) {
#pragma unroll
  for (int i = 0; i < NCOL; i++)
#pragma unroll
    for (int j = 0; j < NCOL; j++) d[i][j] = m[i][j];
}

// determinant by LU with partial pivoting on |re| + |im|
__device__ __forceinline__ double2 mdet(const Mat in) {
  // eliminate on a local copy: the row swaps below then stay in registers
  // (on the caller's array they compile to pointer selects, and from U(5) on
  // the matrix is kept in scratch memory)
  Mat m;
#pragma unroll
  for (int i = 0; i < NCOL; i++)
#pragma unroll
    for (int j = 0; j < NCOL; j++) m[i][j] = in[i][j];
  double2 d = make_double2(1, 0);
#pragma unroll
  for (int k = 0; k < NCOL; k++) {
    int p = k;
    double best = cabs1(m[k][k]);
#pragma unroll
    for (int i = k + 1; i < NCOL; i++) {
      const double v = cabs1(m[i][k]);
      if (v > best) { best = v; p = i; }
    }
    if (p != k) {
#pragma unroll
      for (int j = 0; j < NCOL; j++) {
        const double2 tk = m[k][j];
#pragma unroll
        for (int i = k + 1; i < NCOL; i++)
          if (i == p) { m[k][j] = m[i][j]; m[i][j] = tk; }
      }
      d = make_double2(-d.x, -d.y);
    }
    d = cmul(d, m[k][k]);
    if (m[k][k].x == 0 && m[k][k].y == 0) return make_double2(0, 0);
    const double2 inv = cdiv(make_double2(1, 0), m[k][k]);
#pragma unroll
    for (int i = k + 1; i < NCOL; i++) {
      const double2 f = cmul(m[i][k], inv);
#pragma unroll
      for (int j = k + 1; j < NCOL; j++) {
        const double2 t = cmul(f, m[k][j]);
        m[i][j].x -= t.x;
        m[i][j].y -= t.y;
      }
    }
  }
  return d;
}

// Eigen-decomposition H = W diag(lam) W^dag of a Hermitian matrix by cyclic
// Jacobi sweeps; each rotation J = diag(1, conj(e)) R (e = H_pq / |H_pq|, R
// the real Jacobi rotation of the phase-rotated 2x2 block) zeroes H_pq:
// H <- J^dag H J, W <- W J.  At most 16 sweeps (quadratic convergence: 3-5
// sweeps for N <= 4 to rounding level); lam ascending, the columns of W
// permuted with them.
__device__ void heig(Mat H, Mat W, double (&lam)[NCOL]) {
#pragma unroll
  for (int i = 0; i < NCOL; i++)
#pragma unroll
    for (int j = 0; j < NCOL; j++) W[i][j] = make_double2(i == j ? 1.0 : 0.0, 0.0);
  for (int sweep = 0; sweep < 16; sweep++) {
    double off = 0, dia = 0;
#pragma unroll
    for (int p = 0; p < NCOL; p++) {
      dia += H[p][p].x * H[p][p].x;
#pragma unroll
      for (int q = p + 1; q < NCOL; q++) off += H[p][q].x * H[p][q].x + H[p][q].y * H[p][q].y;
    }
    if (!(off > 1e-32 * dia)) break;
#pragma unroll
    for (int p = 0; p < NCOL - 1; p++) {
#pragma unroll
      for (int q = p + 1; q < NCOL; q++) {
        const double2 h = H[p][q];
        const double m = sqrt(h.x * h.x + h.y * h.y);
        if (m == 0) continue;
        const double2 e = make_double2(h.x / m, h.y / m), eb = cconjd(e);
        const double tau = (H[q][q].x - H[p][p].x) / (2 * m);
        const double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1 + tau * tau));
        const double cs = 1 / sqrt(1 + t * t), sn = t * cs;
#pragma unroll
        for (int r = 0; r < NCOL; r++) {   // H <- H J, W <- W J
          const double2 x = H[r][p], y = cmul(eb, H[r][q]);
          H[r][p] = make_double2(cs * x.x - sn * y.x, cs * x.y - sn * y.y);
          H[r][q] = make_double2(sn * x.x + cs * y.x, sn * x.y + cs * y.y);
          const double2 u = W[r][p], v = cmul(eb, W[r][q]);
          W[r][p] = make_double2(cs * u.x - sn * v.x, cs * u.y - sn * v.y);
          W[r][q] = make_double2(sn * u.x + cs * v.x, sn * u.y + cs * v.y);
        }
#pragma unroll
        for (int k = 0; k < NCOL; k++) {   // H <- J^dag H
          const double2 x = H[p][k], y = cmul(e, H[q][k]);
          H[p][k] = make_double2(cs * x.x - sn * y.x, cs * x.y - sn * y.y);
          H[q][k] = make_double2(sn * x.x + cs * y.x, sn * x.y + cs * y.y);
        }
        H[p][q] = H[q][p] = make_double2(0, 0);
        H[p][p].y = H[q][q].y = 0;
      }
    }
  }
#pragma unroll
  for (int i = 0; i < NCOL; i++) lam[i] = H[i][i].x;
  // insertion sort, ascending (NCOL <= 4)
#pragma unroll
  for (int i = 1; i < NCOL; i++) {
#pragma unroll
    for (int j = i; j > 0; j--) {
      if (lam[j] < lam[j - 1]) {
        const double tl = lam[j];
        lam[j] = lam[j - 1];
        lam[j - 1] = tl;
#pragma unroll
        for (int r = 0; r < NCOL; r++) {
          const double2 tw = W[r][j];
          W[r][j] = W[r][j - 1];
          W[r][j - 1] = tw;
        }
      }
    }
  }
}

// u = (m m^dag)^(-1/2) m, the unitary factor of m = P u (unit.c:53-110)
__device__ void polar_unitary(const Mat m, Mat u) {
  Mat H, W, T;
  double lam[NCOL];
  mzero(H);
  mmadd<false, true>(H, 1.0, m, m);
  heig(H, W, lam);
  // T = W diag(lam^-1/2) W^dag
#pragma unroll
  for (int i = 0; i < NCOL; i++)
#pragma unroll
    for (int j = 0; j < NCOL; j++) {
      double2 acc = make_double2(0, 0);
#pragma unroll
      for (int k = 0; k < NCOL; k++) {
        const double2 t = cmul(W[i][k], cconjd(W[j][k]));
        const double f = 1.0 / sqrt(lam[k]);
        acc.x += f * t.x;
        acc.y += f * t.y;
      }
      T[i][j] = acc;
    }
  mzero(u);
  mmadd<false, false>(u, 1.0, T, m);
}

// per-block sums of NMEAS values in a fixed order (slot q at
// partial[q * nblk + block])
__device__ __forceinline__ void block_sums(double (&v)[NMEAS], double *partial, int nblk) {
  __shared__ double red[NMEAS][MB / 64];
#pragma unroll
  for (int q = 0; q < NMEAS; q++) {
    double x = v[q];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    if ((threadIdx.x & 63) == 0) red[q][threadIdx.x >> 6] = x;
  }
  __syncthreads();
  if (threadIdx.x < NMEAS) {
    double t = 0;
    for (int w = 0; w < MB / 64; w++) t += red[threadIdx.x][w];
    partial[threadIdx.x * nblk + blockIdx.x] = t;
  }
}

// slots 0..4: Tr(U U^dag)/N per direction; 5: sum of its squares;
// 6..10: Re det of the traceless part per direction; 11: sum of squares
__global__ void __launch_bounds__(MB)
k_meas_links(const double2 *__restrict__ U, double *__restrict__ partial, int nblk, int V) {
  const int s = blockIdx.x * MB + threadIdx.x;
  double v[NMEAS];
#pragma unroll
  for (int q = 0; q < NMEAS; q++) v[q] = 0;
  if (s < V) {
    const size_t M = (size_t)NN * V;
    for (int a = 0; a < NUMLINK; a++) {
      Mat u, t;
      mload(u, U + a * M + s, V);
      mzero(t);
      mmadd<false, true>(t, 1.0, u, u);
      double tr = 0;
#pragma unroll
      for (int i = 0; i < NCOL; i++) tr += t[i][i].x;
      tr /= NCOL;
      v[a] = tr;
      v[5] += tr * tr;
#pragma unroll
      for (int i = 0; i < NCOL; i++) t[i][i].x -= tr;
      const double2 d = mdet(t);
      v[6 + a] = d.x;
      v[11] += d.x * d.x;
    }
  }
  block_sums(v, partial, nblk);
}

// ordered product U_a(x) U_a(x+a) ... U_a(x+(len-1)a) from every start site;
// slots 0, 1, 2: Re, Im and modulus of its trace
__global__ void __launch_bounds__(MB)
k_wilson_line(const double2 *__restrict__ U, const int *__restrict__ nbr,
              const int *__restrict__ start, int nstart, int a, int len, int project,
              double *__restrict__ partial, int nblk, int V) {
  const int k = blockIdx.x * MB + threadIdx.x;
  double v[NMEAS];
#pragma unroll
  for (int q = 0; q < NMEAS; q++) v[q] = 0;
  if (k < nstart) {
    const size_t M = (size_t)NN * V;
    int x = start[k];
    Mat p, u, t;
    mload(p, U + a * M + x, V);
    if (project) {
      polar_unitary(p, t);
      mcopy(p, t);
    }
    for (int step = 1; step < len; step++) {
      x = nbr[g_fwd(a) * V + x];
      mload(u, U + a * M + x, V);
      if (project) {
        polar_unitary(u, t);
        mcopy(u, t);
      }
      mzero(t);
      mmadd<false, false>(t, 1.0, p, u);
      mcopy(p, t);
    }
    double2 tr = make_double2(0, 0);
#pragma unroll
    for (int i = 0; i < NCOL; i++) {
      tr.x += p[i][i].x;
      tr.y += p[i][i].y;
    }
    v[0] = tr.x;
    v[1] = tr.y;
    v[2] = sqrt(tr.x * tr.x + tr.y * tr.y);
  }
  block_sums(v, partial, nblk);
}

// per-block sums of NS values and minima / maxima of NM values: slot q at
// partial[q * nblk + block], sums first, then minima, then maxima
template <int NS, int NM>
__device__ __forceinline__ void block_stats(double (&sm)[NS], double (&mn)[NM],
                                            double (&mx)[NM], double *partial, int nblk) {
  __shared__ double red[NS + 2 * NM][MB / 64];
  const int w = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NS; q++) {
    double x = sm[q];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    if ((threadIdx.x & 63) == 0) red[q][w] = x;
  }
#pragma unroll
  for (int q = 0; q < NM; q++) {
    double x = mn[q], y = mx[q];
    for (int off = 32; off > 0; off >>= 1) {
      x = fmin(x, __shfl_xor(x, off, 64));
      y = fmax(y, __shfl_xor(y, off, 64));
    }
    if ((threadIdx.x & 63) == 0) {
      red[NS + q][w] = x;
      red[NS + NM + q][w] = y;
    }
  }
  __syncthreads();
  const int q = threadIdx.x;
  if (q < NS + 2 * NM) {
    double t = red[q][0];
    for (int k = 1; k < MB / 64; k++) {
      const double v = red[q][k];
      t = q < NS ? t + v : q < NS + NM ? fmin(t, v) : fmax(t, v);
    }
    partial[q * nblk + blockIdx.x] = t;
  }
}

// Eigenvalues of U U^dag for every link: sums of the traceless-part
// eigenvalues and of their squares (slots j, NCOL + j), of the polar
// log-eigenvalues 0.5 log lam_j and their squares (2 NCOL + j, 3 NCOL + j);
// minima / maxima of both per index j.  The reference's zheev on the
// traceless part returns lam_j - Tr/N, and on log P (P = (U U^dag)^(1/2),
// matrix_log) 0.5 log lam_j, both ascending in j as lam_j is.
#define EIG_NS (4 * NCOL)
#define EIG_NM (2 * NCOL)
__global__ void __launch_bounds__(MB)
k_meas_eig(const double2 *__restrict__ U, double *__restrict__ partial, int nblk, int V) {
  const int s = blockIdx.x * MB + threadIdx.x;
  double sm[EIG_NS], mn[EIG_NM], mx[EIG_NM];
#pragma unroll
  for (int q = 0; q < EIG_NS; q++) sm[q] = 0;
#pragma unroll
  for (int q = 0; q < EIG_NM; q++) {
    mn[q] = INFINITY;
    mx[q] = -INFINITY;
  }
  if (s < V) {
    const size_t M = (size_t)NN * V;
    for (int a = 0; a < NUMLINK; a++) {
      Mat u, H, W;
      double lam[NCOL];
      mload(u, U + a * M + s, V);
      mzero(H);
      mmadd<false, true>(H, 1.0, u, u);
      double tr = 0;
#pragma unroll
      for (int i = 0; i < NCOL; i++) tr += H[i][i].x;
      tr /= NCOL;
      heig(H, W, lam);
#pragma unroll
      for (int j = 0; j < NCOL; j++) {
        const double e = lam[j] - tr, l = 0.5 * log(lam[j]);
        sm[j] += e;
        sm[NCOL + j] += e * e;
        sm[2 * NCOL + j] += l;
        sm[3 * NCOL + j] += l * l;
        mn[j] = fmin(mn[j], e);
        mx[j] = fmax(mx[j], e);
        mn[NCOL + j] = fmin(mn[NCOL + j], l);
        mx[NCOL + j] = fmax(mx[NCOL + j], l);
      }
    }
  }
  block_stats<EIG_NS, EIG_NM>(sm, mn, mx, partial, nblk);
}

// Plaquette moments.  Slots 0..5: sums of Re Tr P, (Re Tr P)^2, Re det P,
// (Re det P)^2, Im det P, (Im det P)^2 over the 10 a < b,
// P_ab(x) = U_b(x) U_a(x+b) U_b(x+a)^dag U_a(x)^dag (widths.c);
// slots 6..10: sums of Re, Im, Re^2, Im^2 of plaqdet_ab(x) and of
// |plaqdet_ab(x) - 1|^2 over the 10 a > b (determinant.c:93-119).
__global__ void __launch_bounds__(MB)
k_meas_plaq(const double2 *__restrict__ U, const double2 *__restrict__ pd,
            const int *__restrict__ nbr, double *__restrict__ partial, int nblk, int V) {
  const int s = blockIdx.x * MB + threadIdx.x;
  double v[NMEAS];
#pragma unroll
  for (int q = 0; q < NMEAS; q++) v[q] = 0;
  if (s < V) {
    const size_t M = (size_t)NN * V;
    for (int a = 0; a < NUMLINK; a++) {
      const int xa = nbr[g_fwd(a) * V + s];
      for (int b = a + 1; b < NUMLINK; b++) {
        const int xb = nbr[g_fwd(b) * V + s];
        Mat m, t1, t2;
        mload(t2, U + b * M + s, V);
        mload(m, U + a * M + xb, V);
        mzero(t1);
        mmadd<false, false>(t1, 1.0, t2, m);       // U_b(x) U_a(x+b)
        mload(m, U + b * M + xa, V);
        mzero(t2);
        mmadd<false, true>(t2, 1.0, t1, m);        // ... U_b(x+a)^dag
        mload(m, U + a * M + s, V);
        mzero(t1);
        mmadd<false, true>(t1, 1.0, t2, m);        // ... U_a(x)^dag
        double tr = 0;
#pragma unroll
        for (int i = 0; i < NCOL; i++) tr += t1[i][i].x;
        const double2 d = mdet(t1);
        v[0] += tr;
        v[1] += tr * tr;
        v[2] += d.x;
        v[3] += d.x * d.x;
        v[4] += d.y;
        v[5] += d.y * d.y;
        const double2 p = pd[(b * NUMLINK + a) * V + s];   // plaqdet_ba, b > a
        v[6] += p.x;
        v[7] += p.y;
        v[8] += p.x * p.x;
        v[9] += p.y * p.y;
        v[10] += (p.x - 1.0) * (p.x - 1.0) + p.y * p.y;
      }
    }
  }
  block_sums(v, partial, nblk);
}

// downloads nslot x nb block partials of the last measurement kernel
static int read_partials(susy_ctx *c, std::vector<double> &h, int nslot, int nb) {
  if ((size_t)nb * nslot > (size_t)c->npart * 2) {
    set_error("measurement: partial buffer too small");
    return SUSY_E_STATE;
  }
  HIPCHK(hipGetLastError());
  h.assign((size_t)nslot * nb, 0.0);
  HIPCHK(hipMemcpyAsync(h.data(), c->d_part, sizeof(double) * h.size(), hipMemcpyDeviceToHost,
                        c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  return 0;
}

// scalar_eig (scalar_eig.c:14-123): per eigenvalue index j (ascending), the
// average, width, minimum and maximum over all links of the eigenvalues of
// the traceless part of U U^dag (project = 0) or of log (U U^dag)^(1/2)
// (project = 1)
extern "C" int susy_scalar_eig_dev(susy_ctx *c, int project, double *ave, double *width,
                                   double *mn, double *mx) {
  if (!c) return SUSY_E_STATE;
  if (project != 0 && project != 1) {
    set_error("scalar_eig: project must be 0 or 1");
    return SUSY_E_ARG;
  }
  const int V = c->V, nb = (V + MB - 1) / MB;
  if ((size_t)nb * (EIG_NS + 2 * EIG_NM) > (size_t)c->npart * 2) {
    set_error("scalar_eig: partial buffer too small");
    return SUSY_E_STATE;
  }
  hipLaunchKernelGGL(k_meas_eig, dim3(nb), dim3(MB), 0, c->s, c->d_links, c->d_part, nb, V);
  std::vector<double> h;
  int r = read_partials(c, h, EIG_NS + 2 * EIG_NM, nb);
  if (r) return r;
  const double norm = (double)NUMLINK * V;
  const int o = project ? 2 * NCOL : 0, om = project ? NCOL : 0;
  for (int j = 0; j < NCOL; j++) {
    double s1 = 0, s2 = 0, lo = INFINITY, hi = -INFINITY;
    for (int b = 0; b < nb; b++) {
      s1 += h[(size_t)(o + j) * nb + b];
      s2 += h[(size_t)(o + NCOL + j) * nb + b];
      lo = fmin(lo, h[(size_t)(EIG_NS + om + j) * nb + b]);
      hi = fmax(hi, h[(size_t)(EIG_NS + EIG_NM + om + j) * nb + b]);
    }
    ave[j] = s1 / norm;
    width[j] = sqrt(s2 / norm - ave[j] * ave[j]);
    mn[j] = lo;
    mx[j] = hi;
  }
  return 0;
}

// measure_det (determinant.c:93-119): det5 = <Re>, <Im>, <Re^2>, <Im^2> of
// plaqdet_ab over the 10 a > b and all sites, and <|plaqdet_ab - 1|^2>;
// widths (widths.c): w3 = widths sqrt(<x^2> - <x>^2) of Re Tr P_ab,
// Re det P_ab and Im det P_ab over the 10 a < b and all sites
extern "C" int susy_plaq_moments_dev(susy_ctx *c, double *det5, double *w3) {
  if (!c) return SUSY_E_STATE;
  const int V = c->V, nb = (V + MB - 1) / MB;
  if ((size_t)nb * NMEAS > (size_t)c->npart * 2) {
    set_error("plaquette moments: partial buffer too small");
    return SUSY_E_STATE;
  }
  hipLaunchKernelGGL(k_meas_plaq, dim3(nb), dim3(MB), 0, c->s, c->d_links, c->d_pdet, c->d_nbr,
                     c->d_part, nb, V);
  std::vector<double> h;
  int r = read_partials(c, h, NMEAS, nb);
  if (r) return r;
  double sc[NMEAS];
  for (int q = 0; q < NMEAS; q++) {
    double t = 0;
    for (int b = 0; b < nb; b++) t += h[(size_t)q * nb + b];
    sc[q] = t / ((double)V * NPLAQ);
  }
  for (int k = 0; k < 5; k++) det5[k] = sc[6 + k];
  w3[0] = sqrt(sc[1] - sc[0] * sc[0]);
  w3[1] = sqrt(sc[3] - sc[2] * sc[2]);
  w3[2] = sqrt(sc[5] - sc[4] * sc[4]);
  return 0;
}

extern "C" int susy_link_trace_dev(susy_ctx *c, double *linktr, double *ave, double *width,
                                   double *link_det, double *det_ave, double *det_width) {
  if (!c) return SUSY_E_STATE;
  const int V = c->V, nb = (V + MB - 1) / MB;
  if ((size_t)nb * NMEAS > (size_t)c->npart * 2) {
    set_error("link_trace: partial buffer too small");
    return SUSY_E_STATE;
  }
  hipLaunchKernelGGL(k_meas_links, dim3(nb), dim3(MB), 0, c->s, c->d_links, c->d_part, nb, V);
  HIPCHK(hipGetLastError());
  for (int q = 0; q < NMEAS; q++) {
    int r = reduce_partials(c, c->d_part + (size_t)q * nb, nb, q);
    if (r) return r;
  }
  double sc[NMEAS];
  int r = read_scalars(c, sc, NMEAS);
  if (r) return r;
  double av = 0, dav = 0;
  for (int a = 0; a < NUMLINK; a++) {
    linktr[a] = sc[a] / V;
    link_det[a] = sc[6 + a] / V;
    av += linktr[a];
    dav += link_det[a];
  }
  av /= NUMLINK;
  dav /= NUMLINK;
  *ave = av;
  *width = sqrt(sc[5] / ((double)V * NUMLINK) - av * av);
  *det_ave = dav;
  *det_width = sqrt(sc[11] / ((double)V * NUMLINK) - dav * dav);
  return 0;
}

// Polyakov / Wilson line in link direction a (0..3: x, y, z, t; 4: the fifth
// A4* link, stepped nt times from the t = 0 hyperplane), averaged over its
// starting hyperplane; mod = average modulus of the traces; project = 1 uses
// the polar-projected unitary links (ploop.c:16-24)
extern "C" int susy_wilson_line_dev(susy_ctx *c, int a, int project, double *re, double *im,
                                    double *mod) {
  if (!c) return SUSY_E_STATE;
  if (a < 0 || a >= NUMLINK || (project != 0 && project != 1)) {
    set_error("wilson line: direction %d outside 0..4 or project %d not 0/1", a, project);
    return SUSY_E_ARG;
  }
  const Geometry &g = c->geo;
  const int V = c->V;
  const int len = a == 0 ? g.nx : a == 1 ? g.ny : a == 2 ? g.nz : g.nt;
  std::vector<int> start;
  for (int i = 0; i < V; i++) {
    const int j = a == 0 ? g.x[i] : a == 1 ? g.y[i] : a == 2 ? g.z[i] : g.t[i];
    if (j == 0) start.push_back(i);
  }
  const int ns = (int)start.size(), nb = (ns + MB - 1) / MB;
  int *d_start = nullptr;
  HIPCHK(hipMalloc(&d_start, sizeof(int) * ns));
  int r = 0;
  if (hipMemcpy(d_start, start.data(), sizeof(int) * ns, hipMemcpyHostToDevice) != hipSuccess) {
    set_error("wilson line: upload failed");
    r = SUSY_E_HIP;
  }
  if (!r) {
    hipLaunchKernelGGL(k_wilson_line, dim3(nb), dim3(MB), 0, c->s, c->d_links, c->d_nbr,
                       d_start, ns, a, len, project, c->d_part, nb, V);
    if (hipGetLastError() != hipSuccess) {
      set_error("wilson line: launch failed");
      r = SUSY_E_HIP;
    }
  }
  for (int q = 0; q < 3 && !r; q++) r = reduce_partials(c, c->d_part + (size_t)q * nb, nb, q);
  double sc[3];
  if (!r) r = read_scalars(c, sc, 3);
  (void)hipFree(d_start);
  if (r) return r;
  const double norm = (double)len / V;
  *re = sc[0] * norm;
  *im = sc[1] * norm;
  *mod = sc[2] * norm;
  return 0;
}
