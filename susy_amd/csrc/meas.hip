// Per-trajectory measurements printed by the reference's driver after every
// trajectory (control.c:85-120):
//   FLINK   Tr(U_a U_a^dag) / N per direction, their average and width, and
//           the determinant of the traceless part of U_a U_a^dag
//           (link_trace.c:7-55);
//   GMES    the Polyakov loop in the t direction (ploop.c:8-126);
//   LINES   the Wilson lines in all five link directions (ploop.c).
#include <math.h>

#include <vector>

#include "dev_mat.h"

#define MB 256
#define NMEAS 12

// determinant by LU with partial pivoting on |re| + |im|
__device__ __forceinline__ double2 mdet(Mat m) {
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
              const int *__restrict__ start, int nstart, int a, int len,
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
    for (int step = 1; step < len; step++) {
      x = nbr[g_fwd(a) * V + x];
      mload(u, U + a * M + x, V);
      mzero(t);
      mmadd<false, false>(t, 1.0, p, u);
#pragma unroll
      for (int i = 0; i < NCOL; i++)
#pragma unroll
        for (int j = 0; j < NCOL; j++) p[i][j] = t[i][j];
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
// starting hyperplane; mod = average modulus of the traces
extern "C" int susy_wilson_line_dev(susy_ctx *c, int a, double *re, double *im, double *mod) {
  if (!c) return SUSY_E_STATE;
  if (a < 0 || a >= NUMLINK) {
    set_error("wilson line: direction %d outside 0..4", a);
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
                       d_start, ns, a, len, c->d_part, nb, V);
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
