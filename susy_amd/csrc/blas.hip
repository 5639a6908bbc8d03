// Twist_Fermion BLAS-1 (reference: 4d_Q16/susy/library_util.c:53-168),
// deterministic reductions, and the AoS <-> SoA transposes used at the
// host boundary.
#include <string.h>

#include "susy_internal.h"

#define BB 256

size_t tf_elems(const susy_ctx *c) { return (size_t)NTF * NN * c->V; }

// One block sums n partials in a fixed order with Neumaier compensation and
// leaves the result in d_scal[slot].
#define RB 1024
__global__ void __launch_bounds__(RB)
k_reduce(const double *__restrict__ part, int n, double *__restrict__ out) {
  double s = 0, comp = 0;
  for (int i = threadIdx.x; i < n; i += RB) {
    double v = part[i], t = s + v;
    comp += (fabs(s) >= fabs(v)) ? (s - t) + v : (v - t) + s;
    s = t;
  }
  __shared__ double rs[RB], rc[RB];
  rs[threadIdx.x] = s;
  rc[threadIdx.x] = comp;
  __syncthreads();
  for (int w = RB / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      double a = rs[threadIdx.x], b = rs[threadIdx.x + w], t = a + b;
      double e = (fabs(a) >= fabs(b)) ? (a - t) + b : (b - t) + a;
      rs[threadIdx.x] = t;
      rc[threadIdx.x] += rc[threadIdx.x + w] + e;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = rs[0] + rc[0];
}

int reduce_partials(susy_ctx *c, const double *part, int n, int slot) {
  hipLaunchKernelGGL(k_reduce, dim3(1), dim3(RB), 0, c->s, part, n, c->d_scal + slot);
  HIPCHK(hipGetLastError());
  return 0;
}

// Reads back d_scal[0..n) through the pinned staging buffer (a pageable
// destination would make the runtime stage the copy itself).
int read_scalars(susy_ctx *c, double *out, int n) {
  if (n < 1 || n > 64) {
    set_error("read_scalars: %d scalars requested (1..64)", n);
    return SUSY_E_ARG;
  }
  HIPCHK(hipMemcpyAsync(c->h_scal, c->d_scal, sizeof(double) * n, hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  memcpy(out, c->h_scal, sizeof(double) * n);
  return 0;
}

__device__ __forceinline__ void block_sum(double v, double *partial) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  __shared__ double red[BB / 64];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0;
    for (int w = 0; w < BB / 64; w++) t += red[w];
    partial[blockIdx.x] = t;
  }
}

// partial of Re sum conj(a) b over the field (ReTr(a^dag b) per matrix)
__global__ void __launch_bounds__(BB)
k_dot_re(const double2 *__restrict__ a, const double2 *__restrict__ b, size_t n,
         double *__restrict__ partial) {
  double s = 0;
  for (size_t i = (size_t)blockIdx.x * BB + threadIdx.x; i < n; i += (size_t)gridDim.x * BB) {
    double2 x = a[i], y = b[i];
    s += x.x * y.x + x.y * y.y;
  }
  block_sum(s, partial);
}

static int red_blocks(size_t n) {
  size_t nb = (n + BB - 1) / BB;
  return (int)(nb < 2048 ? nb : 2048);
}

int tf_dot_re(susy_ctx *c, const double2 *a, const double2 *b, double *out) {
  size_t n = tf_elems(c);
  int nb = red_blocks(n);
  hipLaunchKernelGGL(k_dot_re, dim3(nb), dim3(BB), 0, c->s, a, b, n, c->d_part);
  HIPCHK(hipGetLastError());
  int r = reduce_partials(c, c->d_part, nb, 0);
  if (r) return r;
  return read_scalars(c, out, 1);
}

// y = a y + b x
__global__ void k_axpby(double2 *__restrict__ y, double a, const double2 *__restrict__ x,
                        double b, size_t n) {
  size_t i = (size_t)blockIdx.x * BB + threadIdx.x;
  if (i < n) {
    double2 u = y[i], v = x[i];
    u.x = a * u.x + b * v.x;
    u.y = a * u.y + b * v.y;
    y[i] = u;
  }
}
int tf_axpby(susy_ctx *c, double2 *y, double a, const double2 *x, double b) {
  size_t n = tf_elems(c);
  hipLaunchKernelGGL(k_axpby, dim3((n + BB - 1) / BB), dim3(BB), 0, c->s, y, a, x, b, n);
  HIPCHK(hipGetLastError());
  return 0;
}
// Plain streaming copy of one Twist_Fermion field (calibration of the
// memory-traffic counters: it reads and writes exactly 16 NCOL^2 V complex).
__global__ void k_stream_copy(double2 *__restrict__ y, const double2 *__restrict__ x, size_t n) {
  size_t i = (size_t)blockIdx.x * BB + threadIdx.x;
  if (i < n) y[i] = x[i];
}
int tf_stream_copy(susy_ctx *c, double2 *dst, const double2 *src) {
  size_t n = tf_elems(c);
  hipLaunchKernelGGL(k_stream_copy, dim3((n + BB - 1) / BB), dim3(BB), 0, c->s, dst, src, n);
  HIPCHK(hipGetLastError());
  return 0;
}
int tf_copy(susy_ctx *c, double2 *dst, const double2 *src) {
  HIPCHK(hipMemcpyAsync(dst, src, sizeof(double2) * tf_elems(c), hipMemcpyDeviceToDevice, c->s));
  return 0;
}
__global__ void k_zero(double2 *__restrict__ y, size_t n) {
  size_t i = (size_t)blockIdx.x * BB + threadIdx.x;
  if (i < n) y[i] = make_double2(0, 0);
}
int tf_zero(susy_ctx *c, double2 *dst) {
  size_t n = tf_elems(c);
  hipLaunchKernelGGL(k_zero, dim3((n + BB - 1) / BB), dim3(BB), 0, c->s, dst, n);
  HIPCHK(hipGetLastError());
  return 0;
}

// AoS [site][k][i][j] (re,im) <-> SoA [k][i][j][site]; one thread per
// (site, complex element): writes of the SoA side are coalesced.
__global__ void k_aos_to_soa(const double2 *__restrict__ aos, double2 *__restrict__ soa,
                             int nmat, int V) {
  size_t i = (size_t)blockIdx.x * BB + threadIdx.x;
  size_t per = (size_t)nmat * NN;
  if (i >= per * V) return;
  size_t e = i / V, s = i - e * V;
  soa[i] = aos[s * per + e];
}
__global__ void k_soa_to_aos(const double2 *__restrict__ soa, double2 *__restrict__ aos,
                             int nmat, int V) {
  size_t i = (size_t)blockIdx.x * BB + threadIdx.x;
  size_t per = (size_t)nmat * NN;
  if (i >= per * V) return;
  size_t e = i / V, s = i - e * V;
  aos[s * per + e] = soa[i];
}
int aos_to_soa(susy_ctx *c, const double *aos, double2 *soa, int nmat) {
  size_t n = (size_t)nmat * NN * c->V;
  hipLaunchKernelGGL(k_aos_to_soa, dim3((n + BB - 1) / BB), dim3(BB), 0, c->s,
                     (const double2 *)aos, soa, nmat, c->V);
  HIPCHK(hipGetLastError());
  return 0;
}
int soa_to_aos(susy_ctx *c, const double2 *soa, double *aos, int nmat) {
  size_t n = (size_t)nmat * NN * c->V;
  hipLaunchKernelGGL(k_soa_to_aos, dim3((n + BB - 1) / BB), dim3(BB), 0, c->s, soa,
                     (double2 *)aos, nmat, c->V);
  HIPCHK(hipGetLastError());
  return 0;
}
