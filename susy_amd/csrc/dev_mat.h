// Register-resident NCOL x NCOL complex matrices (FP64) for the owner-computes
// kernels of the link fields and forces.
#ifndef SUSY_DEV_MAT_H
#define SUSY_DEV_MAT_H
#include "susy_internal.h"

#define FB 256

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cconjd(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double cabs1(double2 a) { return fabs(a.x) + fabs(a.y); }
__device__ __forceinline__ double2 cdiv(double2 a, double2 b) {
  double t = b.x * b.x + b.y * b.y;
  return make_double2((a.x * b.x + a.y * b.y) / t, (a.y * b.x - a.x * b.y) / t);
}

typedef double2 Mat[NCOL][NCOL];

__device__ __forceinline__ void mload(Mat m, const double2 *p, int V) {
#pragma unroll
  for (int i = 0; i < NCOL; i++)
#pragma unroll
    for (int j = 0; j < NCOL; j++) m[i][j] = p[(i * NCOL + j) * V];
}
// (each stored value is held until the store has read its registers)
__device__ __forceinline__ void mstore(double2 *p, const Mat m, int V) {
#pragma unroll
  for (int i = 0; i < NCOL; i++)
#pragma unroll
    for (int j = 0; j < NCOL; j++) p[(i * NCOL + j) * V] = m[i][j];
#pragma unroll
  for (int i = 0; i < NCOL; i++)
#pragma unroll
    for (int j = 0; j < NCOL; j++) hold(m[i][j]);
}
__device__ __forceinline__ void mzero(Mat m) {
#pragma unroll
  for (int i = 0; i < NCOL; i++)
#pragma unroll
    for (int j = 0; j < NCOL; j++) m[i][j] = make_double2(0, 0);
}
// c += s op(a) op(b)
template <bool TA, bool TB>
__device__ __forceinline__ void mmadd(Mat c, double s, const Mat a, const Mat b) {
#pragma unroll
  for (int i = 0; i < NCOL; i++)
#pragma unroll
    for (int k = 0; k < NCOL; k++) {
      double2 acc = make_double2(0, 0);
#pragma unroll
      for (int j = 0; j < NCOL; j++) {
        double2 x = TA ? cconjd(a[j][i]) : a[i][j];
        double2 y = TB ? cconjd(b[k][j]) : b[j][k];
        acc.x += x.x * y.x - x.y * y.y;
        acc.y += x.x * y.y + x.y * y.x;
      }
      c[i][k].x += s * acc.x;
      c[i][k].y += s * acc.y;
    }
}
__device__ __forceinline__ double retrace_dag(const Mat a, const Mat b) {  // ReTr(a^dag b)
  double s = 0;
#pragma unroll
  for (int i = 0; i < NCOL; i++)
#pragma unroll
    for (int j = 0; j < NCOL; j++) s += a[i][j].x * b[i][j].x + a[i][j].y * b[i][j].y;
  return s;
}

__device__ __forceinline__ void block_partial(double v, double *partial, int slot_stride,
                                              int slot) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  __shared__ double red[FB / 64][4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][slot] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0;
    for (int w = 0; w < (int)(blockDim.x / 64); w++) t += red[w][slot];
    st_keep(partial + slot * slot_stride + blockIdx.x, t);
  }
  __syncthreads();
}

#endif  // SUSY_DEV_MAT_H
