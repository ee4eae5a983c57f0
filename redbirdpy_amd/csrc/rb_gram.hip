// K8: Gauss-Newton Gram matrix G = D (JT^T JT) D, JT: K x m row-major
// (node-major Jacobian, pairs contiguous), D = diag(d).  redbirdpy computes
// this as `Amat @ Amat.T` (recon.py:474).  Only tiles on or above the
// diagonal are computed; each is mirrored.
//
// mode 0: float64 FMA, LDS-tiled 64x64 output tiles, 4x4 per thread.
#include "rb_common.h"
#include "../../include/rbhip.h"

#define GT 64   // output tile
#define GK 16   // K slab

template <typename TI>
__global__ void __launch_bounds__(256) k_gram_f64(int64_t K, int64_t m, const TI* __restrict__ JT, int64_t ldj,
                                                  const double* __restrict__ d, double* __restrict__ G) {
  __shared__ double As[GK][GT + 1];
  __shared__ double Bs[GK][GT + 1];
  // upper-triangular tile index -> (ti, tj), ti <= tj
  int64_t nt = (m + GT - 1) / GT;
  int64_t t = blockIdx.x;
  int64_t ti = 0;
  while (t >= nt - ti) {
    t -= nt - ti;
    ++ti;
  }
  int64_t tj = ti + t;
  int64_t i0 = ti * GT, j0 = tj * GT;
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  double acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
  for (int64_t k0 = 0; k0 < K; k0 += GK) {
    for (int i = threadIdx.x; i < GK * GT; i += 256) {
      int kk = i / GT, c = i - kk * GT;
      int64_t k = k0 + kk;
      As[kk][c] = (k < K && i0 + c < m) ? (double)JT[k * ldj + i0 + c] : 0.0;
      Bs[kk][c] = (k < K && j0 + c < m) ? (double)JT[k * ldj + j0 + c] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < GK; ++kk) {
      double a[4], b[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        a[q] = As[kk][ty * 4 + q];
        b[q] = Bs[kk][tx * 4 + q];
      }
#pragma unroll
      for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[p][q] = fma(a[p], b[q], acc[p][q]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    int64_t i = i0 + ty * 4 + p;
    if (i >= m) continue;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      int64_t j = j0 + tx * 4 + q;
      if (j >= m) continue;
      double v = acc[p][q];
      if (d) v *= d[i] * d[j];
      G[i * m + j] = v;
      G[j * m + i] = v;
    }
  }
}

extern "C" int rb_gram(int64_t K, int64_t m, const void* JT, int64_t ldj, int in_dtype, const double* d, int mode,
                       double* G, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  if (K <= 0 || m <= 0 || !JT || !G) {
    rb_set_error("rb_gram: bad arguments");
    return RB_ERR_ARG;
  }
  int64_t nt = (m + GT - 1) / GT;
  unsigned tiles = (unsigned)(nt * (nt + 1) / 2);
  if (mode != 0) {
    rb_set_error("rb_gram: mode %d not available", mode);
    return RB_ERR_ARG;
  }
  if (in_dtype == RB_F64)
    hipLaunchKernelGGL(k_gram_f64<double>, dim3(tiles), dim3(256), 0, s, K, m, (const double*)JT, ldj, d, G);
  else if (in_dtype == RB_F32)
    hipLaunchKernelGGL(k_gram_f64<float>, dim3(tiles), dim3(256), 0, s, K, m, (const float*)JT, ldj, d, G);
  else {
    rb_set_error("rb_gram: in_dtype must be RB_F64 or RB_F32");
    return RB_ERR_ARG;
  }
  RB_CHECK_LAUNCH();
  return RB_OK;
}
