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

// ---------------------------------------------------------------------------
// mode 1 (bf16 hi*hi) and mode 3 (bf16x3: hi*hi + hi*lo + lo*hi) on MFMA
// v_mfma_f32_32x32x16_bf16.  Each float is split x = hi + lo with hi, lo
// bf16 (round to nearest even), giving ~16 mantissa bits per product.
// Block: 256 threads = 4 waves, 128x128 output tile, each wave 64x64 as
// 2x2 MFMA tiles.  K is streamed in slabs of 32 through LDS stored [i][k]
// (k contiguous, so a lane's 8 k-values are one 16-byte read); fp32 MFMA
// accumulators are flushed into fp64 registers every 8 slabs; K is split
// over S chunks whose fp64 partial tiles are summed in a fixed order by
// k_gram_reduce.
// ---------------------------------------------------------------------------
typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(16))) float f32x16;

#define MT 128       // output tile
#define MKS 32       // K slab
#define MLD (MKS + 8)  // LDS row stride in bf16 (padding against bank conflicts)
#define FLUSH 8      // slabs between fp32 -> fp64 flushes

template <typename TI>
__device__ inline void split_bf16(TI x, __bf16& hi, __bf16& lo) {
  float xf = (float)x;
  hi = (__bf16)xf;
  lo = (__bf16)(float)((double)x - (double)(float)hi);
}

// upper-triangular tile t -> (ti, tj), ti <= tj
__device__ inline void tri_tile(int64_t t, int64_t nt, int64_t& ti, int64_t& tj) {
  ti = 0;
  while (t >= nt - ti) {
    t -= nt - ti;
    ++ti;
  }
  tj = ti + t;
}

template <typename TI, int NTERM>
__global__ void __launch_bounds__(256) k_gram_mfma(int64_t K, int64_t m, const TI* __restrict__ JT, int64_t ldj,
                                                   int nsplit, double* __restrict__ part) {
  __shared__ __align__(16) __bf16 Ahi[MT * MLD];
  __shared__ __align__(16) __bf16 Alo[MT * MLD];
  __shared__ __align__(16) __bf16 Bhi[MT * MLD];
  __shared__ __align__(16) __bf16 Blo[MT * MLD];
  const int64_t nt = (m + MT - 1) / MT;
  const int64_t ntri = nt * (nt + 1) / 2;
  const int64_t tile = blockIdx.x % ntri;
  const int split = (int)(blockIdx.x / ntri);
  int64_t ti, tj;
  tri_tile(tile, nt, ti, tj);
  const int64_t i0 = ti * MT, j0 = tj * MT;
  const int64_t kchunk = ((K + nsplit - 1) / nsplit + MKS - 1) / MKS * MKS;
  const int64_t kb = (int64_t)split * kchunk;
  const int64_t ke = min(K, kb + kchunk);
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wr = (wave >> 1) * 64, wc = (wave & 1) * 64;
  f32x16 acc[2][2];
  double acc64[2][2][16];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        acc[a][b][r] = 0.f;
        acc64[a][b][r] = 0.0;
      }
  int slab = 0;
  for (int64_t k0 = kb; k0 < ke; k0 += MKS, ++slab) {
    // global -> bf16 split -> LDS [i][k]; 256 threads x 16 values per operand
    for (int q = threadIdx.x; q < MKS * MT; q += 256) {
      int kk = q / MT, c = q - kk * MT;
      int64_t k = k0 + kk;
      TI va = (k < ke && i0 + c < m) ? JT[k * ldj + i0 + c] : (TI)0;
      TI vb = (k < ke && j0 + c < m) ? JT[k * ldj + j0 + c] : (TI)0;
      __bf16 h, l;
      split_bf16(va, h, l);
      Ahi[c * MLD + kk] = h;
      Alo[c * MLD + kk] = l;
      split_bf16(vb, h, l);
      Bhi[c * MLD + kk] = h;
      Blo[c * MLD + kk] = l;
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < MKS; ks += 16) {
      const int kof = ks + (lane >> 5) * 8;
      bf16x8 ah[2], al[2], bh[2], bl[2];
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        int ra = wr + t * 32 + (lane & 31);
        int rb = wc + t * 32 + (lane & 31);
        ah[t] = *reinterpret_cast<const bf16x8*>(&Ahi[ra * MLD + kof]);
        bh[t] = *reinterpret_cast<const bf16x8*>(&Bhi[rb * MLD + kof]);
        if (NTERM > 1) {
          al[t] = *reinterpret_cast<const bf16x8*>(&Alo[ra * MLD + kof]);
          bl[t] = *reinterpret_cast<const bf16x8*>(&Blo[rb * MLD + kof]);
        }
      }
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
          acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[a], bh[b], acc[a][b], 0, 0, 0);
          if (NTERM > 1) {
            acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[a], bl[b], acc[a][b], 0, 0, 0);
            acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[a], bh[b], acc[a][b], 0, 0, 0);
          }
        }
    }
    __syncthreads();
    if ((slab + 1) % FLUSH == 0) {
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
          for (int r = 0; r < 16; ++r) {
            acc64[a][b][r] += (double)acc[a][b][r];
            acc[a][b][r] = 0.f;
          }
    }
  }
  double* out = part + (int64_t)split * m * m;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        // 32x32 MFMA output layout: col = lane % 32, row = 8*(r/4) + 4*(lane/32) + r%4
        int64_t i = i0 + wr + a * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        int64_t j = j0 + wc + b * 32 + (lane & 31);
        if (i < m && j < m) out[i * m + j] = acc64[a][b][r] + (double)acc[a][b][r];
      }
}

// G = diag(d) (sum_s part_s) diag(d), upper tiles mirrored; fixed summation order
__global__ void k_gram_reduce(int64_t m, int nsplit, const double* __restrict__ part, const double* __restrict__ d,
                              double* __restrict__ G) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= m * m) return;
  int64_t i = idx / m, j = idx - i * m;
  int64_t src = (i <= j) ? idx : j * m + i;  // upper triangle (upper tiles were written): exact symmetry
  double v = 0.0;
  for (int s = 0; s < nsplit; ++s) v += part[(int64_t)s * m * m + src];
  if (d) v *= d[i] * d[j];
  G[idx] = v;
}

extern "C" int rb_gram_workspace(int64_t K, int64_t m, int mode, size_t* bytes) {
  if (!bytes) return RB_ERR_ARG;
  int nsplit = (mode == 0) ? 0 : 4;
  *bytes = (size_t)nsplit * m * m * sizeof(double);
  return RB_OK;
}

extern "C" int rb_gram_mfma(int64_t K, int64_t m, const void* JT, int64_t ldj, int in_dtype, const double* d,
                            int mode, double* G, void* work, size_t work_bytes, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  const int nsplit = 4;
  if ((size_t)nsplit * m * m * sizeof(double) > work_bytes || !work) {
    rb_set_error("rb_gram_mfma: workspace too small");
    return RB_ERR_SIZE;
  }
  int64_t nt = (m + MT - 1) / MT;
  unsigned blocks = (unsigned)(nt * (nt + 1) / 2 * nsplit);
  double* part = (double*)work;
#define LG(TI, NT) \
  hipLaunchKernelGGL((k_gram_mfma<TI, NT>), dim3(blocks), dim3(256), 0, s, K, m, (const TI*)JT, ldj, nsplit, part)
  if (in_dtype == RB_F32) {
    if (mode == 1)
      LG(float, 1);
    else
      LG(float, 3);
  } else if (in_dtype == RB_F64) {
    if (mode == 1)
      LG(double, 1);
    else
      LG(double, 3);
  } else {
    rb_set_error("rb_gram_mfma: in_dtype must be RB_F64 or RB_F32");
    return RB_ERR_ARG;
  }
#undef LG
  RB_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_gram_reduce, dim3((unsigned)ceil_div(m * m, 256)), dim3(256), 0, s, m, nsplit, part, d, G);
  RB_CHECK_LAUNCH();
  return RB_OK;
}

// ---------------------------------------------------------------------------
// K9: Jacobian matvecs with float64 accumulation (JT: K x M node-major, the
// M columns grouped in nw blocks of np pairs).
//   rows: t[k][w] = sum_p JT[k][w*np + p] v[w*np + p]           (J_w^T v_w)
//   cols: u[j]    = sum_k JT[k][j] z[k][j / np]  (mode 0)        (J_w z_w)
//         u[j]    = sum_k JT[k][j]^2             (mode 1)        (diag of J J^T)
// ---------------------------------------------------------------------------
template <typename TI>
__global__ void __launch_bounds__(256) k_jt_rows(int64_t K, int64_t M, const TI* __restrict__ JT, int64_t ldj,
                                                 int nw, int64_t np, const double* __restrict__ v,
                                                 double* __restrict__ t) {
  const int lane = threadIdx.x & 63;
  const int64_t k = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (k >= K) return;
  const TI* row = JT + k * ldj;
  for (int w = 0; w < nw; ++w) {
    double acc = 0.0;
    for (int64_t p = lane; p < np; p += 64) acc += (double)row[w * np + p] * v[w * np + p];
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
    if (lane == 0) t[k * nw + w] = acc;
  }
}

#define JCR 256  // rows per block of the column kernel
template <typename TI>
__global__ void __launch_bounds__(256) k_jt_cols(int64_t K, int64_t M, const TI* __restrict__ JT, int64_t ldj,
                                                 int nw, int64_t np, const double* __restrict__ z, int mode,
                                                 double* __restrict__ part) {
  const int64_t k0 = (int64_t)blockIdx.y * JCR;
  const int64_t k1 = min(K, k0 + JCR);
  for (int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x; j < M; j += (int64_t)gridDim.x * 256) {
    const int64_t w = j / np;
    double acc = 0.0;
    for (int64_t k = k0; k < k1; ++k) {
      double a = (double)JT[k * ldj + j];
      acc += mode ? a * a : a * z[k * nw + w];
    }
    part[(int64_t)blockIdx.y * M + j] = acc;
  }
}

__global__ void k_sum_parts(int64_t M, int nparts, const double* __restrict__ part, double* __restrict__ u) {
  int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= M) return;
  double acc = 0.0;
  for (int b = 0; b < nparts; ++b) acc += part[(int64_t)b * M + j];
  u[j] = acc;
}

extern "C" int rb_jt_rows(int64_t K, int64_t M, const void* JT, int64_t ldj, int in_dtype, int nw, int64_t np,
                          const double* v, double* t, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  if (K <= 0 || nw * np > M) return RB_ERR_ARG;
  unsigned g = (unsigned)ceil_div(K, 4);
  if (in_dtype == RB_F64)
    hipLaunchKernelGGL(k_jt_rows<double>, dim3(g), dim3(256), 0, s, K, M, (const double*)JT, ldj, nw, np, v, t);
  else if (in_dtype == RB_F32)
    hipLaunchKernelGGL(k_jt_rows<float>, dim3(g), dim3(256), 0, s, K, M, (const float*)JT, ldj, nw, np, v, t);
  else
    return RB_ERR_ARG;
  RB_CHECK_LAUNCH();
  return RB_OK;
}

extern "C" int rb_jt_cols_workspace(int64_t K, int64_t M, size_t* bytes) {
  if (!bytes) return RB_ERR_ARG;
  *bytes = (size_t)ceil_div(K, JCR) * M * sizeof(double);
  return RB_OK;
}

extern "C" int rb_jt_cols(int64_t K, int64_t M, const void* JT, int64_t ldj, int in_dtype, int nw, int64_t np,
                          const double* z, int mode, double* u, void* work, size_t work_bytes, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  int64_t nparts = ceil_div(K, JCR);
  if (K <= 0 || nw * np > M || (size_t)nparts * M * sizeof(double) > work_bytes || !work) {
    rb_set_error("rb_jt_cols: bad sizes or workspace");
    return RB_ERR_ARG;
  }
  dim3 g((unsigned)ceil_div(M, 256), (unsigned)nparts);
  if (in_dtype == RB_F64)
    hipLaunchKernelGGL(k_jt_cols<double>, g, dim3(256), 0, s, K, M, (const double*)JT, ldj, nw, np, z, mode,
                       (double*)work);
  else if (in_dtype == RB_F32)
    hipLaunchKernelGGL(k_jt_cols<float>, g, dim3(256), 0, s, K, M, (const float*)JT, ldj, nw, np, z, mode,
                       (double*)work);
  else
    return RB_ERR_ARG;
  RB_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_sum_parts, dim3((unsigned)ceil_div(M, 256)), dim3(256), 0, s, M, (int)nparts,
                     (const double*)work, u);
  RB_CHECK_LAUNCH();
  return RB_OK;
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
