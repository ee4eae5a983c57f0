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
  // thread: rows ty + 16 p, columns tx + 16 q of the tile (interleaved: a
  // wavefront's LDS reads of Bs fall on consecutive banks)
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
        a[q] = As[kk][ty + 16 * q];
        b[q] = Bs[kk][tx + 16 * q];
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
    int64_t i = i0 + ty + 16 * p;
    if (i >= m) continue;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      int64_t j = j0 + tx + 16 * q;
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
// v_mfma_f32_32x32x16_bf16.  Each value is split x = hi + lo with hi, lo
// bf16 (round to nearest even), ~16 mantissa bits per product.
//
// 1. k_split_t: JT (K x m, fp32/fp64) -> Hi, Lo (m x Kp bf16, transposed,
//    zero padded to Kp = K rounded up to 64), through 64x64 LDS tiles; on
//    request also the exact float64 column sums of squares (the Gram
//    diagonal) from the same read.
// 2. k_gram_bf16w: 512 threads = 8 waves per 256x256 upper-triangular
//    output tile and K chunk (split-K, about 2048 workgroups in all).  fp32
//    accumulation inside a chunk; the fp32 partial tiles are summed in fp64
//    in a fixed order by k_gram_reduce.
// ---------------------------------------------------------------------------
typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(16))) float f32x16;

// upper-triangular tile t -> (ti, tj), ti <= tj
__device__ inline void tri_tile(int64_t t, int64_t nt, int64_t& ti, int64_t& tj) {
  ti = 0;
  while (t >= nt - ti) {
    t -= nt - ti;
    ++ti;
  }
  tj = ti + t;
}

#define SPLIT_GK 64  // K-strided workgroups per 64-column strip of the split

// blockIdx.y: 64-column strip; blockIdx.x walks the 64-k tiles kt = x, x + GK, ...
// colpart[x][c] (optional) = sum over those tiles of JT[k][c]^2 in float64,
// each thread's k in ascending order, the four thread partials in a fixed order
template <typename TI>
__global__ void __launch_bounds__(256) k_split_t(int64_t K, int64_t m, int64_t Kp, const TI* __restrict__ JT,
                                                 int64_t ldj, __bf16* __restrict__ Hi, __bf16* __restrict__ Lo,
                                                 double* __restrict__ colpart) {
  __shared__ float th[64][65];
  __shared__ float tl[64][65];
  __shared__ double sq[4][64];
  const int64_t c0 = (int64_t)blockIdx.y * 64;
  const int cc0 = threadIdx.x & 63;  // this thread's column in the load loop (q % 64)
  double acc = 0.0;
  for (int64_t kt = blockIdx.x; kt * 64 < Kp; kt += gridDim.x) {
    const int64_t k0 = kt * 64;
#pragma unroll
    for (int it = 0; it < 64 * 64 / 256; ++it) {  // all 16 loads in flight together
      const int q = threadIdx.x + 256 * it;
      int kk = q / 64, cc = q % 64;
      int64_t k = k0 + kk, c = c0 + cc;
      double x = (k < K && c < m) ? (double)JT[k * ldj + c] : 0.0;
      acc += x * x;
      __bf16 h = (__bf16)(float)x;
      th[cc][kk] = (float)h;
      tl[cc][kk] = (float)(x - (double)(float)h);
    }
    __syncthreads();
    // eight consecutive k of one column per thread, one 16-byte store each
    // (Kp and k0 are multiples of 64: every store is aligned and in the row)
    for (int q = threadIdx.x; q < 64 * 8; q += 256) {
      const int cc = q >> 3, kk = (q & 7) * 8;
      const int64_t c = c0 + cc;
      if (c < m) {
        unsigned int hw[4], lw[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          hw[i] = (unsigned int)__builtin_bit_cast(unsigned short, (__bf16)th[cc][kk + 2 * i]) |
                  ((unsigned int)__builtin_bit_cast(unsigned short, (__bf16)th[cc][kk + 2 * i + 1]) << 16);
          lw[i] = (unsigned int)__builtin_bit_cast(unsigned short, (__bf16)tl[cc][kk + 2 * i]) |
                  ((unsigned int)__builtin_bit_cast(unsigned short, (__bf16)tl[cc][kk + 2 * i + 1]) << 16);
        }
        *reinterpret_cast<uint4*>(Hi + c * Kp + k0 + kk) = make_uint4(hw[0], hw[1], hw[2], hw[3]);
        *reinterpret_cast<uint4*>(Lo + c * Kp + k0 + kk) = make_uint4(lw[0], lw[1], lw[2], lw[3]);
      }
    }
    __syncthreads();
  }
  if (colpart) {
    sq[threadIdx.x >> 6][cc0] = acc;
    __syncthreads();
    if (threadIdx.x < 64 && c0 + cc0 < m)
      colpart[(int64_t)blockIdx.x * m + c0 + cc0] = ((sq[0][cc0] + sq[1][cc0]) + sq[2][cc0]) + sq[3][cc0];
  }
}

__global__ void k_colsq_finish(int64_t m, int gk, const double* __restrict__ colpart, double* __restrict__ colsq) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= m) return;
  double v = 0.0;
  for (int g = 0; g < gk; ++g) v += colpart[(int64_t)g * m + c];
  colsq[c] = v;
}

// Wide variant: 512 threads = 8 waves per 256x256 upper-triangular output
// tile and K chunk.  Slabs of 32 k are double-buffered in LDS (4 operand
// arrays x 256 rows x 32 k, 64 KB per buffer): the next slab's global loads
// are in flight in registers while the MFMAs run on the current one, one
// barrier per slab.  Each row's four 16-byte chunks are stored XOR-swizzled
// by (row >> 2) & 3, so the MFMA operand reads (32 rows, one chunk) hit
// distinct banks.  Each wave owns a 128x64 block = 4x2 MFMA tiles.
#define WT 256   // output tile
#define WKS 32   // K slab
#define WARR (WT * WKS)  // bf16 per operand array per buffer
#define GXR 32           // consecutive work items per XCD (its CUs' resident workgroups)
#define GNT 512          // threads per workgroup: 8 waves of 128x64 (256: 4 waves of 128x128, measured slower)

__device__ __forceinline__ int wswz(int row, int c) { return row * WKS + ((c ^ ((row >> 2) & 3)) << 3); }

template <int NTERM, int NT>
__global__ void __launch_bounds__(NT) k_gram_bf16w(int64_t m, int64_t Kp, int nsplit, const __bf16* __restrict__ Hi,
                                                   const __bf16* __restrict__ Lo, float* __restrict__ part) {
  constexpr int WV = NT / 64;            // waves: 2 rows x WV/2 columns of wave tiles
  constexpr int TC = 16 / WV;            // 32-column MFMA tiles per wave (4 row tiles each)
  static_assert(2 * 128 == WT && (WV / 2) * 32 * TC == WT, "the waves cover the WT x WT tile");
  constexpr int HPT = WARR / 8 / NT;     // 16-byte chunks per operand array and thread
  static_assert(HPT * NT * 8 == WARR && NT / 4 * HPT == WT, "one slab: WT rows x 4 chunks per array");
  __shared__ __align__(16) __bf16 lds[2 * 4 * WARR];
  const int64_t nt = (m + WT - 1) / WT;
  const int64_t ntri = nt * (nt + 1) / 2;
  // workgroups go to the 8 XCDs round-robin: give each XCD runs of GXR
  // consecutive (split, tile) items, so the tiles resident on one XCD at a
  // time share operand rows in its L2 (the grid is padded to whole rounds)
  const int64_t slot = blockIdx.x / 8;
  const int64_t g = (slot / GXR) * (8 * GXR) + (int64_t)(blockIdx.x % 8) * GXR + slot % GXR;
  if (g >= ntri * nsplit) return;
  const int64_t tile = g % ntri;
  const int split = (int)(g / ntri);
  int64_t ti, tj;
  tri_tile(tile, nt, ti, tj);
  const int64_t i0 = ti * WT, j0 = tj * WT;
  const int64_t kchunk = ((Kp + nsplit - 1) / nsplit + WKS - 1) / WKS * WKS;
  const int64_t kb = (int64_t)split * kchunk;
  const int64_t ke = min(Kp, kb + kchunk);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = (wave / (WV / 2)) * 128, wc = (wave % (WV / 2)) * (32 * TC);
  // operand arrays: 0 A hi, 1 A lo, 2 B hi, 3 B lo; thread loads chunks q = tid + NT * h
  // chunk q = tid + NT*h is row q >> 2 (tid >> 2 + NT/4 * h), 16-byte piece q & 3 (tid & 3)
  const int lr0 = tid >> 2, lc = tid & 3;
  bf16x8 pre[4][HPT];
  auto load = [&](int64_t k0) {
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      if (NTERM == 1 && (a & 1)) continue;
      const __bf16* base = ((a & 1) ? Lo : Hi) + k0 + lc * 8;
      const int64_t r0 = ((a < 2) ? i0 : j0) + lr0;
#pragma unroll
      for (int h = 0; h < HPT; ++h) {
        const int64_t r = r0 + (NT / 4) * h;
        bf16x8 z;
#pragma unroll
        for (int e = 0; e < 8; ++e) z[e] = (__bf16)0.f;
        pre[a][h] = r < m ? *reinterpret_cast<const bf16x8*>(base + r * Kp) : z;
      }
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      if (NTERM == 1 && (a & 1)) continue;
#pragma unroll
      for (int h = 0; h < HPT; ++h)
        *reinterpret_cast<bf16x8*>(&lds[(buf * 4 + a) * WARR + wswz(lr0 + (NT / 4) * h, lc)]) = pre[a][h];
    }
  };
  f32x16 acc[4][TC];
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int u = 0; u < TC; ++u)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[t][u][r] = 0.f;
  if (kb < ke) {
    load(kb);
    store(0);
  }
  __syncthreads();
  int cur = 0;
  for (int64_t k0 = kb; k0 < ke; k0 += WKS) {
    const bool more = k0 + WKS < ke;
    if (more) load(k0 + WKS);
    const __bf16* L = lds + cur * 4 * WARR;
#pragma unroll
    for (int ks = 0; ks < WKS; ks += 16) {
      const int c = (ks >> 3) + (lane >> 5);
      bf16x8 ah[4], al[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int ra = wr + t * 32 + (lane & 31);
        ah[t] = *reinterpret_cast<const bf16x8*>(&L[0 * WARR + wswz(ra, c)]);
        if (NTERM > 1) al[t] = *reinterpret_cast<const bf16x8*>(&L[1 * WARR + wswz(ra, c)]);
      }
#pragma unroll
      for (int u = 0; u < TC; ++u) {  // B fragments one column tile at a time
        const int rb = wc + u * 32 + (lane & 31);
        const bf16x8 bh = *reinterpret_cast<const bf16x8*>(&L[2 * WARR + wswz(rb, c)]);
        bf16x8 bl;
        if (NTERM > 1) bl = *reinterpret_cast<const bf16x8*>(&L[3 * WARR + wswz(rb, c)]);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          acc[t][u] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[t], bh, acc[t][u], 0, 0, 0);
          if (NTERM > 1) {
            acc[t][u] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[t], bl, acc[t][u], 0, 0, 0);
            acc[t][u] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[t], bh, acc[t][u], 0, 0, 0);
          }
        }
      }
    }
    if (more) store(cur ^ 1);
    __syncthreads();
    cur ^= 1;
  }
  float* out = part + (int64_t)split * m * m;
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int u = 0; u < TC; ++u)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int64_t i = i0 + wr + t * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        const int64_t j = j0 + wc + u * 32 + (lane & 31);
        if (i < m && j < m) out[i * m + j] = acc[t][u][r];
      }
}

// K chunks of the wide kernel: about 2048 workgroups in all (8 per CU)
static int gram_nsplit(int64_t m) {
  const int64_t nt = (m + WT - 1) / WT;
  const int64_t ntri = nt * (nt + 1) / 2;
  int64_t s = (2048 + ntri / 2) / ntri;
  return (int)(s < 1 ? 1 : s > 32 ? 32 : s);
}

// G = diag(d) (sum_s part_s) diag(d) in fp64 over 64x64 tiles on or above
// the diagonal: the partials are read once, coalesced, summed in a fixed
// order, and each tile is written twice (as is and transposed through LDS);
// in a diagonal tile the upper entries are mirrored, so G is exactly symmetric
__global__ void __launch_bounds__(256) k_gram_reduce(int64_t m, int nsplit, const float* __restrict__ part,
                                                     const double* __restrict__ d, double* __restrict__ G) {
  __shared__ double t[64][65];
  const int64_t nt = (m + 63) / 64;
  int64_t ti, tj;
  tri_tile(blockIdx.x, nt, ti, tj);
  const int64_t i0 = ti * 64, j0 = tj * 64;
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    const int r = e >> 6, c = e & 63;
    const int64_t i = i0 + r, j = j0 + c;
    double v = 0.0;
    if (i < m && j < m) {
      for (int s = 0; s < nsplit; ++s) v += (double)part[(int64_t)s * m * m + i * m + j];
      if (d) v *= d[i] * d[j];
    }
    t[r][c] = v;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    const int r = e >> 6, c = e & 63;
    const int64_t i = i0 + r, j = j0 + c;
    if (i < m && j < m) G[i * m + j] = (ti == tj && r > c) ? t[c][r] : t[r][c];
    if (ti != tj) {
      const int64_t jj = j0 + r, ii = i0 + c;
      if (jj < m && ii < m) G[jj * m + ii] = t[c][r];
    }
  }
}

static int64_t pad64(int64_t K) { return (K + 63) / 64 * 64; }

extern "C" int rb_gram_workspace(int64_t K, int64_t m, int mode, size_t* bytes) {
  if (!bytes) return RB_ERR_ARG;
  if (mode == 0) {
    *bytes = 0;
    return RB_OK;
  }
  size_t split_bytes = (size_t)2 * m * pad64(K) * sizeof(__bf16);
  size_t part_bytes = (size_t)gram_nsplit(m) * m * m * sizeof(float);
  size_t col_bytes = (size_t)SPLIT_GK * m * sizeof(double);
  *bytes = ((split_bytes + 255) & ~(size_t)255) + ((part_bytes + 255) & ~(size_t)255) + col_bytes;
  return RB_OK;
}

extern "C" int rb_gram_mfma(int64_t K, int64_t m, const void* JT, int64_t ldj, int in_dtype, const double* d,
                            int mode, double* G, double* colsq, void* work, size_t work_bytes, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  size_t need = 0;
  rb_gram_workspace(K, m, mode, &need);
  if (need > work_bytes || !work) {
    rb_set_error("rb_gram_mfma: workspace %zu < %zu bytes", work_bytes, need);
    return RB_ERR_SIZE;
  }
  const int64_t Kp = pad64(K);
  __bf16* Hi = (__bf16*)work;
  __bf16* Lo = Hi + m * Kp;
  size_t split_bytes = (size_t)2 * m * Kp * sizeof(__bf16);
  float* part = (float*)((char*)work + ((split_bytes + 255) & ~(size_t)255));
  const int nsplit = gram_nsplit(m);
  size_t part_bytes = (size_t)nsplit * m * m * sizeof(float);
  double* colpart = colsq ? (double*)((char*)part + ((part_bytes + 255) & ~(size_t)255)) : nullptr;
  const int64_t nkt = Kp / 64;
  const int gk = (int)(nkt < SPLIT_GK ? nkt : SPLIT_GK);
  dim3 gs((unsigned)gk, (unsigned)((m + 63) / 64));
  if (in_dtype == RB_F32)
    hipLaunchKernelGGL(k_split_t<float>, gs, dim3(256), 0, s, K, m, Kp, (const float*)JT, ldj, Hi, Lo, colpart);
  else if (in_dtype == RB_F64)
    hipLaunchKernelGGL(k_split_t<double>, gs, dim3(256), 0, s, K, m, Kp, (const double*)JT, ldj, Hi, Lo, colpart);
  else {
    rb_set_error("rb_gram_mfma: in_dtype must be RB_F64 or RB_F32");
    return RB_ERR_ARG;
  }
  RB_CHECK_LAUNCH();
  if (colsq) {
    hipLaunchKernelGGL(k_colsq_finish, dim3((unsigned)ceil_div(m, 256)), dim3(256), 0, s, m, gk,
                       (const double*)colpart, colsq);
    RB_CHECK_LAUNCH();
  }
  const int64_t nt = (m + WT - 1) / WT;
  unsigned blocks = (unsigned)ceil_div(nt * (nt + 1) / 2 * nsplit, 8 * GXR) * (8 * GXR);
  if (mode == 1)
    hipLaunchKernelGGL((k_gram_bf16w<1, GNT>), dim3(blocks), dim3(GNT), 0, s, m, Kp, nsplit, Hi, Lo, part);
  else
    hipLaunchKernelGGL((k_gram_bf16w<3, GNT>), dim3(blocks), dim3(GNT), 0, s, m, Kp, nsplit, Hi, Lo, part);
  RB_CHECK_LAUNCH();
  const int64_t nt64 = (m + 63) / 64;
  hipLaunchKernelGGL(k_gram_reduce, dim3((unsigned)(nt64 * (nt64 + 1) / 2)), dim3(256), 0, s, m, nsplit, part, d,
                     G);
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
  if (K <= 0 || nw <= 0 || np <= 0 || nw * np > M || ldj < M || !JT || !v || !t) {
    rb_set_error("rb_jt_rows: bad sizes (K %lld, M %lld, ldj %lld, nw %d, np %lld)", (long long)K, (long long)M,
                 (long long)ldj, nw, (long long)np);
    return RB_ERR_ARG;
  }
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
  // mode 0 reads z[k][j / np] for every column j < M: M must be exactly nw blocks of np
  if (K <= 0 || M <= 0 || nw <= 0 || np <= 0 || nw * np > M || (mode == 0 && (nw * np != M || !z)) ||
      (mode != 0 && mode != 1) || ldj < M || !JT || !u) {
    rb_set_error("rb_jt_cols: bad sizes (K %lld, M %lld, ldj %lld, nw %d, np %lld, mode %d)", (long long)K,
                 (long long)M, (long long)ldj, nw, (long long)np, mode);
    return RB_ERR_ARG;
  }
  if ((size_t)nparts * M * sizeof(double) > work_bytes || !work) {
    rb_set_error("rb_jt_cols: workspace of %zu bytes, %zu needed", work_bytes, (size_t)nparts * M * sizeof(double));
    return RB_ERR_SIZE;
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

// ---------------------------------------------------------------------------
// K9b: the normal-equation product out = s2 * J W (J^T v) + lam * v in ONE read
// of a float64 JT (K x M, M = nb blocks of np pairs, np a multiple of 64).
// Per node k: t[b] = sum_p JT[k][b*np+p] v[b*np+p], z = t W, u[j] += JT[k][j] z[j/np].
// A 256-thread workgroup holds NB node rows at a time, J columns per thread
// (column 64*wave + 256*j + lane, so every 64-column segment lies in one
// block); per-workgroup u goes to part[blockIdx][M] and a finish kernel sums
// the parts in a fixed order.
// ---------------------------------------------------------------------------
#define HV_MAXB 16
#define HV_MAXG 512

__device__ __forceinline__ double hv_wave_sum(double a) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o, 64);
  return a;
}

// does wave w own a 64-column segment of block b (segments s = w + 4j)
__device__ __forceinline__ bool hv_wave_has(int w, int b, int64_t np) {
  const int64_t q = np >> 6;
  if (q >= 4) return true;
  for (int64_t s = b * q; s < (b + 1) * q; ++s)
    if ((s & 3) == w) return true;
  return false;
}

template <int J, int NB>
__global__ void __launch_bounds__(256) k_jt_hv(int64_t K, int64_t M, const double* __restrict__ JT, int64_t ldj,
                                               int nb, int64_t np, const double* __restrict__ W,
                                               const double* __restrict__ v, double* __restrict__ part) {
  __shared__ double red[NB][4][HV_MAXB];
  __shared__ double zs[NB][HV_MAXB];
  const int t = threadIdx.x, lane = t & 63;
  const int wv = __builtin_amdgcn_readfirstlane(t >> 6);
  double vv[J], u[J];
  int bj[J];  // block of slot j (wave-uniform), -1 past M
#pragma unroll
  for (int j = 0; j < J; ++j) {
    const int64_t c0 = 64 * wv + 256 * j;
    bj[j] = c0 < M ? (int)(c0 / np) : -1;
    vv[j] = c0 < M ? v[c0 + lane] : 0.0;
    u[j] = 0.0;
  }
  for (int64_t n0 = (int64_t)blockIdx.x * NB; n0 < K; n0 += (int64_t)gridDim.x * NB) {
    double x[NB][J];
#pragma unroll
    for (int i = 0; i < NB; ++i)
#pragma unroll
      for (int j = 0; j < J; ++j)
        x[i][j] = (n0 + i < K && bj[j] >= 0) ? JT[(n0 + i) * ldj + 64 * wv + 256 * j + lane] : 0.0;
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      double acc = 0.0;
#pragma unroll
      for (int j = 0; j < J; ++j) {
        acc += x[i][j] * vv[j];
        if (j == J - 1 || bj[j + 1] != bj[j]) {  // end of this wave's run in block bj[j]
          if (bj[j] >= 0) {
            const double r = hv_wave_sum(acc);
            if (lane == 0) red[i][wv][bj[j]] = r;
          }
          acc = 0.0;
        }
      }
    }
    __syncthreads();
    if (t < NB * nb) {
      const int i = t / nb, b2 = t - (t / nb) * nb;
      double z = 0.0;
      for (int b = 0; b < nb; ++b) {
        double tb = 0.0;
        for (int w = 0; w < 4; ++w)
          if (hv_wave_has(w, b, np)) tb += red[i][w][b];
        z += tb * W[b * nb + b2];
      }
      zs[i][b2] = z;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NB; ++i)
#pragma unroll
      for (int j = 0; j < J; ++j)
        if (bj[j] >= 0) u[j] += x[i][j] * zs[i][bj[j]];
  }
#pragma unroll
  for (int j = 0; j < J; ++j)
    if (bj[j] >= 0) part[(int64_t)blockIdx.x * M + 64 * wv + 256 * j + lane] = u[j];
}

// out[p] = s2 * sum_g part[g][p] + lam * v[p]; 64 columns per workgroup, four
// strided partial sums combined in a fixed order
__global__ void __launch_bounds__(256) k_jt_hv_finish(int64_t M, int G, const double* __restrict__ part,
                                                      const double* __restrict__ v, double s2, double lam,
                                                      double* __restrict__ out) {
  __shared__ double sh[4][64];
  const int c = threadIdx.x & 63, g = threadIdx.x >> 6;
  const int64_t p = (int64_t)blockIdx.x * 64 + c;
  double acc = 0.0;
  if (p < M)
    for (int b = g; b < G; b += 4) acc += part[(int64_t)b * M + p];
  sh[g][c] = acc;
  __syncthreads();
  if (g == 0 && p < M) out[p] = s2 * (((sh[0][c] + sh[1][c]) + sh[2][c]) + sh[3][c]) + lam * v[p];
}

static int hv_cols(int64_t M) {  // columns per thread: the smallest J with 256 J >= M
  return M <= 1024 ? 4 : M <= 2048 ? 8 : M <= 4096 ? 16 : M <= 8192 ? 32 : 0;
}

static unsigned hv_grid(int64_t K, int64_t M) {
  const int64_t nbn = 64 / hv_cols(M);  // rows per workgroup pass
  const int64_t g = ceil_div(K, nbn);
  return (unsigned)(g < HV_MAXG ? g : HV_MAXG);
}

extern "C" int rb_jt_hv_workspace(int64_t K, int64_t M, size_t* bytes) {
  if (!bytes || K <= 0 || M <= 0 || !hv_cols(M)) return RB_ERR_ARG;
  *bytes = (size_t)hv_grid(K, M) * M * sizeof(double);
  return RB_OK;
}

extern "C" int rb_jt_hv(int64_t K, int64_t M, const double* JT, int64_t ldj, int nb, int64_t np, const double* W,
                        const double* v, double s2, double lam, double* out, void* work, size_t work_bytes,
                        void* stream) {
  hipStream_t s = (hipStream_t)stream;
  if (K <= 0 || np <= 0 || np % 64 || nb < 1 || nb > HV_MAXB || (int64_t)nb * np != M || !hv_cols(M) || ldj < M) {
    rb_set_error("rb_jt_hv: needs M = nb * np, np a multiple of 64, nb <= %d, M <= 8192", HV_MAXB);
    return RB_ERR_ARG;
  }
  const unsigned g = hv_grid(K, M);
  if (!work || work_bytes < (size_t)g * M * sizeof(double)) {
    rb_set_error("rb_jt_hv: workspace too small");
    return RB_ERR_ARG;
  }
  double* part = (double*)work;
  switch (hv_cols(M)) {
    case 4: hipLaunchKernelGGL((k_jt_hv<4, 16>), dim3(g), dim3(256), 0, s, K, M, JT, ldj, nb, np, W, v, part); break;
    case 8: hipLaunchKernelGGL((k_jt_hv<8, 8>), dim3(g), dim3(256), 0, s, K, M, JT, ldj, nb, np, W, v, part); break;
    case 16: hipLaunchKernelGGL((k_jt_hv<16, 4>), dim3(g), dim3(256), 0, s, K, M, JT, ldj, nb, np, W, v, part); break;
    default: hipLaunchKernelGGL((k_jt_hv<32, 2>), dim3(g), dim3(256), 0, s, K, M, JT, ldj, nb, np, W, v, part); break;
  }
  RB_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_jt_hv_finish, dim3((unsigned)(M / 64)), dim3(256), 0, s, M, (int)g, (const double*)part, v,
                     s2, lam, out);
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

// ---------------------------------------------------------------------------
// Jacobian columns onto a second (reconstruction) mesh: Y[r][:] =
// sum_k w[k] X[idx[k]][:] for k in [ptr[r], ptr[r+1]), in k order.  One
// workgroup per output row, lanes over the M contiguous columns (coalesced
// row reads); deterministic.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_jt_remap(int64_t nrows, const int32_t* __restrict__ ptr,
                                                  const int32_t* __restrict__ idx, const double* __restrict__ w,
                                                  int64_t M, const double* __restrict__ X, int64_t ldx,
                                                  double* __restrict__ Y, int64_t ldy) {
  const int64_t r = blockIdx.x;
  if (r >= nrows) return;
  const int32_t k0 = ptr[r], k1 = ptr[r + 1];
  for (int64_t c = threadIdx.x; c < M; c += blockDim.x) {
    double acc = 0.0;
    for (int32_t k = k0; k < k1; ++k) acc += X[(int64_t)idx[k] * ldx + c] * w[k];
    Y[r * ldy + c] = acc;
  }
}

extern "C" int rb_jt_remap(int64_t nrows, const int32_t* ptr, const int32_t* idx, const double* w, int64_t M,
                           const double* X, int64_t ldx, double* Y, int64_t ldy, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  if (nrows < 0 || M < 0 || ldx < M || ldy < M || (nrows > 0 && (!ptr || !Y)) || nrows > 0x7fffffff) {
    rb_set_error("rb_jt_remap: bad arguments");
    return RB_ERR_ARG;
  }
  if (nrows == 0 || M == 0) return RB_OK;
  hipLaunchKernelGGL(k_jt_remap, dim3((unsigned)nrows), dim3(256), 0, s, nrows, ptr, idx, w, M, X, ldx, Y, ldy);
  RB_CHECK_LAUNCH();
  return RB_OK;
}
