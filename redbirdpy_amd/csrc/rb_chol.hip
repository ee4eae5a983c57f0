// K10: Cholesky factorisation A = L L^T of a dense float64 SPD matrix, blocked
// right-looking in 64 x 64 blocks, in place in the lower triangle.  Per block
// column kb:
//   k_potrf_diag  four wavefronts factor A_kk (lane i: row i; 8-column panels
//                 factored in registers, their rank-8 updates through LDS) and
//                 form D = L_kk^-1 by forward substitution (lane j: column j)
//   k_potrf_panel L_ik = A_ik D^T for every block row i below (a workgroup each)
//   k_potrf_syrk  A_ij -= L_ik L_jk^T for kb < j <= i (a workgroup per block);
//                 the workgroup of block (kb+1, kb+1) then factors it
// The Gauss-Newton step of redbirdpy solves (J J^T + lambda I) x = r with a
// dense direct solve (recon.py:429-509); this factor replaces that solve.
#include <math.h>
#include "rb_common.h"
#include "../../include/rbhip.h"

#define CB 64   // block size
#define TP 66   // column stride of the transposed block in LDS (even: 16-byte aligned column segments)
#define PW 8    // panel width inside the diagonal factor
#define LPS 10  // row stride of the panel broadcast buffer

// LDS of one diagonal-block factorisation
struct DiagLds {
  double T[CB * TP];   // T[c * TP + r] = block[r][c]; ends as L (lower part)
  double X[CB * CB];   // X[r * CB + j] = (L^-1)[r][j]
  double Lp[CB * LPS]; // Lp[j * LPS + c]: the current panel's L[j][kb*PW + c]
  double rv[CB];       // 1 / L[k][k]
};

__device__ __forceinline__ double rl64(double x, int lane) {
  const long long u = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)(u & 0xffffffffll), lane);
  const int hi = __builtin_amdgcn_readlane((int)(u >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// orders LDS accesses between the lanes of one wavefront
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Factor one diagonal block held transposed in z.T with the NW wavefronts of
// the workgroup (t: thread index; lane l: row l of the block, column l of its
// inverse).  Blocked in panels of PW columns: every wavefront factors the
// panel's columns in registers (the same arithmetic in each, pivot rows read
// across lanes with readlane), then the wavefronts share the rank-PW update
// of the columns to its right (JU columns per pass, every load of a pass
// issued before its stores).  The inverse D = L^-1 is formed PW rows at a
// time: the wavefronts split the rows' sums over the earlier rows, the first
// wavefront finishes the block by forward substitution.  Writes L over the
// block in global memory (zero above the diagonal), D (row-major, CB x CB)
// and the first bad pivot into *info.  The loops stay rolled: a fully
// unrolled 64-step factor is larger than the instruction cache.  Every
// thread of the workgroup must call it (it holds workgroup barriers).
#define NW 4  // wavefronts per diagonal factor
__device__ __forceinline__ void diag_factor(DiagLds& z, double* __restrict__ blk, int64_t lda, int64_t k0,
                                            double* __restrict__ D, int* __restrict__ info, int t) {
  const int l = t & 63;
  const int w = __builtin_amdgcn_readfirstlane(t >> 6);
  int bad = -1;
  for (int kb = 0; kb < CB / PW; ++kb) {
    const int c0 = kb * PW;
    double p[PW];  // row l, columns c0 .. c0 + PW - 1
#pragma unroll
    for (int c = 0; c < PW; ++c) p[c] = z.T[(c0 + c) * TP + l];
#pragma unroll
    for (int c = 0; c < PW; ++c) {
      const double dkk = rl64(p[c], c0 + c);
      if (bad < 0 && !(dkk > 0.0)) bad = c0 + c;
      // 1/sqrt: hardware estimate plus two Newton steps (a few ulp; the
      // correctly rounded sqrt and divide are long dependent sequences)
      double ri = __builtin_amdgcn_rsq(dkk);
      ri = ri * (1.5 - 0.5 * dkk * ri * ri);
      ri = ri * (1.5 - 0.5 * dkk * ri * ri);
      p[c] *= ri;  // rows l >= c0 + c: L[l][c0 + c] (rows above hold unused upper entries)
      if (t == 0) z.rv[c0 + c] = ri;
#pragma unroll
      for (int d = c + 1; d < PW; ++d) p[d] -= p[c] * rl64(p[c], c0 + d);
    }
    if (w == 0) {
#pragma unroll
      for (int c = 0; c < PW; ++c) {
        z.T[(c0 + c) * TP + l] = p[c];
        z.Lp[l * LPS + c] = p[c];
      }
    }
    __syncthreads();
    // columns right of the panel: block[l][j] -= sum_c L[l][c0+c] L[j][c0+c]
    constexpr int JU = 4;  // divides CB - c0 - PW (a multiple of PW)
    for (int j = c0 + PW + JU * w; j < CB; j += JU * NW) {
      double a[JU], lp[JU][PW];
#pragma unroll
      for (int u = 0; u < JU; ++u) {
        a[u] = z.T[(j + u) * TP + l];
#pragma unroll
        for (int c = 0; c < PW; ++c) lp[u][c] = z.Lp[(j + u) * LPS + c];
      }
#pragma unroll
      for (int u = 0; u < JU; ++u)
#pragma unroll
        for (int c = 0; c < PW; ++c) a[u] -= p[c] * lp[u][c];
#pragma unroll
      for (int u = 0; u < JU; ++u) z.T[(j + u) * TP + l] = a[u];
    }
    __syncthreads();
  }
  // D = L^-1, lane l solves column l, PW rows at a time; wavefront w sums
  // rows r0 + RW*w .. r0 + RW*w + RW - 1 over the earlier rows into z.Lp
  constexpr int RW = PW / NW;
  static_assert(RW * NW == PW && PW * CB <= CB * LPS, "inverse row split");
  for (int rb = 0; rb < CB / PW; ++rb) {
    const int r0 = rb * PW;
    double sm[RW];
#pragma unroll
    for (int i = 0; i < RW; ++i) sm[i] = 0.0;
#pragma unroll 4
    for (int k = 0; k < r0; ++k) {
      const double xk = z.X[k * CB + l];
#pragma unroll
      for (int i = 0; i < RW; ++i) sm[i] += z.T[k * TP + r0 + RW * w + i] * xk;
    }
#pragma unroll
    for (int i = 0; i < RW; ++i) z.Lp[(RW * w + i) * CB + l] = sm[i];
    __syncthreads();
    if (w == 0) {
      double x[PW];
#pragma unroll
      for (int i = 0; i < PW; ++i) x[i] = ((l == r0 + i) ? 1.0 : 0.0) - z.Lp[i * CB + l];
#pragma unroll
      for (int i = 0; i < PW; ++i) {
#pragma unroll
        for (int c = 0; c < i; ++c) x[i] -= z.T[(r0 + c) * TP + r0 + i] * x[c];
        x[i] *= z.rv[r0 + i];
      }
#pragma unroll
      for (int i = 0; i < PW; ++i) {  // stores after the block's loads
        z.X[(r0 + i) * CB + l] = x[i];
        D[(r0 + i) * CB + l] = x[i];
      }
    }
    __syncthreads();
  }
#pragma unroll 4
  for (int r = w; r < CB; r += NW) {
    const double v = z.T[l * TP + r];  // read unconditionally: a guarded load would wait per row
    blk[r * lda + l] = (l <= r) ? v : 0.0;
  }
  if (t == 0) {
    int cur = (k0 == 0) ? 0 : *info;
    if (cur == 0 && bad >= 0) cur = (int)(k0 + bad + 1);
    *info = cur;
  }
}

// the first diagonal block (the others are factored by the trailing update
// that finishes them)
__global__ void __launch_bounds__(64 * NW) k_potrf_diag(double* __restrict__ A, int64_t lda, int64_t k0,
                                                        double* __restrict__ D, int* __restrict__ info) {
  __shared__ __align__(16) DiagLds z;
  const int t = threadIdx.x, l = t & 63;
  double* blk = A + k0 * lda + k0;
  for (int r = t >> 6; r < CB; r += NW) z.T[l * TP + r] = blk[r * lda + l];
  __syncthreads();
  diag_factor(z, blk, lda, k0, D, info, t);
}

// L_ik = A_ik D^T, D = L_kk^-1; block row i = kb + 1 + blockIdx.x
__global__ void __launch_bounds__(256) k_potrf_panel(double* __restrict__ A, int64_t lda, int64_t k0,
                                                     const double* __restrict__ D) {
  __shared__ double As[CB][CB + 1];  // As[m][r] = A_ik[r][m]
  __shared__ double Ds[CB][CB + 1];  // Ds[m][c] = D[c][m]
  const int64_t r0 = k0 + CB * (1 + (int64_t)blockIdx.x);
  double* blk = A + r0 * lda + k0;
  const int t = threadIdx.x;
  for (int e = t; e < CB * CB; e += 256) {
    const int r = e >> 6, m = e & 63;
    As[m][r] = blk[r * lda + m];
    Ds[m][r] = D[e];
  }
  __syncthreads();
  const int tr = t >> 4, tc = t & 15;  // rows tr + 16 p, columns tc + 16 q (interleaved)
  double acc[4][4] = {};
  for (int m = 0; m < CB; ++m) {
    double av[4], dv[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      av[q] = As[m][tr + 16 * q];
      dv[q] = Ds[m][tc + 16 * q];
    }
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[p][q] += av[p] * dv[q];
  }
#pragma unroll
  for (int p = 0; p < 4; ++p)
#pragma unroll
    for (int q = 0; q < 4; ++q) blk[(tr + 16 * p) * lda + tc + 16 * q] = acc[p][q];
}

// A_ij -= L_ik L_jk^T over the lower blocks kb < j <= i (triangular index).
// Workgroup 0 holds block (kb+1, kb+1), which this update finishes: its four
// wavefronts then factor it (diag_factor) while the other workgroups are
// still updating, so the next step's diagonal factor is off the critical
// path whenever the update takes longer.
static_assert(64 * NW == 256, "k_potrf_syrk's workgroup runs diag_factor");
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) k_potrf_syrk(double* __restrict__ A, int64_t lda, int64_t k0,
                                                    double* __restrict__ D, int* __restrict__ info) {
  constexpr size_t kUpd = 2 * CB * (CB + 1) * sizeof(double);
  __shared__ __align__(16) unsigned char smem[kUpd > sizeof(DiagLds) ? kUpd : sizeof(DiagLds)];
  double* sm = (double*)smem;
  double (*Ls)[CB + 1] = (double (*)[CB + 1])sm;              // Ls[m][r] = L_ik[r][m]
  double (*Rs)[CB + 1] = (double (*)[CB + 1])(sm + CB * (CB + 1));  // Rs[m][c] = L_jk[c][m]
  const int64_t e = blockIdx.x;
  int64_t bi = (int64_t)((sqrt(8.0 * (double)e + 1.0) - 1.0) * 0.5);
  while (bi * (bi + 1) / 2 > e) --bi;
  while ((bi + 1) * (bi + 2) / 2 <= e) ++bi;
  const int64_t bj = e - bi * (bi + 1) / 2;
  const int64_t i0 = k0 + CB * (1 + bi), j0 = k0 + CB * (1 + bj);
  const int t = threadIdx.x;
  // thread t: rows (t >> 4) + 16 p, columns (t & 15) + 16 q of the block
  // (interleaved: a wavefront's LDS reads hit consecutive banks, its global
  // accesses consecutive addresses)
  const int tr = t >> 4, tc = t & 15;
  double* c = A + i0 * lda + j0;
  double cv[4][4];  // this thread's entries of block (i, j), loaded with the L blocks (one latency)
#pragma unroll
  for (int p = 0; p < 4; ++p)
#pragma unroll
    for (int q = 0; q < 4; ++q) cv[p][q] = c[(tr + 16 * p) * lda + tc + 16 * q];
  for (int q = t; q < CB * CB; q += 256) {
    const int r = q >> 6, m = q & 63;
    Ls[m][r] = A[(i0 + r) * lda + k0 + m];
    Rs[m][r] = A[(j0 + r) * lda + k0 + m];
  }
  __syncthreads();
  double acc[4][4] = {};
#pragma unroll 4
  for (int m = 0; m < CB; ++m) {
    double lv[4], rv[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      lv[q] = Ls[m][tr + 16 * q];
      rv[q] = Rs[m][tc + 16 * q];
    }
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[p][q] += lv[p] * rv[q];
  }
  const bool fuse = (e == 0);  // block (kb+1, kb+1)
  DiagLds& z = *(DiagLds*)smem;  // reuses Ls/Rs once they are read
  if (fuse) __syncthreads();
#pragma unroll
  for (int p = 0; p < 4; ++p)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const double v = cv[p][q] - acc[p][q];
      c[(tr + 16 * p) * lda + tc + 16 * q] = v;
      if (fuse) z.T[(tc + 16 * q) * TP + tr + 16 * p] = v;
    }
  if (fuse) {
    __syncthreads();
    diag_factor(z, c, lda, k0 + CB, D, info, t);  // all 256 threads
  }
}

extern "C" int rb_potrf_workspace(int64_t n, size_t* bytes) {
  if (!bytes || n <= 0 || n % CB) return RB_ERR_ARG;
  *bytes = (size_t)CB * CB * sizeof(double);
  return RB_OK;
}

extern "C" int rb_potrf(int64_t n, double* A, int64_t lda, void* work, size_t work_bytes, int* info,
                        void* stream) {
  hipStream_t s = (hipStream_t)stream;
  if (n <= 0 || n % CB || lda < n || !A || !info) {
    rb_set_error("rb_potrf: needs n a positive multiple of %d and lda >= n", CB);
    return RB_ERR_ARG;
  }
  if (!work || work_bytes < (size_t)CB * CB * sizeof(double)) {
    rb_set_error("rb_potrf: workspace too small");
    return RB_ERR_ARG;
  }
  double* D = (double*)work;
  const int64_t nbk = n / CB;
  hipLaunchKernelGGL(k_potrf_diag, dim3(1), dim3(64 * NW), 0, s, A, lda, (int64_t)0, D, info);
  RB_CHECK_LAUNCH();
  for (int64_t kb = 0; kb + 1 < nbk; ++kb) {
    const int64_t k0 = kb * CB, rest = nbk - kb - 1;
    hipLaunchKernelGGL(k_potrf_panel, dim3((unsigned)rest), dim3(256), 0, s, A, lda, k0, (const double*)D);
    RB_CHECK_LAUNCH();
    // also factors diagonal block kb + 1 (into D, for the next panel)
    hipLaunchKernelGGL(k_potrf_syrk, dim3((unsigned)(rest * (rest + 1) / 2)), dim3(256), 0, s, A, lda, k0, D,
                       info);
    RB_CHECK_LAUNCH();
  }
  return RB_OK;
}
