// K10: Cholesky factorisation A = L L^T of a dense float64 SPD matrix, blocked
// right-looking in 64 x 64 blocks, in place in the lower triangle.  Per block
// column kb:
//   k_potrf_diag  one wavefront factors A_kk (lane i holds row i in registers,
//                 each step's column broadcast through LDS) and forms
//                 D = L_kk^-1 by forward substitution (lane j solves column j)
//   k_potrf_panel L_ik = A_ik D^T for every block row i below (a workgroup each)
//   k_potrf_syrk  A_ij -= L_ik L_jk^T for kb < j <= i (a workgroup per block);
//                 the workgroup of block (kb+1, kb+1) then factors it
// The Gauss-Newton step of redbirdpy solves (J J^T + lambda I) x = r with a
// dense direct solve (recon.py:429-509); this factor replaces that solve.
#include <math.h>
#include "rb_common.h"
#include "../../include/rbhip.h"

#define CB 64  // block size

__device__ __forceinline__ double rl64(double x, int lane) {
  const long long u = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)(u & 0xffffffffll), lane);
  const int hi = __builtin_amdgcn_readlane((int)(u >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Factor one diagonal block held in LDS (s: its 64 rows; cv: 64 doubles of
// scratch) with one wavefront: lane i keeps row i in registers, each step's
// column is broadcast through LDS.  Writes L_kk over the block (zero above
// the diagonal), D = L_kk^-1 and the first bad pivot into *info.  Only this
// wavefront runs it: LDS traffic inside one wavefront needs no s_barrier (its
// LDS instructions complete in order); wave_barrier keeps the compiler from
// moving memory operations across the exchange points.
__device__ __forceinline__ void diag_factor(double (*s)[CB + 2], double* cv, double* __restrict__ blk, int64_t lda,
                                            int64_t k0, double* __restrict__ D, int* __restrict__ info, int l) {
  double a[CB];
#pragma unroll
  for (int j = 0; j < CB; ++j) a[j] = s[l][j];
  int bad = -1;
  double myri = 0.0;  // lane l: 1 / L_ll
#pragma unroll
  for (int k = 0; k < CB; ++k) {
    const double dkk = rl64(a[k], k);
    if (bad < 0 && !(dkk > 0.0)) bad = k;
    // 1/sqrt: hardware estimate plus two Newton steps (a few ulp; the
    // correctly rounded sqrt and divide are long dependent sequences)
    double ri = __builtin_amdgcn_rsq(dkk);
    ri = ri * (1.5 - 0.5 * dkk * ri * ri);
    ri = ri * (1.5 - 0.5 * dkk * ri * ri);
    a[k] *= ri;  // lane k: L_kk; lanes i > k: L_ik (lanes i < k hold unused upper entries)
    cv[l] = a[k];
    if (l == k) myri = ri;
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int j = k + 1; j < CB; ++j) a[j] -= a[k] * cv[j];
    __builtin_amdgcn_wave_barrier();
  }
#pragma unroll
  for (int j = 0; j < CB; ++j) s[l][j] = (j <= l) ? a[j] : 0.0;  // row l of L_kk
  __builtin_amdgcn_wave_barrier();
  // D = L_kk^-1: lane j solves column j (four partial sums shorten the chain)
  double x[CB];
#pragma unroll
  for (int r = 0; r < CB; ++r) {
    double acc[4] = {(r == l) ? 1.0 : 0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < r; ++k) acc[k & 3] -= s[r][k] * x[k];
    x[r] = ((acc[0] + acc[1]) + (acc[2] + acc[3])) * rl64(myri, r);
    __builtin_amdgcn_sched_barrier(0);  // keep row r's loads from piling up ahead
  }
#pragma unroll
  for (int r = 0; r < CB; ++r) D[r * CB + l] = x[r];
#pragma unroll 16
  for (int r = 0; r < CB; ++r) blk[r * lda + l] = s[r][l];
  if (l == 0) {
    int cur = (k0 == 0) ? 0 : *info;
    if (cur == 0 && bad >= 0) cur = (int)(k0 + bad + 1);
    *info = cur;
  }
}

// the first diagonal block (the others are factored by the trailing update
// that finishes them)
__global__ void __launch_bounds__(64) k_potrf_diag(double* __restrict__ A, int64_t lda, int64_t k0,
                                                   double* __restrict__ D, int* __restrict__ info) {
  __shared__ __align__(16) double s[CB][CB + 2];  // even row stride: 16-byte aligned rows
  __shared__ __align__(16) double cv[CB];
  const int l = threadIdx.x;
  double* blk = A + k0 * lda + k0;
#pragma unroll 16
  for (int r = 0; r < CB; ++r) s[r][l] = blk[r * lda + l];
  __syncthreads();
  diag_factor(s, cv, blk, lda, k0, D, info, l);
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
  const int tr = (t >> 4) * 4, tc = (t & 15) * 4;
  double acc[4][4] = {};
  for (int m = 0; m < CB; ++m) {
    double av[4], dv[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      av[q] = As[m][tr + q];
      dv[q] = Ds[m][tc + q];
    }
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[p][q] += av[p] * dv[q];
  }
#pragma unroll
  for (int p = 0; p < 4; ++p)
#pragma unroll
    for (int q = 0; q < 4; ++q) blk[(tr + p) * lda + tc + q] = acc[p][q];
}

// A_ij -= L_ik L_jk^T over the lower blocks kb < j <= i (triangular index).
// Workgroup 0 holds block (kb+1, kb+1), which this update finishes: its
// first wavefront then factors it (diag_factor) while the other workgroups
// are still updating, so the next step's diagonal factor is off the
// critical path whenever the update takes longer.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) k_potrf_syrk(double* __restrict__ A, int64_t lda, int64_t k0,
                                                    double* __restrict__ D, int* __restrict__ info) {
  __shared__ __align__(16) double sm[2 * CB * (CB + 1)];
  double (*Ls)[CB + 1] = (double (*)[CB + 1])sm;              // Ls[m][r] = L_ik[r][m]
  double (*Rs)[CB + 1] = (double (*)[CB + 1])(sm + CB * (CB + 1));  // Rs[m][c] = L_jk[c][m]
  const int64_t e = blockIdx.x;
  int64_t bi = (int64_t)((sqrt(8.0 * (double)e + 1.0) - 1.0) * 0.5);
  while (bi * (bi + 1) / 2 > e) --bi;
  while ((bi + 1) * (bi + 2) / 2 <= e) ++bi;
  const int64_t bj = e - bi * (bi + 1) / 2;
  const int64_t i0 = k0 + CB * (1 + bi), j0 = k0 + CB * (1 + bj);
  const int t = threadIdx.x;
  for (int q = t; q < CB * CB; q += 256) {
    const int r = q >> 6, m = q & 63;
    Ls[m][r] = A[(i0 + r) * lda + k0 + m];
    Rs[m][r] = A[(j0 + r) * lda + k0 + m];
  }
  __syncthreads();
  const int tr = (t >> 4) * 4, tc = (t & 15) * 4;
  double acc[4][4] = {};
  for (int m = 0; m < CB; ++m) {
    double lv[4], rv[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      lv[q] = Ls[m][tr + q];
      rv[q] = Rs[m][tc + q];
    }
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[p][q] += lv[p] * rv[q];
  }
  double* c = A + i0 * lda + j0;
  const bool fuse = (e == 0);  // block (kb+1, kb+1)
  double (*S)[CB + 2] = (double (*)[CB + 2])sm;  // reuses Ls/Rs once they are read
  double* cv = sm + CB * (CB + 2);
  if (fuse) __syncthreads();
#pragma unroll
  for (int p = 0; p < 4; ++p)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const double v = c[(tr + p) * lda + tc + q] - acc[p][q];
      c[(tr + p) * lda + tc + q] = v;
      if (fuse) S[tr + p][tc + q] = v;
    }
  if (fuse) {
    __syncthreads();
    if (t < 64) diag_factor(S, cv, c, lda, k0 + CB, D, info, t);
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
  hipLaunchKernelGGL(k_potrf_diag, dim3(1), dim3(64), 0, s, A, lda, (int64_t)0, D, info);
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
