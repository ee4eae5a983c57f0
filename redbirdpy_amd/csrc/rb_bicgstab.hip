// Multi-RHS BiCGSTAB (van der Vorst 1992) with right Jacobi preconditioning,
// float64 / complex128, for matrices that are not (complex) symmetric and
// as the fallback when COCG breaks down (redbirdpy solver.py:791-824,
// method 'bicgstab').  Every column runs its own scalar recurrence; all
// columns share one launch sequence per iteration:
//
//   p  = r + beta (p - omega v),  ph = D^-1 p
//   v  = A ph                                        (rb_spmm)
//   alpha = rho / (rh, v)
//   s  = r - alpha v,  sh = D^-1 s                   (s overwrites r)
//   t  = A sh                                        (rb_spmm)
//   omega = (t, s) / (t, t)
//   x += alpha ph + omega sh,  r = s - omega t
//   rho' = (rh, r), beta = (rho' / rho) (alpha / omega), stop at |r| <= tol |b|
//
// Inner products are conjugated ((a, b) = sum conj(a) b).  Right
// preconditioning keeps r the true residual of A x = b.  Reductions are
// per-block partials summed by one finish block per column in a fixed
// order, so results are deterministic.
#include "rb_common.h"
#include "../../include/rbhip.h"
#include <vector>
#include <algorithm>
#include <math.h>

#define BS_KB 256      // threads per block: 64 columns x 4 rows
#define BS_NB 1024     // blocks per vector kernel (grid-stride over rows)
#define BS_FB 256      // threads per finish block
#define BS_NDOT 3

namespace {

__device__ inline double cdot(double a, double b) { return a * b; }
__device__ inline cplx<double> cdot(cplx<double> a, cplx<double> b) {
  return mk<double>(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);  // conj(a) b
}

template <typename T>
struct BiState {
  T* rho;
  T* alpha;
  T* omega;
  T* beta;
  double* rr;
  double* bb;
  int* active;
  int* iters;
  int* flag;
  int* nactive;
  T* part;  // [d][c][block]
  int nb;
  double tol2;
  int maxiter;
};

// per-thread column sums pv[d] (column c of this thread) -> part[d][c][block]
template <typename T>
__device__ void bs_partials(const BiState<T>& st, const T (&pv)[BS_NDOT], int nd, int C, int c) {
  __shared__ T red[BS_NDOT][4][64];
  const int lane = threadIdx.x & 63, rg = threadIdx.x >> 6;
  for (int d = 0; d < nd; ++d) red[d][rg][lane] = pv[d];
  __syncthreads();
  if (rg == 0 && c < C) {
    for (int d = 0; d < nd; ++d) {
      T a = red[d][0][lane];
      for (int r = 1; r < 4; ++r) a = a + red[d][r][lane];
      st.part[((int64_t)d * C + c) * st.nb + blockIdx.x] = a;
    }
  }
}

// v[d] = sum over blocks of part[d][c][.] in block order
template <typename T>
__device__ void bs_finish(const BiState<T>& st, int nd, int C, int c, T* v) {
  __shared__ T red[BS_FB];
  for (int d = 0; d < nd; ++d) {
    T a = sc<T>::zero();
    const T* pp = st.part + ((int64_t)d * C + c) * st.nb;
    for (int b = threadIdx.x; b < st.nb; b += BS_FB) a = a + pp[b];
    red[threadIdx.x] = a;
    __syncthreads();
    for (int o = BS_FB / 2; o > 0; o >>= 1) {
      if (threadIdx.x < o) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + o];
      __syncthreads();
    }
    v[d] = red[0];
    __syncthreads();
  }
}

// grid: (BS_NB, column chunks of 64); thread (lane, rg) walks rows rg, rg+4, ...
#define BS_ROWS(n)                                                                  \
  const int lane = threadIdx.x & 63, rg = threadIdx.x >> 6;                         \
  const int c = blockIdx.y * 64 + lane;                                             \
  const bool live = c < C && st.active[c];                                          \
  for (int64_t row = (int64_t)blockIdx.x * 4 + rg; row < (n); row += (int64_t)gridDim.x * 4)

template <typename T>
__global__ void k_bs_dinv(int64_t n, const T* vals, const int32_t* diag, T* dinv) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  dinv[i] = tdiv(one_of(T()), vals[diag[i]]);
}

// r = b - Ax (AX may be null: x = 0), rh = r, p = v = 0; partials (rh,r), (r,r), (b,b)
template <typename T>
__global__ void __launch_bounds__(BS_KB) k_bs_init(int64_t n, int C, const T* B, const T* AX, T* X, T* R, T* RH,
                                                   T* P, T* V, BiState<T> st) {
  const int lane = threadIdx.x & 63, rg = threadIdx.x >> 6;
  const int c = blockIdx.y * 64 + lane;
  T pv[BS_NDOT] = {sc<T>::zero(), sc<T>::zero(), sc<T>::zero()};
  if (c < C) {
    for (int64_t row = (int64_t)blockIdx.x * 4 + rg; row < n; row += (int64_t)gridDim.x * 4) {
      const int64_t o = row * C + c;
      const T b = B[o];
      const T r = AX ? b - AX[o] : b;
      if (!AX) X[o] = sc<T>::zero();
      R[o] = r;
      RH[o] = r;
      P[o] = sc<T>::zero();
      V[o] = sc<T>::zero();
      pv[0] = pv[0] + cdot(r, r);
      pv[1] = pv[1] + acc_from_real(sc<T>::norm2(r), T());
      pv[2] = pv[2] + acc_from_real(sc<T>::norm2(b), T());
    }
  }
  bs_partials(st, pv, 3, C, c);
}

template <typename T>
__global__ void __launch_bounds__(BS_FB) k_bs_fin_init(int C, BiState<T> st) {
  const int c = blockIdx.x;
  T v[3];
  bs_finish(st, 3, C, c, v);
  if (threadIdx.x == 0) {
    const double rr = acc_re(v[1]), bb = acc_re(v[2]);
    st.rho[c] = v[0];
    st.rr[c] = rr;
    st.bb[c] = bb;
    st.alpha[c] = one_of(T());
    st.omega[c] = one_of(T());
    st.beta[c] = sc<T>::zero();
    st.iters[c] = 0;
    st.flag[c] = 0;
    st.active[c] = bb > 0.0 && rr > st.tol2 * bb;
  }
}

// p = r + beta (p - omega v), ph = dinv p
template <typename T>
__global__ void __launch_bounds__(BS_KB) k_bs_p(int64_t n, int C, const T* dinv, const T* R, T* P, const T* V,
                                                T* PH, BiState<T> st) {
  const int lane = threadIdx.x & 63, rg = threadIdx.x >> 6;
  const int c = blockIdx.y * 64 + lane;
  if (c >= C || !st.active[c]) return;
  const T be = st.beta[c], om = st.omega[c];
  for (int64_t row = (int64_t)blockIdx.x * 4 + rg; row < n; row += (int64_t)gridDim.x * 4) {
    const int64_t o = row * C + c;
    const T p = R[o] + be * (P[o] - om * V[o]);
    P[o] = p;
    PH[o] = dinv[row] * p;
  }
}

// partial (rh, v)
template <typename T>
__global__ void __launch_bounds__(BS_KB) k_bs_dot_rv(int64_t n, int C, const T* RH, const T* V, BiState<T> st) {
  T pv[BS_NDOT] = {sc<T>::zero(), sc<T>::zero(), sc<T>::zero()};
  BS_ROWS(n) {
    if (live) {
      const int64_t o = row * C + c;
      pv[0] = pv[0] + cdot(RH[o], V[o]);
    }
  }
  bs_partials(st, pv, 1, C, c);
}

template <typename T>
__global__ void __launch_bounds__(BS_FB) k_bs_fin_alpha(int C, BiState<T> st) {
  const int c = blockIdx.x;
  if (!st.active[c]) return;
  T v[1];
  bs_finish(st, 1, C, c, v);
  if (threadIdx.x == 0) {
    if (acc_abs2(v[0]) == 0.0) {  // (rh, v) = 0: breakdown
      st.active[c] = 0;
      st.flag[c] = 4;
      st.alpha[c] = sc<T>::zero();
    } else {
      st.alpha[c] = tdiv(st.rho[c], v[0]);
    }
  }
}

// s = r - alpha v (into R), sh = dinv s
template <typename T>
__global__ void __launch_bounds__(BS_KB) k_bs_s(int64_t n, int C, const T* dinv, T* R, const T* V, T* SH,
                                                BiState<T> st) {
  const int lane = threadIdx.x & 63, rg = threadIdx.x >> 6;
  const int c = blockIdx.y * 64 + lane;
  if (c >= C || !st.active[c]) return;
  const T al = st.alpha[c];
  for (int64_t row = (int64_t)blockIdx.x * 4 + rg; row < n; row += (int64_t)gridDim.x * 4) {
    const int64_t o = row * C + c;
    const T s = R[o] - al * V[o];
    R[o] = s;
    SH[o] = dinv[row] * s;
  }
}

// partials (t, s), (t, t)
template <typename T>
__global__ void __launch_bounds__(BS_KB) k_bs_dot_ts(int64_t n, int C, const T* Tv, const T* S, BiState<T> st) {
  T pv[BS_NDOT] = {sc<T>::zero(), sc<T>::zero(), sc<T>::zero()};
  BS_ROWS(n) {
    if (live) {
      const int64_t o = row * C + c;
      const T t = Tv[o];
      pv[0] = pv[0] + cdot(t, S[o]);
      pv[1] = pv[1] + acc_from_real(sc<T>::norm2(t), T());
    }
  }
  bs_partials(st, pv, 2, C, c);
}

template <typename T>
__global__ void __launch_bounds__(BS_FB) k_bs_fin_omega(int C, BiState<T> st) {
  const int c = blockIdx.x;
  if (!st.active[c]) return;
  T v[2];
  bs_finish(st, 2, C, c, v);
  if (threadIdx.x == 0) {
    const double tt = acc_re(v[1]);
    st.omega[c] = tt > 0.0 ? tdiv(v[0], acc_from_real(tt, T())) : sc<T>::zero();
  }
}

// x += alpha ph + omega sh, r = s - omega t; partials (rh, r), (r, r)
template <typename T>
__global__ void __launch_bounds__(BS_KB) k_bs_x(int64_t n, int C, T* X, T* R, const T* RH, const T* PH,
                                                const T* SH, const T* Tv, BiState<T> st) {
  T pv[BS_NDOT] = {sc<T>::zero(), sc<T>::zero(), sc<T>::zero()};
  BS_ROWS(n) {
    if (live) {
      const int64_t o = row * C + c;
      const T al = st.alpha[c], om = st.omega[c];
      X[o] = X[o] + al * PH[o] + om * SH[o];
      const T r = R[o] - om * Tv[o];
      R[o] = r;
      pv[0] = pv[0] + cdot(RH[o], r);
      pv[1] = pv[1] + acc_from_real(sc<T>::norm2(r), T());
    }
  }
  bs_partials(st, pv, 2, C, c);
}

template <typename T>
__global__ void __launch_bounds__(BS_FB) k_bs_fin_beta(int C, BiState<T> st) {
  const int c = blockIdx.x;
  if (!st.active[c]) return;
  T v[2];
  bs_finish(st, 2, C, c, v);
  if (threadIdx.x == 0) {
    const double rr = acc_re(v[1]);
    st.rr[c] = rr;
    const int it = st.iters[c] + 1;
    st.iters[c] = it;
    if (rr <= st.tol2 * st.bb[c]) {
      st.active[c] = 0;
    } else if (it >= st.maxiter) {
      st.active[c] = 0;
      st.flag[c] = 1;
    } else if (acc_abs2(v[0]) == 0.0 || acc_abs2(st.omega[c]) == 0.0) {
      st.active[c] = 0;
      st.flag[c] = 4;
    } else {
      st.beta[c] = tdiv(v[0], st.rho[c]) * tdiv(st.alpha[c], st.omega[c]);
      st.rho[c] = v[0];
    }
  }
}

__global__ void k_bs_count(const int* active, int C, int* out) {
  __shared__ int s;
  if (threadIdx.x == 0) s = 0;
  __syncthreads();
  for (int c = threadIdx.x; c < C; c += blockDim.x)
    if (active[c]) atomicAdd(&s, 1);
  __syncthreads();
  if (threadIdx.x == 0) *out = s;
}

struct BsLayout {
  size_t off = 0;
  size_t take(size_t b) {
    size_t o = off;
    off += (b + 255) & ~(size_t)255;
    return o;
  }
};

template <typename T>
static size_t bs_layout(int64_t n, int C, size_t* o) {
  BsLayout L;
  for (int v = 0; v < 7; ++v) o[v] = L.take((size_t)n * C * sizeof(T));  // r rh p ph v sh t
  o[7] = L.take((size_t)n * sizeof(T));                                    // dinv
  o[8] = L.take((size_t)BS_NDOT * C * BS_NB * sizeof(T));                  // partials
  for (int v = 0; v < 4; ++v) o[9 + v] = L.take(C * sizeof(T));            // rho alpha omega beta
  o[13] = L.take(C * sizeof(double));                                      // rr
  o[14] = L.take(C * sizeof(double));                                      // bb
  for (int v = 0; v < 3; ++v) o[15 + v] = L.take(C * sizeof(int));         // active iters flag
  o[18] = L.take(sizeof(int));                                             // nactive
  return L.off;
}

template <typename T>
static int bs_solve(int64_t n, int64_t nnz, const int32_t* row_ptr, const int32_t* col_idx, const int32_t* diag,
                    int C, const T* vals, const T* B, T* X, double tol, int maxiter, int use_x0, int check_every,
                    char* base, int* iters_out, double* relres_out, int* flag_out, hipStream_t s) {
  size_t o[19];
  bs_layout<T>(n, C, o);
  T* R = (T*)(base + o[0]);
  T* RH = (T*)(base + o[1]);
  T* P = (T*)(base + o[2]);
  T* PH = (T*)(base + o[3]);
  T* V = (T*)(base + o[4]);
  T* SH = (T*)(base + o[5]);
  T* Tv = (T*)(base + o[6]);
  T* dinv = (T*)(base + o[7]);
  BiState<T> st;
  st.part = (T*)(base + o[8]);
  st.rho = (T*)(base + o[9]);
  st.alpha = (T*)(base + o[10]);
  st.omega = (T*)(base + o[11]);
  st.beta = (T*)(base + o[12]);
  st.rr = (double*)(base + o[13]);
  st.bb = (double*)(base + o[14]);
  st.active = (int*)(base + o[15]);
  st.iters = (int*)(base + o[16]);
  st.flag = (int*)(base + o[17]);
  st.nactive = (int*)(base + o[18]);
  st.nb = BS_NB;
  st.tol2 = tol * tol;
  st.maxiter = maxiter;
  const int dt = sc<T>::is_complex ? RB_C128 : RB_F64;
  const dim3 vg(BS_NB, (unsigned)ceil_div(C, 64));
  hipLaunchKernelGGL(k_bs_dinv<T>, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s, n, vals, diag, dinv);
  RB_CHECK_LAUNCH();
  const T* AX = nullptr;
  if (use_x0) {
    int rc = rb_spmm(n, nnz, row_ptr, col_idx, 1, C, dt, vals, X, V, s);
    if (rc) return rc;
    AX = V;
  }
  hipLaunchKernelGGL(k_bs_init<T>, vg, dim3(BS_KB), 0, s, n, C, B, AX, X, R, RH, P, V, st);
  // the kernel reads AX (= V) before zeroing V row by row: each element is read, then written, by one thread
  hipLaunchKernelGGL(k_bs_fin_init<T>, dim3(C), dim3(BS_FB), 0, s, C, st);
  RB_CHECK_LAUNCH();
  int* hn = nullptr;
  RB_HIP(hipHostMalloc((void**)&hn, sizeof(int), hipHostMallocDefault));
  auto poll = [&](int* nact) -> int {
    hipLaunchKernelGGL(k_bs_count, dim3(1), dim3(256), 0, s, st.active, C, st.nactive);
    RB_CHECK_LAUNCH();
    RB_HIP(hipMemcpyAsync(hn, st.nactive, sizeof(int), hipMemcpyDeviceToHost, s));
    RB_HIP(hipStreamSynchronize(s));
    *nact = *hn;
    return RB_OK;
  };
  int nact = 0, rc = poll(&nact), it = 0;
  while (rc == RB_OK && nact > 0 && it < maxiter) {
    const int chunk = std::min(check_every, maxiter - it);
    for (int j = 0; j < chunk && rc == RB_OK; ++j) {
      hipLaunchKernelGGL(k_bs_p<T>, vg, dim3(BS_KB), 0, s, n, C, dinv, R, P, V, PH, st);
      if ((rc = rb_spmm(n, nnz, row_ptr, col_idx, 1, C, dt, vals, PH, V, s))) break;
      hipLaunchKernelGGL(k_bs_dot_rv<T>, vg, dim3(BS_KB), 0, s, n, C, RH, V, st);
      hipLaunchKernelGGL(k_bs_fin_alpha<T>, dim3(C), dim3(BS_FB), 0, s, C, st);
      hipLaunchKernelGGL(k_bs_s<T>, vg, dim3(BS_KB), 0, s, n, C, dinv, R, V, SH, st);
      if ((rc = rb_spmm(n, nnz, row_ptr, col_idx, 1, C, dt, vals, SH, Tv, s))) break;
      hipLaunchKernelGGL(k_bs_dot_ts<T>, vg, dim3(BS_KB), 0, s, n, C, Tv, R, st);
      hipLaunchKernelGGL(k_bs_fin_omega<T>, dim3(C), dim3(BS_FB), 0, s, C, st);
      hipLaunchKernelGGL(k_bs_x<T>, vg, dim3(BS_KB), 0, s, n, C, X, R, RH, PH, SH, Tv, st);
      hipLaunchKernelGGL(k_bs_fin_beta<T>, dim3(C), dim3(BS_FB), 0, s, C, st);
    }
    if (rc) break;
    if (hipGetLastError() != hipSuccess) {
      rb_set_error("rb_bicgstab_solve: launch failed");
      rc = RB_ERR_HIP;
      break;
    }
    it += chunk;
    rc = poll(&nact);
  }
  if (rc == RB_OK) {
    std::vector<double> hrr(C), hbb(C);
    std::vector<int> hfl(C), hit(C);
    if (hipMemcpyAsync(hrr.data(), st.rr, C * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemcpyAsync(hbb.data(), st.bb, C * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemcpyAsync(hfl.data(), st.flag, C * sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemcpyAsync(hit.data(), st.iters, C * sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
      rb_set_error("rb_bicgstab_solve: reading the column states failed");
      rc = RB_ERR_HIP;
    } else {
      int fl = 0, mx = 0;
      double rel = 0;
      for (int c = 0; c < C; ++c) {
        fl = std::max(fl, hfl[c]);
        mx = std::max(mx, hit[c]);
        if (hbb[c] > 0) rel = std::max(rel, sqrt(hrr[c] / hbb[c]));
      }
      if (iters_out) *iters_out = mx;
      if (relres_out) *relres_out = rel;
      if (flag_out) *flag_out = fl;
    }
  }
  (void)hipHostFree(hn);
  return rc;
}

}  // namespace

extern "C" int rb_bicgstab_workspace(int64_t n, int ncol, int is_complex, size_t* bytes) {
  if (!bytes || n <= 0 || ncol <= 0) return RB_ERR_ARG;
  size_t o[19];
  *bytes = is_complex ? bs_layout<cplx<double>>(n, ncol, o) : bs_layout<double>(n, ncol, o);
  return RB_OK;
}

extern "C" int rb_bicgstab_solve(int64_t n, int64_t nnz, const int32_t* row_ptr, const int32_t* col_idx,
                                 const int32_t* diag, int ncol, int is_complex, const void* vals, const void* B,
                                 void* X, double tol, int maxiter, int use_x0, int check_every, void* work,
                                 size_t work_bytes, int* iters, double* relres, int* flag, void* stream) {
  if (n <= 0 || ncol <= 0 || ncol > 512 || !row_ptr || !col_idx || !diag || !vals || !B || !X || !work ||
      maxiter <= 0 || !(tol > 0)) {
    rb_set_error("rb_bicgstab_solve: bad arguments");
    return RB_ERR_ARG;
  }
  size_t need = 0;
  rb_bicgstab_workspace(n, ncol, is_complex, &need);
  if (work_bytes < need) {
    rb_set_error("rb_bicgstab_solve: workspace of %zu bytes, %zu needed", work_bytes, need);
    return RB_ERR_SIZE;
  }
  hipStream_t s = (hipStream_t)stream;
  check_every = check_every > 0 ? check_every : 16;
  if (is_complex)
    return bs_solve<cplx<double>>(n, nnz, row_ptr, col_idx, diag, ncol, (const cplx<double>*)vals,
                                  (const cplx<double>*)B, (cplx<double>*)X, tol, maxiter, use_x0, check_every,
                                  (char*)work, iters, relres, flag, s);
  return bs_solve<double>(n, nnz, row_ptr, col_idx, diag, ncol, (const double*)vals, (const double*)B, (double*)X,
                          tol, maxiter, use_x0, check_every, (char*)work, iters, relres, flag, s);
}
