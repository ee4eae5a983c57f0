// K2-K4: multi-RHS Krylov solves over one shared CSR pattern.
//
// Every column of the n x C right-hand-side block runs its own Jacobi-
// preconditioned CG (real SPD) or COCG (complex symmetric, unconjugated
// inner products) recurrence; all columns share one SpMM per iteration,
// which streams the CSR arrays once for all sources and detectors.  Column
// c belongs to system c / k, so several matrices with the same pattern
// (wavelengths) are solved in the same launches.
//
// Layout: X, R, P, Q are row-major n x C (ld = C), so one CSR row gathers
// contiguous C-wide rows of P; a group of G lanes (G = 64 when C > 16)
// covers the columns of one row in NCH chunks.  Per-column dot products are
// reduced in double precision without atomics or fences: lanes -> LDS ->
// per-block partials (plain stores) -> a finish kernel with one block per
// column that sums the block partials in a fixed tree order at the next
// kernel boundary.  Every run gives the same bits, and no kernel issues an
// agent-scope fence (which on gfx950 writes back and invalidates the whole
// L2 that the SpMM gather relies on).
//
// Precision modes: plain (everything in the 64-bit type) and mixed (inner
// solves in the 32-bit type on a 32-bit copy of A; outer FP64 residual
// r = b - A x and refinement x += d until ||r_c|| <= tol ||b_c||).
#include <hip/hip_ext.h>
#include "rb_common.h"
#include "../../include/rbhip.h"
#include <vector>
#include <cstring>
#include <atomic>
#include <mutex>
#include <stdlib.h>
#include <algorithm>
#include <type_traits>
#include <math.h>

#define KB 256       // threads per block
#define MAXND 3      // dot products per kernel
#define MAXCHUNK 8   // 64-column chunks of one launch (C <= 512)
#define NBMAX 2048   // blocks per launch (grid-stride over rows)
#define FB 512       // threads of a finish block
#define UNR 8        // CSR entries gathered per step

template <typename A>
struct ColState {
  A* rho;
  A* alpha;
  A* beta;
  double* bb;
  double* rr;
  int* active;
  int* iters;
  int* flag;
  int* nactive;
  A* part;  // [MAXND][C][nblocks]
  int nb;   // producer blocks whose partials are valid
  double tol2;
  int maxiter;
  // mixed precision (reliable updates): rho before the last beta, squared
  // true residual at the last replacement, and the replacement request flag
  A* rho_prev;
  double* rmax2;
  // replacement requests, one flag per 64-column chunk, double-buffered by
  // iteration parity: iteration k's beta step raises ru_req[MAXCHUNK*par + ch]
  // and iteration k+1 (parity par^1 on its launches) serves them
  int* ru_req;
  int ru_par;
  double delta2;  // request a replacement when rr <= delta2 * rmax2 (0: plain CG)
  double tol2ru;  // ... or when every active column of the chunk has rr <= tol2ru * bb
  // per chunk: [0, MAXCHUNK) columns arrived in this k_fin_beta launch,
  // [MAXCHUNK, 2 MAXCHUNK) active columns above the tolerance, [2 MAXCHUNK,
  // 3 MAXCHUNK) active columns below it; the chunk's last column resets them
  int* ru_cnt;
  int poly;  // polynomial preconditioner: a replacement keeps rho (the next z brings the new residual in)
};

template <typename A>
__device__ inline int* ru_cur(const ColState<A>& st) { return st.ru_req + MAXCHUNK * st.ru_par; }
template <typename A>
__device__ inline const int* ru_prev(const ColState<A>& st) { return st.ru_req + MAXCHUNK * (st.ru_par ^ 1); }

struct Csr {
  const int32_t* row_ptr;
  const int32_t* col_idx;
  const int32_t* diag;
  int64_t n, nnz;
};

// per-thread partials pv[ch][d] (column lane + ch*G) -> part[d][c][blockIdx]
template <typename A, int G, int NCH>
__device__ void block_partials(const ColState<A>& st, A (&pv)[NCH][MAXND], int nd, int C) {
  __shared__ A lds[KB * MAXND];
  const int lane = threadIdx.x % G;
  const int rg = threadIdx.x / G;
  constexpr int RPB = KB / G;
#pragma unroll
  for (int ch = 0; ch < NCH; ++ch) {
    for (int d = 0; d < nd; ++d) lds[(rg * MAXND + d) * G + lane] = pv[ch][d];
    __syncthreads();
    for (int i = threadIdx.x; i < nd * G; i += KB) {
      int d = i / G, l = i - d * G;
      int c = l + ch * G;
      if (c < C) {
        A acc = lds[d * G + l];
        for (int r = 1; r < RPB; ++r) acc = acc_add(acc, lds[(r * MAXND + d) * G + l]);
        st.part[(int64_t)(d * C + c) * st.nb + blockIdx.x] = acc;
      }
    }
    __syncthreads();
  }
}

// true when some column is still active (read by every producer block)
__device__ inline bool block_any_active(const int* active, int C) {
  __shared__ int s_any;
  if (threadIdx.x == 0) s_any = 0;
  __syncthreads();
  for (int c = threadIdx.x; c < C; c += blockDim.x)
    if (active[c]) s_any = 1;
  __syncthreads();
  return s_any != 0;
}

// sum over the 64 lanes of a wavefront; the xor butterfly adds the same two
// operands on both lanes of every pair, so all lanes end with the same bits
__device__ inline double wave_sum(double x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  return x;
}
__device__ inline cplx<double> wave_sum(cplx<double> x) { return mk<double>(wave_sum(x.x), wave_sum(x.y)); }

// finish-block helper: v[d] = sum over producer blocks of part[b][d][c].
// Every thread issues all of its partial loads (FB apart) before adding, the
// wavefronts reduce by butterfly and wave 0 adds the FB/64 wave sums in
// order: a fixed summation order for a given producer grid.
template <typename A>
__device__ void finish_reduce(const ColState<A>& st, int nd, int C, int c, A* v) {
  constexpr int NW = FB / 64;
  constexpr int PER = 4;  // partials per thread and dot held in flight
  __shared__ A red[MAXND][NW];
  const int nb = st.nb;
  const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
  A acc[MAXND];
#pragma unroll
  for (int d = 0; d < MAXND; ++d) acc[d] = acc_zero(A());
  for (int b0 = 0; b0 < nb; b0 += PER * FB) {
    A x[MAXND][PER];
#pragma unroll
    for (int d = 0; d < MAXND; ++d) {
      const A* pp = st.part + (int64_t)(d * C + c) * nb;
#pragma unroll
      for (int u = 0; u < PER; ++u) {
        const int b = b0 + u * FB + threadIdx.x;
        x[d][u] = (d < nd && b < nb) ? pp[b] : acc_zero(A());
      }
    }
#pragma unroll
    for (int d = 0; d < MAXND; ++d)
#pragma unroll
      for (int u = 0; u < PER; ++u) acc[d] = acc_add(acc[d], x[d][u]);
  }
#pragma unroll
  for (int d = 0; d < MAXND; ++d) {
    A t = wave_sum(acc[d]);
    if (lane == 0 && d < nd) red[d][w] = t;
  }
  __syncthreads();
  for (int d = 0; d < nd; ++d) {
    A t = red[d][0];
    for (int i = 1; i < NW; ++i) t = acc_add(t, red[d][i]);
    v[d] = t;
  }
  __syncthreads();
}

// y[ch] = sum_j A_sj x[col_j, lane + ch*G]; UNR column indices and values are
// fetched before their gathers so the loads of one step are in flight together
template <typename T, int G, int NCH>
__device__ inline void row_gather(const int32_t* __restrict__ col_idx, const T* __restrict__ vals, int32_t rs,
                                  int32_t re, const T* __restrict__ X, int64_t ld, int lane, int C, T (&y)[NCH]) {
#pragma unroll
  for (int ch = 0; ch < NCH; ++ch) y[ch] = sc<T>::zero();
  int32_t j = rs;
  for (; j + UNR <= re; j += UNR) {
    int32_t cc[UNR];
    T a[UNR];
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      cc[u] = col_idx[j + u];
      a[u] = vals[j + u];
    }
#pragma unroll
    for (int ch = 0; ch < NCH; ++ch) {
      int c = lane + ch * G;
      if (c < C) {
        T x[UNR];
#pragma unroll
        for (int u = 0; u < UNR; ++u) x[u] = X[(int64_t)cc[u] * ld + c];
#pragma unroll
        for (int u = 0; u < UNR; ++u) y[ch] = y[ch] + a[u] * x[u];
      }
    }
  }
  for (; j < re; ++j) {
    int32_t cj = col_idx[j];
    T aj = vals[j];
#pragma unroll
    for (int ch = 0; ch < NCH; ++ch) {
      int c = lane + ch * G;
      if (c < C) y[ch] = y[ch] + aj * X[(int64_t)cj * ld + c];
    }
  }
}

__device__ inline float rl_bcast(float v, int l) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l)); }
__device__ inline double rl_bcast(double v, int l) {
  int2 b = __builtin_bit_cast(int2, v);
  b.x = __builtin_amdgcn_readlane(b.x, l);
  b.y = __builtin_amdgcn_readlane(b.y, l);
  return __builtin_bit_cast(double, b);
}
template <typename R>
__device__ inline cplx<R> rl_bcast(cplx<R> v, int l) {
  return mk<R>(rl_bcast(v.x, l), rl_bcast(v.y, l));
}

#define RG 16  // gathers in flight per row step

// one row per wavefront: lanes load up to 64 (col, val) entries of the row
// at once, then readlane broadcasts feed RG masked gathers per step (a padded
// slot re-reads the row's own entry with weight 0)
template <typename T, int NCH>
__device__ inline void row_gather_w64(const int32_t* __restrict__ col_idx, const T* __restrict__ vals, int32_t rs,
                                      int32_t re, int64_t row, const T* __restrict__ X, int64_t ld, int lane,
                                      int C, T (&y)[NCH]) {
#pragma unroll
  for (int ch = 0; ch < NCH; ++ch) y[ch] = sc<T>::zero();
  for (int32_t base = rs; base < re; base += 64) {
    const int len = min(64, re - base);
    const int32_t cl = (lane < len) ? col_idx[base + lane] : (int32_t)row;
    const T vl = (lane < len) ? vals[base + lane] : sc<T>::zero();
    for (int j = 0; j < len; j += RG) {
      int32_t cc[RG];
      T a[RG];
#pragma unroll
      for (int u = 0; u < RG; ++u) {
        const int jj = min(j + u, 63);
        cc[u] = __builtin_amdgcn_readlane(cl, jj);
        a[u] = (j + u < len) ? rl_bcast(vl, jj) : sc<T>::zero();
      }
#pragma unroll
      for (int ch = 0; ch < NCH; ++ch) {
        const int c = lane + ch * 64;
        if (c < C) {
          T x[RG];
#pragma unroll
          for (int u = 0; u < RG; ++u) x[u] = X[(int64_t)cc[u] * ld + c];
#pragma unroll
          for (int u = 0; u < RG; ++u) y[ch] = y[ch] + a[u] * x[u];
        }
      }
    }
  }
}

// XCD-aware row partition: workgroups are dispatched round-robin over the
// NXCD accelerator dies (workgroup b runs on die b % NXCD), each with its own
// L2.  Die x owns the contiguous row segment x, so the neighbours a CSR row
// gathers (x/y/z lattice neighbours for a structured mesh) mostly sit in the
// same die's L2 instead of being fetched by all eight.  The grid is a
// multiple of NXCD.
#define NXCD 8
#define ROW_LOOP(nrows)                                                                    \
  const int lane = threadIdx.x % G;                                                        \
  const int rg = threadIdx.x / G;                                                          \
  constexpr int RPB = KB / G;                                                              \
  const int64_t seg_ = ((nrows) + NXCD - 1) / NXCD;                                        \
  const int64_t r0_ = (int64_t)(blockIdx.x % NXCD) * seg_;                                 \
  const int64_t r1_ = min((int64_t)(nrows), r0_ + seg_);                                   \
  const int64_t nbx_ = gridDim.x / NXCD;                                                   \
  for (int64_t row = r0_ + (int64_t)(blockIdx.x / NXCD) * RPB + rg; row < r1_; row += nbx_ * RPB)

// ---------------------------------------------------------------------------
// producer kernels
// ---------------------------------------------------------------------------
// r = b - A x0 (or b, with x = 0), z = dinv r, p = z ; dots rz, rr, bb
template <typename T, int G, int NCH>
__global__ void __launch_bounds__(KB) k_cg_init(Csr m, const T* __restrict__ vals, const T* __restrict__ dinv,
                                                const T* __restrict__ B, T* __restrict__ X, int use_x0,
                                                T* __restrict__ R, T* __restrict__ P, int C, int k,
                                                ColState<typename sc<T>::acc> st) {
  typedef typename sc<T>::acc A;
  A pv[NCH][MAXND];
#pragma unroll
  for (int ch = 0; ch < NCH; ++ch)
    for (int d = 0; d < MAXND; ++d) pv[ch][d] = acc_zero(A());
  ROW_LOOP(m.n) {
    T ax[NCH];
    if (use_x0) {
      int32_t rs = m.row_ptr[row], re = m.row_ptr[row + 1];
      row_gather<T, G, NCH>(m.col_idx, vals + 0, rs, re, X, C, lane, C, ax);
    }
#pragma unroll
    for (int ch = 0; ch < NCH; ++ch) {
      int c = lane + ch * G;
      if (c < C) {
        int s = c / k;
        int64_t o = row * C + c;
        T b = B[o];
        T r = b;
        if (use_x0) {
          if (s != 0) {  // multi-system: redo the row with this system's values
            int32_t rs = m.row_ptr[row], re = m.row_ptr[row + 1];
            T acc = sc<T>::zero();
            for (int32_t j = rs; j < re; ++j) acc = acc + vals[(int64_t)s * m.nnz + j] * X[(int64_t)m.col_idx[j] * C + c];
            ax[ch] = acc;
          }
          r = b - ax[ch];
        } else {
          X[o] = sc<T>::zero();
        }
        T z = dinv[(int64_t)s * m.n + row] * r;
        R[o] = r;
        P[o] = z;
        pv[ch][0] = acc_add(pv[ch][0], amul(r, z));
        pv[ch][1] = acc_add(pv[ch][1], acc_from_real(sc<T>::norm2(r), A()));
        pv[ch][2] = acc_add(pv[ch][2], acc_from_real(sc<T>::norm2(b), A()));
      }
    }
  }
  block_partials<A, G, NCH>(st, pv, 3, C);
}

// q = A_s p ; partial p.q
template <typename T, int G, int NCH, bool MULTI>
__global__ void __launch_bounds__(KB) k_cg_spmm(Csr m, const T* __restrict__ vals, const T* __restrict__ P,
                                                T* __restrict__ Q, int C, int k, ColState<typename sc<T>::acc> st) {
  typedef typename sc<T>::acc A;
  // this iteration's replacement requests start cleared
  if (st.ru_req && blockIdx.x == 0 && threadIdx.x < MAXCHUNK) ru_cur(st)[threadIdx.x] = 0;
  if (!block_any_active(st.active, C)) return;
  A pv[NCH][MAXND];
#pragma unroll
  for (int ch = 0; ch < NCH; ++ch) pv[ch][0] = acc_zero(A());
  ROW_LOOP(m.n) {
    int32_t rs = m.row_ptr[row], re = m.row_ptr[row + 1];
    T q[NCH];
    if (!MULTI && G == 64) {
      row_gather_w64<T, NCH>(m.col_idx, vals, rs, re, row, P, C, lane, C, q);
    } else if (!MULTI) {
      row_gather<T, G, NCH>(m.col_idx, vals, rs, re, P, C, lane, C, q);
    } else if (G == 64 && k % 64 == 0) {  // every chunk of 64 columns belongs to one system
#pragma unroll
      for (int ch = 0; ch < NCH; ++ch) {
        q[ch] = sc<T>::zero();
        if (ch * G >= C) continue;  // a chunk past the last column has no system (and no values)
        T q1[1];
        row_gather_w64<T, 1>(m.col_idx, vals + (int64_t)((ch * G) / k) * m.nnz, rs, re, row, P + ch * G, C, lane,
                             C - ch * G, q1);
        q[ch] = q1[0];
      }
    } else {
#pragma unroll
      for (int ch = 0; ch < NCH; ++ch) {
        int c = lane + ch * G;
        q[ch] = sc<T>::zero();
        if (c < C) {
          const T* v = vals + (int64_t)(c / k) * m.nnz;
          for (int32_t j = rs; j < re; ++j) q[ch] = q[ch] + v[j] * P[(int64_t)m.col_idx[j] * C + c];
        }
      }
    }
#pragma unroll
    for (int ch = 0; ch < NCH; ++ch) {
      int c = lane + ch * G;
      if (c < C) {
        int64_t o = row * C + c;
        Q[o] = q[ch];
        pv[ch][0] = acc_add(pv[ch][0], amul(P[o], q[ch]));
      }
    }
  }
  block_partials<A, G, NCH>(st, pv, 1, C);
}

// q = A_s p ; partial p.q -- the die-wide sweep.  Workgroups run round-robin
// on the NXCD dies (workgroup b on die b % NXCD) and die x owns the
// contiguous row segment x.  The segment is cut into chunks of SRC rows and
// the W wavefronts of a die take chunks w, w+W, w+2W, ...: at any moment a
// die works on a narrow band of consecutive chunks, so a P row fetched for
// one chunk is still in the die's L2 when the neighbouring chunks gather it
// (with a locality-preserving node order each P row leaves HBM about once
// per die; contiguous per-wavefront runs keep the whole segment live at once
// and re-fetch it several times).  Lane c owns column c of a 64-column
// chunk.  A chunk's CSR entries are taken in pieces of up to 64: the next
// piece's (column, value) pairs are loaded while the current one is used and
// parked in the wavefront's LDS double buffer (padding slots repeat the
// chunk's first row with value 0); each entry becomes one coalesced
// 64-column gather of P addressed by a scalar row offset through a buffer
// descriptor.  Gathers are double-buffered across batch and piece
// boundaries.  Row ends are scalar compares against the chunk's row
// pointers (rows of any length, empty ones included); the row's own p (for
// p.q) is taken from the gather of its diagonal entry.  A wavefront only
// reads LDS it wrote itself and LDS operations of one wavefront complete in
// order, so no barrier is needed.  Requires every 64-column chunk inside one
// system (nsys == 1 or k % 64 == 0) and the P block < 2 GiB.
#define SRC 4   // rows per chunk
#define SGC 12  // chunks whose row pointers one lane-parallel load covers
static_assert(SGC * (SRC + 1) <= 64, "row pointers of a chunk group must fit one wavefront");
#define BUF_WORD3 0x00020000  // raw (untyped) buffer descriptor for gfx9 parts

// gathered rows are addressed through a buffer descriptor: the row's byte
// offset is a scalar (soffset) and the lane's column offset one fixed VGPR
template <typename T>
__device__ __forceinline__ T buf_ld(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
  } else if constexpr (sizeof(T) == 8) {
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
  } else {
    static_assert(sizeof(T) == 16, "unsupported element size");
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
  }
}

// operands of one sweep: MODE 0 is q = A p (gathered vector p, output q,
// partial p.q); MODE 1 is the float64 true residual of a mixed solve,
// r32 = (TD)(b - A (x64 + d32)) with partials |r|^2, |b|^2 and r32.(dinv r32),
// for the 64-column chunks whose gate flag is set (all when gate is null)
template <typename T, typename TD>
struct SweepOps {
  const T* vals;
  const T* P;      // gathered: p (MODE 0) or x64 (MODE 1)
  const TD* D;     // MODE 1: 32-bit correction added to x64 (may be null)
  T* Q;            // MODE 0: q
  const T* B;      // MODE 1: right-hand sides
  const TD* dinv;  // MODE 1: 32-bit inverse diagonal, n per system
  TD* R32;         // MODE 1: replaced residual
  const int* gate; // MODE 1: one flag per 64-column chunk
};

// FRG: gathers per batch (two batches in flight)
template <typename T, typename TD, int NCH, int MODE, int FRG = 8>
__device__ __forceinline__ void sweep_body(const Csr& m, const SweepOps<T, TD>& op, int C, int k,
                                           const ColState<typename sc<T>::acc>& st) {
  typedef typename sc<T>::acc A;
  __shared__ int32_t s_col[KB / 64][2][64];
  __shared__ T s_val[KB / 64][2][64];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform: bookkeeping stays scalar
  const int64_t seg = (m.n + NXCD - 1) / NXCD;
  const int64_t s0 = (int64_t)(blockIdx.x % NXCD) * seg;
  const int64_t s1 = min(m.n, s0 + seg);
  const int64_t W = (int64_t)(gridDim.x / NXCD) * (KB / 64);           // wavefronts per die
  const int64_t wx = (int64_t)(blockIdx.x / NXCD) * (KB / 64) + wave;  // rank on the die
  const int64_t nq = (s1 > s0) ? (s1 - s0 + SRC - 1) / SRC : 0;        // chunks of the segment
  const int64_t nk = (nq > wx) ? (nq - wx + W - 1) / W : 0;            // chunks of this wavefront
  const uint32_t rowb = (uint32_t)C * (uint32_t)sizeof(T);
  const uint32_t rowd = (uint32_t)C * (uint32_t)sizeof(TD);
  const __amdgpu_buffer_rsrc_t rsrc =
      __builtin_amdgcn_make_buffer_rsrc((void*)op.P, (short)0, (int)(m.n * rowb), BUF_WORD3);
  const bool has_d = MODE == 1 && op.D != nullptr;
  const __amdgpu_buffer_rsrc_t rsrd =
      __builtin_amdgcn_make_buffer_rsrc((void*)(has_d ? op.D : (const TD*)op.P), (short)0,
                                        has_d ? (int)(m.n * rowd) : 0, BUF_WORD3);
  A pv[NCH][MAXND];
#pragma unroll
  for (int ch = 0; ch < NCH; ++ch)
    for (int d = 0; d < MAXND; ++d) pv[ch][d] = acc_zero(A());
#pragma unroll
  for (int ch = 0; ch < NCH; ++ch) {
    if (ch * 64 >= C) break;
    if (MODE == 1 && op.gate && !op.gate[ch]) continue;  // chunk not replaced this time
    if constexpr (MODE == 0) {
      // a chunk whose columns have all converged needs no q (its alpha is 0,
      // its p.q partials stay 0 and its stale q is never read)
      const int cc = lane + ch * 64;
      if (!__any(cc < C && st.active[cc])) continue;
    }
    const int c = lane + ch * 64;
    const bool live = c < C;
    const uint32_t voff = (live ? c : 0) * (uint32_t)sizeof(T);
    const uint32_t voffd = (live ? c : 0) * (uint32_t)sizeof(TD);
    const int sys = (ch * 64) / k;
    const T* v = op.vals + (int64_t)sys * m.nnz;
    for (int64_t g0 = 0; g0 < nk; g0 += SGC) {
      const int ng = (int)min((int64_t)SGC, nk - g0);
      // lane l holds row pointer l % (SRC+1) of chunk l / (SRC+1) of the group
      int32_t rp = 0;
      {
        const int kq = lane / (SRC + 1), o = lane % (SRC + 1);
        if (kq < ng) rp = m.row_ptr[min(s0 + (wx + (g0 + kq) * W) * SRC + o, s1)];
      }
      // current chunk (scalar state)
      int kc = 0;
      int64_t cr0 = 0;
      int cnr = 0, row = 0;
      int32_t ce0 = 0, ce1 = 0, rend = 0;
      T acc = sc<T>::zero(), pown = sc<T>::zero();
      auto chunk_at = [&](int kq, int64_t& r0, int& nr, int32_t& e0, int32_t& e1) {
        r0 = s0 + (wx + (g0 + kq) * W) * SRC;
        nr = (int)min((int64_t)SRC, s1 - r0);
        e0 = __builtin_amdgcn_readlane(rp, kq * (SRC + 1));
        e1 = __builtin_amdgcn_readlane(rp, kq * (SRC + 1) + nr);
      };
      // rows of the current chunk that end at entry e (empty rows included)
      auto drain = [&](int32_t e) {
        while (row < cnr && rend == e) {
          if (live) {
            const int64_t o = (cr0 + row) * C + c;
            if constexpr (MODE == 0) {
              op.Q[o] = acc;
              pv[ch][0] = acc_add(pv[ch][0], amul(pown, acc));
            } else {
              const T b = op.B[o];
              const T r = b - acc;
              const TD r32 = narrow(r, TD());
              const TD z = op.dinv[(int64_t)sys * m.n + cr0 + row] * r32;
              op.R32[o] = r32;
              pv[ch][0] = acc_add(pv[ch][0], acc_from_real(sc<T>::norm2(r), A()));
              pv[ch][1] = acc_add(pv[ch][1], acc_from_real(sc<T>::norm2(b), A()));
              pv[ch][2] = acc_add(pv[ch][2], amul(r32, z));
            }
          }
          acc = sc<T>::zero();
          pown = sc<T>::zero();
          ++row;
          rend = (row < cnr) ? __builtin_amdgcn_readlane(rp, kc * (SRC + 1) + row + 1) : -1;
        }
      };
      auto enter_chunk = [&](int kq) {
        kc = kq;
        chunk_at(kq, cr0, cnr, ce0, ce1);
        row = 0;
        rend = __builtin_amdgcn_readlane(rp, kq * (SRC + 1) + 1);
        drain(ce0);
      };
      auto gather = [&](T(&x)[FRG], TD(&xd)[FRG], int sl, int jb) {
#pragma unroll
        for (int u = 0; u < FRG; ++u) {
          const uint32_t cu = __builtin_amdgcn_readfirstlane(s_col[wave][sl][jb * FRG + u]);  // uniform
          x[u] = buf_ld<T>(rsrc, voff, cu * rowb);
          if constexpr (MODE == 1) {
            if (has_d) xd[u] = buf_ld<TD>(rsrd, voffd, cu * rowd);
          }
        }
      };
      auto consume = [&](const T(&x)[FRG], const TD(&xd)[FRG], int sl, int jb, int32_t pb) {
#pragma unroll
        for (int u = 0; u < FRG; ++u) {
          T xv = x[u];
          if constexpr (MODE == 1) {
            if (has_d) xv = xv + widen(xd[u], T());
          }
          acc = acc + s_val[wave][sl][jb * FRG + u] * xv;
          if constexpr (MODE == 0) {
            const int32_t cu = __builtin_amdgcn_readfirstlane(s_col[wave][sl][jb * FRG + u]);
            if (cu == (int32_t)(cr0 + row)) pown = xv;
          }
          drain(pb + jb * FRG + u + 1);
        }
      };
      enter_chunk(0);
      int32_t pb = ce0;
      int sl = 0;
      {
        const int32_t e = pb + lane;
        s_col[wave][0][lane] = (e < ce1) ? m.col_idx[e] : (int32_t)cr0;
        s_val[wave][0][lane] = (e < ce1) ? v[e] : sc<T>::zero();
      }
      T xa[FRG], xb[FRG];
      TD da[FRG], db[FRG];
      gather(xa, da, 0, 0);
      for (;;) {
        // the piece after this one: the rest of this chunk, or the next chunk
        int kn = kc;
        int64_t nr0 = cr0;
        int nnr = cnr;
        int32_t npb = pb + 64, ne0 = ce0, ne1 = ce1;
        const bool more = npb < ce1;
        const bool has_next = more || kc + 1 < ng;
        if (!more && has_next) {
          kn = kc + 1;
          chunk_at(kn, nr0, nnr, ne0, ne1);
          npb = ne0;
        }
        // its (column, value) pairs, parked in registers until the last pair
        const int32_t en = min(npb + lane, max(ne1 - 1, 0));
        const bool vn = has_next && npb + lane < ne1;
        const int32_t tcol = m.col_idx[en];
        const T tval = v[en];
        // the piece's 64 slots as 4 pairs of batches, straight-line
#pragma unroll
        for (int pr = 0; pr < 64 / (2 * FRG); ++pr) {
          gather(xb, db, sl, 2 * pr + 1);
          consume(xa, da, sl, 2 * pr, pb);
          if (pr + 1 < 64 / (2 * FRG)) {
            gather(xa, da, sl, 2 * pr + 2);
          } else {
            s_col[wave][sl ^ 1][lane] = vn ? tcol : (int32_t)nr0;
            s_val[wave][sl ^ 1][lane] = vn ? tval : sc<T>::zero();
            gather(xa, da, sl ^ 1, 0);  // first batch of the next piece (padding when there is none)
          }
          consume(xb, db, sl, 2 * pr + 1, pb);
        }
        if (!has_next) break;
        sl ^= 1;
        pb = npb;
        if (kn != kc) enter_chunk(kn);
      }
    }
  }
  block_partials<A, 64, NCH>(st, pv, MODE == 0 ? 1 : 3, C);
}

// Real solves of 64-column systems (one system per chunk): one sweep covers
// a pair of chunks, lane l gathering columns 2l and 2l+1 of the pair's 128
// (8 bytes, so a 512-byte row per gather as for complex64); lanes 0-31
// scale by the first system's matrix values and lanes 32-63 by the second's.
// Schedule, LDS staging and row-end handling as in sweep_body (MODE 0).
typedef __attribute__((ext_vector_type(2))) float fl2;
typedef __attribute__((ext_vector_type(2))) double dl2;
template <typename TS> struct pair_of;
template <> struct pair_of<float> { typedef fl2 type; };
template <> struct pair_of<double> { typedef dl2 type; };

template <int NCH>
__device__ void block_partials_pair(const ColState<double>& st, double (&pv)[NCH / 2][2], int C) {
  __shared__ double lds[KB / 64][128];
  const int lane = threadIdx.x & 63, rg = threadIdx.x >> 6;
  constexpr int RPB = KB / 64;
#pragma unroll
  for (int cp = 0; cp < NCH / 2; ++cp) {
    lds[rg][2 * lane] = pv[cp][0];
    lds[rg][2 * lane + 1] = pv[cp][1];
    __syncthreads();
    for (int i = threadIdx.x; i < 128; i += KB) {
      const int c = cp * 128 + i;
      if (c < C) {
        double acc = lds[0][i];
        for (int r = 1; r < RPB; ++r) acc = acc + lds[r][i];
        st.part[(int64_t)c * st.nb + blockIdx.x] = acc;
      }
    }
    __syncthreads();
  }
}

template <typename TS, int NCH, int FRG = 8>
__device__ __forceinline__ void sweep_pair_body(const Csr& m, const TS* __restrict__ vals,
                                                const TS* __restrict__ P, TS* __restrict__ Q, int C,
                                                const ColState<double>& st) {
  typedef typename pair_of<TS>::type fl2;
  __shared__ int32_t s_col[KB / 64][2][64];
  __shared__ fl2 s_val[KB / 64][2][64];  // (first system, second system) values of an entry
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t seg = (m.n + NXCD - 1) / NXCD;
  const int64_t s0 = (int64_t)(blockIdx.x % NXCD) * seg;
  const int64_t s1 = min(m.n, s0 + seg);
  const int64_t W = (int64_t)(gridDim.x / NXCD) * (KB / 64);
  const int64_t wx = (int64_t)(blockIdx.x / NXCD) * (KB / 64) + wave;
  const int64_t nq = (s1 > s0) ? (s1 - s0 + SRC - 1) / SRC : 0;
  const int64_t nk = (nq > wx) ? (nq - wx + W - 1) / W : 0;
  const uint32_t rowb = (uint32_t)C * (uint32_t)sizeof(TS);
  const __amdgpu_buffer_rsrc_t rsrc =
      __builtin_amdgcn_make_buffer_rsrc((void*)P, (short)0, (int)(m.n * rowb), BUF_WORD3);
  const bool hi = lane >= 32;
  double pv[NCH / 2][2];
#pragma unroll
  for (int cp = 0; cp < NCH / 2; ++cp) pv[cp][0] = pv[cp][1] = 0.0;
#pragma unroll
  for (int cp = 0; cp < NCH / 2; ++cp) {
    const int c = cp * 128 + 2 * lane;  // this lane's first column (C = NCH * 64)
    if (!__any(st.active[c] || st.active[c + 1])) continue;  // both chunks converged
    const uint32_t voff = (uint32_t)c * (uint32_t)sizeof(TS);
    const TS* vA = vals + (int64_t)(2 * cp) * m.nnz;  // system = chunk
    const TS* vB = vA + m.nnz;
    for (int64_t g0 = 0; g0 < nk; g0 += SGC) {
      const int ng = (int)min((int64_t)SGC, nk - g0);
      int32_t rp = 0;
      {
        const int kq = lane / (SRC + 1), o = lane % (SRC + 1);
        if (kq < ng) rp = m.row_ptr[min(s0 + (wx + (g0 + kq) * W) * SRC + o, s1)];
      }
      int kc = 0;
      int64_t cr0 = 0;
      int cnr = 0, row = 0;
      int32_t ce0 = 0, ce1 = 0, rend = 0;
      fl2 acc = {0, 0}, pown = {0, 0};
      auto chunk_at = [&](int kq, int64_t& r0, int& nr, int32_t& e0, int32_t& e1) {
        r0 = s0 + (wx + (g0 + kq) * W) * SRC;
        nr = (int)min((int64_t)SRC, s1 - r0);
        e0 = __builtin_amdgcn_readlane(rp, kq * (SRC + 1));
        e1 = __builtin_amdgcn_readlane(rp, kq * (SRC + 1) + nr);
      };
      auto drain = [&](int32_t e) {
        while (row < cnr && rend == e) {
          const int64_t o = (cr0 + row) * C + c;
          *reinterpret_cast<fl2*>(Q + o) = acc;
          pv[cp][0] += (double)pown.x * (double)acc.x;
          pv[cp][1] += (double)pown.y * (double)acc.y;
          acc = fl2{0, 0};
          pown = fl2{0, 0};
          ++row;
          rend = (row < cnr) ? __builtin_amdgcn_readlane(rp, kc * (SRC + 1) + row + 1) : -1;
        }
      };
      auto enter_chunk = [&](int kq) {
        kc = kq;
        chunk_at(kq, cr0, cnr, ce0, ce1);
        row = 0;
        rend = __builtin_amdgcn_readlane(rp, kq * (SRC + 1) + 1);
        drain(ce0);
      };
      auto gather = [&](fl2(&x)[FRG], int sl, int jb) {
#pragma unroll
        for (int u = 0; u < FRG; ++u) {
          const uint32_t cu = __builtin_amdgcn_readfirstlane(s_col[wave][sl][jb * FRG + u]);
          x[u] = buf_ld<fl2>(rsrc, voff, cu * rowb);
        }
      };
      auto consume = [&](const fl2(&x)[FRG], int sl, int jb, int32_t pb) {
#pragma unroll
        for (int u = 0; u < FRG; ++u) {
          const fl2 sv = s_val[wave][sl][jb * FRG + u];
          const TS v = hi ? sv.y : sv.x;
          acc = acc + v * x[u];
          const int32_t cu = __builtin_amdgcn_readfirstlane(s_col[wave][sl][jb * FRG + u]);
          if (cu == (int32_t)(cr0 + row)) pown = x[u];
          drain(pb + jb * FRG + u + 1);
        }
      };
      enter_chunk(0);
      int32_t pb = ce0;
      int sl = 0;
      {
        const int32_t e = pb + lane;
        s_col[wave][0][lane] = (e < ce1) ? m.col_idx[e] : (int32_t)cr0;
        s_val[wave][0][lane] = (e < ce1) ? fl2{vA[e], vB[e]} : fl2{0, 0};
      }
      fl2 xa[FRG], xb[FRG];
      gather(xa, 0, 0);
      for (;;) {
        int kn = kc;
        int64_t nr0 = cr0;
        int nnr = cnr;
        int32_t npb = pb + 64, ne0 = ce0, ne1 = ce1;
        const bool more = npb < ce1;
        const bool has_next = more || kc + 1 < ng;
        if (!more && has_next) {
          kn = kc + 1;
          chunk_at(kn, nr0, nnr, ne0, ne1);
          npb = ne0;
        }
        const int32_t en = min(npb + lane, max(ne1 - 1, 0));
        const bool vn = has_next && npb + lane < ne1;
        const int32_t tcol = m.col_idx[en];
        const fl2 tval = fl2{vA[en], vB[en]};
#pragma unroll
        for (int pr = 0; pr < 64 / (2 * FRG); ++pr) {
          gather(xb, sl, 2 * pr + 1);
          consume(xa, sl, 2 * pr, pb);
          if (pr + 1 < 64 / (2 * FRG)) {
            gather(xa, sl, 2 * pr + 2);
          } else {
            s_col[wave][sl ^ 1][lane] = vn ? tcol : (int32_t)nr0;
            s_val[wave][sl ^ 1][lane] = vn ? tval : fl2{0, 0};
            gather(xa, sl ^ 1, 0);
          }
          consume(xb, sl, 2 * pr + 1, pb);
        }
        if (!has_next) break;
        sl ^= 1;
        pb = npb;
        if (kn != kc) enter_chunk(kn);
      }
    }
  }
  block_partials_pair<NCH>(st, pv, C);
}

// Float32 solves of 64-column systems four at a time: one sweep covers a
// quad of chunks (256 columns), lane l gathering columns 4l..4l+3 (16 bytes,
// so one gather instruction brings a whole 1 KB row of the four systems,
// half the instructions of two pair sweeps).  Lanes 16q..16q+15 belong to
// system q of the quad and read that system's matrix value of each entry
// (an LDS read of one float at a per-quarter address).  Schedule, LDS
// staging and row-end handling as in sweep_pair_body.
typedef __attribute__((ext_vector_type(4))) float fl4;

template <int NCH>
__device__ void block_partials_quad(const ColState<double>& st, double (&pv)[NCH / 4][4], int C) {
  __shared__ double lds[KB / 64][256];
  const int lane = threadIdx.x & 63, rg = threadIdx.x >> 6;
  constexpr int RPB = KB / 64;
#pragma unroll
  for (int cq = 0; cq < NCH / 4; ++cq) {
#pragma unroll
    for (int e = 0; e < 4; ++e) lds[rg][4 * lane + e] = pv[cq][e];
    __syncthreads();
    for (int i = threadIdx.x; i < 256; i += KB) {
      const int c = cq * 256 + i;
      if (c < C) {
        double acc = lds[0][i];
        for (int r = 1; r < RPB; ++r) acc = acc + lds[r][i];
        st.part[(int64_t)c * st.nb + blockIdx.x] = acc;
      }
    }
    __syncthreads();
  }
}

// FRG 4: eight 1 KB gathers in flight per wavefront, the pair sweep's 8 KB
template <int NCH, int FRG = 4>
__device__ __forceinline__ void sweep_quad_body(const Csr& m, const float* __restrict__ vals,
                                                const float* __restrict__ P, float* __restrict__ Q, int C,
                                                const ColState<double>& st) {
  static_assert(NCH % 4 == 0, "quad sweeps cover whole groups of four 64-column systems");
  __shared__ int32_t s_col[KB / 64][2][64];
  __shared__ fl4 s_val[KB / 64][2][64];  // the four systems' values of an entry
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int q4 = lane >> 4;  // this lane's system within the quad
  const int64_t seg = (m.n + NXCD - 1) / NXCD;
  const int64_t s0 = (int64_t)(blockIdx.x % NXCD) * seg;
  const int64_t s1 = min(m.n, s0 + seg);
  const int64_t W = (int64_t)(gridDim.x / NXCD) * (KB / 64);
  const int64_t wx = (int64_t)(blockIdx.x / NXCD) * (KB / 64) + wave;
  const int64_t nq = (s1 > s0) ? (s1 - s0 + SRC - 1) / SRC : 0;
  const int64_t nk = (nq > wx) ? (nq - wx + W - 1) / W : 0;
  const uint32_t rowb = (uint32_t)C * (uint32_t)sizeof(float);
  const __amdgpu_buffer_rsrc_t rsrc =
      __builtin_amdgcn_make_buffer_rsrc((void*)P, (short)0, (int)(m.n * rowb), BUF_WORD3);
  double pv[NCH / 4][4];
#pragma unroll
  for (int cq = 0; cq < NCH / 4; ++cq) pv[cq][0] = pv[cq][1] = pv[cq][2] = pv[cq][3] = 0.0;
#pragma unroll
  for (int cq = 0; cq < NCH / 4; ++cq) {
    const int c = cq * 256 + 4 * lane;  // this lane's first column (C = NCH * 64)
    if (!__any(st.active[c] || st.active[c + 1] || st.active[c + 2] || st.active[c + 3])) continue;
    const uint32_t voff = (uint32_t)c * (uint32_t)sizeof(float);
    const float* vA = vals + (int64_t)(4 * cq) * m.nnz;  // system = chunk
    const float* vB = vA + m.nnz;
    const float* vC = vB + m.nnz;
    const float* vD = vC + m.nnz;
    for (int64_t g0 = 0; g0 < nk; g0 += SGC) {
      const int ng = (int)min((int64_t)SGC, nk - g0);
      int32_t rp = 0;
      {
        const int kq = lane / (SRC + 1), o = lane % (SRC + 1);
        if (kq < ng) rp = m.row_ptr[min(s0 + (wx + (g0 + kq) * W) * SRC + o, s1)];
      }
      int kc = 0;
      int64_t cr0 = 0;
      int cnr = 0, row = 0;
      int32_t ce0 = 0, ce1 = 0, rend = 0;
      fl4 acc = {0, 0, 0, 0}, pown = {0, 0, 0, 0};
      auto chunk_at = [&](int kq, int64_t& r0, int& nr, int32_t& e0, int32_t& e1) {
        r0 = s0 + (wx + (g0 + kq) * W) * SRC;
        nr = (int)min((int64_t)SRC, s1 - r0);
        e0 = __builtin_amdgcn_readlane(rp, kq * (SRC + 1));
        e1 = __builtin_amdgcn_readlane(rp, kq * (SRC + 1) + nr);
      };
      auto drain = [&](int32_t e) {
        while (row < cnr && rend == e) {
          const int64_t o = (cr0 + row) * C + c;
          *reinterpret_cast<fl4*>(Q + o) = acc;
          pv[cq][0] += (double)pown.x * (double)acc.x;
          pv[cq][1] += (double)pown.y * (double)acc.y;
          pv[cq][2] += (double)pown.z * (double)acc.z;
          pv[cq][3] += (double)pown.w * (double)acc.w;
          acc = fl4{0, 0, 0, 0};
          pown = fl4{0, 0, 0, 0};
          ++row;
          rend = (row < cnr) ? __builtin_amdgcn_readlane(rp, kc * (SRC + 1) + row + 1) : -1;
        }
      };
      auto enter_chunk = [&](int kq) {
        kc = kq;
        chunk_at(kq, cr0, cnr, ce0, ce1);
        row = 0;
        rend = __builtin_amdgcn_readlane(rp, kq * (SRC + 1) + 1);
        drain(ce0);
      };
      auto gather = [&](fl4(&x)[FRG], int sl, int jb) {
#pragma unroll
        for (int u = 0; u < FRG; ++u) {
          const uint32_t cu = __builtin_amdgcn_readfirstlane(s_col[wave][sl][jb * FRG + u]);
          x[u] = buf_ld<fl4>(rsrc, voff, cu * rowb);
        }
      };
      auto consume = [&](const fl4(&x)[FRG], int sl, int jb, int32_t pb) {
#pragma unroll
        for (int u = 0; u < FRG; ++u) {
          const float v = reinterpret_cast<const float*>(&s_val[wave][sl][jb * FRG + u])[q4];
          acc = acc + v * x[u];
          const int32_t cu = __builtin_amdgcn_readfirstlane(s_col[wave][sl][jb * FRG + u]);
          if (cu == (int32_t)(cr0 + row)) pown = x[u];
          drain(pb + jb * FRG + u + 1);
        }
      };
      enter_chunk(0);
      int32_t pb = ce0;
      int sl = 0;
      {
        const int32_t e = pb + lane;
        s_col[wave][0][lane] = (e < ce1) ? m.col_idx[e] : (int32_t)cr0;
        s_val[wave][0][lane] = (e < ce1) ? fl4{vA[e], vB[e], vC[e], vD[e]} : fl4{0, 0, 0, 0};
      }
      fl4 xa[FRG], xb[FRG];
      gather(xa, 0, 0);
      for (;;) {
        int kn = kc;
        int64_t nr0 = cr0;
        int nnr = cnr;
        int32_t npb = pb + 64, ne0 = ce0, ne1 = ce1;
        const bool more = npb < ce1;
        const bool has_next = more || kc + 1 < ng;
        if (!more && has_next) {
          kn = kc + 1;
          chunk_at(kn, nr0, nnr, ne0, ne1);
          npb = ne0;
        }
        const int32_t en = min(npb + lane, max(ne1 - 1, 0));
        const bool vn = has_next && npb + lane < ne1;
        const int32_t tcol = m.col_idx[en];
        const fl4 tval = fl4{vA[en], vB[en], vC[en], vD[en]};
#pragma unroll
        for (int pr = 0; pr < 64 / (2 * FRG); ++pr) {
          gather(xb, sl, 2 * pr + 1);
          consume(xa, sl, 2 * pr, pb);
          if (pr + 1 < 64 / (2 * FRG)) {
            gather(xa, sl, 2 * pr + 2);
          } else {
            s_col[wave][sl ^ 1][lane] = vn ? tcol : (int32_t)nr0;
            s_val[wave][sl ^ 1][lane] = vn ? tval : fl4{0, 0, 0, 0};
            gather(xa, sl ^ 1, 0);
          }
          consume(xb, sl, 2 * pr + 1, pb);
        }
        if (!has_next) break;
        sl ^= 1;
        pb = npb;
        if (kn != kc) enter_chunk(kn);
      }
    }
  }
  block_partials_quad<NCH>(st, pv, C);
}

template <typename T, int NCH, bool PAIR = false>
__global__ void __launch_bounds__(KB) k_cg_spmm_sweep(Csr m, const T* __restrict__ vals, const T* __restrict__ P,
                                                      T* __restrict__ Q, int C, int k,
                                                      ColState<typename sc<T>::acc> st) {
  if (st.ru_req && blockIdx.x == 0 && threadIdx.x < MAXCHUNK) ru_cur(st)[threadIdx.x] = 0;
  if (!block_any_active(st.active, C)) return;
  if constexpr (PAIR) {
    static_assert((std::is_same<T, float>::value || std::is_same<T, double>::value) && NCH % 2 == 0,
                  "pair sweeps are for real 64-column systems");
    sweep_pair_body<T, NCH>(m, vals, P, Q, C, st);
  } else {
    SweepOps<T, T> op{vals, P, nullptr, Q, nullptr, nullptr, nullptr, nullptr};
    sweep_body<T, T, NCH, 0>(m, op, C, k, st);
  }
}

// mixed solves: q = A32 p (partials in st) and, for the 64-column chunks
// whose replacement the previous iteration requested, the float64 true
// residual r32 = b - A64 (x64 + d32) (partials |r|^2, |b|^2, r.z in so) in
// the same launch; the residual pass runs smaller batches so the launch
// keeps the SpMM's occupancy.  Clears this iteration's request flags.
template <typename T64, typename T32, int NCH, bool PAIR = false>
__global__ void __launch_bounds__(KB) __attribute__((amdgpu_waves_per_eu(5))) k_cg_spmm_sweep_ru(Csr m, const T32* __restrict__ vals32,
                                                         const T32* __restrict__ P, T32* __restrict__ Q,
                                                         const T64* __restrict__ vals64, const T64* __restrict__ B,
                                                         const T64* __restrict__ X, const T32* __restrict__ D,
                                                         const T32* __restrict__ dinv32, T32* __restrict__ R32, int C,
                                                         int k, ColState<typename sc<T32>::acc> st,
                                                         ColState<typename sc<T64>::acc> so) {
  if (blockIdx.x == 0 && threadIdx.x < MAXCHUNK) ru_cur(st)[threadIdx.x] = 0;
  if (!block_any_active(st.active, C)) return;
  if constexpr (PAIR) {
    static_assert(std::is_same<T32, float>::value && NCH % 2 == 0, "pair sweeps are for real 64-column systems");
    if constexpr (NCH % 4 == 0)
      sweep_quad_body<NCH>(m, vals32, P, Q, C, st);  // four float32 systems per 16-byte gather
    else
      sweep_pair_body<float, NCH>(m, vals32, P, Q, C, st);
  } else {
    SweepOps<T32, T32> op{vals32, P, nullptr, Q, nullptr, nullptr, nullptr, nullptr};
    sweep_body<T32, T32, NCH, 0>(m, op, C, k, st);
  }
  const int* gate = ru_prev(st);
  int any = 0;
  for (int ch = 0; ch < NCH; ++ch) any |= gate[ch];
  if (!any) return;
  SweepOps<T64, T32> oo{vals64, X, D, nullptr, B, dinv32, R32, gate};
  sweep_body<T64, T32, NCH, 1, 4>(m, oo, C, k, so);
}

// ---------------------------------------------------------------------------
// LDS-staged SpMM for 512-byte rows (64 complex64 or float64 columns of one
// system).  A workgroup owns a sequence of plan blocks (consecutive CSR rows
// whose columns fall in at most RB_LDS_UMAX distinct rows).  Producer
// wavefronts copy the next block's neighbour rows of p and its entry records
// from global memory straight into one of two LDS buffers
// (global_load_lds_dwordx4, 1 KB per instruction: two 512-byte rows, or a
// 1 KB run of records) while the consumer wavefronts compute the current
// block from the other buffer, one row at a time: each entry is a broadcast
// LDS read of its record (offset, value) and one 512-byte LDS read of its
// column's row.  Each p row therefore crosses the L2 once per block whose
// union holds it (about 4 times for a tetrahedral mesh in Morton order)
// instead of once per CSR entry (about 14.5 times).  Producers and consumers
// run separate loops with one bare barrier per block each (a __syncthreads()
// would make the producers drain their in-flight copies at every block),
// and every producer issues the same number of copies per block (spare
// copies land in a scratch KB), so one s_waitcnt vmcnt(0) before the barrier
// hands a complete block over.  Consumers also form the per-column partials
// of p.q from the row's own p (its union slot).
// ---------------------------------------------------------------------------
#define LDS_NW 12  // consumer wavefronts
#define LDS_NP 4   // producer wavefronts
#define LDS_RINB 1024  // (first record, own-row offset) of up to 128 rows of a block

// block limits and buffer ring of one kernel configuration; a plan is built
// for one (UMAX, RECMAX) pair and runs on the kernel configured for it
template <int UMAX_, int RECMAX_, int NBUF_, int VB_ = 8>
struct LdsCfg {
  static constexpr int UMAX = UMAX_;      // neighbour rows per block
  static constexpr int RECMAX = RECMAX_;  // records per block
  static constexpr int NBUF = NBUF_;      // block buffers in the ring (copies run NBUF-1 blocks ahead)
  static constexpr int VB = VB_;          // bytes per value: a row is 64 of them
  static constexpr int RB = 64 * VB;      // row bytes (512, or 256 for 4-byte values)
  static constexpr int RPC = 1024 / RB;   // rows per 1 KB copy
  static constexpr int LSH = RB == 256 ? 1 : 0;  // plan offsets count 512-byte row slots
  static constexpr int ROWB = (UMAX + RPC) * RB;  // neighbour rows (+ the rows a last copy may fill)
  static constexpr int LOFB = RECMAX * 4;         // record offsets
  static constexpr int VALB = RECMAX * VB;        // record values
  static constexpr int BUFB = ROWB + LOFB + VALB + LDS_RINB;
  static constexpr int TRASH = NBUF * BUFB;      // 1 KB written by spare copies, never read
  static constexpr int HDR = TRASH + 1024;       // (first row, rows) of each buffer's block
  static constexpr int BYTES = HDR + 8 * NBUF;
  static constexpr int JB = (UMAX / RPC + LDS_NP - 1) / LDS_NP;  // row copies per producer and block
  static constexpr int JR = (LOFB + VALB + LDS_RINB) / 1024;     // record copies per block
  static constexpr int JRP = (JR + LDS_NP - 1) / LDS_NP;         // ... per producer
  static_assert(VB == 8 || VB == 4, "rows of 64 values of 8 or 4 bytes");
  static_assert(BYTES <= 160 * 1024, "the block buffers must fit the 160 KB LDS of a gfx950 CU");
  static_assert((LOFB + VALB) % 1024 == 0, "records are copied in 1 KB pieces");
  static_assert(2 * (RB_LDS_RMAX + 1) * 4 <= LDS_RINB, "row table of a block");
  static_assert(RPC * LDS_NP * JB <= UMAX + RB_LDS_UPAD, "producers read union index runs within the padding");
  static_assert((JB + JRP) * (NBUF - 2) < 64, "copies in flight behind the awaited block fit vmcnt");
};
// RB_LDS_UMAX / RB_LDS_RECMAX, two buffers: each p row crosses the L2 ~4 times per SpMM.  A ring of
// three buffers with blocks of 80 rows (copies two blocks ahead) measured slower in the solve (68.8
// against 63.5 us per launch at 115k nodes): the smaller unions load ~35% more rows.
typedef LdsCfg<RB_LDS_UMAX, RB_LDS_RECMAX, 2> LdsWide;
// one real float32 system of 64 columns: 256-byte rows, four to a copy, on the same plans (a third
// buffer fits in the LDS here but measured 4.5% slower on the 2.6M-node batch, two runs on one box)
typedef LdsCfg<RB_LDS_UMAX, RB_LDS_RECMAX, 2, 4> LdsHalf;

// s_waitcnt vmcnt(N), the other counters untouched (gfx9 encoding: vmcnt in bits 3:0 and 15:14)
template <int N>
__device__ __forceinline__ void wait_vmcnt() {
  static_assert(N >= 0 && N < 64, "vmcnt is 6 bits");
  __builtin_amdgcn_s_waitcnt((N & 15) | (7 << 4) | (15 << 8) | ((N >> 4) << 14));
}

struct LdsArgs {
  int64_t nblk;
  int umax;
  const int32_t* brow;
  const int32_t* bu;
  const int32_t* ulist;
  const int32_t* rinfo;
  const int32_t* loff;
  const void* recv;
  const void* recv_ad;  // the records of A D^-1 (the polynomial preconditioner's product), or null
};

// real type of a column type, and a column value scaled by it
template <typename T>
struct RealOf;
template <>
struct RealOf<float> {
  typedef float type;
};
template <>
struct RealOf<double> {
  typedef double type;
};
template <>
struct RealOf<cplx<float>> {
  typedef float type;
};
template <>
struct RealOf<cplx<double>> {
  typedef double type;
};
__device__ inline float rscale(float v, float s) { return v * s; }
__device__ inline double rscale(double v, double s) { return v * s; }
__device__ inline cplx<float> rscale(cplx<float> v, float s) { return mk<float>(v.x * s, v.y * s); }
__device__ inline cplx<double> rscale(cplx<double> v, double s) { return mk<double>(v.x * s, v.y * s); }

// slope of the polynomial preconditioner p(x) = (3 / L)(1 - POLY_K x / L)
#define POLY_K 0.8
// the degree-2 one: p(x) = (P2_0 - P2_1 t + P2_2 t^2) / L with t = x / L, from the Chebyshev
// residual polynomial 1 - t p(t) = T3((1.03 - 2t) / 0.97) / T3(1.03 / 0.97) (smallest on [0.03, 1]);
// p stays positive on [0, 1] (its minimum 0.49 at t = 0.77)
#define P2_0 13.540030
#define P2_1 33.781406
#define P2_2 21.864988
// the degree-3 one, from the Chebyshev residual polynomial of degree 4 on [0.03, 1] (minimum 0.53)
#define P3_0 20.445082
#define P3_1 91.179713
#define P3_2 138.505086
#define P3_3 67.235478

typedef __attribute__((address_space(3))) void lds_void_t;

// one direct 16-byte-per-lane copy from global memory into LDS: lane l's 16
// bytes land at dst + 16 l
__device__ __forceinline__ void dma16(const void* src, char* dst) {
  __builtin_amdgcn_global_load_lds(src, (lds_void_t*)dst, 16, 0, 0);
}

// partial-sum type of an LDS SpMM launch: the column type's accumulator, or for a pair of
// real float32 systems (RPAIR) double, two columns per lane
template <typename T, bool RPAIR>
struct LdsAcc {
  typedef typename sc<T>::acc type;
};
template <>
struct LdsAcc<cplx<float>, true> {
  typedef double type;
};

// (the plan arrays come as separate read-only kernel arguments so their
// uniform reads are scalar loads, outside the vector counter the copies use)
//
// Each staged row is 512 bytes at P + row * rs + coff (T units): one system's 64 columns of 8
// bytes, or with RPAIR the 128 float32 columns of a pair of real 64-column systems, lane l
// holding columns 2l and 2l+1 (lanes 0-31 the first system's, scaled by the record's .x,
// lanes 32-63 the second's, by .y).  The launch's columns are [cbase, cbase + ncol) of st.
//
// POLY: the product of the polynomial preconditioner with its epilogue fused: P holds the
// residual r and recv the records of A D^-1, so the row's w = (A D^-1 r)_row and its own staged
// r give z = D^-1 (c0 r + c1 w) (c0, c1 from the system's Gershgorin bound in lbound); z goes to
// Q, and the partials are r.z and r.r in st's slots 0 and 1 of C columns, as k_cg_poly_z writes
// them.  dinvp is the system's D^-1 (T), or with RPAIR the float D^-1 of every system, system s
// at dinvp + s * dstride (the lane's system is cbase / 64 + its half of the pair).
template <typename T, typename CF, bool RPAIR, bool POLY>
__global__ void __launch_bounds__((LDS_NW + LDS_NP) * 64)
k_cg_spmm_lds(int64_t nblk, const int32_t* __restrict__ brow, const int32_t* __restrict__ bu,
              const int32_t* __restrict__ ulist, const int32_t* __restrict__ rinfo,
              const int32_t* __restrict__ rinfo_m, const int32_t* __restrict__ loff,
              const T* __restrict__ recv, const T* __restrict__ P, T* __restrict__ Q, int64_t rs, int64_t coff,
              int cbase, ColState<typename LdsAcc<T, RPAIR>::type> st, const void* __restrict__ dinvp,
              int64_t dstride, const float* __restrict__ lbound, int C, const T* __restrict__ Rg, int pdeg,
              const T* __restrict__ W1g) {
  static_assert(sizeof(T) == CF::VB, "rows of 64 values of the configuration's size");
  static_assert(!RPAIR || std::is_same<T, cplx<float>>::value, "a real pair is staged as 8-byte (float, float)");
  typedef typename sc<T>::acc A;
  typedef typename LdsAcc<T, RPAIR>::type AS;
  constexpr int NC = RPAIR ? 128 : 64;  // columns of this launch
  constexpr int JB = CF::JB, JR = CF::JR, JRP = CF::JRP, NBUF = CF::NBUF;
  extern __shared__ __attribute__((aligned(16))) char lds[];
  if (st.ru_req && blockIdx.x == 0 && threadIdx.x < MAXCHUNK) ru_cur(st)[threadIdx.x] = 0;
  if (!block_any_active(st.active + cbase, NC)) return;
  const int lane = threadIdx.x & 63;
  const bool second = lane >= 32;  // RPAIR: this lane's columns belong to the pair's second system
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t seg = (nblk + NXCD - 1) / NXCD;
  const int64_t b0 = (int64_t)(blockIdx.x % NXCD) * seg;
  const int64_t b1 = min(nblk, b0 + seg);
  const int64_t stride = gridDim.x / NXCD;
  const int64_t bfirst = b0 + blockIdx.x / NXCD;
  const int nmine = bfirst < b1 ? (int)((b1 - bfirst + stride - 1) / stride) : 0;
  A pv = acc_zero(A());  // RPAIR: .x column 2l, .y column 2l+1; POLY: r.z
  A pv2 = acc_zero(A());  // POLY: r.r
  typedef typename RealOf<T>::type R1;
  R1 pc0 = 0, pc1 = 0, pc2 = 0, pc3 = 0;
  const int psys = RPAIR ? cbase / 64 + (second ? 1 : 0) : 0;  // POLY: this lane's system
  if constexpr (POLY) {
    const R1 L = (R1)lbound[psys];
    if (pdeg == 3) {
      pc0 = (R1)P3_0 / L;
      pc1 = (R1)(-P3_1) / (L * L);
      pc2 = (R1)P3_2 / (L * L * L);
      pc3 = (R1)(-P3_3) / (L * L * L * L);
    } else if (pdeg == 2) {
      pc0 = (R1)P2_0 / L;
      pc1 = (R1)(-P2_1) / (L * L);
      pc2 = (R1)P2_2 / (L * L * L);
    } else {
      pc0 = (R1)3.0 / L;
      pc1 = (R1)(-3.0 * POLY_K) / (L * L);
    }
  }
  if (wave >= LDS_NW) {
    // producer: block k of this workgroup into buffer k % NBUF, JB + JRP copies every time.  A
    // block's metadata costs two dependent scalar-load round trips (its rows and union start,
    // then the union's row indices and first record); they are made one block ahead, while the
    // previous copies are in flight, so a block's copies start right after its barrier
    const int pw = wave - LDS_NW;
    constexpr int RPC = CF::RPC;
    struct Meta {
      int ni, r0, nr, e0;
      int g[RPC * JB];
    };
    auto meta = [&](int kb) {
      Meta mt;
      const bool real = kb < nmine;
      const int64_t bb = real ? bfirst + (int64_t)kb * stride : 0;
      const int u0 = real ? bu[bb] : 0;
      mt.ni = real ? (bu[bb + 1] - u0 + RPC - 1) / RPC : 0;  // row copies (RPC rows each)
      mt.r0 = real ? brow[bb] : 0;
      mt.nr = real ? brow[bb + 1] - mt.r0 : 0;
      // rinfo_m is rinfo again: rinfo is also a source of the copies, which the compiler takes as
      // possibly written, so its own loads could not use the scalar cache and would make the wave
      // wait on its copies in flight
      mt.e0 = real ? rinfo_m[2 * mt.r0] : 0;
      const int* ix = ulist + u0 + RPC * pw * JB;  // the padding keeps these runs readable
      // the row indices are uniform: scalar loads, kept apart from the copies' counter
#pragma unroll
      for (int j = 0; j < RPC * JB; ++j) mt.g[j] = __builtin_amdgcn_readfirstlane(ix[j]);
      return mt;
    };
    auto issue = [&](const Meta& mt, int kb) {
      char* base = lds + (kb % NBUF) * CF::BUFB;
      // the block's rows for the consumers (a plain LDS store, ordered by the next barrier; the
      // consumers' own loads of brow would wait on their outstanding q stores)
      if (pw == 0 && lane == 0) {
        int* h = (int*)(lds + CF::HDR) + 2 * (kb % NBUF);
        h[0] = mt.r0;
        h[1] = mt.nr;
      }
#pragma unroll
      for (int j = 0; j < JRP; ++j) {
        const int q = pw * JRP + j;  // 1 KB piece q: offsets, values, then the row table
        const char* src;
        if (q < CF::LOFB / 1024) src = (const char*)(loff + mt.e0) + q * 1024;
        else if (q < (CF::LOFB + CF::VALB) / 1024) src = (const char*)(recv + mt.e0) + (q - CF::LOFB / 1024) * 1024;
        else src = (const char*)(rinfo + 2 * mt.r0);
        dma16(src + lane * 16, q < JR ? base + CF::ROWB + q * 1024 : lds + CF::TRASH);
      }
#pragma unroll
      for (int j = 0; j < JB; ++j) {
        const int i = pw * JB + j;
        // lanes [q * 64 / RPC, (q + 1) * 64 / RPC) copy row q of the RPC rows of this copy
        int g;
        if constexpr (RPC == 2) {
          g = lane < 32 ? mt.g[2 * j] : mt.g[2 * j + 1];
        } else {
          const int ga = lane < 16 ? mt.g[4 * j] : mt.g[4 * j + 1];
          const int gb = lane < 48 ? mt.g[4 * j + 2] : mt.g[4 * j + 3];
          g = lane < 32 ? ga : gb;
        }
        const char* src = (const char*)(P + (int64_t)g * rs + coff) + (lane & (64 / RPC - 1)) * 16;
        dma16(src, i < mt.ni ? base + i * 1024 : base + CF::UMAX * CF::RB);
      }
    };
    for (int kb = 0; kb < NBUF - 1; ++kb) issue(meta(kb), kb);
    Meta nxt = meta(NBUF - 1);
    for (int kb = 0; kb < nmine; ++kb) {
      wait_vmcnt<(JB + JRP) * (NBUF - 2)>();  // block kb has landed (the next NBUF-2 may still fly)
      asm volatile("s_barrier" ::: "memory");
      issue(nxt, kb + NBUF - 1);
      nxt = meta(kb + NBUF);
    }
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");  // spare copies land before the buffers are reused
  } else {
    for (int kb = 0; kb < nmine; ++kb) {
      asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
      const char* buf = lds + (kb % NBUF) * CF::BUFB;
      const int* h = (const int*)(lds + CF::HDR) + 2 * (kb % NBUF);
      const int* lo = (const int*)(buf + CF::ROWB);
      const T* lv = (const T*)(buf + CF::ROWB + CF::LOFB);
      const int* ri = (const int*)(buf + CF::ROWB + CF::LOFB + CF::VALB);
      const char* cb = buf + lane * (int)sizeof(T);
      const int r0 = h[0], nr = h[1];
      const int e0 = ri[0];
      for (int i = wave; i < nr; i += LDS_NW) {
        const int ea = ri[2 * i] - e0, ez = ri[2 * i + 2] - e0;
        T acc = sc<T>::zero();
        for (int e = ea; e < ez; e += 8) {
          int la[8];
          T va[8], pa[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            la[u] = lo[e + u];
            va[u] = lv[e + u];
          }
#pragma unroll
          for (int u = 0; u < 8; ++u) pa[u] = *(const T*)(cb + (la[u] >> CF::LSH));
          if constexpr (RPAIR) {
            // select between values read into registers (a select between the fields' addresses
            // would index va dynamically and put it in scratch memory)
#pragma unroll
            for (int u = 0; u < 8; ++u) {
              const float vx = va[u].x, vy = va[u].y;
              const float v = second ? vy : vx;
              acc.x += v * pa[u].x;
              acc.y += v * pa[u].y;
            }
          } else {
#pragma unroll
            for (int u = 0; u < 8; ++u) acc = acc + va[u] * pa[u];
          }
        }
        const T pown = *(const T*)(cb + (ri[2 * i + 1] >> CF::LSH));
        if constexpr (POLY && RPAIR) {  // columns 2l, 2l+1 of this lane's system
          const int64_t ro = (int64_t)(r0 + i) * rs + coff + lane;
          const float d = ((const float*)dinvp)[(int64_t)psys * dstride + r0 + i];
          T z, rv;
          if (pdeg == 3) {  // pown: w2 = (A D^-1 w1), acc: (A D^-1 w2); r and w1 from HBM
            rv = Rg[ro];
            const T w1 = W1g[ro];
            z.x = d * (rv.x * pc0 + w1.x * pc1 + pown.x * pc2 + acc.x * pc3);
            z.y = d * (rv.y * pc0 + w1.y * pc1 + pown.y * pc2 + acc.y * pc3);
          } else if (pdeg == 2) {  // pown: this row's w1 = (A D^-1 r), acc: (A D^-1 w1); r from HBM
            rv = Rg[ro];
            z.x = d * (rv.x * pc0 + pown.x * pc1 + acc.x * pc2);
            z.y = d * (rv.y * pc0 + pown.y * pc1 + acc.y * pc2);
          } else {  // pown: this row's r, acc: (A D^-1 r)
            rv = pown;
            z.x = d * (pown.x * pc0 + acc.x * pc1);
            z.y = d * (pown.y * pc0 + acc.y * pc1);
          }
          Q[ro] = z;
          pv.x += (double)rv.x * (double)z.x;
          pv.y += (double)rv.y * (double)z.y;
          pv2.x += (double)rv.x * (double)rv.x;
          pv2.y += (double)rv.y * (double)rv.y;
          continue;
        } else if constexpr (POLY) {
          const int64_t ro = (int64_t)(r0 + i) * rs + coff + lane;
          const T d = ((const T*)dinvp)[r0 + i];
          T z, rv;
          if (pdeg == 3) {  // pown: w2 = (A D^-1 w1), acc: (A D^-1 w2); r and w1 from HBM
            rv = Rg[ro];
            z = d * (rscale(rv, pc0) + rscale(W1g[ro], pc1) + rscale(pown, pc2) + rscale(acc, pc3));
          } else if (pdeg == 2) {  // pown is this row's w1 = (A D^-1 r), acc its (A D^-1 w1); r from HBM
            rv = Rg[ro];
            z = d * (rscale(rv, pc0) + rscale(pown, pc1) + rscale(acc, pc2));
          } else {  // pown is this row's r, acc its (A D^-1 r)
            rv = pown;
            z = d * (rscale(pown, pc0) + rscale(acc, pc1));
          }
          Q[ro] = z;
          pv = acc_add(pv, amul(rv, z));
          pv2 = acc_add(pv2, acc_from_real(sc<T>::norm2(rv), A()));
          continue;
        }
        Q[(int64_t)(r0 + i) * rs + coff + lane] = acc;
        if constexpr (RPAIR) {
          pv.x += (double)pown.x * (double)acc.x;
          pv.y += (double)pown.y * (double)acc.y;
        } else {
          pv = acc_add(pv, amul(pown, acc));
        }
      }
    }
  }
  // per-column partials of this workgroup, consumer wavefronts summed in order
  __syncthreads();
  AS* red = (AS*)lds;
  if (wave < LDS_NW) {
    if constexpr (RPAIR) {
      red[wave * NC + 2 * lane] = pv.x;
      red[wave * NC + 2 * lane + 1] = pv.y;
    } else {
      red[wave * NC + lane] = pv;
    }
  }
  __syncthreads();
  if (threadIdx.x < NC) {
    AS t = red[threadIdx.x];
    for (int w = 1; w < LDS_NW; ++w) t = acc_add(t, red[w * NC + threadIdx.x]);
    st.part[(int64_t)(cbase + threadIdx.x) * st.nb + blockIdx.x] = t;
  }
  if constexpr (POLY) {  // the r.r partials in slot 1
    __syncthreads();
    if (wave < LDS_NW) {
      if constexpr (RPAIR) {
        red[wave * NC + 2 * lane] = pv2.x;
        red[wave * NC + 2 * lane + 1] = pv2.y;
      } else {
        red[wave * NC + lane] = pv2;
      }
    }
    __syncthreads();
    if (threadIdx.x < NC) {
      AS t = red[threadIdx.x];
      for (int w = 1; w < LDS_NW; ++w) t = acc_add(t, red[w * NC + threadIdx.x]);
      st.part[(int64_t)(C + cbase + threadIdx.x) * st.nb + blockIdx.x] = t;
    }
  }
}

// records of a pair of real float32 systems: recv[i] = (va[rmap[i]], vb[rmap[i]]) (0 for padding)
__global__ void k_lds_pack_pair(int64_t nrec, const int32_t* __restrict__ rmap, const float* __restrict__ va,
                                const float* __restrict__ vb, cplx<float>* __restrict__ recv) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nrec; i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t e = rmap[i];
    recv[i] = e >= 0 ? mk(va[e], vb[e]) : mk(0.0f, 0.0f);
  }
}

// record values of the matrix being solved: recv[i] = vals[rmap[i]] (0 for padding)
template <typename T>
__global__ void k_lds_pack(int64_t nrec, const int32_t* __restrict__ rmap, const T* __restrict__ vals,
                           T* __restrict__ recv) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nrec; i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t e = rmap[i];
    recv[i] = e >= 0 ? vals[e] : sc<T>::zero();
  }
}

// records of A D^-1 for a pair of real float32 systems: (va[e] / a_cc, vb[e] / b_cc) for slot
// e = rmap[i] in column c, da and db the two systems' D^-1 (0 for padding)
__global__ void k_lds_pack_pair_ad(int64_t nrec, const int32_t* __restrict__ rmap, const float* __restrict__ va,
                                   const float* __restrict__ vb, const int32_t* __restrict__ col_idx,
                                   const float* __restrict__ da, const float* __restrict__ db,
                                   cplx<float>* __restrict__ recv) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nrec; i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t e = rmap[i];
    cplx<float> v{0.0f, 0.0f};
    if (e >= 0) {
      const int32_t c = col_idx[e];
      v.x = va[e] * da[c];
      v.y = vb[e] * db[c];
    }
    recv[i] = v;
  }
}

// records of A D^-1 (system 0): recv[i] = a_e / a_cc for slot e = rmap[i] in column c (0 for padding)
template <typename T>
__global__ void k_lds_pack_ad(int64_t nrec, const int32_t* __restrict__ rmap, const T* __restrict__ vals,
                              const int32_t* __restrict__ col_idx, const T* __restrict__ dinv,
                              T* __restrict__ recv) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nrec; i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t e = rmap[i];
    recv[i] = e >= 0 ? vals[e] * dinv[col_idx[e]] : sc<T>::zero();
  }
}

// mixed solves whose SpMM runs in its own launch: the float64 true residual
// of the 64-column chunks whose replacement the previous iteration requested
// (the second half of k_cg_spmm_sweep_ru).  Exits at once when none asked.
template <typename T64, typename T32, int NCH>
__global__ void __launch_bounds__(KB) __attribute__((amdgpu_waves_per_eu(5)))
k_cg_resid_sweep(Csr m, const T64* __restrict__ vals64, const T64* __restrict__ B, const T64* __restrict__ X,
                 const T32* __restrict__ D, const T32* __restrict__ dinv32, T32* __restrict__ R32, int C, int k,
                 ColState<typename sc<T32>::acc> st, ColState<typename sc<T64>::acc> so) {
  const int* gate = ru_prev(st);
  int any = 0;
  for (int ch = 0; ch < NCH; ++ch) any |= gate[ch];
  if (!any) return;
  if (!block_any_active(st.active, C)) return;
  SweepOps<T64, T32> oo{vals64, X, D, nullptr, B, dinv32, R32, gate};
  sweep_body<T64, T32, NCH, 1, 4>(m, oo, C, k, so);
}

// q = A_s p ; partial p.q -- one wavefront per contiguous run of rows.
// The run's row pointers come in with one lane-parallel load; rows are taken
// two at a time, each with its (col, val) entries loaded by the lanes and
// broadcast by readlane, so up to 2*RG gathers of 64-column rows are in
// flight per wavefront.  Consecutive rows share most of their neighbours,
// which the run order keeps in the CU's L1.  Requires C <= 64*NCH with every
// 64-column chunk inside one system (nsys == 1 or k % 64 == 0).
template <typename T, int NCH>
__global__ void __launch_bounds__(KB) k_cg_spmm_w64(Csr m, const T* __restrict__ vals, const T* __restrict__ P,
                                                    T* __restrict__ Q, int C, int k,
                                                    ColState<typename sc<T>::acc> st) {
  typedef typename sc<T>::acc A;
  if (!block_any_active(st.active, C)) return;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform: row bookkeeping stays scalar
  const int64_t seg = (m.n + NXCD - 1) / NXCD;
  const int64_t s0 = (int64_t)(blockIdx.x % NXCD) * seg;
  const int64_t s1 = min(m.n, s0 + seg);
  const int64_t wpx = (int64_t)(gridDim.x / NXCD) * (KB / 64);  // wavefronts per die
  const int64_t wid = (int64_t)(blockIdx.x / NXCD) * (KB / 64) + wave;
  const int64_t rpw = (s1 - s0 + wpx - 1) / wpx;
  const int64_t w0 = s0 + wid * rpw;
  const int64_t w1 = min(s1, w0 + rpw);
  A pv[NCH][MAXND];
#pragma unroll
  for (int ch = 0; ch < NCH; ++ch) pv[ch][0] = acc_zero(A());
  for (int64_t b0 = w0; b0 < w1; b0 += 63) {
    const int nr = (int)min((int64_t)63, w1 - b0);
    const int32_t rp = (lane <= nr) ? m.row_ptr[b0 + lane] : 0;
    for (int i = 0; i < nr; i += 2) {
      const bool two = (i + 1 < nr);
      const int32_t ra = __builtin_amdgcn_readlane(rp, i), rb = __builtin_amdgcn_readlane(rp, i + 1);
      const int32_t rc = two ? __builtin_amdgcn_readlane(rp, i + 2) : rb;
      const int la = min(64, rb - ra), lb = min(64, rc - rb);
      const int32_t cla = (lane < la) ? m.col_idx[ra + lane] : (int32_t)(b0 + i);
      const int32_t clb = (lane < lb) ? m.col_idx[rb + lane] : (int32_t)(b0 + i);
#pragma unroll
      for (int ch = 0; ch < NCH; ++ch) {
        const int c = lane + ch * 64;
        const T* v = vals + (int64_t)((ch * 64) / k) * m.nnz;
        const T vla = (lane < la) ? v[ra + lane] : sc<T>::zero();
        const T vlb = (lane < lb) ? v[rb + lane] : sc<T>::zero();
        T ya = sc<T>::zero(), yb = sc<T>::zero();
        if (c < C) {
          const int lmax = max(la, lb);
          for (int j = 0; j < lmax; j += RG) {
            T xa[RG], xb[RG];
#pragma unroll
            for (int u = 0; u < RG; ++u) {
              const int jj = min(j + u, 63);
              xa[u] = P[(int64_t)__builtin_amdgcn_readlane(cla, jj) * C + c];
              xb[u] = P[(int64_t)__builtin_amdgcn_readlane(clb, jj) * C + c];
            }
#pragma unroll
            for (int u = 0; u < RG; ++u) {
              const int jj = min(j + u, 63);
              const T aa = (j + u < la) ? rl_bcast(vla, jj) : sc<T>::zero();
              const T ab = (j + u < lb) ? rl_bcast(vlb, jj) : sc<T>::zero();
              ya = ya + aa * xa[u];
              yb = yb + ab * xb[u];
            }
          }
          const int64_t oa = (b0 + i) * C + c;
          Q[oa] = ya;
          pv[ch][0] = acc_add(pv[ch][0], amul(P[oa], ya));
          if (two) {
            const int64_t ob = oa + C;
            Q[ob] = yb;
            pv[ch][0] = acc_add(pv[ch][0], amul(P[ob], yb));
          }
        }
      }
    }
    // rows longer than 64 entries: finish the remainder one row at a time
    for (int i = 0; i < nr; ++i) {
      const int32_t ra = __builtin_amdgcn_readlane(rp, i), rb = __builtin_amdgcn_readlane(rp, i + 1);
      if (rb - ra <= 64) continue;
#pragma unroll
      for (int ch = 0; ch < NCH; ++ch) {
        const int c = lane + ch * 64;
        if (c >= C) continue;
        const T* v = vals + (int64_t)((ch * 64) / k) * m.nnz;
        T y = sc<T>::zero();
        for (int32_t j = ra + 64; j < rb; ++j) y = y + v[j] * P[(int64_t)m.col_idx[j] * C + c];
        const int64_t oa = (b0 + i) * C + c;
        const T q0 = Q[oa];
        Q[oa] = q0 + y;
        pv[ch][0] = acc_add(pv[ch][0], amul(P[oa], y));
      }
    }
  }
  // block reduction of the per-wave partials (same layout as block_partials with G = 64)
  block_partials<A, 64, NCH>(st, pv, 1, C);
}

// r -= alpha q ; z = dinv r ; partials rz, rr.  With the polynomial preconditioner (st.poly)
// z and the partials come later: Y given, y = D^-1 r goes there; else only r is updated.
template <typename T, int G, int NCH>
__global__ void __launch_bounds__(KB) k_cg_update(int64_t n, const T* __restrict__ dinv, T* __restrict__ R,
                                                  const T* __restrict__ Q, int C, int k,
                                                  ColState<typename sc<T>::acc> st, T* __restrict__ Y) {
  typedef typename sc<T>::acc A;
  if (!block_any_active(st.active, C)) return;
  A pv[NCH][MAXND];
  A al[NCH];
  bool act[NCH];  // converged columns (alpha 0) keep r and need no partials
#pragma unroll
  for (int ch = 0; ch < NCH; ++ch) {
    pv[ch][0] = acc_zero(A());
    pv[ch][1] = acc_zero(A());
    int c = (int)(threadIdx.x % G) + ch * G;
    al[ch] = (c < C) ? st.alpha[c] : acc_zero(A());
    act[ch] = c < C && (st.active[c] || acc_abs2(al[ch]) != 0.0);
  }
  ROW_LOOP(n) {
#pragma unroll
    for (int ch = 0; ch < NCH; ++ch) {
      int c = lane + ch * G;
      if (act[ch]) {
        int s = c / k;
        int64_t o = row * C + c;
        T r = R[o];
        if (acc_abs2(al[ch]) != 0.0) {
          r = r - scale_to(Q[o], al[ch]);
          R[o] = r;
        }
        if (Y) {  // y = D^-1 r for the preconditioner's product; k_cg_poly_z forms z
          Y[o] = dinv[(int64_t)s * n + row] * r;
        } else if (!st.poly) {
          const T z = dinv[(int64_t)s * n + row] * r;
          pv[ch][0] = acc_add(pv[ch][0], amul(r, z));
          pv[ch][1] = acc_add(pv[ch][1], acc_from_real(sc<T>::norm2(r), A()));
        }
      }
    }
  }
  if (!Y && !st.poly) block_partials<A, G, NCH>(st, pv, 2, C);
}


// Polynomial (degree-1, Jacobi-based) preconditioner: z = p(D^-1 A) D^-1 r with
// p(x) = c0 + c1 x = (3 / L)(1 - 0.8 x / L), L a Gershgorin bound of D^-1 A's
// spectrum per system (p > 0 on [0, L], so M = p(D^-1 A) D^-1 stays positive
// for SPD A).  On the 3-D FEM systems it cuts CG iterations ~1.8x for one more
// SpMM per iteration.  YZ holds y = D^-1 r on entry and z on return, W = A y;
// partial r.z (unconjugated) and r.r go to st as k_cg_update's do.
template <typename T, int G, int NCH>
__global__ void __launch_bounds__(KB) k_cg_poly_z(int64_t n, const T* __restrict__ dinv, const T* __restrict__ R,
                                                  T* YZ, const T* __restrict__ W, const float* __restrict__ lbound,
                                                  int C, int k, ColState<typename sc<T>::acc> st) {
  typedef typename sc<T>::acc A;
  typedef typename RealOf<T>::type R1;
  if (!block_any_active(st.active, C)) return;
  A pv[NCH][MAXND];
  bool act[NCH];
  R1 c0[NCH], c1[NCH];
#pragma unroll
  for (int ch = 0; ch < NCH; ++ch) {
    pv[ch][0] = acc_zero(A());
    pv[ch][1] = acc_zero(A());
    int c = (int)(threadIdx.x % G) + ch * G;
    act[ch] = c < C && st.active[c];
    const R1 L = (R1)(c < C ? lbound[c / k] : 2.0f);
    c0[ch] = (R1)3.0 / L;
    c1[ch] = (R1)(-3.0 * POLY_K) / (L * L);
  }
  ROW_LOOP(n) {
#pragma unroll
    for (int ch = 0; ch < NCH; ++ch) {
      int c = lane + ch * G;
      if (act[ch]) {
        int s = c / k;
        int64_t o = row * C + c;
        const T r = R[o];
        const T z = rscale(YZ[o], c0[ch]) + rscale(dinv[(int64_t)s * n + row] * W[o], c1[ch]);
        YZ[o] = z;
        pv[ch][0] = acc_add(pv[ch][0], amul(r, z));
        pv[ch][1] = acc_add(pv[ch][1], acc_from_real(sc<T>::norm2(r), A()));
      }
    }
  }
  block_partials<A, G, NCH>(st, pv, 2, C);
}

// Gershgorin bound of D^-1 A per system: max over rows of sum_j |a_rj| / |a_rr|; workgroup
// (x, s) takes 256 rows of system s, and its maximum goes out by one atomic max on the bit
// pattern (monotonic for non-negative floats); out[s] starts at 0
template <typename T>
__global__ void __launch_bounds__(256) k_gersh(Csr m, const T* __restrict__ vals, unsigned* __restrict__ out) {
  const int sy = blockIdx.y;
  const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  float v = 0.0f;
  if (row < m.n) {
    const T* a = vals + (int64_t)sy * m.nnz;
    double sum = 0.0;
    for (int32_t j = m.row_ptr[row]; j < m.row_ptr[row + 1]; ++j) sum += sqrt((double)sc<T>::norm2(a[j]));
    const int32_t d = m.diag[row];
    const double ad = d >= 0 ? sqrt((double)sc<T>::norm2(a[d])) : 0.0;
    v = ad > 0.0 ? (float)(sum / ad) * 1.0001f : 0.0f;  // rounded up past the float conversion
  }
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
  __shared__ float wmax[4];
  if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) atomicMax(out + sy, __float_as_uint(fmaxf(fmaxf(wmax[0], wmax[1]), fmaxf(wmax[2], wmax[3]))));
}

// y = D^-1 r over every column (the start of a polynomial-preconditioned solve)
template <typename T>
__global__ void k_poly_y(int64_t n, const T* __restrict__ dinv, const T* __restrict__ R, T* __restrict__ Y, int C,
                         int k) {
  const int64_t total = n * C;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t row = i / C;
    const int c = (int)(i - row * C);
    Y[i] = dinv[(int64_t)(c / k) * n + row] * R[i];
  }
}

// start of a polynomial-preconditioned solve: rho = r.z of the first z (k_cg_poly_z's
// partials), alpha = beta = 0 (the next p-update takes p = z)
template <typename A>
__global__ void __launch_bounds__(FB) k_fin_rho(int C, ColState<A> st) {
  int c = blockIdx.x;
  if (!st.active[c]) return;
  A v[1];
  finish_reduce(st, 1, C, c, v);
  if (threadIdx.x == 0) {
    st.rho[c] = v[0];
    st.alpha[c] = acc_zero(A());
    st.beta[c] = acc_zero(A());
  }
}

// true when some column is active or still moves by alpha p (a column that
// stopped in this iteration's k_fin_beta takes its last step here)
template <typename A>
__device__ inline bool block_any_live(const ColState<A>& st, int C) {
  __shared__ int s_any;
  if (threadIdx.x == 0) s_any = 0;
  __syncthreads();
  for (int c = threadIdx.x; c < C; c += blockDim.x)
    if (st.active[c] || acc_abs2(st.alpha[c]) != 0.0) s_any = 1;
  __syncthreads();
  return s_any != 0;
}

// x += alpha p (before p changes), then p = dinv r + beta p on active
// columns.  Mixed solves (X64 given): X is the 32-bit correction d, and on
// the 64-column chunks whose residual this iteration's SpMM launch replaced
// (ru_prev) the replacement was taken at x64 + d, so d moves into x64 and
// restarts at alpha p.
template <typename T, typename TW, int G, int NCH>
__global__ void __launch_bounds__(KB) k_cg_pupdate(int64_t n, const T* __restrict__ dinv, T* __restrict__ X,
                                                   const T* __restrict__ R, T* __restrict__ P, int C, int k,
                                                   TW* __restrict__ X64, ColState<typename sc<T>::acc> st,
                                                   const T* __restrict__ Z) {
  typedef typename sc<T>::acc A;
  if (!block_any_live(st, C)) return;
  A al[NCH], be[NCH];
  bool act[NCH], fold[NCH];
#pragma unroll
  for (int ch = 0; ch < NCH; ++ch) {
    int c = (int)(threadIdx.x % G) + ch * G;
    al[ch] = (c < C) ? st.alpha[c] : acc_zero(A());
    be[ch] = (c < C) ? st.beta[c] : acc_zero(A());
    act[ch] = c < C && st.active[c];
    fold[ch] = X64 && st.ru_req && c < C && ru_prev(st)[c >> 6];
  }
  ROW_LOOP(n) {
#pragma unroll
    for (int ch = 0; ch < NCH; ++ch) {
      int c = lane + ch * G;
      if (c < C) {
        const bool move = acc_abs2(al[ch]) != 0.0;
        if (!(move || act[ch] || fold[ch])) continue;
        int s = c / k;
        int64_t o = row * C + c;
        T p = (move || act[ch]) ? P[o] : sc<T>::zero();
        if (fold[ch]) {
          X64[o] = X64[o] + widen(X[o], TW());
          X[o] = move ? scale_to(p, al[ch]) : sc<T>::zero();
        } else if (move) {
          X[o] = X[o] + scale_to(p, al[ch]);
        }
        if (act[ch]) P[o] = (Z ? Z[o] : dinv[(int64_t)s * n + row] * R[o]) + scale_to(p, be[ch]);
      }
    }
  }
}

// Float32 systems of 64 columns each in chunk pairs (the pair sweep's
// layout): lane l handles columns 2l and 2l+1 of each 128-column pair with
// 8-byte accesses; per column the same arithmetic as k_cg_update /
// k_cg_pupdate (lanes 0-31 are the pair's first system, 32-63 its second).
template <int NCH>
__device__ void block_partials_pair2(const ColState<double>& st, double (&pv)[NCH / 2][2][2], int C) {
  __shared__ double lds[KB / 64][2][128];
  const int lane = threadIdx.x & 63, rg = threadIdx.x >> 6;
  constexpr int RPB = KB / 64;
#pragma unroll
  for (int cp = 0; cp < NCH / 2; ++cp) {
#pragma unroll
    for (int d = 0; d < 2; ++d) {
      lds[rg][d][2 * lane] = pv[cp][0][d];
      lds[rg][d][2 * lane + 1] = pv[cp][1][d];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 2 * 128; i += KB) {
      const int d = i >> 7, j = i & 127;
      const int c = cp * 128 + j;
      double acc = lds[0][d][j];
      for (int r = 1; r < RPB; ++r) acc = acc + lds[r][d][j];
      st.part[(int64_t)(d * C + c) * st.nb + blockIdx.x] = acc;
    }
    __syncthreads();
  }
}

template <typename TS, int NCH>
__global__ void __launch_bounds__(KB) k_cg_update_pair(int64_t n, const TS* __restrict__ dinv,
                                                       TS* __restrict__ R, const TS* __restrict__ Q, int C,
                                                       ColState<double> st) {
  typedef typename pair_of<TS>::type fl2;
  constexpr int G = 64;
  constexpr int NP = NCH / 2;
  if (!block_any_active(st.active, C)) return;
  double pv[NP][2][2], al[NP][2];
  bool act[NP][2];
#pragma unroll
  for (int cp = 0; cp < NP; ++cp)
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      const int c = cp * 128 + 2 * (int)(threadIdx.x & 63) + e;
      al[cp][e] = st.alpha[c];
      act[cp][e] = st.active[c] || al[cp][e] != 0.0;
      pv[cp][e][0] = pv[cp][e][1] = 0.0;
    }
  ROW_LOOP(n) {
#pragma unroll
    for (int cp = 0; cp < NP; ++cp) {
      if (!(act[cp][0] || act[cp][1])) continue;
      const int sys = 2 * cp + (lane >> 5);
      const int64_t o = row * C + cp * 128 + 2 * lane;
      fl2 r = *reinterpret_cast<const fl2*>(R + o);
      if (al[cp][0] != 0.0 || al[cp][1] != 0.0) {
        const fl2 q = *reinterpret_cast<const fl2*>(Q + o);
        if (al[cp][0] != 0.0) r.x = r.x - scale_to(q.x, al[cp][0]);
        if (al[cp][1] != 0.0) r.y = r.y - scale_to(q.y, al[cp][1]);
        *reinterpret_cast<fl2*>(R + o) = r;
      }
      if (st.poly) continue;  // the preconditioner's product forms z and the partials
      const TS d = dinv[(int64_t)sys * n + row];
      if (act[cp][0]) {
        pv[cp][0][0] += amul(r.x, (TS)(d * r.x));
        pv[cp][0][1] += sc<TS>::norm2(r.x);
      }
      if (act[cp][1]) {
        pv[cp][1][0] += amul(r.y, (TS)(d * r.y));
        pv[cp][1][1] += sc<TS>::norm2(r.y);
      }
    }
  }
  if (!st.poly) block_partials_pair2<NCH>(st, pv, C);
}

template <typename TS, int NCH>
__global__ void __launch_bounds__(KB) k_cg_pupdate_pair(int64_t n, const TS* __restrict__ dinv,
                                                        TS* __restrict__ X, const TS* __restrict__ R,
                                                        TS* __restrict__ P, int C, double* __restrict__ X64,
                                                        ColState<double> st, const TS* __restrict__ Z) {
  typedef typename pair_of<TS>::type fl2;
  constexpr int G = 64;
  constexpr int NP = NCH / 2;
  if (!block_any_live(st, C)) return;
  double al[NP][2], be[NP][2];
  bool act[NP][2], mv[NP][2], fold[NP];
#pragma unroll
  for (int cp = 0; cp < NP; ++cp) {
    const int c0 = cp * 128 + 2 * (int)(threadIdx.x & 63);
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      al[cp][e] = st.alpha[c0 + e];
      be[cp][e] = st.beta[c0 + e];
      act[cp][e] = st.active[c0 + e];
      mv[cp][e] = al[cp][e] != 0.0;
    }
    fold[cp] = X64 && st.ru_req && ru_prev(st)[c0 >> 6];  // both columns lie in one 64-column chunk
  }
  ROW_LOOP(n) {
#pragma unroll
    for (int cp = 0; cp < NP; ++cp) {
      const bool use0 = mv[cp][0] || act[cp][0], use1 = mv[cp][1] || act[cp][1];
      if (!(use0 || use1 || fold[cp])) continue;
      const int sys = 2 * cp + (lane >> 5);
      const int64_t o = row * C + cp * 128 + 2 * lane;
      const fl2 p = (use0 || use1) ? *reinterpret_cast<const fl2*>(P + o) : fl2{0, 0};
      if (fold[cp]) {
        fl2 x = *reinterpret_cast<const fl2*>(X + o);
        dl2 x64 = *reinterpret_cast<const dl2*>(X64 + o);
        x64.x = x64.x + (double)x.x;
        x64.y = x64.y + (double)x.y;
        *reinterpret_cast<dl2*>(X64 + o) = x64;
        x.x = mv[cp][0] ? scale_to(p.x, al[cp][0]) : (TS)0;
        x.y = mv[cp][1] ? scale_to(p.y, al[cp][1]) : (TS)0;
        *reinterpret_cast<fl2*>(X + o) = x;
      } else if (mv[cp][0] || mv[cp][1]) {
        fl2 x = *reinterpret_cast<const fl2*>(X + o);
        if (mv[cp][0]) x.x = x.x + scale_to(p.x, al[cp][0]);
        if (mv[cp][1]) x.y = x.y + scale_to(p.y, al[cp][1]);
        *reinterpret_cast<fl2*>(X + o) = x;
      }
      if (act[cp][0] || act[cp][1]) {
        fl2 z;  // the preconditioned residual: given (polynomial), or D^-1 r
        if (Z) {
          z = *reinterpret_cast<const fl2*>(Z + o);
        } else {
          const fl2 r = *reinterpret_cast<const fl2*>(R + o);
          const TS d = dinv[(int64_t)sys * n + row];
          z.x = d * r.x;
          z.y = d * r.y;
        }
        fl2 pn = p;
        if (act[cp][0]) pn.x = z.x + scale_to(p.x, be[cp][0]);
        if (act[cp][1]) pn.y = z.y + scale_to(p.y, be[cp][1]);
        *reinterpret_cast<fl2*>(P + o) = pn;
      }
    }
  }
}


// y = sum_j A_sj (x[col_j, c] + d[col_j, c]) for one row, one wavefront per
// row (lane c of a 64-column chunk): the row's (col, val) entries come in
// with one lane-parallel load and are broadcast by readlane, so RG gathers
// of x and of the widened 32-bit correction d (d may be null) are in flight
template <typename T64, typename T32>
__device__ inline T64 row_gather_w64_xd(const int32_t* __restrict__ col_idx, const T64* __restrict__ vals,
                                        int32_t rs, int32_t re, int64_t row, const T64* __restrict__ X,
                                        const T32* __restrict__ D, int64_t ld, int lane, int c, bool live) {
  T64 y = sc<T64>::zero();
  for (int32_t base = rs; base < re; base += 64) {
    const int len = min(64, re - base);
    const int32_t cl = (lane < len) ? col_idx[base + lane] : (int32_t)row;
    const T64 vl = (lane < len) ? vals[base + lane] : sc<T64>::zero();
    for (int j = 0; j < len; j += RG) {
      T64 x[RG];
#pragma unroll
      for (int u = 0; u < RG; ++u) {
        const int64_t o = (int64_t)__builtin_amdgcn_readlane(cl, min(j + u, 63)) * ld + (live ? c : 0);
        x[u] = D ? X[o] + widen(D[o], T64()) : X[o];
      }
#pragma unroll
      for (int u = 0; u < RG; ++u) {
        const T64 a = (j + u < len) ? rl_bcast(vl, min(j + u, 63)) : sc<T64>::zero();
        y = y + a * x[u];
      }
    }
  }
  return y;
}

// y = sum_j A_sj (x[col_j] + d[col_j]) for one row and the lane's column
// (any column group width; d may be null)
template <typename T64, typename T32>
__device__ inline T64 row_gather_xd(const int32_t* __restrict__ col_idx, const T64* __restrict__ vals, int32_t rs,
                                    int32_t re, const T64* __restrict__ X, const T32* __restrict__ D, int64_t ld,
                                    int c, bool live) {
  T64 y = sc<T64>::zero();
  if (!live) return y;
  for (int32_t j = rs; j < re; ++j) {
    const int64_t o = (int64_t)col_idx[j] * ld + c;
    y = y + vals[j] * (D ? X[o] + widen(D[o], T64()) : X[o]);
  }
  return y;
}

// true residual of the mixed solve: r32 = (T32)(b - A64 (x64 + d32)) into
// the inner residual buffer; partials rr64, bb64 and rz32 = r32 . (dinv32 r32).
// Gated by *gate when given; launched on one resident round of workgroups so
// a gated-off launch costs one flag read per workgroup.
template <typename T64, typename T32, int G, int NCH>
__global__ void __launch_bounds__(KB) k_outer_resid(Csr m, const T64* __restrict__ vals, const T64* __restrict__ B,
                                                    const T64* __restrict__ X, const T32* __restrict__ D, int use_x,
                                                    const T32* __restrict__ dinv32, T32* __restrict__ R32, int C,
                                                    int k, const int* __restrict__ gate,
                                                    ColState<typename sc<T64>::acc> st) {
  typedef typename sc<T64>::acc A;
  if (gate) {
    int any = 0;
    for (int i = 0; i < MAXCHUNK; ++i) any |= gate[i];
    if (!any) return;
  }
  A pv[NCH][MAXND];
#pragma unroll
  for (int ch = 0; ch < NCH; ++ch) {
    pv[ch][0] = acc_zero(A());
    pv[ch][1] = acc_zero(A());
    pv[ch][2] = acc_zero(A());
  }
  ROW_LOOP(m.n) {
    int32_t rs = m.row_ptr[row], re = m.row_ptr[row + 1];
    T64 ax[NCH];
    if (use_x) {
#pragma unroll
      for (int ch = 0; ch < NCH; ++ch) {
        const int c = lane + ch * G;
        const T64* v = vals + (int64_t)((c < C ? c : 0) / k) * m.nnz;
        ax[ch] = sc<T64>::zero();
        if (ch * G >= C) continue;  // a chunk past the last column has no system (and no values)
        if (G == 64 && (k % 64 == 0 || k == C))  // each 64-column chunk within one system
          ax[ch] = row_gather_w64_xd<T64, T32>(m.col_idx, vals + (int64_t)((ch * G) / k) * m.nnz, rs, re, row, X,
                                               D, C, lane, c, c < C);
        else
          ax[ch] = row_gather_xd<T64, T32>(m.col_idx, v, rs, re, X, D, C, c, c < C);
      }
    }
#pragma unroll
    for (int ch = 0; ch < NCH; ++ch) {
      int c = lane + ch * G;
      if (c < C && (!gate || gate[c >> 6])) {  // only the requested 64-column chunks
        int s = c / k;
        int64_t o = row * C + c;
        T64 b = B[o];
        T64 r = use_x ? b - ax[ch] : b;
        T32 r32 = narrow(r, T32());
        T32 z = dinv32[(int64_t)s * m.n + row] * r32;
        R32[o] = r32;
        pv[ch][0] = acc_add(pv[ch][0], acc_from_real(sc<T64>::norm2(r), A()));
        pv[ch][1] = acc_add(pv[ch][1], acc_from_real(sc<T64>::norm2(b), A()));
        pv[ch][2] = acc_add(pv[ch][2], amul(r32, z));
      }
    }
  }
  block_partials<A, G, NCH>(st, pv, 3, C);
}

// ---------------------------------------------------------------------------
// finish kernels: one block per column
// ---------------------------------------------------------------------------
template <typename A>
__global__ void __launch_bounds__(FB) k_fin_init(int C, ColState<A> st) {
  int c = blockIdx.x;
  A v[3];
  finish_reduce(st, 3, C, c, v);
  if (threadIdx.x == 0) {
    double bb = acc_re(v[2]);
    double rr = acc_re(v[1]);
    st.bb[c] = bb;
    st.rr[c] = rr;
    st.rho[c] = v[0];
    st.iters[c] = 0;
    st.alpha[c] = acc_zero(A());
    st.beta[c] = acc_zero(A());
    st.flag[c] = 0;
    st.active[c] = (bb > 0.0) && (rr > st.tol2 * bb);
  }
}

template <typename A>
__global__ void __launch_bounds__(FB) k_fin_alpha(int C, ColState<A> st) {
  int c = blockIdx.x;
  if (!st.active[c]) {
    if (threadIdx.x == 0) st.alpha[c] = acc_zero(A());
    return;
  }
  A v[1];
  finish_reduce(st, 1, C, c, v);
  if (threadIdx.x == 0) {
    A pq = v[0];
    if (acc_abs2(pq) == 0.0) {
      st.flag[c] = 4;  // breakdown
      st.active[c] = 0;
      st.alpha[c] = acc_zero(A());
    } else {
      st.alpha[c] = acc_div(st.rho[c], pq);
    }
  }
}

// one column's arrival at the chunk-wide tolerance trigger (thread 0 of its k_fin_beta
// block): the chunk's last column to arrive requests a replacement when no active
// column is above the tolerance and at least one is below it, then clears the counters
template <typename A>
__device__ void ru_arrive(const ColState<A>& st, int C, int c, bool live, bool below) {
  const int ch = c >> 6;
  int* cnt = st.ru_cnt;
  if (live) atomicAdd(cnt + (below ? 2 : 1) * MAXCHUNK + ch, 1);
  __threadfence();
  const int ncol = min(64, C - (ch << 6));
  if (atomicAdd(cnt + ch, 1) == ncol - 1) {
    __threadfence();
    const int above = atomicAdd(cnt + MAXCHUNK + ch, 0);
    const int nbelow = atomicAdd(cnt + 2 * MAXCHUNK + ch, 0);
    if (above == 0 && nbelow > 0) ru_cur(st)[ch] = 1;
    atomicExch(cnt + ch, 0);
    atomicExch(cnt + MAXCHUNK + ch, 0);
    atomicExch(cnt + 2 * MAXCHUNK + ch, 0);
  }
}

template <typename A>
__global__ void __launch_bounds__(FB) k_fin_beta(int C, ColState<A> st) {
  int c = blockIdx.x;
  if (!st.active[c]) {
    if (threadIdx.x == 0) {
      st.beta[c] = acc_zero(A());
      if (st.delta2 > 0.0) ru_arrive(st, C, c, false, false);
    }
    return;
  }
  A v[2];
  finish_reduce(st, 2, C, c, v);
  if (threadIdx.x == 0) {
    double rr = acc_re(v[1]);
    st.rr[c] = rr;
    int it = st.iters[c] + 1;
    st.iters[c] = it;
    A rz = v[0];
    if (st.delta2 > 0.0) {  // mixed: the float64 true residual decides (next iteration's k_fin_alpha_ru)
      // a column below the tolerance keeps iterating until every active column of its chunk is
      // (one replacement confirms them together); above it, a drop by delta since the last
      // replacement requests one at once
      const bool below = rr <= st.tol2ru * st.bb[c];
      if (!below && rr <= st.delta2 * st.rmax2[c]) ru_cur(st)[c >> 6] = 1;
      bool live = false;
      if (it >= st.maxiter) {
        st.active[c] = 0;
        st.flag[c] = 1;
        st.beta[c] = acc_zero(A());
      } else if (acc_abs2(rz) == 0.0 || acc_abs2(st.rho[c]) == 0.0) {
        st.active[c] = 0;
        st.flag[c] = 4;
        st.beta[c] = acc_zero(A());
      } else {
        st.rho_prev[c] = st.rho[c];
        st.beta[c] = acc_div(rz, st.rho[c]);
        st.rho[c] = rz;
        live = true;
      }
      ru_arrive(st, C, c, live, below);
    } else if (rr <= st.tol2 * st.bb[c]) {
      st.active[c] = 0;
      st.beta[c] = acc_zero(A());
    } else if (it >= st.maxiter) {
      st.active[c] = 0;
      st.flag[c] = 1;
      st.beta[c] = acc_zero(A());
    } else if (acc_abs2(rz) == 0.0 || acc_abs2(st.rho[c]) == 0.0) {
      st.active[c] = 0;
      st.flag[c] = 4;
      st.beta[c] = acc_zero(A());
    } else {
      st.beta[c] = acc_div(rz, st.rho[c]);
      st.rho[c] = rz;
    }
  }
}

// mixed solves, before the first iteration: the float64 true residual of
// x0 (partials in so) initialises the inner recurrence of every column
template <typename A>
__global__ void __launch_bounds__(FB) k_fin_ru(int C, double tol2, ColState<A> so, ColState<A> si) {
  int c = blockIdx.x;
  A v[3];
  finish_reduce(so, 3, C, c, v);
  if (threadIdx.x == 0) {
    double rr = acc_re(v[0]), bb = acc_re(v[1]);
    so.rr[c] = rr;
    so.bb[c] = bb;
    si.bb[c] = bb;
    si.rmax2[c] = rr;
    si.rho[c] = v[2];
    si.rho_prev[c] = v[2];
    si.beta[c] = acc_zero(A());
    si.alpha[c] = acc_zero(A());
    si.iters[c] = 0;
    si.flag[c] = 0;
    si.active[c] = !(!(bb > 0.0) || rr <= tol2 * bb);
  }
}

// mixed solves: alpha = rho / p.q, and for the columns whose chunk was
// replaced in this iteration's SpMM launch (the previous iteration's
// request) first the reliable update (Sleijpen & van der Vorst): the float64
// true residual decides convergence and restarts rmax, and rho becomes
// r32.(dinv r32) of the replaced residual.  The search direction is kept.
template <typename A>
__global__ void __launch_bounds__(FB) k_fin_alpha_ru(int C, double tol2, ColState<A> si, ColState<A> so) {
  int c = blockIdx.x;
  if (!si.active[c]) {
    if (threadIdx.x == 0) si.alpha[c] = acc_zero(A());
    return;
  }
  const bool ru = ru_prev(si)[c >> 6] != 0;
  A v[3];
  if (ru) finish_reduce(so, 3, C, c, v);
  A pq[1];
  finish_reduce(si, 1, C, c, pq);
  if (threadIdx.x == 0) {
    if (ru) {
      double rr = acc_re(v[0]), bb = acc_re(v[1]);
      so.rr[c] = rr;
      so.bb[c] = bb;
      si.bb[c] = bb;
      si.rmax2[c] = rr;
      if (!(bb > 0.0) || rr <= tol2 * bb) {  // converged
        si.active[c] = 0;
        si.alpha[c] = acc_zero(A());
        return;
      }
      if (acc_abs2(v[2]) == 0.0) {
        si.active[c] = 0;
        si.flag[c] = 4;
        si.alpha[c] = acc_zero(A());
        return;
      }
      if (!si.poly) si.rho[c] = v[2];  // the polynomial z of the new residual comes with the next update
    }
    if (acc_abs2(pq[0]) == 0.0) {
      si.flag[c] = 4;  // breakdown
      si.active[c] = 0;
      si.alpha[c] = acc_zero(A());
    } else {
      si.alpha[c] = acc_div(si.rho[c], pq[0]);
    }
  }
}


template <typename T32>
__global__ void k_mask_cols(int64_t total, int C, const int* __restrict__ active, T32* __restrict__ R32) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stp = (int64_t)gridDim.x * blockDim.x;
  for (; i < total; i += stp) {
    int c = (int)(i % C);
    if (!active[c]) R32[i] = sc<T32>::zero();
  }
}

// x64 += d32 ; d32 = 0   (skipped unless *gate when gate is given)
template <typename T64, typename T32>
__global__ void k_accum(int64_t total, T32* __restrict__ D, T64* __restrict__ X, const int* __restrict__ gate) {
  if (gate && !*gate) return;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stp = (int64_t)gridDim.x * blockDim.x;
  for (; i < total; i += stp) {
    X[i] = X[i] + widen(D[i], T64());
    D[i] = sc<T32>::zero();
  }
}

template <typename T>
__global__ void k_dinv(Csr m, const T* __restrict__ vals, int nsys, T* __restrict__ dinv) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.n * nsys) return;
  int64_t s = i / m.n, row = i - s * m.n;
  int32_t d = m.diag[row];
  dinv[i] = (d >= 0) ? tdiv(one_of(T()), vals[s * m.nnz + d]) : sc<T>::zero();
}

template <typename T, int G, int NCH>
__global__ void __launch_bounds__(KB) k_spmm(Csr m, const T* __restrict__ vals, const T* __restrict__ X,
                                             T* __restrict__ Y, int C, int k) {
  ROW_LOOP(m.n) {
    int32_t rs = m.row_ptr[row], re = m.row_ptr[row + 1];
#pragma unroll
    for (int ch = 0; ch < NCH; ++ch) {
      int c = lane + ch * G;
      if (c < C) {
        const T* v = vals + (int64_t)(c / k) * m.nnz;
        T acc = sc<T>::zero();
        for (int32_t j = rs; j < re; ++j) acc = acc + v[j] * X[(int64_t)m.col_idx[j] * C + c];
        Y[row * C + c] = acc;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static int pick_G(int C) { return C <= 4 ? 4 : (C <= 16 ? 16 : 64); }
static int pick_NCH(int C) {
  int g = pick_G(C);
  int n = (C + g - 1) / g;
  return n <= 1 ? 1 : (n <= 2 ? 2 : (n <= 4 ? 4 : 8));
}
static int64_t grid_for(int64_t n, int C) {
  int64_t rpb = KB / pick_G(C);
  int64_t g = ceil_div(ceil_div(n, NXCD), rpb) * NXCD;  // whole row segments per die
  if (g > NBMAX) g = NBMAX;
  if (g < NXCD) g = NXCD;
  return g;
}

// dispatch a kernel template over (G, NCH)
#define DISPATCH_GN(C, KERNEL, ...)                                       \
  do {                                                                    \
    int _g = pick_G(C), _n = pick_NCH(C);                                 \
    if (_g == 4) {                                                        \
      KERNEL(4, 1, __VA_ARGS__);                                          \
    } else if (_g == 16) {                                                \
      KERNEL(16, 1, __VA_ARGS__);                                         \
    } else if (_n == 1) {                                                 \
      KERNEL(64, 1, __VA_ARGS__);                                         \
    } else if (_n == 2) {                                                 \
      KERNEL(64, 2, __VA_ARGS__);                                         \
    } else if (_n == 4) {                                                 \
      KERNEL(64, 4, __VA_ARGS__);                                         \
    } else {                                                              \
      KERNEL(64, 8, __VA_ARGS__);                                         \
    }                                                                     \
  } while (0)

struct Layout {
  size_t off = 0;
  size_t take(size_t bytes) {
    size_t o = off;
    off += (bytes + 255) & ~(size_t)255;
    return o;
  }
};

template <typename A>
static void state_layout(Layout& L, int64_t C, int64_t nblocks, size_t* offs) {
  offs[0] = L.take(C * sizeof(A));                    // rho
  offs[1] = L.take(C * sizeof(A));                    // alpha
  offs[2] = L.take(C * sizeof(A));                    // beta
  offs[3] = L.take(C * sizeof(double));               // bb
  offs[4] = L.take(C * sizeof(double));               // rr
  offs[5] = L.take(C * sizeof(int));                  // active
  offs[6] = L.take(C * sizeof(int));                  // iters
  offs[7] = L.take(C * sizeof(int));                  // flag
  offs[8] = L.take(64);                               // nactive
  offs[9] = L.take(nblocks * MAXND * C * sizeof(A));  // part
  offs[10] = L.take(C * sizeof(A));                   // rho_prev
  offs[11] = L.take(C * sizeof(double));              // rmax2
}

template <typename A>
static ColState<A> state_at(char* base, const size_t* o, int64_t nblocks) {
  ColState<A> s;
  s.rho = (A*)(base + o[0]);
  s.alpha = (A*)(base + o[1]);
  s.beta = (A*)(base + o[2]);
  s.bb = (double*)(base + o[3]);
  s.rr = (double*)(base + o[4]);
  s.active = (int*)(base + o[5]);
  s.iters = (int*)(base + o[6]);
  s.flag = (int*)(base + o[7]);
  s.nactive = (int*)(base + o[8]);
  s.part = (A*)(base + o[9]);
  s.rho_prev = (A*)(base + o[10]);
  s.rmax2 = (double*)(base + o[11]);
  s.ru_req = nullptr;
  s.ru_cnt = nullptr;
  s.poly = 0;
  s.ru_par = 0;
  s.nb = (int)nblocks;
  s.tol2 = 0;
  s.maxiter = 0;
  s.delta2 = 0;
  s.tol2ru = 0;
  return s;
}

template <typename T>
struct SolveBufs {
  T* R;
  T* P;
  T* Q;
  T* dinv;
};

// workspace: plain -> R,P,Q,dinv (T64) + one state; mixed -> R,P,Q,dinv,D
// (T32) + inner and outer states
// the preconditioner of the CG/COCG solves (rb_krylov_set_precond): 0 Jacobi, 1 polynomial
static std::atomic<int> g_precond{0};

extern "C" int rb_krylov_set_precond(int kind) {
  if (kind != RB_PRECOND_JACOBI && kind != RB_PRECOND_POLY && kind != RB_PRECOND_POLY2 &&
      kind != RB_PRECOND_POLY3) {
    rb_set_error("rb_krylov_set_precond: unknown preconditioner %d", kind);
    return RB_ERR_ARG;
  }
  g_precond.store(kind);
  return RB_OK;
}

extern "C" int rb_krylov_get_precond(void) { return g_precond.load(); }

template <typename T64, typename T32>
// (precond: the preconditioner setting the solve runs with, read once per solve)
static size_t ws_layout(int64_t n, int64_t C, int nsys, int mixed, size_t* o, int precond) {
  typedef typename sc<T64>::acc A;
  Layout L;
  int64_t nb = grid_for(n, (int)C);
  if (!mixed) {
    o[0] = L.take(n * C * sizeof(T64));
    o[1] = L.take(n * C * sizeof(T64));
    o[2] = L.take(n * C * sizeof(T64));
    o[3] = L.take(n * nsys * sizeof(T64));
    state_layout<A>(L, C, nb, o + 8);
    if (precond != RB_PRECOND_JACOBI) {
      o[5] = L.take(n * C * sizeof(T64) * (precond == RB_PRECOND_POLY3 ? 2 : 1));  // y / w1 (, w2)
      o[7] = L.take(std::max<int64_t>(64, nsys * 4));   // Gershgorin bound per system
    }
  } else {
    o[0] = L.take(n * C * sizeof(T32));
    o[1] = L.take(n * C * sizeof(T32));
    o[2] = L.take(n * C * sizeof(T32));
    o[3] = L.take(n * nsys * sizeof(T32));
    o[4] = L.take(n * C * sizeof(T32));  // D (inner solution)
    state_layout<A>(L, C, nb, o + 8);
    state_layout<A>(L, C, nb, o + 20);
    o[6] = L.take(5 * MAXCHUNK * sizeof(int));  // replacement requests (2 parities) + chunk counters
    if (precond != RB_PRECOND_JACOBI) {
      o[5] = L.take(n * C * sizeof(T32) * (precond == RB_PRECOND_POLY3 ? 2 : 1));
      o[7] = L.take(std::max<int64_t>(64, nsys * 4));
    }
  }
  return L.off;
}

static unsigned ew_grid(int64_t total) {
  int64_t g = ceil_div(total, 256);
  if (g > 8192) g = 8192;
  if (g < 1) g = 1;
  return (unsigned)g;
}

static int* g_pinned = nullptr;  // host flag for convergence polls

// set while rb_krylov_solve_noalloc runs on this thread: no host or device
// allocation (no pinned poll buffer, no timing events)
static thread_local bool t_noalloc = false;

// active columns and the largest rr / (tol2 bb) among them (one block): the poll the
// iteration loops size their next chunk from
__global__ void k_poll(const int* __restrict__ active, const double* __restrict__ rr, const double* __restrict__ bb,
                       double tol2, int C, int* out) {
  __shared__ int s_n[256];
  __shared__ double s_r[256];
  int n = 0;
  double r = 0.0;
  for (int c = threadIdx.x; c < C; c += blockDim.x) {
    if (!active[c]) continue;
    ++n;
    const double den = tol2 * bb[c];
    const double q = den > 0.0 ? rr[c] / den : 0.0;
    r = q > r ? q : r;
  }
  s_n[threadIdx.x] = n;
  s_r[threadIdx.x] = r;
  __syncthreads();
  for (int h = blockDim.x / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      s_n[threadIdx.x] += s_n[threadIdx.x + h];
      s_r[threadIdx.x] = fmax(s_r[threadIdx.x], s_r[threadIdx.x + h]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = s_n[0];
    *(double*)(out + 2) = s_r[0];
  }
}

// poll: active columns and the largest ratio (out_ratio; 0 when none is known)
template <typename A>
static int poll_state(const ColState<A>& st, int C, double tol2, hipStream_t s, int* out, double* out_ratio) {
  hipLaunchKernelGGL(k_poll, dim3(1), dim3(256), 0, s, st.active, st.rr, st.bb, tol2, C, st.nactive);
  RB_CHECK_LAUNCH();
  if (!g_pinned && !t_noalloc) RB_HIP(hipHostMalloc((void**)&g_pinned, 64, hipHostMallocDefault));
  int buf[4] = {0, 0, 0, 0};
  int* dst = g_pinned ? g_pinned : buf;  // a pageable destination completes before the copy returns
  RB_HIP(hipMemcpyAsync(dst, st.nactive, 16, hipMemcpyDeviceToHost, s));
  RB_HIP(hipStreamSynchronize(s));
  *out = dst[0];
  double r;
  memcpy(&r, dst + 2, sizeof(double));
  *out_ratio = r;
  return RB_OK;
}

// iterations to run before the next poll: check_every, or fewer when the decay of the largest
// residual ratio since the last poll says the active columns reach the tolerance sooner (a
// converged solve then stops within a few iterations instead of running up to check_every - 1
// rounds of gated launches); at least 2
static int next_chunk(int check_every, double ratio, double prev_ratio, int prev_chunk) {
  const int lo = check_every < 2 ? check_every : 2;
  if (!(ratio > 0.0) || !(prev_ratio > 0.0) || prev_chunk <= 0) return check_every;
  if (ratio <= 1.0) return lo;  // below the tolerance: the confirming replacement is next
  if (!(prev_ratio > ratio)) return check_every;
  const double per_it = log(prev_ratio / ratio) / prev_chunk;
  const double need = log(ratio) / per_it;
  if (!(need < (double)check_every)) return check_every;
  const int c = (int)ceil(need) + 1;
  return c < lo ? lo : (c > check_every ? check_every : c);
}

// ---------------------------------------------------------------------------
// optional per-kernel timing: hipEvents recorded around each launch on the
// launch stream, read back by rb_prof_read (RB_PROF_* ids in rbhip.h)
// ---------------------------------------------------------------------------
static int g_prof = 0;
static double g_prof_ms[RB_PROF_NKERNELS];
static int64_t g_prof_cnt[RB_PROF_NKERNELS];
static std::vector<std::pair<hipEvent_t, hipEvent_t>> g_prof_pending[RB_PROF_NKERNELS];
static std::vector<hipEvent_t> g_prof_pool;

static hipEvent_t prof_event() {
  hipEvent_t e = nullptr;
  if (!g_prof_pool.empty()) {
    e = g_prof_pool.back();
    g_prof_pool.pop_back();
  } else if (hipEventCreate(&e) != hipSuccess) {
    e = nullptr;
  }
  return e;
}

// id of the open ProfScope (-1: none).  Launches made through plaunch while
// a scope is open carry their start/stop events in the dispatch packet
// itself (hipExtLaunchKernel), so timing adds no packets to the stream.
static int g_prof_open = -1;

struct ProfScope {
  int prev;
  ProfScope(int id, hipStream_t) : prev(g_prof_open) {
    if (g_prof && !t_noalloc) g_prof_open = id;
  }
  ~ProfScope() { g_prof_open = prev; }
};

template <typename F, typename... Args>
static inline void plaunch(F kernel, dim3 grid, dim3 block, uint32_t shm, hipStream_t s, Args... args) {
  hipEvent_t a = nullptr, b = nullptr;
  if (g_prof && g_prof_open >= 0 && !t_noalloc) {
    a = prof_event();
    b = prof_event();
    if (!a || !b) a = b = nullptr;
  }
  hipExtLaunchKernelGGL(kernel, grid, block, shm, s, a, b, 0, args...);
  if (a) g_prof_pending[g_prof_open].push_back({a, b});
}

// the fused SpMM + residual launch, in its pair form for real 64-column systems
template <typename T64, typename T32, int N, typename... Args>
static inline void launch_spmm_ru(bool pair, dim3 grid, hipStream_t s, Args... args) {
  if constexpr (std::is_same<T32, float>::value && N % 2 == 0) {
    if (pair) {
      plaunch((k_cg_spmm_sweep_ru<T64, T32, N, true>), grid, dim3(KB), 0, s, args...);
      return;
    }
  }
  plaunch((k_cg_spmm_sweep_ru<T64, T32, N>), grid, dim3(KB), 0, s, args...);
}

// the plain sweep SpMM, in its pair form for real 64-column systems
template <typename T, int N, typename... Args>
static inline void launch_spmm_sweep(bool pair, dim3 grid, hipStream_t s, Args... args) {
  if constexpr ((std::is_same<T, float>::value || std::is_same<T, double>::value) && N % 2 == 0) {
    if (pair) {
      plaunch((k_cg_spmm_sweep<T, N, true>), grid, dim3(KB), 0, s, args...);
      return;
    }
  }
  plaunch((k_cg_spmm_sweep<T, N>), grid, dim3(KB), 0, s, args...);
}

// real systems of 64 columns each in whole chunk pairs: one sweep per pair of systems
template <typename T>
static bool pair_sweep(int nsys, int C, int k) {
  return (std::is_same<T, float>::value || std::is_same<T, double>::value) && k == 64 && C == nsys * 64 &&
         (C == 128 || C == 256 || C == 512);  // C == 64 * NCH, NCH even
}

// the vector updates in their pair form (real types; false: not taken)
template <typename T32, typename A>
static bool launch_update_pair(int C, int64_t nb, hipStream_t s, int64_t n, const T32* dinv, T32* R, const T32* Q,
                               ColState<A> st) {
  if constexpr (std::is_same<T32, float>::value || std::is_same<T32, double>::value) {
    const int nch = C / 64;
    if (nch == 2) plaunch(k_cg_update_pair<T32, 2>, dim3(nb), dim3(KB), 0, s, n, dinv, R, Q, C, st);
    else if (nch == 4) plaunch(k_cg_update_pair<T32, 4>, dim3(nb), dim3(KB), 0, s, n, dinv, R, Q, C, st);
    else plaunch(k_cg_update_pair<T32, 8>, dim3(nb), dim3(KB), 0, s, n, dinv, R, Q, C, st);
    return true;
  }
  return false;
}

template <typename T32, typename A, typename TW>
static bool launch_pupdate_pair(int C, int64_t nb, hipStream_t s, int64_t n, const T32* dinv, T32* X, const T32* R,
                                T32* P, TW* X64, ColState<A> st, const T32* Z = nullptr) {
  if constexpr (std::is_same<T32, float>::value || std::is_same<T32, double>::value) {
    const int nch = C / 64;
    double* x64 = (double*)X64;  // the float64 solution (mixed) or null
    if (nch == 2) plaunch(k_cg_pupdate_pair<T32, 2>, dim3(nb), dim3(KB), 0, s, n, dinv, X, R, P, C, x64, st, Z);
    else if (nch == 4) plaunch(k_cg_pupdate_pair<T32, 4>, dim3(nb), dim3(KB), 0, s, n, dinv, X, R, P, C, x64, st, Z);
    else plaunch(k_cg_pupdate_pair<T32, 8>, dim3(nb), dim3(KB), 0, s, n, dinv, X, R, P, C, x64, st, Z);
    return true;
  }
  return false;
}

template <typename T64, typename T32>
static const void* spmm_ru_fn(int nch, bool pair) {
  if constexpr (std::is_same<T32, float>::value) {
    if (pair) {
      return nch == 2 ? (const void*)k_cg_spmm_sweep_ru<T64, T32, 2, true>
                      : nch == 4 ? (const void*)k_cg_spmm_sweep_ru<T64, T32, 4, true>
                                 : (const void*)k_cg_spmm_sweep_ru<T64, T32, 8, true>;
    }
  }
  return nch == 1   ? (const void*)k_cg_spmm_sweep_ru<T64, T32, 1>
         : nch == 2 ? (const void*)k_cg_spmm_sweep_ru<T64, T32, 2>
         : nch == 4 ? (const void*)k_cg_spmm_sweep_ru<T64, T32, 4>
                    : (const void*)k_cg_spmm_sweep_ru<T64, T32, 8>;
}

extern "C" int rb_prof_enable(int on) {
  g_prof = on ? 1 : 0;
  return RB_OK;
}

extern "C" int rb_prof_reset(void) {
  for (int i = 0; i < RB_PROF_NKERNELS; ++i) {
    for (auto& pr : g_prof_pending[i]) {
      (void)hipEventSynchronize(pr.second);
      g_prof_pool.push_back(pr.first);
      g_prof_pool.push_back(pr.second);
    }
    g_prof_pending[i].clear();
    g_prof_ms[i] = 0;
    g_prof_cnt[i] = 0;
  }
  return RB_OK;
}

extern "C" int rb_prof_read(int id, double* total_ms, int64_t* count) {
  if (id < 0 || id >= RB_PROF_NKERNELS || !total_ms || !count) return RB_ERR_ARG;
  for (auto& pr : g_prof_pending[id]) {
    RB_HIP(hipEventSynchronize(pr.second));
    float ms = 0;
    RB_HIP(hipEventElapsedTime(&ms, pr.first, pr.second));
    g_prof_ms[id] += ms;
    g_prof_cnt[id] += 1;
    g_prof_pool.push_back(pr.first);
    g_prof_pool.push_back(pr.second);
  }
  g_prof_pending[id].clear();
  *total_ms = g_prof_ms[id];
  *count = g_prof_cnt[id];
  return RB_OK;
}

#define L_INIT(G_, N_, ...) plaunch((k_cg_init<T, G_, N_>), dim3(nb), dim3(KB), 0, s, __VA_ARGS__)
// q = A p launch over the solve's columns: the sweep kernel when every
// 64-column chunk lies in one system and P fits its buffer descriptor
// (< 2 GiB), on a grid of exactly one resident round of workgroups; the
// row-per-group kernel otherwise.  `flat` and `nbs` come from spmm_plan; `spmm_pair`
// (real 64-column systems in chunk pairs) is the caller's layout choice.
#define L_SPMM(G_, N_, ...)                                                                                 \
  do {                                                                                                    \
    if (G_ == 64 && flat)                                                                                 \
      launch_spmm_sweep<T, N_>(spmm_pair, dim3(nbs), s, __VA_ARGS__);                                       \
    else if (nsys == 1)                                                                                   \
      plaunch((k_cg_spmm<T, G_, N_, false>), dim3(nbs), dim3(KB), 0, s, __VA_ARGS__);                      \
    else                                                                                                  \
      plaunch((k_cg_spmm<T, G_, N_, true>), dim3(nbs), dim3(KB), 0, s, __VA_ARGS__);                       \
  } while (0)

// workgroups of kernel fn resident at once on the current device (one
// dispatch round), rounded down to a multiple of NXCD and at most nb;
// computed once per kernel and device
static int64_t resident_grid(const void* fn, int64_t nb) {
  struct Ent {
    const void* fn;
    int dev;
    int64_t g;
  };
  static std::mutex mu;
  static Ent cache[128];
  static int ncache = 0;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) dev = 0;
  std::lock_guard<std::mutex> lk(mu);
  int64_t g = 0;
  for (int i = 0; i < ncache; ++i)
    if (cache[i].fn == fn && cache[i].dev == dev) g = cache[i].g;
  if (!g) {
    int per = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, KB, 0) != hipSuccess || per < 1) per = 1;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) cus = NXCD;
    g = std::max<int64_t>(NXCD, (int64_t)per * cus / NXCD * NXCD);
    if (ncache < 128) cache[ncache++] = {fn, dev, g};
  }
  return std::min(g, nb);
}

// the column sweep's grid cap set by rb_krylov_set_sweep_grid (0: one resident round)
static std::atomic<int64_t> g_sweep_grid{0};

extern "C" int rb_krylov_set_sweep_grid(int64_t cap) {
  if (cap < 0) {
    rb_set_error("rb_krylov_set_sweep_grid: negative cap %lld", (long long)cap);
    return RB_ERR_ARG;
  }
  g_sweep_grid.store(cap);
  return RB_OK;
}

extern "C" int64_t rb_krylov_get_sweep_grid(void) { return g_sweep_grid.load(); }

template <typename T>
static void spmm_plan(const Csr& m, int nsys, int C, int k, int64_t nb, bool* flat, int64_t* nbs) {
  // the sweep addresses P through a buffer descriptor: under 2 GiB (RB_SPMM_FLAT_MAX_BYTES lowers the
  // limit, so tests reach the row-per-group kernels of larger solves on small meshes)
  int64_t lim = (int64_t)1 << 31;
  if (const char* e = getenv("RB_SPMM_FLAT_MAX_BYTES")) lim = std::min(lim, (int64_t)atoll(e));
  *flat = pick_G(C) == 64 && (nsys == 1 || k % 64 == 0) && m.nnz > 0 &&
          m.n * (int64_t)C * (int64_t)sizeof(T) < lim;
  *nbs = nb;
  if (!*flat) return;
  const int nch = pick_NCH(C);
  const void* fn = nch == 1   ? (const void*)k_cg_spmm_sweep<T, 1>
                   : nch == 2 ? (const void*)k_cg_spmm_sweep<T, 2>
                   : nch == 4 ? (const void*)k_cg_spmm_sweep<T, 4>
                              : (const void*)k_cg_spmm_sweep<T, 8>;
  *nbs = resident_grid(fn, nb);
  // RB_SWEEP_MAX_GRID caps the grid (tests use it to give each wavefront
  // many chunk groups on a small matrix)
  if (const char* e = getenv("RB_SWEEP_MAX_GRID")) {
    const int64_t cap = atoll(e) / NXCD * NXCD;
    if (cap >= NXCD) *nbs = std::min(*nbs, cap);
  }
  const int64_t tuned = g_sweep_grid.load() / NXCD * NXCD;
  if (tuned >= NXCD) *nbs = std::min(*nbs, tuned);
}
// one workgroup per CU (each holds both LDS block buffers), a multiple of NXCD
static int64_t lds_grid(int64_t nb) {
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess) dev = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < NXCD) cus = NXCD;
  return std::max<int64_t>(NXCD, std::min<int64_t>(nb, cus) / NXCD * NXCD);
}

// the gated float64 residual sweep of an LDS-staged solve, by 64-column chunks per launch
template <typename T64, typename T32>
static const void* resid_sweep_fn(int nch) {
  return nch == 1   ? (const void*)k_cg_resid_sweep<T64, T32, 1>
         : nch == 2 ? (const void*)k_cg_resid_sweep<T64, T32, 2>
         : nch == 4 ? (const void*)k_cg_resid_sweep<T64, T32, 4>
                    : (const void*)k_cg_resid_sweep<T64, T32, 8>;
}
template <typename T64, typename T32, typename A32, typename A64>
static void launch_resid_sweep(int nch, dim3 grid, hipStream_t s, const Csr& m, const T64* A64v, const T64* B,
                               const T64* X, const T32* D, const T32* dinv, T32* R, int C, int k, ColState<A32> ss,
                               ColState<A64> sol) {
  if (nch == 1)
    plaunch(k_cg_resid_sweep<T64, T32, 1>, grid, dim3(KB), 0, s, m, A64v, B, X, D, dinv, R, C, k, ss, sol);
  else if (nch == 2)
    plaunch(k_cg_resid_sweep<T64, T32, 2>, grid, dim3(KB), 0, s, m, A64v, B, X, D, dinv, R, C, k, ss, sol);
  else if (nch == 4)
    plaunch(k_cg_resid_sweep<T64, T32, 4>, grid, dim3(KB), 0, s, m, A64v, B, X, D, dinv, R, C, k, ss, sol);
  else
    plaunch(k_cg_resid_sweep<T64, T32, 8>, grid, dim3(KB), 0, s, m, A64v, B, X, D, dinv, R, C, k, ss, sol);
}

// the LDS-staged SpMM (its dynamic LDS opt-in is per kernel, set before each launch)
template <typename T, typename CF, bool RPAIR, typename A, bool POLY = false>
static void launch_spmm_lds_cf(const LdsArgs& pl, dim3 grid, hipStream_t s, const T* recv, const T* P, T* Q,
                               int64_t rs, int64_t coff, int cbase, ColState<A> st, const void* dinvp = nullptr,
                               int64_t dstride = 0, const float* lbound = nullptr, int C = 64,
                               const T* Rg = nullptr, int pdeg = 1, const T* W1g = nullptr) {
  (void)hipFuncSetAttribute((const void*)k_cg_spmm_lds<T, CF, RPAIR, POLY>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, CF::BYTES);
  plaunch(k_cg_spmm_lds<T, CF, RPAIR, POLY>, grid, dim3((LDS_NW + LDS_NP) * 64), CF::BYTES, s, pl.nblk, pl.brow,
          pl.bu, pl.ulist, pl.rinfo, pl.rinfo, pl.loff, recv, P, Q, rs, coff, cbase, st, dinvp, dstride, lbound, C,
          Rg, pdeg, W1g);
}
// the polynomial preconditioner's product with z formed in the epilogue (one 64-column system,
// pl.recv_ad packed): the residual R in, z out to Z, r.z and r.r partials into st.  W (degree 2):
// first w1 = A D^-1 r into W (a plain launch on the A D^-1 records; its p.q partials are
// overwritten by the second launch's), then the epilogue's launch on the staged w1 rows
template <typename T, typename A>
static void launch_spmm_lds_poly(const LdsArgs& pl, dim3 grid, hipStream_t s, const T* R, T* Z, const T* dinv,
                                 const unsigned* lbound, ColState<A> st, T* W = nullptr, T* W2 = nullptr) {
  typedef typename std::conditional<sizeof(T) == 8, LdsWide, LdsHalf>::type CF;
  if constexpr (sizeof(T) == 8 || sizeof(T) == 4) {
    if (W && W2) {  // degree 3: w1, w2 = A D^-1 w1, then the epilogue's launch on w2 (r, w1 from HBM)
      launch_spmm_lds_cf<T, CF, false, A, false>(pl, grid, s, (const T*)pl.recv_ad, R, W, 64, 0, 0, st);
      launch_spmm_lds_cf<T, CF, false, A, false>(pl, grid, s, (const T*)pl.recv_ad, (const T*)W, W2, 64, 0, 0, st);
      launch_spmm_lds_cf<T, CF, false, A, true>(pl, grid, s, (const T*)pl.recv_ad, (const T*)W2, Z, 64, 0, 0, st,
                                                (const void*)dinv, 0, (const float*)lbound, 64, R, 3,
                                                (const T*)W);
    } else if (W) {
      launch_spmm_lds_cf<T, CF, false, A, false>(pl, grid, s, (const T*)pl.recv_ad, R, W, 64, 0, 0, st);
      launch_spmm_lds_cf<T, CF, false, A, true>(pl, grid, s, (const T*)pl.recv_ad, (const T*)W, Z, 64, 0, 0, st,
                                                (const void*)dinv, 0, (const float*)lbound, 64, R, 2);
    } else {
      launch_spmm_lds_cf<T, CF, false, A, true>(pl, grid, s, (const T*)pl.recv_ad, R, Z, 64, 0, 0, st,
                                                (const void*)dinv, 0, (const float*)lbound, 64);
    }
  }
}
// one 64-column system with 8-byte values (512-byte rows) or 4-byte values (256-byte rows)
template <typename T, typename A>
static void launch_spmm_lds(const LdsArgs& pl, dim3 grid, hipStream_t s, const T* P, T* Q, ColState<A> st) {
  if constexpr (sizeof(T) == 8) {
    launch_spmm_lds_cf<T, LdsWide, false>(pl, grid, s, (const T*)pl.recv, P, Q, 64, 0, 0, st);
  } else if constexpr (sizeof(T) == 4) {
    launch_spmm_lds_cf<T, LdsHalf, false>(pl, grid, s, (const T*)pl.recv, P, Q, 64, 0, 0, st);
  }
}
// real float32 systems of 64 columns each, in pairs: one launch per pair, record set p for pair p
template <typename A>
static void launch_spmm_lds_pairs(const LdsArgs& pl, int64_t nrec_set, dim3 grid, hipStream_t s, int nsys,
                                  const float* P, float* Q, ColState<A> st) {
  const int C = 64 * nsys;
  for (int p = 0; p < nsys / 2; ++p)
    launch_spmm_lds_cf<cplx<float>, LdsWide, true>(pl, grid, s, (const cplx<float>*)pl.recv + p * nrec_set,
                                                   (const cplx<float>*)P, (cplx<float>*)Q, C / 2, 64 * p, 128 * p,
                                                   st);
}

// ... and the polynomial preconditioner's product of such pairs (pl.recv_ad: the pairs' A D^-1
// records, set p for pair p): the residual R in, z out to Z, r.z and r.r partials into st; dinv
// holds every system's float D^-1, system s at dinv + s * n
// (W: the degree-2 form, w1 = A D^-1 r of every pair into W first, as launch_spmm_lds_poly does)
template <typename A>
static void launch_spmm_lds_pairs_poly(const LdsArgs& pl, int64_t nrec_set, dim3 grid, hipStream_t s, int nsys,
                                       int64_t n, const float* R, float* Z, const float* dinv,
                                       const unsigned* lbound, ColState<A> st, float* W = nullptr,
                                       float* W2 = nullptr) {
  typedef cplx<float> c2;
  const int C = 64 * nsys;
  for (int p = 0; p < nsys / 2; ++p) {
    const c2* rec = (const c2*)pl.recv_ad + p * nrec_set;
    if (W && W2) {  // degree 3
      launch_spmm_lds_cf<c2, LdsWide, true, A, false>(pl, grid, s, rec, (const c2*)R, (c2*)W, C / 2, 64 * p,
                                                      128 * p, st);
      launch_spmm_lds_cf<c2, LdsWide, true, A, false>(pl, grid, s, rec, (const c2*)W, (c2*)W2, C / 2, 64 * p,
                                                      128 * p, st);
      launch_spmm_lds_cf<c2, LdsWide, true, A, true>(pl, grid, s, rec, (const c2*)W2, (c2*)Z, C / 2, 64 * p,
                                                     128 * p, st, (const void*)dinv, n, (const float*)lbound, C,
                                                     (const c2*)R, 3, (const c2*)W);
    } else if (W) {
      launch_spmm_lds_cf<c2, LdsWide, true, A, false>(pl, grid, s, rec, (const c2*)R, (c2*)W, C / 2, 64 * p,
                                                      128 * p, st);
      launch_spmm_lds_cf<c2, LdsWide, true, A, true>(pl, grid, s, rec, (const c2*)W, (c2*)Z, C / 2, 64 * p, 128 * p,
                                                     st, (const void*)dinv, n, (const float*)lbound, C,
                                                     (const c2*)R, 2);
    } else {
      launch_spmm_lds_cf<c2, LdsWide, true, A, true>(pl, grid, s, rec, (const c2*)R, (c2*)Z, C / 2, 64 * p, 128 * p,
                                                     st, (const void*)dinv, n, (const float*)lbound, C);
    }
  }
}

// a plan applies to 64-column solves of one system with 8-byte or 4-byte values
static bool lds_plan_ok(const rb_spmm_plan* plan) {
  const bool cfg_ok = plan && plan->umax == LdsWide::UMAX && plan->recmax == LdsWide::RECMAX;
  return cfg_ok && plan->nblk > 0 && plan->brow && plan->bu && plan->ulist && plan->rinfo && plan->rmap &&
         plan->loff && plan->recv && plan->recv_sets >= 1;
}
static bool lds_eligible(const rb_spmm_plan* plan, int nsys, int C, size_t value_bytes) {
  return lds_plan_ok(plan) && nsys == 1 && C == 64 && (value_bytes == 8 || value_bytes == 4);
}
// ... and to real float32 systems of 64 columns each in pairs, one record set per pair
static bool lds_pairs_eligible(const rb_spmm_plan* plan, int nsys, int k, size_t value_bytes) {
  return lds_plan_ok(plan) && k == 64 && nsys >= 2 && nsys % 2 == 0 && value_bytes == 4 &&
         plan->recv_sets >= nsys / 2;
}

static LdsArgs lds_args(const rb_spmm_plan* plan) {
  return LdsArgs{plan->nblk, plan->umax, plan->brow, plan->bu, plan->ulist, plan->rinfo, plan->loff, plan->recv,
                 nullptr};
}

#define L_UPD(G_, N_, ...) plaunch((k_cg_update<T, G_, N_>), dim3(nb), dim3(KB), 0, s, __VA_ARGS__)
#define L_PZ(G_, N_, ...) plaunch((k_cg_poly_z<T, G_, N_>), dim3(nb), dim3(KB), 0, s, __VA_ARGS__)
#define L_PUP(G_, N_, ...) plaunch((k_cg_pupdate<T, TW, G_, N_>), dim3(nb), dim3(KB), 0, s, __VA_ARGS__)

// one CG/COCG solve of all columns; X is the solution (and x0 if use_x0)
template <typename T>
static int run_cg(const Csr& m, int nsys, const T* vals, const T* dinv, const T* B, T* X, int use_x0,
                  SolveBufs<T> w, int C, int k, ColState<typename sc<T>::acc> st, int check_every, hipStream_t s,
                  int* iters_done, const LdsArgs* lds = nullptr, T* YZ = nullptr, unsigned* lb = nullptr,
                  int precond = RB_PRECOND_JACOBI) {
  typedef typename sc<T>::acc A;
  typedef T TW;
  int64_t nb = grid_for(m.n, C);
  st.nb = (int)nb;
  bool flat;
  int64_t nbs;
  spmm_plan<T>(m, nsys, C, k, nb, &flat, &nbs);
  if (lds) nbs = lds_grid(nb);
  ColState<A> ss = st;  // the SpMM's partials come from its own grid
  ss.nb = (int)nbs;
  // polynomial preconditioner (real 64-column systems then run in the one-column layout rather
  // than in chunk pairs); with the LDS plan's A D^-1 records its product stages r itself and
  // forms z in its epilogue
  const bool poly = YZ && lb;
  const bool spmm_pair = pair_sweep<T>(nsys, C, k) && !poly;
  const bool lds_poly = poly && lds && lds->recv_ad;
  // the degree-2 polynomial where the fused product applies (its w1 = A D^-1 r goes to YZ)
  const int pdeg = lds_poly ? (precond == RB_PRECOND_POLY3 ? 3 : precond == RB_PRECOND_POLY2 ? 2 : 1) : 1;
  T* W2 = pdeg >= 2 ? YZ : nullptr;                  // w1
  T* W3 = pdeg == 3 ? YZ + (int64_t)m.n * C : nullptr;  // w2
  st.poly = ss.poly = poly ? 1 : 0;
  auto spmm = [&](const T* Pin, T* Qout) {
    if (lds)
      launch_spmm_lds<T>(*lds, dim3(nbs), s, Pin, Qout, ss);
    else
      DISPATCH_GN(C, L_SPMM, m, vals, Pin, Qout, C, k, ss);
  };
  {
    ProfScope ps(RB_PROF_INIT, s);
    DISPATCH_GN(C, L_INIT, m, vals, dinv, B, X, use_x0, w.R, w.P, C, k, st);
    hipLaunchKernelGGL(k_fin_init<A>, dim3(C), dim3(FB), 0, s, C, st);
    if (poly) {  // z0 = M r0, rho = r0.z0, p0 = z0
      int rcf = rb_fill_i32((int32_t*)lb, 0, nsys, s);
      if (rcf) return rcf;
      hipLaunchKernelGGL(k_gersh<T>, dim3((unsigned)ceil_div(m.n, 256), nsys), dim3(256), 0, s, m, vals, lb);
      if (lds_poly) {  // z formed in the LDS product's epilogue, into the spent q
        launch_spmm_lds_poly<T>(*lds, dim3(nbs), s, (const T*)w.R, w.Q, dinv, lb, ss, W2, W3);
        hipLaunchKernelGGL(k_fin_rho<A>, dim3(C), dim3(FB), 0, s, C, ss);
        DISPATCH_GN(C, L_PUP, m.n, dinv, X, w.R, w.P, C, k, (TW*)nullptr, st, (const T*)w.Q);
      } else {
        hipLaunchKernelGGL(k_poly_y<T>, dim3(ew_grid(m.n * C)), dim3(256), 0, s, m.n, dinv, (const T*)w.R, YZ, C,
                           k);
        spmm(YZ, w.Q);
        DISPATCH_GN(C, L_PZ, m.n, dinv, (const T*)w.R, YZ, (const T*)w.Q, (const float*)lb, C, k, st);
        hipLaunchKernelGGL(k_fin_rho<A>, dim3(C), dim3(FB), 0, s, C, st);
        DISPATCH_GN(C, L_PUP, m.n, dinv, X, w.R, w.P, C, k, (TW*)nullptr, st, (const T*)YZ);
      }
    }
  }
  RB_CHECK_LAUNCH();
  int nact = 0;
  double ratio = 0.0, prev_ratio = 0.0;
  int rc = poll_state(st, C, st.tol2, s, &nact, &ratio);
  if (rc) return rc;
  int it = 0, last = 0;
  while (nact > 0 && it < st.maxiter) {
    int chunk = std::min(next_chunk(check_every, ratio, prev_ratio, last), st.maxiter - it);
    for (int j = 0; j < chunk; ++j) {
      {
        ProfScope ps(RB_PROF_SPMM, s);
        spmm(w.P, w.Q);
      }
      hipLaunchKernelGGL(k_fin_alpha<A>, dim3(C), dim3(FB), 0, s, C, ss);
      {
        ProfScope ps(RB_PROF_UPDATE, s);
        if (!spmm_pair || !launch_update_pair<T>(C, nb, s, m.n, dinv, w.R, w.Q, st))
          DISPATCH_GN(C, L_UPD, m.n, dinv, w.R, w.Q, C, k, st, poly && !lds_poly ? YZ : (T*)nullptr);
      }
      const T* zin = nullptr;  // the polynomial preconditioner's z for the p-update
      bool zlds = false;       // its r.z, r.r partials on the LDS grid
      if (lds_poly) {          // w = A D^-1 r with z formed in the epilogue, into the spent q
        ProfScope ps(RB_PROF_POLY, s);
        launch_spmm_lds_poly<T>(*lds, dim3(nbs), s, (const T*)w.R, w.Q, dinv, lb, ss, W2, W3);
        zin = w.Q;
        zlds = true;
      } else if (poly) {  // w = A y into the spent q, then z and the r.z, r.r partials
        ProfScope ps(RB_PROF_POLY, s);
        spmm(YZ, w.Q);
        DISPATCH_GN(C, L_PZ, m.n, dinv, (const T*)w.R, YZ, (const T*)w.Q, (const float*)lb, C, k, st);
        zin = YZ;
      }
      hipLaunchKernelGGL(k_fin_beta<A>, dim3(C), dim3(FB), 0, s, C, zlds ? ss : st);
      {
        ProfScope ps(RB_PROF_PUPDATE, s);
        if (!spmm_pair || !launch_pupdate_pair<T>(C, nb, s, m.n, dinv, X, w.R, w.P, (TW*)nullptr, st))
          DISPATCH_GN(C, L_PUP, m.n, dinv, X, w.R, w.P, C, k, (TW*)nullptr, st, zin);
      }
    }
    RB_CHECK_LAUNCH();
    it += chunk;
    last = chunk;
    prev_ratio = ratio;
    rc = poll_state(st, C, st.tol2, s, &nact, &ratio);
    if (rc) return rc;
  }
  *iters_done = it;
  return RB_OK;
}

template <typename T64, typename T32>
static int solve_impl(const Csr& m, int nsys, int k, const void* vals64, const void* vals32, const void* Bv, void* Xv,
                      int precision, double tol, int maxiter, int use_x0, int check_every, void* work,
                      size_t work_bytes, int* iters_out, double* relres_out, int* flag_out, hipStream_t s,
                      const rb_spmm_plan* plan = nullptr) {
  typedef typename sc<T64>::acc A;
  const int C = nsys * k;
  const int64_t n = m.n;
  size_t o[32];
  // the preconditioner setting, once: the workspace layout and every choice below follow it even
  // when another thread changes the setting during the solve
  const int precond = g_precond.load();
  size_t need = ws_layout<T64, T32>(n, C, nsys, precision, o, precond);
  if (work_bytes < need || !work) {
    rb_set_error("rb_krylov_solve: workspace %zu bytes < required %zu", work_bytes, need);
    return RB_ERR_SIZE;
  }
  char* base = (char*)work;
  int64_t nb = grid_for(n, C);
  const T64* B = (const T64*)Bv;
  T64* X = (T64*)Xv;
  const double tol2 = tol * tol;
  int rc;
  // per-column results (C <= 512, checked by the entry points); stack arrays, no allocation
  double hrr[512], hbb[512];
  int hflag[512], hit[512];

  if (!precision) {
    SolveBufs<T64> w{(T64*)(base + o[0]), (T64*)(base + o[1]), (T64*)(base + o[2]), (T64*)(base + o[3])};
    ColState<A> st = state_at<A>(base, o + 8, nb);
    st.tol2 = tol2;
    st.maxiter = maxiter;
    const T64* A64 = (const T64*)vals64;
    hipLaunchKernelGGL(k_dinv<T64>, dim3((unsigned)ceil_div(n * nsys, 256)), dim3(256), 0, s, m, A64, nsys, w.dinv);
    RB_CHECK_LAUNCH();
    int its = 0;
    LdsArgs la{};
    const bool use_lds = lds_eligible(plan, nsys, C, sizeof(T64));
    if (use_lds) {  // this matrix's values into the plan's records
      la = lds_args(plan);
      const int64_t nr = plan->nrec + RB_LDS_RECPAD;
      hipLaunchKernelGGL(k_lds_pack<T64>, dim3(ew_grid(nr)), dim3(256), 0, s, nr, plan->rmap, A64, (T64*)plan->recv);
      RB_CHECK_LAUNCH();
    }
    const bool poly = precond != RB_PRECOND_JACOBI;
    if (use_lds && poly && plan->recv_sets >= 2) {  // the second record set: A D^-1
      const int64_t nr = plan->nrec + RB_LDS_RECPAD;
      la.recv_ad = (const T64*)plan->recv + nr;
      hipLaunchKernelGGL(k_lds_pack_ad<T64>, dim3(ew_grid(nr)), dim3(256), 0, s, nr, plan->rmap, A64, m.col_idx,
                         (const T64*)w.dinv, (T64*)la.recv_ad);
      RB_CHECK_LAUNCH();
    }
    rc = run_cg<T64>(m, nsys, A64, w.dinv, B, X, use_x0, w, C, k, st, check_every, s, &its, use_lds ? &la : nullptr,
                     poly ? (T64*)(base + o[5]) : nullptr, poly ? (unsigned*)(base + o[7]) : nullptr, precond);
    if (rc) return rc;
    RB_HIP(hipMemcpyAsync(hrr, st.rr, C * sizeof(double), hipMemcpyDeviceToHost, s));
    RB_HIP(hipMemcpyAsync(hbb, st.bb, C * sizeof(double), hipMemcpyDeviceToHost, s));
    RB_HIP(hipMemcpyAsync(hflag, st.flag, C * sizeof(int), hipMemcpyDeviceToHost, s));
    RB_HIP(hipMemcpyAsync(hit, st.iters, C * sizeof(int), hipMemcpyDeviceToHost, s));
    RB_HIP(hipStreamSynchronize(s));
    int mx = 0, fl = 0;
    double rel = 0;
    for (int c = 0; c < C; ++c) {
      mx = std::max(mx, hit[c]);
      fl = std::max(fl, hflag[c]);
      if (hbb[c] > 0) rel = std::max(rel, sqrt(hrr[c] / hbb[c]));
    }
    if (iters_out) *iters_out = mx;
    if (relres_out) *relres_out = rel;
    if (flag_out) *flag_out = fl;
    return RB_OK;
  }

  // mixed precision: CG/COCG in T32 on vals32 with reliable updates
  // (Sleijpen & van der Vorst): when a column's recursive residual has
  // dropped by RU_DELTA since its last replacement (or reaches the
  // tolerance), x64 += d32 and every residual is replaced by the float64
  // true residual b - A64 x64, which alone decides convergence.  The search
  // direction is kept, so the iteration count stays that of one float64 solve.
  if (!vals32) {
    rb_set_error("rb_krylov_solve: mixed precision needs the 32-bit matrix values");
    return RB_ERR_ARG;
  }
  // RB_RU_DELTA overrides the drop factor (tuning runs)
  static const double RU_DELTA = [] {
    const char* e = getenv("RB_RU_DELTA");
    const double d = e ? atof(e) : 0.0;
    return d > 0.0 && d < 1.0 ? d : 0.02;
  }();
  SolveBufs<T32> w{(T32*)(base + o[0]), (T32*)(base + o[1]), (T32*)(base + o[2]), (T32*)(base + o[3])};
  T32* D = (T32*)(base + o[4]);
  int* ru_req = (int*)(base + o[6]);
  ColState<A> si = state_at<A>(base, o + 8, nb);
  ColState<A> so = state_at<A>(base, o + 20, nb);
  si.tol2 = 0.0;
  si.maxiter = maxiter;
  si.delta2 = RU_DELTA * RU_DELTA;
  si.tol2ru = tol2;
  si.ru_req = ru_req;
  si.ru_cnt = ru_req + 2 * MAXCHUNK;
  // the polynomial preconditioner (real 64-column systems then run in the one-column layout
  // rather than in chunk pairs)
  const bool poly = precond != RB_PRECOND_JACOBI;
  si.poly = poly ? 1 : 0;
  T32* YZ = poly ? (T32*)(base + o[5]) : nullptr;
  unsigned* lb = poly ? (unsigned*)(base + o[7]) : nullptr;
  so.tol2 = tol2;
  so.maxiter = maxiter;
  const T64* A64 = (const T64*)vals64;
  const T32* A32 = (const T32*)vals32;
  hipLaunchKernelGGL(k_dinv<T32>, dim3((unsigned)ceil_div(n * nsys, 256)), dim3(256), 0, s, m, A32, nsys, w.dinv);
  RB_CHECK_LAUNCH();
  rc = rb_fill_bytes0(w.P, (size_t)n * C * sizeof(T32), s);
  if (rc) return rc;
  rc = rb_fill_bytes0(D, (size_t)n * C * sizeof(T32), s);
  if (rc) return rc;
  rc = rb_fill_i32(ru_req, 0, 5 * MAXCHUNK, s);
  if (rc) return rc;
  if (!use_x0) {
    rc = rb_fill_bytes0(X, (size_t)n * C * sizeof(T64), s);
    if (rc) return rc;
  }
  // the float64 true residual: once of x0 here (row-per-group kernel on the
  // regular grid), then in the loop for the chunks whose replacement the
  // previous iteration requested -- inside the SpMM launch when the sweep
  // applies, else a row-per-group launch gated by those flags
  const int64_t nbo = nb;
  so.nb = (int)nbo;
#define L_ORES(G_, N_, ...) \
  plaunch((k_outer_resid<T64, T32, G_, N_>), dim3(nbo), dim3(KB), 0, s, __VA_ARGS__)
  {
    ProfScope ps(RB_PROF_RESID, s);
    DISPATCH_GN(C, L_ORES, m, A64, B, X, (const T32*)nullptr, use_x0 ? 1 : 0, w.dinv, w.R, C, k,
                (const int*)nullptr, so);
    hipLaunchKernelGGL(k_fin_ru<A>, dim3(C), dim3(FB), 0, s, C, tol2, so, si);
  }
  RB_CHECK_LAUNCH();
#define T T32
  bool flat;
  int64_t nbs;
  spmm_plan<T32>(m, nsys, C, k, nb, &flat, &nbs);
  {
    // the fused residual pass addresses x64 through a buffer descriptor as well: its float64
    // rows must fit the same limit (else both kernels take the row-per-group path)
    bool flat64;
    int64_t unused;
    spmm_plan<T64>(m, nsys, C, k, nb, &flat64, &unused);
    flat = flat && flat64;
  }
  // real systems of 64 columns each, whole chunk pairs: one sweep per pair of systems (under the
  // polynomial preconditioner only where its pair-form LDS product applies: a plan holding the
  // pairs' A D^-1 records as well, sets nsys/2 .. nsys-1)
  const bool pair_shape = std::is_same<T32, float>::value && pair_sweep<T32>(nsys, C, k);
  const bool pair_poly = poly && pair_shape && flat && lds_pairs_eligible(plan, nsys, k, sizeof(T32)) &&
                         plan->recv_sets >= nsys;
  const bool pair = pair_shape && (!poly || pair_poly);
  const bool spmm_pair = pair;
  // the LDS-staged SpMM (a plan for this 64-column system): its own launch, then the
  // float64 residual of the requested chunks in a gated sweep launch
  // (or, for real float32 systems of 64 columns in pairs, one LDS launch per pair)
  const bool lds_one = flat && lds_eligible(plan, nsys, C, sizeof(T32));
  const bool lds_pairs = flat && pair && lds_pairs_eligible(plan, nsys, k, sizeof(T32));
  const bool use_lds = lds_one || lds_pairs;
  const int rnch = pick_NCH(C);
  LdsArgs la{};
  int64_t nbr = 0, nrec_set = 0;
  if (use_lds) {
    la = lds_args(plan);
    nrec_set = plan->nrec + RB_LDS_RECPAD;
    if (lds_one) {
      hipLaunchKernelGGL(k_lds_pack<T32>, dim3(ew_grid(nrec_set)), dim3(256), 0, s, nrec_set, plan->rmap, A32,
                         (T32*)plan->recv);
    } else if constexpr (std::is_same<T32, float>::value) {
      for (int pp = 0; pp < nsys / 2; ++pp)
        hipLaunchKernelGGL(k_lds_pack_pair, dim3(ew_grid(nrec_set)), dim3(256), 0, s, nrec_set, plan->rmap,
                           A32 + (int64_t)(2 * pp) * m.nnz, A32 + (int64_t)(2 * pp + 1) * m.nnz,
                           (cplx<float>*)plan->recv + pp * nrec_set);
    }
    RB_CHECK_LAUNCH();
    nbs = lds_grid(nb);
    nbr = resident_grid(resid_sweep_fn<T64, T32>(rnch), nb);
  } else if (flat) {  // the fused SpMM + residual launch has its own occupancy
    nbs = std::min(nbs, resident_grid(spmm_ru_fn<T64, T32>(pick_NCH(C), pair), nb));
  }
  ColState<A> ss = si;  // the SpMM's partials come from its own grid
  ss.nb = (int)nbs;
  ColState<A> sol = so;  // the in-loop residual's partials: its sweep's grid, or the regular one
  sol.nb = (int)(use_lds ? nbr : flat ? nbs : nbo);
  typedef T64 TW;
#define L_SPMMRU(G_, N_, ...)                                                    \
  do {                                                                         \
    if (G_ == 64) launch_spmm_ru<T64, T32, N_>(pair, dim3(nbs), s, __VA_ARGS__); \
  } while (0)
  // w = A y (the polynomial preconditioner's product), into the spent q
  auto spmm2 = [&](const T32* Pin, T32* Qout) {
    if (use_lds)
      launch_spmm_lds<T32>(la, dim3(nbs), s, Pin, Qout, ss);
    else
      DISPATCH_GN(C, L_SPMM, m, A32, Pin, Qout, C, k, ss);
  };
  // with the plan's second record set (A D^-1) the preconditioner's product stages r itself
  const bool lds_poly = poly && lds_one && plan->recv_sets >= 2;
  // the degree-2 polynomial where the fused product applies (its w1 = A D^-1 r goes to YZ)
  const int pdeg = lds_poly ? (precond == RB_PRECOND_POLY3 ? 3 : precond == RB_PRECOND_POLY2 ? 2 : 1) : 1;
  T32* W2 = pdeg >= 2 ? YZ : nullptr;              // w1
  T32* W3 = pdeg == 3 ? YZ + n * C : nullptr;      // w2
  if (lds_poly) {
    la.recv_ad = (const T32*)plan->recv + nrec_set;
    hipLaunchKernelGGL(k_lds_pack_ad<T32>, dim3(ew_grid(nrec_set)), dim3(256), 0, s, nrec_set, plan->rmap, A32,
                       m.col_idx, (const T32*)w.dinv, (T32*)la.recv_ad);
  }
  if constexpr (std::is_same<T32, float>::value) {
    if (pair_poly) {  // the pairs' A D^-1 records after their A records
      la.recv_ad = (const cplx<float>*)plan->recv + (int64_t)(nsys / 2) * nrec_set;
      for (int pp = 0; pp < nsys / 2; ++pp)
        hipLaunchKernelGGL(k_lds_pack_pair_ad, dim3(ew_grid(nrec_set)), dim3(256), 0, s, nrec_set, plan->rmap,
                           A32 + (int64_t)(2 * pp) * m.nnz, A32 + (int64_t)(2 * pp + 1) * m.nnz, m.col_idx,
                           (const float*)w.dinv + (int64_t)(2 * pp) * n,
                           (const float*)w.dinv + (int64_t)(2 * pp + 1) * n,
                           (cplx<float>*)la.recv_ad + pp * nrec_set);
    }
  }
  // the pair-form product of the polynomial preconditioner (float32 systems only; degrees 2 and 3
  // with w1 in YZ and w2 after it)
  auto pair_poly_product = [&](ColState<A> sz) {
    if constexpr (std::is_same<T32, float>::value)
      launch_spmm_lds_pairs_poly(la, nrec_set, dim3(nbs), s, nsys, n, (const float*)w.R, (float*)w.Q,
                                 (const float*)w.dinv, lb, sz,
                                 precond != RB_PRECOND_POLY ? (float*)YZ : (float*)nullptr,
                                 precond == RB_PRECOND_POLY3 ? (float*)YZ + n * C : (float*)nullptr);
  };
  if (poly) {  // z0 = M r0, rho = r0.z0, p0 = z0 (P is zero, alpha and beta 0)
    rc = rb_fill_i32((int32_t*)lb, 0, nsys, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_gersh<T32>, dim3((unsigned)ceil_div(n, 256), nsys), dim3(256), 0, s, m, A32, lb);
    if (lds_poly) {  // z formed in the LDS product's epilogue, into the spent q
      ColState<A> sz = si;
      sz.nb = (int)nbs;
      launch_spmm_lds_poly<T32>(la, dim3(nbs), s, (const T32*)w.R, w.Q, (const T32*)w.dinv, lb, sz, W2, W3);
      hipLaunchKernelGGL(k_fin_rho<A>, dim3(C), dim3(FB), 0, s, C, sz);
      DISPATCH_GN(C, L_PUP, m.n, w.dinv, D, w.R, w.P, C, k, X, si, (const T32*)w.Q);
    } else if (pair_poly) {  // the same with the pairs' product
      ColState<A> sz = si;
      sz.nb = (int)nbs;
      pair_poly_product(sz);
      hipLaunchKernelGGL(k_fin_rho<A>, dim3(C), dim3(FB), 0, s, C, sz);
      DISPATCH_GN(C, L_PUP, m.n, w.dinv, D, w.R, w.P, C, k, X, si, (const T32*)w.Q);
    } else {
      hipLaunchKernelGGL(k_poly_y<T32>, dim3(ew_grid(n * C)), dim3(256), 0, s, n, (const T32*)w.dinv,
                         (const T32*)w.R, YZ, C, k);
      spmm2(YZ, w.Q);
      DISPATCH_GN(C, L_PZ, m.n, w.dinv, (const T32*)w.R, YZ, (const T32*)w.Q, (const float*)lb, C, k, si);
      hipLaunchKernelGGL(k_fin_rho<A>, dim3(C), dim3(FB), 0, s, C, si);
      DISPATCH_GN(C, L_PUP, m.n, w.dinv, D, w.R, w.P, C, k, X, si, (const T32*)YZ);
    }
  } else {
    DISPATCH_GN(C, L_PUP, m.n, w.dinv, D, w.R, w.P, C, k, X, si,
                (const T32*)nullptr);  // p = z (P is zero, alpha 0, no flags)
  }
  RB_CHECK_LAUNCH();
  int nact = 0;
  // (the inner recursive residual against the float64 tolerance sizes the chunks; the float64
  // replacements alone decide convergence)
  double ratio = 0.0, prev_ratio = 0.0;
  rc = poll_state(si, C, tol2, s, &nact, &ratio);
  if (rc) return rc;
  int it = 0, last = 0;
  while (nact > 0 && it < maxiter) {
    int chunk = std::min(next_chunk(check_every, ratio, prev_ratio, last), maxiter - it);
    for (int j = 0; j < chunk; ++j) {
      // request flags: this iteration raises half (it+j)&1 and serves the other
      si.ru_par = ss.ru_par = sol.ru_par = (it + j) & 1;
      {
        ProfScope ps(RB_PROF_SPMM, s);
        if (use_lds) {
          if (lds_pairs) {
            if constexpr (std::is_same<T32, float>::value)
              launch_spmm_lds_pairs(la, nrec_set, dim3(nbs), s, nsys, w.P, w.Q, ss);
          } else {
            launch_spmm_lds<T32>(la, dim3(nbs), s, w.P, w.Q, ss);
          }
          ProfScope pr(RB_PROF_RESID, s);
          launch_resid_sweep<T64, T32>(rnch, dim3(nbr), s, m, A64, B, X, D, w.dinv, w.R, C, k, ss, sol);
        } else if (flat) {
          DISPATCH_GN(C, L_SPMMRU, m, A32, w.P, w.Q, A64, B, X, D, w.dinv, w.R, C, k, ss, sol);
        } else {
          DISPATCH_GN(C, L_SPMM, m, A32, w.P, w.Q, C, k, ss);
          DISPATCH_GN(C, L_ORES, m, A64, B, X, D, 1, w.dinv, w.R, C, k,
                      (const int*)(ru_req + MAXCHUNK * (si.ru_par ^ 1)), sol);
        }
      }
      hipLaunchKernelGGL(k_fin_alpha_ru<A>, dim3(C), dim3(FB), 0, s, C, tol2, ss, sol);
      {
        ProfScope ps(RB_PROF_UPDATE, s);
        if (!pair || !launch_update_pair<T32>(C, nb, s, m.n, w.dinv, w.R, w.Q, si))
          DISPATCH_GN(C, L_UPD, m.n, w.dinv, w.R, w.Q, C, k, si,
                      poly && !lds_poly && !pair_poly ? YZ : (T32*)nullptr);
      }
      const T32* zin = nullptr;  // the polynomial preconditioner's z for the p-update
      ColState<A> sz = si;       // ... and where its r.z, r.r partials are
      if (lds_poly) {            // w = A D^-1 r with z formed in the epilogue, into the spent q
        sz.nb = (int)nbs;
        ProfScope ps(RB_PROF_POLY, s);
        launch_spmm_lds_poly<T32>(la, dim3(nbs), s, (const T32*)w.R, w.Q, (const T32*)w.dinv, lb, sz, W2, W3);
        zin = w.Q;
      } else if (pair_poly) {  // the same, one launch per system pair
        sz.nb = (int)nbs;
        ProfScope ps(RB_PROF_POLY, s);
        pair_poly_product(sz);
        zin = w.Q;
      } else if (poly) {  // w = A y into the spent q, then z and the r.z, r.r partials
        ProfScope ps(RB_PROF_POLY, s);
        spmm2(YZ, w.Q);
        DISPATCH_GN(C, L_PZ, m.n, w.dinv, (const T32*)w.R, YZ, (const T32*)w.Q, (const float*)lb, C, k, si);
        zin = YZ;
      }
      hipLaunchKernelGGL(k_fin_beta<A>, dim3(C), dim3(FB), 0, s, C, sz);
      {
        ProfScope ps(RB_PROF_PUPDATE, s);
        if (!pair || !launch_pupdate_pair<T32>(C, nb, s, m.n, w.dinv, D, w.R, w.P, X, si, zin))
          DISPATCH_GN(C, L_PUP, m.n, w.dinv, D, w.R, w.P, C, k, X, si, zin);
      }
    }
    RB_CHECK_LAUNCH();
    it += chunk;
    last = chunk;
    prev_ratio = ratio;
    rc = poll_state(si, C, tol2, s, &nact, &ratio);
    if (rc) return rc;
  }
#undef T
#undef L_ORES
#undef L_SPMMRU
  // fold in the correction since the last replacement (columns stopped by maxiter)
  hipLaunchKernelGGL((k_accum<T64, T32>), dim3(ew_grid(n * C)), dim3(256), 0, s, (int64_t)n * C, D, X,
                     (const int*)nullptr);
  RB_CHECK_LAUNCH();
  RB_HIP(hipMemcpyAsync(hrr, so.rr, C * sizeof(double), hipMemcpyDeviceToHost, s));
  RB_HIP(hipMemcpyAsync(hbb, so.bb, C * sizeof(double), hipMemcpyDeviceToHost, s));
  RB_HIP(hipMemcpyAsync(hflag, si.flag, C * sizeof(int), hipMemcpyDeviceToHost, s));
  RB_HIP(hipMemcpyAsync(hit, si.iters, C * sizeof(int), hipMemcpyDeviceToHost, s));
  RB_HIP(hipStreamSynchronize(s));
  int fl = 0, mx = 0;
  double rel = 0;
  for (int c = 0; c < C; ++c) {
    fl = std::max(fl, hflag[c]);
    mx = std::max(mx, hit[c]);
    if (hbb[c] > 0) rel = std::max(rel, sqrt(hrr[c] / hbb[c]));
  }
  if (iters_out) *iters_out = mx;
  if (relres_out) *relres_out = rel;
  if (flag_out) *flag_out = fl;
  return RB_OK;
}

extern "C" int rb_krylov_workspace(int64_t n, int nsys, int k, int is_complex, int precision, size_t* bytes) {
  if (!bytes || n <= 0 || nsys <= 0 || k <= 0) return RB_ERR_ARG;
  size_t o[32];
  if (is_complex)
    *bytes = ws_layout<cplx<double>, cplx<float>>(n, (int64_t)nsys * k, nsys, precision, o, g_precond.load());
  else
    *bytes = ws_layout<double, float>(n, (int64_t)nsys * k, nsys, precision, o, g_precond.load());
  return RB_OK;
}

extern "C" int rb_krylov_solve(int64_t n, int64_t nnz, const int32_t* row_ptr, const int32_t* col_idx,
                               const int32_t* diag_slot, int nsys, int k, int is_complex, const void* vals64,
                               const void* vals32, const void* B, void* X, int precision, double tol, int maxiter,
                               int use_x0, int check_every, void* work, size_t work_bytes, int* iters_out,
                               double* relres_out, int* flag_out, void* stream) {
  if (n <= 0 || nsys <= 0 || k <= 0 || !vals64 || !B || !X) {
    rb_set_error("rb_krylov_solve: bad arguments");
    return RB_ERR_ARG;
  }
  int64_t C = (int64_t)nsys * k;
  if (C > 512) {
    rb_set_error("rb_krylov_solve: %lld columns exceed the 512-column limit of one launch", (long long)C);
    return RB_ERR_SIZE;
  }
  if (check_every <= 0) check_every = 16;
  if (maxiter <= 0) maxiter = 1000;
  Csr m{row_ptr, col_idx, diag_slot, n, nnz};
  hipStream_t s = (hipStream_t)stream;
  if (is_complex)
    return solve_impl<cplx<double>, cplx<float>>(m, nsys, k, vals64, vals32, B, X, precision, tol, maxiter, use_x0,
                                                 check_every, work, work_bytes, iters_out, relres_out, flag_out, s);
  return solve_impl<double, float>(m, nsys, k, vals64, vals32, B, X, precision, tol, maxiter, use_x0, check_every,
                                   work, work_bytes, iters_out, relres_out, flag_out, s);
}

// ---------------------------------------------------------------------------
// LDS SpMM plan (host): greedy blocks of consecutive rows, each closed when
// one more row would take its union of columns past RB_LDS_UMAX, its records
// past RB_LDS_RECMAX or its rows past RB_LDS_RMAX.  A block's union always
// holds its own rows (the p.q partials read each row's own p there).
// ---------------------------------------------------------------------------
namespace {
struct PlanScan {
  int64_t n;
  int umax = RB_LDS_UMAX, recmax = RB_LDS_RECMAX;
  const int32_t* rp;
  const int32_t* ci;
  std::vector<int64_t> mark;  // block id + 1 that holds the column in its union
  std::vector<int32_t> slot;  // the column's slot in that block's union
  std::vector<int32_t> uni;
  PlanScan(int64_t n_, const int32_t* rp_, const int32_t* ci_) : n(n_), rp(rp_), ci(ci_), mark(n_, 0), slot(n_, 0) {}
  static int64_t pad8(int64_t x) { return (x + 7) / 8 * 8; }
  // rows [r0, r1) of block blk; returns r1 (r0 when a single row does not fit)
  int64_t grow(int64_t r0, int64_t blk, int64_t* recs) {
    uni.clear();
    int64_t r = r0, rec = 0;
    while (r < n && r - r0 < RB_LDS_RMAX) {
      int add = mark[r] == blk + 1 ? 0 : 1;  // the row itself
      for (int32_t e = rp[r]; e < rp[r + 1]; ++e)
        if (mark[ci[e]] != blk + 1 && ci[e] != r) ++add;
      const int64_t pr = pad8(rp[r + 1] - rp[r]);
      if ((int64_t)uni.size() + add > umax || rec + pr > recmax) break;
      if (mark[r] != blk + 1) {
        mark[r] = blk + 1;
        uni.push_back((int32_t)r);
      }
      for (int32_t e = rp[r]; e < rp[r + 1]; ++e)
        if (mark[ci[e]] != blk + 1) {
          mark[ci[e]] = blk + 1;
          uni.push_back(ci[e]);
        }
      rec += pr;
      ++r;
    }
    *recs = rec;
    std::sort(uni.begin(), uni.end());
    for (size_t i = 0; i < uni.size(); ++i) slot[uni[i]] = (int32_t)i;
    return r;
  }
};

int plan_check(int64_t n, const int32_t* rp, const int32_t* ci) {
  if (n <= 0 || !rp || !ci) return RB_ERR_ARG;
  if (n >= (1LL << 31)) return RB_ERR_SIZE;
  for (int64_t r = 0; r < n; ++r)
    if (rp[r + 1] < rp[r]) return RB_ERR_ARG;
  for (int64_t e = 0; e < rp[n]; ++e)
    if (ci[e] < 0 || ci[e] >= n) return RB_ERR_ARG;
  return RB_OK;
}
}  // namespace

static bool plan_limits_ok(int umax, int recmax) {
  return umax == LdsWide::UMAX && recmax == LdsWide::RECMAX;
}

extern "C" int rb_spmm_plan_count(int64_t n, const int32_t* row_ptr, const int32_t* col_idx, int umax, int recmax,
                                  int64_t* nblk, int64_t* nunion, int64_t* nrec) {
  if (!nblk || !nunion || !nrec) return RB_ERR_ARG;
  if (!plan_limits_ok(umax, recmax)) {
    rb_set_error("rb_spmm_plan_count: no kernel for block limits (%d, %d)", umax, recmax);
    return RB_ERR_ARG;
  }
  int rc = plan_check(n, row_ptr, col_idx);
  if (rc) {
    rb_set_error("rb_spmm_plan_count: bad pattern (n %lld)", (long long)n);
    return rc;
  }
  PlanScan ps(n, row_ptr, col_idx);
  ps.umax = umax;
  ps.recmax = recmax;
  int64_t r = 0, blk = 0, nu = 0, nr = 0;
  while (r < n) {
    int64_t recs = 0;
    const int64_t r1 = ps.grow(r, blk, &recs);
    if (r1 == r) {
      rb_set_error("rb_spmm_plan_count: row %lld alone exceeds the block limits (%d neighbour rows, %d records)",
                   (long long)r, umax, recmax);
      return RB_ERR_SIZE;
    }
    nu += (int64_t)ps.uni.size();
    nr += recs;
    r = r1;
    ++blk;
  }
  *nblk = blk;
  *nunion = nu;
  *nrec = nr;
  return RB_OK;
}

extern "C" int rb_spmm_plan_fill(int64_t n, const int32_t* row_ptr, const int32_t* col_idx, int umax, int recmax,
                                 int32_t* brow, int32_t* bu, int32_t* ulist, int32_t* rinfo, int32_t* rmap,
                                 int32_t* loff) {
  if (!brow || !bu || !ulist || !rinfo || !rmap || !loff) return RB_ERR_ARG;
  if (!plan_limits_ok(umax, recmax)) {
    rb_set_error("rb_spmm_plan_fill: no kernel for block limits (%d, %d)", umax, recmax);
    return RB_ERR_ARG;
  }
  int rc = plan_check(n, row_ptr, col_idx);
  if (rc) {
    rb_set_error("rb_spmm_plan_fill: bad pattern (n %lld)", (long long)n);
    return rc;
  }
  PlanScan ps(n, row_ptr, col_idx);
  ps.umax = umax;
  ps.recmax = recmax;
  int64_t r = 0, blk = 0, nu = 0, nr = 0;
  brow[0] = 0;
  bu[0] = 0;
  while (r < n) {
    int64_t recs = 0;
    const int64_t r1 = ps.grow(r, blk, &recs);
    if (r1 == r) {
      rb_set_error("rb_spmm_plan_fill: row %lld alone exceeds the block limits", (long long)r);
      return RB_ERR_SIZE;
    }
    for (size_t i = 0; i < ps.uni.size(); ++i) ulist[nu + (int64_t)i] = ps.uni[i];
    nu += (int64_t)ps.uni.size();
    for (int64_t q = r; q < r1; ++q) {
      rinfo[2 * q] = (int32_t)nr;
      rinfo[2 * q + 1] = ps.slot[q] * 512;
      const int32_t e0 = row_ptr[q], e1 = row_ptr[q + 1];
      const int64_t pr = PlanScan::pad8(e1 - e0);
      for (int64_t j = 0; j < pr; ++j) {
        const bool real = e0 + j < e1;
        rmap[nr + j] = real ? (int32_t)(e0 + j) : -1;
        loff[nr + j] = real ? ps.slot[col_idx[e0 + j]] * 512 : 0;
      }
      nr += pr;
    }
    r = r1;
    ++blk;
    brow[blk] = (int32_t)r;
    bu[blk] = (int32_t)nu;
  }
  rinfo[2 * n] = (int32_t)nr;
  rinfo[2 * n + 1] = 0;
  for (int i = 0; i < RB_LDS_RPAD; ++i) rinfo[2 * (n + 1) + i] = 0;
  for (int i = 0; i < RB_LDS_UPAD; ++i) ulist[nu + i] = 0;
  for (int i = 0; i < RB_LDS_RECPAD; ++i) {
    rmap[nr + i] = -1;
    loff[nr + i] = 0;
  }
  return RB_OK;
}

extern "C" int rb_krylov_solve_planned(int64_t n, int64_t nnz, const int32_t* row_ptr, const int32_t* col_idx,
                                       const int32_t* diag_slot, int nsys, int k, int is_complex,
                                       const void* vals64, const void* vals32, const void* B, void* X, int precision,
                                       double tol, int maxiter, int use_x0, int check_every, void* work,
                                       size_t work_bytes, const rb_spmm_plan* plan, int* iters_out,
                                       double* relres_out, int* flag_out, void* stream) {
  if (n <= 0 || nsys <= 0 || k <= 0 || !vals64 || !B || !X) {
    rb_set_error("rb_krylov_solve_planned: bad arguments");
    return RB_ERR_ARG;
  }
  if ((int64_t)nsys * k > 512) {
    rb_set_error("rb_krylov_solve_planned: %lld columns exceed the 512-column limit of one launch",
                 (long long)nsys * k);
    return RB_ERR_SIZE;
  }
  if (plan && (plan->nblk <= 0 || plan->nrec <= 0)) {
    rb_set_error("rb_krylov_solve_planned: empty plan");
    return RB_ERR_ARG;
  }
  if (check_every <= 0) check_every = 16;
  if (maxiter <= 0) maxiter = 1000;
  Csr m{row_ptr, col_idx, diag_slot, n, nnz};
  hipStream_t s = (hipStream_t)stream;
  if (is_complex)
    return solve_impl<cplx<double>, cplx<float>>(m, nsys, k, vals64, vals32, B, X, precision, tol, maxiter, use_x0,
                                                 check_every, work, work_bytes, iters_out, relres_out, flag_out, s,
                                                 plan);
  return solve_impl<double, float>(m, nsys, k, vals64, vals32, B, X, precision, tol, maxiter, use_x0, check_every,
                                   work, work_bytes, iters_out, relres_out, flag_out, s, plan);
}

// ---------------------------------------------------------------------------
// rb_krylov_solve_noalloc: the same solve on a fixed per-device scratch array
// instead of a caller workspace, for callers that cannot allocate.  The
// array below is one copy per device, held from module load to process end.
// Host code never uses its name as an address: the device address comes from
// hipGetSymbolAddress, once per device with that device current, and only
// that pointer reaches kernels.  Calls that overlap on one device share the
// array and race; every buffer the solve reads is written by the solve first
// (fills, the initial residual, the state kernels), as with a fresh workspace.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[RB_NOALLOC_SCRATCH_BYTES];

#define RB_SCRATCH_MAXDEV 64
static std::mutex g_scratch_mu;
static void* g_scratch_dev[RB_SCRATCH_MAXDEV];

static int scratch_for_current_device(void** out) {
  int dev = 0;
  RB_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= RB_SCRATCH_MAXDEV) {
    rb_set_error("rb_krylov_solve_noalloc: device %d beyond the %d devices with scratch slots", dev,
                 RB_SCRATCH_MAXDEV);
    return RB_ERR_ARG;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      rb_set_error("rb_krylov_solve_noalloc: hipGetSymbolAddress(scratch) on device %d -> %s", dev,
                   hipGetErrorString(e));
      return RB_ERR_HIP;
    }
    g_scratch_dev[dev] = p;
  }
  *out = g_scratch_dev[dev];
  return RB_OK;
}

struct NoAllocScope {
  bool prev;
  NoAllocScope() : prev(t_noalloc) { t_noalloc = true; }
  ~NoAllocScope() { t_noalloc = prev; }
};

extern "C" int rb_krylov_solve_noalloc(int64_t n, int64_t nnz, const int32_t* row_ptr, const int32_t* col_idx,
                                       const int32_t* diag_slot, int nsys, int k, int is_complex,
                                       const void* vals64, const void* vals32, const void* B, void* X, int precision,
                                       double tol, int maxiter, int use_x0, int check_every, int* iters_out,
                                       double* relres_out, int* flag_out, void* stream) {
  if (n <= 0 || nsys <= 0 || k <= 0 || !vals64 || !B || !X) {
    rb_set_error("rb_krylov_solve_noalloc: bad arguments");
    return RB_ERR_ARG;
  }
  if ((int64_t)nsys * k > 512) {
    rb_set_error("rb_krylov_solve_noalloc: %lld columns exceed the 512-column limit of one launch",
                 (long long)nsys * k);
    return RB_ERR_SIZE;
  }
  size_t need = 0;
  int rc = rb_krylov_workspace(n, nsys, k, is_complex, precision, &need);
  if (rc) return rc;
  if (need > (size_t)RB_NOALLOC_SCRATCH_BYTES) {
    rb_set_error("rb_krylov_solve_noalloc: this solve needs %zu bytes of scratch, the fixed array holds %zu "
                 "(use rb_krylov_solve with a workspace)", need, (size_t)RB_NOALLOC_SCRATCH_BYTES);
    return RB_ERR_SIZE;
  }
  void* work = nullptr;
  rc = scratch_for_current_device(&work);
  if (rc) return rc;
  NoAllocScope guard;
  return rb_krylov_solve(n, nnz, row_ptr, col_idx, diag_slot, nsys, k, is_complex, vals64, vals32, B, X, precision,
                         tol, maxiter, use_x0, check_every, work, (size_t)RB_NOALLOC_SCRATCH_BYTES, iters_out,
                         relres_out, flag_out, stream);
}

// Y = A X for every system (dtype: 0 f64, 1 c128, 2 f32, 3 c64)
extern "C" int rb_spmm(int64_t n, int64_t nnz, const int32_t* row_ptr, const int32_t* col_idx, int nsys, int k,
                       int dtype, const void* vals, const void* X, void* Y, void* stream) {
  if (n <= 0 || nsys <= 0 || k <= 0) return RB_ERR_ARG;
  int C = nsys * k;
  if (C > 512) {
    rb_set_error("rb_spmm: too many columns");
    return RB_ERR_SIZE;
  }
  Csr m{row_ptr, col_idx, nullptr, n, nnz};
  hipStream_t s = (hipStream_t)stream;
  int64_t nb = grid_for(n, C);
#define L_SP(G_, N_, TT)                                                                              \
  hipLaunchKernelGGL((k_spmm<TT, G_, N_>), dim3(nb), dim3(KB), 0, s, m, (const TT*)vals, (const TT*)X, \
                     (TT*)Y, C, k)
  switch (dtype) {
    case 0:
      DISPATCH_GN(C, L_SP, double);
      break;
    case 1:
      DISPATCH_GN(C, L_SP, cplx<double>);
      break;
    case 2:
      DISPATCH_GN(C, L_SP, float);
      break;
    case 3:
      DISPATCH_GN(C, L_SP, cplx<float>);
      break;
    default:
      rb_set_error("rb_spmm: bad dtype %d", dtype);
      return RB_ERR_ARG;
  }
#undef L_SP
  RB_CHECK_LAUNCH();
  return RB_OK;
}
