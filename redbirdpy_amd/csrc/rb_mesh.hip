// Mesh-side integer kernels: incidence lists, CSR pattern, assembly slot map,
// point location, RHS columns.  All outputs are deterministic: where atomics
// are used to place entries, each list is sorted afterwards.
#include "rb_common.h"
#include "../../include/rbhip.h"
#include <stdarg.h>
#include <stdio.h>

static thread_local char g_err[1024] = "";

void rb_set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

extern "C" const char* rb_last_error(void) { return g_err; }
extern "C" int rb_version(void) { return RB_ABI_VERSION; }

extern "C" int rb_device_count(int* n) {
  if (!n) return RB_ERR_ARG;
  RB_HIP(hipGetDeviceCount(n));
  return RB_OK;
}

// ---------------------------------------------------------------------------
// fills and scan
// ---------------------------------------------------------------------------
__global__ void k_fill_i32(int32_t* p, int32_t v, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t st = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += st) p[i] = v;
}

__global__ void k_zero_u32(uint32_t* p, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t st = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += st) p[i] = 0u;
}

__global__ void k_zero_u8(uint8_t* p, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

int rb_fill_i32(int32_t* p, int32_t v, int64_t n, hipStream_t s) {
  if (n <= 0) return RB_OK;
  int64_t g = ceil_div(n, 256);
  if (g > 65536) g = 65536;
  hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)g), dim3(256), 0, s, p, v, n);
  RB_CHECK_LAUNCH();
  return RB_OK;
}

int rb_fill_bytes0(void* p, size_t bytes, hipStream_t s) {
  if (bytes == 0) return RB_OK;
  size_t words = bytes / 4;
  if (words) {
    int64_t g = ceil_div((int64_t)words, 256);
    if (g > 65536) g = 65536;
    hipLaunchKernelGGL(k_zero_u32, dim3((unsigned)g), dim3(256), 0, s, (uint32_t*)p, (int64_t)words);
    RB_CHECK_LAUNCH();
  }
  size_t tail = bytes - words * 4;
  if (tail) {
    hipLaunchKernelGGL(k_zero_u8, dim3(1), dim3(64), 0, s, (uint8_t*)p + words * 4, (int64_t)tail);
    RB_CHECK_LAUNCH();
  }
  return RB_OK;
}

extern "C" int rb_zero(void* p, size_t bytes, void* stream) {
  return rb_fill_bytes0(p, bytes, (hipStream_t)stream);
}

#define SCAN_BS 256
#define SCAN_PER 8
#define SCAN_TILE (SCAN_BS * SCAN_PER)

__global__ void k_scan_tile_sum(const int32_t* in, int64_t n, int32_t* bsum) {
  __shared__ int32_t red[SCAN_BS];
  int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
  int32_t s = 0;
  for (int k = 0; k < SCAN_PER; ++k) {
    int64_t i = base + (int64_t)k * SCAN_BS + threadIdx.x;
    if (i < n) s += in[i];
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = SCAN_BS / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) bsum[blockIdx.x] = red[0];
}

// exclusive scan in place of nb block sums with one block; also total
__global__ void k_scan_bsum(int32_t* bsum, int64_t nb, int32_t* total) {
  __shared__ int32_t buf[SCAN_BS];
  __shared__ int32_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int64_t b0 = 0; b0 < nb; b0 += SCAN_BS) {
    int64_t i = b0 + threadIdx.x;
    int32_t v = (i < nb) ? bsum[i] : 0;
    buf[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < SCAN_BS; o <<= 1) {
      int32_t t = (threadIdx.x >= (unsigned)o) ? buf[threadIdx.x - o] : 0;
      __syncthreads();
      buf[threadIdx.x] += t;
      __syncthreads();
    }
    int32_t incl = buf[threadIdx.x];
    if (i < nb) bsum[i] = carry + incl - v;
    __syncthreads();
    if (threadIdx.x == SCAN_BS - 1) carry += incl;
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = carry;
}

__global__ void k_scan_tile_apply(const int32_t* in, int64_t n, const int32_t* boff, int32_t* out) {
  __shared__ int32_t buf[SCAN_BS];
  __shared__ int32_t carry;
  int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
  if (threadIdx.x == 0) carry = boff[blockIdx.x];
  __syncthreads();
  for (int k = 0; k < SCAN_PER; ++k) {
    int64_t i = base + (int64_t)k * SCAN_BS + threadIdx.x;
    int32_t v = (i < n) ? in[i] : 0;
    buf[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < SCAN_BS; o <<= 1) {
      int32_t t = (threadIdx.x >= (unsigned)o) ? buf[threadIdx.x - o] : 0;
      __syncthreads();
      buf[threadIdx.x] += t;
      __syncthreads();
    }
    int32_t incl = buf[threadIdx.x];
    if (i < n) out[i] = carry + incl - v;
    __syncthreads();
    if (threadIdx.x == SCAN_BS - 1) carry += incl;
    __syncthreads();
  }
}

int rb_scan_i32(const int32_t* in, int32_t* out, int64_t n, hipStream_t s) {
  int64_t nb = ceil_div(n, SCAN_TILE);
  if (nb < 1) nb = 1;
  int32_t* bsum = nullptr;
  RB_HIP(hipMallocAsync((void**)&bsum, (size_t)nb * sizeof(int32_t), s));
  hipLaunchKernelGGL(k_scan_tile_sum, dim3((unsigned)nb), dim3(SCAN_BS), 0, s, in, n, bsum);
  RB_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_scan_bsum, dim3(1), dim3(SCAN_BS), 0, s, bsum, nb, out + n);
  RB_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_scan_tile_apply, dim3((unsigned)nb), dim3(SCAN_BS), 0, s, in, n, bsum, out);
  RB_CHECK_LAUNCH();
  RB_HIP(hipFreeAsync(bsum, s));
  return RB_OK;
}

// ---------------------------------------------------------------------------
// node -> element incidence (entries are e*nv + local), sorted per node
// ---------------------------------------------------------------------------
__global__ void k_inc_count(const int32_t* el, int64_t ne, int nv, int64_t nn, int32_t* cnt, int32_t* bad) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ne * nv) return;
  int32_t v = el[t];
  if (v < 0 || v >= nn) {
    atomicOr(bad, 1);
    return;
  }
  atomicAdd(&cnt[v], 1);
}

__global__ void k_inc_fill(const int32_t* el, int64_t ne, int nv, int64_t nn, int32_t* cursor, int32_t* list) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ne * nv) return;
  int32_t v = el[t];
  if (v < 0 || v >= nn) return;
  int32_t pos = atomicAdd(&cursor[v], 1);
  list[pos] = (int32_t)t;
}

__global__ void k_sort_segments(const int32_t* ptr, int32_t* list, int64_t nseg) {
  int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nseg) return;
  int32_t a = ptr[n], b = ptr[n + 1];
  for (int32_t i = a + 1; i < b; ++i) {
    int32_t v = list[i];
    int32_t j = i - 1;
    while (j >= a && list[j] > v) {
      list[j + 1] = list[j];
      --j;
    }
    list[j + 1] = v;
  }
}

extern "C" int rb_incidence(const int32_t* el, int64_t ne, int nv, int64_t nn, int32_t* inc_ptr,
                            int32_t* inc_list, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  if (ne < 0 || nn <= 0 || nv <= 0 || (ne * nv) >= INT32_MAX) {
    rb_set_error("rb_incidence: bad sizes ne=%lld nv=%d nn=%lld", (long long)ne, nv, (long long)nn);
    return RB_ERR_ARG;
  }
  int32_t* cnt = nullptr;
  RB_HIP(hipMallocAsync((void**)&cnt, (size_t)(nn + 1) * sizeof(int32_t), s));
  int rc = rb_fill_i32(cnt, 0, nn + 1, s);
  if (rc) return rc;
  int64_t tot = ne * nv;
  if (tot > 0) {
    hipLaunchKernelGGL(k_inc_count, dim3((unsigned)ceil_div(tot, 256)), dim3(256), 0, s, el, ne, nv, nn, cnt,
                       cnt + nn);
    RB_CHECK_LAUNCH();
  }
  int32_t hbad = 0;
  RB_HIP(hipMemcpyAsync(&hbad, cnt + nn, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  RB_HIP(hipStreamSynchronize(s));
  if (hbad) {
    (void)hipFreeAsync(cnt, s);
    rb_set_error("rb_incidence: element index out of range [0, %lld)", (long long)nn);
    return RB_ERR_ARG;
  }
  rc = rb_scan_i32(cnt, inc_ptr, nn, s);
  if (rc) return rc;
  RB_HIP(hipMemcpyAsync(cnt, inc_ptr, (size_t)nn * sizeof(int32_t), hipMemcpyDeviceToDevice, s));
  if (tot > 0) {
    hipLaunchKernelGGL(k_inc_fill, dim3((unsigned)ceil_div(tot, 256)), dim3(256), 0, s, el, ne, nv, nn, cnt,
                       inc_list);
    RB_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_sort_segments, dim3((unsigned)ceil_div(nn, 256)), dim3(256), 0, s, inc_ptr, inc_list, nn);
    RB_CHECK_LAUNCH();
  }
  RB_HIP(hipFreeAsync(cnt, s));
  return RB_OK;
}

// ---------------------------------------------------------------------------
// CSR pattern: row n holds every node sharing a tet with n (n included),
// columns ascending.  Found by repeated "smallest candidate above the last
// one" over the incident tets, so no per-row scratch is needed.
// ---------------------------------------------------------------------------
template <bool FILL>
__global__ void k_row_pattern(const int32_t* el, const int32_t* inc_ptr, const int32_t* inc_list, int64_t nn,
                              int32_t* row_cnt, const int32_t* row_ptr, int32_t* col_idx) {
  int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nn) return;
  int32_t a = inc_ptr[n], b = inc_ptr[n + 1];
  int32_t last = -1;
  int32_t cnt = 0;
  int32_t out = FILL ? row_ptr[n] : 0;
  while (true) {
    int32_t best = INT32_MAX;
    for (int32_t t = a; t < b; ++t) {
      int32_t e = inc_list[t] >> 2;
      const int32_t* q = el + 4 * (int64_t)e;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        int32_t c = q[j];
        if (c > last && c < best) best = c;
      }
    }
    if (best == INT32_MAX) break;
    if (FILL) col_idx[out + cnt] = best;
    ++cnt;
    last = best;
  }
  if (!FILL) row_cnt[n] = cnt;
}

__device__ inline int32_t find_slot(const int32_t* row_ptr, const int32_t* col_idx, int32_t r, int32_t c) {
  int32_t lo = row_ptr[r], hi = row_ptr[r + 1] - 1;
  while (lo <= hi) {
    int32_t mid = (lo + hi) >> 1;
    int32_t v = col_idx[mid];
    if (v == c) return mid;
    if (v < c)
      lo = mid + 1;
    else
      hi = mid - 1;
  }
  return -1;
}

// diagonal slot of each row
__global__ void k_diag_slot(const int32_t* row_ptr, const int32_t* col_idx, int64_t nn, int32_t* diag) {
  int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nn) return;
  diag[n] = find_slot(row_ptr, col_idx, (int32_t)n, (int32_t)n);
}

extern "C" int rb_csr_pattern(const int32_t* el, int64_t ne, int64_t nn, const int32_t* inc_ptr,
                              const int32_t* inc_list, int32_t* row_ptr, int32_t* col_idx, int64_t col_cap,
                              int64_t* nnz_out, int32_t* diag_slot, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  if (!nnz_out) return RB_ERR_ARG;
  int32_t* cnt = nullptr;
  RB_HIP(hipMallocAsync((void**)&cnt, (size_t)nn * sizeof(int32_t), s));
  unsigned g = (unsigned)ceil_div(nn, 256);
  hipLaunchKernelGGL((k_row_pattern<false>), dim3(g), dim3(256), 0, s, el, inc_ptr, inc_list, nn, cnt,
                     (const int32_t*)nullptr, (int32_t*)nullptr);
  RB_CHECK_LAUNCH();
  int rc = rb_scan_i32(cnt, row_ptr, nn, s);
  if (rc) return rc;
  RB_HIP(hipFreeAsync(cnt, s));
  int32_t hn = 0;
  RB_HIP(hipMemcpyAsync(&hn, row_ptr + nn, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  RB_HIP(hipStreamSynchronize(s));
  *nnz_out = hn;
  if (col_idx == nullptr) return RB_OK;  // size query
  if (hn > col_cap) {
    rb_set_error("rb_csr_pattern: nnz %d exceeds capacity %lld", hn, (long long)col_cap);
    return RB_ERR_SIZE;
  }
  hipLaunchKernelGGL((k_row_pattern<true>), dim3(g), dim3(256), 0, s, el, inc_ptr, inc_list, nn, (int32_t*)nullptr,
                     row_ptr, col_idx);
  RB_CHECK_LAUNCH();
  if (diag_slot) {
    hipLaunchKernelGGL(k_diag_slot, dim3(g), dim3(256), 0, s, row_ptr, col_idx, nn, diag_slot);
    RB_CHECK_LAUNCH();
  }
  return RB_OK;
}

// ---------------------------------------------------------------------------
// Assembly slot map.  Contribution codes: tet (e,i,j) -> e*16 + i*4 + j,
// face (f,i,j) -> 16*ne + f*9 + i*3 + j, where i is the local index of the
// row node.  Each slot lists its codes in ascending order (tets by element
// index, then faces), which fixes the summation order of the assembly.
// ---------------------------------------------------------------------------
template <bool FILL>
__global__ void k_slot_map(const int32_t* el, int64_t ne, const int32_t* fc, const int32_t* tinc_ptr,
                           const int32_t* tinc_list, const int32_t* finc_ptr, const int32_t* finc_list,
                           const int32_t* row_ptr, const int32_t* col_idx, int64_t nn, int32_t* slot_cnt,
                           const int32_t* slot_ptr, int32_t* codes) {
  int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nn) return;
  int32_t r0 = row_ptr[n], r1 = row_ptr[n + 1];
  for (int32_t sl = r0; sl < r1; ++sl) {
    int32_t m = col_idx[sl];
    int32_t c = 0;
    int32_t o = FILL ? slot_ptr[sl] : 0;
    for (int32_t t = tinc_ptr[n]; t < tinc_ptr[n + 1]; ++t) {
      int32_t code = tinc_list[t];
      int32_t e = code >> 2, i = code & 3;
      const int32_t* q = el + 4 * (int64_t)e;
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (q[j] == m) {
          if (FILL) codes[o + c] = e * 16 + i * 4 + j;
          ++c;
        }
    }
    if (fc) {
      for (int32_t t = finc_ptr[n]; t < finc_ptr[n + 1]; ++t) {
        int32_t code = finc_list[t];
        int32_t f = code / 3, i = code - 3 * (code / 3);
        const int32_t* q = fc + 3 * (int64_t)f;
#pragma unroll
        for (int j = 0; j < 3; ++j)
          if (q[j] == m) {
            if (FILL) codes[o + c] = (int32_t)(16 * ne) + f * 9 + i * 3 + j;
            ++c;
          }
      }
    }
    if (!FILL) slot_cnt[sl] = c;
  }
}

// every face pair must exist in the tet pattern
__global__ void k_check_faces(const int32_t* fc, int64_t nf, const int32_t* row_ptr, const int32_t* col_idx,
                              int64_t nn, int32_t* bad) {
  int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  const int32_t* q = fc + 3 * f;
  for (int i = 0; i < 3; ++i) {
    if (q[i] < 0 || q[i] >= nn) {
      atomicOr(bad, 1);
      return;
    }
  }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      if (find_slot(row_ptr, col_idx, q[i], q[j]) < 0) atomicOr(bad, 2);
}

extern "C" int rb_slot_map(const int32_t* el, int64_t ne, const int32_t* fc, int64_t nf, int64_t nn,
                           const int32_t* tinc_ptr, const int32_t* tinc_list, const int32_t* finc_ptr,
                           const int32_t* finc_list, const int32_t* row_ptr, const int32_t* col_idx, int64_t nnz,
                           int32_t* slot_ptr, int32_t* codes, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  if (16 * ne + 9 * nf >= (int64_t)INT32_MAX) {
    rb_set_error("rb_slot_map: mesh too large for int32 contribution codes");
    return RB_ERR_SIZE;
  }
  if (nf > 0 && fc) {
    int32_t* bad = nullptr;
    RB_HIP(hipMallocAsync((void**)&bad, sizeof(int32_t), s));
    int rc = rb_fill_i32(bad, 0, 1, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_check_faces, dim3((unsigned)ceil_div(nf, 256)), dim3(256), 0, s, fc, nf, row_ptr, col_idx,
                       nn, bad);
    RB_CHECK_LAUNCH();
    int32_t hb = 0;
    RB_HIP(hipMemcpyAsync(&hb, bad, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    RB_HIP(hipStreamSynchronize(s));
    RB_HIP(hipFreeAsync(bad, s));
    if (hb) {
      rb_set_error("rb_slot_map: boundary face with nodes outside the tet pattern (code %d)", hb);
      return RB_ERR_PATTERN;
    }
  }
  int32_t* cnt = nullptr;
  RB_HIP(hipMallocAsync((void**)&cnt, (size_t)(nnz > 0 ? nnz : 1) * sizeof(int32_t), s));
  unsigned g = (unsigned)ceil_div(nn, 128);
  const int32_t* fcp = (nf > 0) ? fc : nullptr;
  hipLaunchKernelGGL((k_slot_map<false>), dim3(g), dim3(128), 0, s, el, ne, fcp, tinc_ptr, tinc_list, finc_ptr,
                     finc_list, row_ptr, col_idx, nn, cnt, (const int32_t*)nullptr, (int32_t*)nullptr);
  RB_CHECK_LAUNCH();
  int rc = rb_scan_i32(cnt, slot_ptr, nnz, s);
  if (rc) return rc;
  hipLaunchKernelGGL((k_slot_map<true>), dim3(g), dim3(128), 0, s, el, ne, fcp, tinc_ptr, tinc_list, finc_ptr,
                     finc_list, row_ptr, col_idx, nn, (int32_t*)nullptr, slot_ptr, codes);
  RB_CHECK_LAUNCH();
  RB_HIP(hipFreeAsync(cnt, s));
  return RB_OK;
}

// ---------------------------------------------------------------------------
// Point location: lowest-index tet whose barycentric coordinates are all
// >= -tol.  One thread per tet tests every point; atomicMin picks the winner.
// ---------------------------------------------------------------------------
__device__ inline void tet_bary(const double* node, const int32_t* q, const double* x, double* lam) {
  double p0[3], a[3], b[3], c[3], d[3];
  for (int k = 0; k < 3; ++k) {
    p0[k] = node[3 * (int64_t)q[0] + k];
    a[k] = node[3 * (int64_t)q[1] + k] - p0[k];
    b[k] = node[3 * (int64_t)q[2] + k] - p0[k];
    c[k] = node[3 * (int64_t)q[3] + k] - p0[k];
    d[k] = x[k] - p0[k];
  }
  // Cramer's rule on [a b c] * (l1,l2,l3) = d
  double bc0 = b[1] * c[2] - b[2] * c[1], bc1 = b[2] * c[0] - b[0] * c[2], bc2 = b[0] * c[1] - b[1] * c[0];
  double det = a[0] * bc0 + a[1] * bc1 + a[2] * bc2;
  double l1 = (d[0] * bc0 + d[1] * bc1 + d[2] * bc2) / det;
  double ca0 = c[1] * a[2] - c[2] * a[1], ca1 = c[2] * a[0] - c[0] * a[2], ca2 = c[0] * a[1] - c[1] * a[0];
  double l2 = (d[0] * ca0 + d[1] * ca1 + d[2] * ca2) / det;
  double ab0 = a[1] * b[2] - a[2] * b[1], ab1 = a[2] * b[0] - a[0] * b[2], ab2 = a[0] * b[1] - a[1] * b[0];
  double l3 = (d[0] * ab0 + d[1] * ab1 + d[2] * ab2) / det;
  lam[0] = 1.0 - l1 - l2 - l3;
  lam[1] = l1;
  lam[2] = l2;
  lam[3] = l3;
}

__global__ void k_tsearch(const double* node, const int32_t* el, int64_t ne, const double* pts, int npts, double tol,
                          int32_t* loc) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  const int32_t* q = el + 4 * e;
  double lo[3], hi[3];
  for (int k = 0; k < 3; ++k) {
    lo[k] = hi[k] = node[3 * (int64_t)q[0] + k];
    for (int j = 1; j < 4; ++j) {
      double v = node[3 * (int64_t)q[j] + k];
      lo[k] = v < lo[k] ? v : lo[k];
      hi[k] = v > hi[k] ? v : hi[k];
    }
  }
  for (int p = 0; p < npts; ++p) {
    const double* x = pts + 3 * p;
    double pad = 1e-9 * (1.0 + fabs(x[0]) + fabs(x[1]) + fabs(x[2]));
    if (x[0] < lo[0] - pad || x[0] > hi[0] + pad || x[1] < lo[1] - pad || x[1] > hi[1] + pad || x[2] < lo[2] - pad ||
        x[2] > hi[2] + pad)
      continue;
    double lam[4];
    tet_bary(node, q, x, lam);
    if (lam[0] >= -tol && lam[1] >= -tol && lam[2] >= -tol && lam[3] >= -tol) atomicMin(&loc[p], (int32_t)e);
  }
}

__global__ void k_tsearch_bary(const double* node, const int32_t* el, const double* pts, int npts,
                               const int32_t* loc, double* bary) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npts) return;
  int32_t e = loc[p];
  if (e == INT32_MAX) {
    for (int k = 0; k < 4; ++k) bary[4 * p + k] = __builtin_nan("");
    return;
  }
  double lam[4];
  tet_bary(node, el + 4 * (int64_t)e, pts + 3 * p, lam);
  for (int k = 0; k < 4; ++k) bary[4 * p + k] = lam[k];
}

extern "C" int rb_tsearch(const double* node, const int32_t* el, int64_t ne, const double* pts, int npts, double tol,
                          int32_t* loc, double* bary, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  if (npts <= 0) return RB_OK;
  int rc = rb_fill_i32(loc, INT32_MAX, npts, s);
  if (rc) return rc;
  if (ne > 0) {
    hipLaunchKernelGGL(k_tsearch, dim3((unsigned)ceil_div(ne, 256)), dim3(256), 0, s, node, el, ne, pts, npts, tol,
                       loc);
    RB_CHECK_LAUNCH();
  }
  hipLaunchKernelGGL(k_tsearch_bary, dim3((unsigned)ceil_div(npts, 64)), dim3(64), 0, s, node, el, pts, npts, loc,
                     bary);
  RB_CHECK_LAUNCH();
  return RB_OK;
}

// ---------------------------------------------------------------------------
// RHS block: dense n x ld (row-major), point optode column c gets the four
// barycentric weights of its tet; a column whose optode is outside stays 0.
// ---------------------------------------------------------------------------
__global__ void k_rhs_points(const int32_t* el, const int32_t* loc, const double* bary, int npts, const int32_t* cols,
                             int64_t ld, int cplx, double* B) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npts) return;
  int32_t e = loc[p];
  if (e == INT32_MAX || e < 0) return;
  int64_t c = cols[p];
  for (int k = 0; k < 4; ++k) {
    int64_t r = el[4 * (int64_t)e + k];
    if (cplx) {
      B[2 * (r * ld + c)] = bary[4 * p + k];
      B[2 * (r * ld + c) + 1] = 0.0;
    } else {
      B[r * ld + c] = bary[4 * p + k];
    }
  }
}

extern "C" int rb_rhs_points(const int32_t* el, const int32_t* loc, const double* bary, int npts, const int32_t* cols,
                             int64_t ld, int is_complex, double* B, void* stream) {
  if (npts <= 0) return RB_OK;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_rhs_points, dim3((unsigned)ceil_div(npts, 64)), dim3(64), 0, s, el, loc, bary, npts, cols, ld,
                     is_complex, B);
  RB_CHECK_LAUNCH();
  return RB_OK;
}
