// K5-K7: detector readout and the adjoint mua Jacobian.
//
// redbirdpy forward.py:488-504/:562-567 define, for pair p=(s,d) and tet e,
//   Jelem[p,e] = -0.05 V_e (sum_i ps_i pd_i + S_s S_d),  S = sum over e's nodes,
//   Jnode[p,n] = 0.25 sum_{e ni n} Jelem[p,e].
// Summing per node instead of per element gives the bilinear form
//   Jnode[.,n] = -0.0125 Phi_S^T Q_n Phi_D,   Q_n = sum_{e ni n} V_e (I + 1 1^T)_e
// over the node's patch (its CSR row).  One wavefront per node builds Q_n in
// LDS (one lane per patch row, tets in ascending order: deterministic),
// stages the patch rows of phi for 32 sources x 32 detectors at a time,
// forms Y = Q_n Phi_D and writes the 32x32 tile of J from 4x4 register
// tiles.  J is written node-major (JT[n][p], pairs contiguous), the layout
// the Gram kernel streams.
#include "rb_common.h"
#include "../../include/rbhip.h"

#define JW 64   // tets staged at a time when building a node's patch matrix
#define JB 256  // threads per node workgroup
#define JC 32   // sources / detectors per staged chunk

template <typename TO, typename T>
__device__ inline TO cast_out(T v);
template <>
__device__ inline double cast_out<double, double>(double v) { return v; }
template <>
__device__ inline float cast_out<float, double>(double v) { return (float)v; }
template <>
__device__ inline cplx<double> cast_out<cplx<double>, cplx<double>>(cplx<double> v) { return v; }
template <>
__device__ inline cplx<float> cast_out<cplx<float>, cplx<double>>(cplx<double> v) {
  return mk<float>((float)v.x, (float)v.y);
}

__device__ inline double jmul(double a, double b) { return a * b; }
__device__ inline cplx<double> jmul(cplx<double> a, cplx<double> b) { return a * b; }
__device__ inline double jscale(double s, double v) { return s * v; }
__device__ inline cplx<double> jscale(double s, cplx<double> v) { return mk<double>(s * v.x, s * v.y); }

// Q_n (np x np, row-major in LDS) from the node's incident tets, JW tets at
// a time: a thread per staged tet loads it (vertices, volume) and finds its
// vertices' positions in the sorted patch (tl, tv in LDS); then a thread per
// entry (k, l) of Q adds V (2V on the diagonal) of every staged tet holding
// both patch nodes, tets in ascending order (so Q is bit-reproducible).
__device__ inline void build_patch_q(int64_t n, const int32_t* el, const double* evol, const int32_t* inc_ptr,
                                     const int32_t* inc_list, const int32_t* cols, int np, double* Q,
                                     int32_t* tl /* JW x 4 */, double* tv /* JW */) {
  const int t = threadIdx.x;
  for (int i = t; i < np * np; i += JB) Q[i] = 0.0;
  const int32_t t0 = inc_ptr[n], t1 = inc_ptr[n + 1];
  for (int32_t c0 = t0; c0 < t1; c0 += JW) {
    const int nt = min(JW, t1 - c0);
    if (t < nt) {
      const int32_t e = inc_list[c0 + t] >> 2;
      const int32_t* q = el + 4 * (int64_t)e;
      int32_t v[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) v[j] = q[j];
      tv[t] = evol[e];
#pragma unroll
      for (int j = 0; j < 4; ++j) {  // position of v[j] in the sorted patch (every tet vertex is a neighbour)
        int lo = 0, hi = np - 1, pos = -1;
        while (lo <= hi) {
          const int mid = (lo + hi) >> 1;
          const int32_t cv = cols[mid];
          if (cv == v[j]) {
            pos = mid;
            break;
          }
          if (cv < v[j])
            lo = mid + 1;
          else
            hi = mid - 1;
        }
        tl[t * 4 + j] = pos;
      }
    }
    __syncthreads();
    for (int i = t; i < np * np; i += JB) {
      const int k = i / np, l = i - (i / np) * np;
      double acc = Q[i];
      for (int u = 0; u < nt; ++u) {
        const int32_t l0 = tl[u * 4], l1 = tl[u * 4 + 1], l2 = tl[u * 4 + 2], l3 = tl[u * 4 + 3];
        const bool hk = l0 == k || l1 == k || l2 == k || l3 == k;
        const bool hl = l0 == l || l1 == l || l2 == l || l3 == l;
        if (hk && hl) acc += (k == l) ? 2.0 * tv[u] : tv[u];
      }
      Q[i] = acc;
    }
    __syncthreads();
  }
}

// J_n = -0.0125 PS^T Q_n PD for every (s, d) chunk of JC x JC pairs, one
// node per JB-thread workgroup: the node's patch rows of phi are staged in
// LDS (PS, PD), Y = Q PD, then each thread forms JC*JC/JB consecutive
// sources of one detector, so a wavefront's stores cover contiguous runs of
// the node's JT row.
#define JPT (JC * JC / JB)  // pairs per thread
static_assert(JC % JPT == 0, "a thread's pairs must lie in one detector row");
template <typename T, typename TO>
__global__ void __launch_bounds__(JB) k_jac_node(int64_t nn, const int32_t* __restrict__ el,
                                                 const double* __restrict__ evol, const int32_t* __restrict__ inc_ptr,
                                                 const int32_t* __restrict__ inc_list,
                                                 const int32_t* __restrict__ row_ptr,
                                                 const int32_t* __restrict__ col_idx, const T* __restrict__ phi,
                                                 int64_t ldphi, int Ns, int Nd, const int32_t* __restrict__ scol,
                                                 const int32_t* __restrict__ dcol, const int32_t* __restrict__ pmap,
                                                 TO* __restrict__ JT, int64_t ldj, int maxnp,
                                                 const int64_t* __restrict__ qptr, const double* __restrict__ Qc) {
  extern __shared__ __align__(16) unsigned char smem[];
  T* PS = (T*)smem;                        // np x JC
  T* PD = PS + maxnp * JC;                 // np x JC
  T* Y = PD + maxnp * JC;                  // np x JC
  double* Q = (double*)(Y + maxnp * JC);   // np x np
  double* tv = Q + maxnp * maxnp;          // JW staged tet volumes
  int32_t* tl = (int32_t*)(tv + JW);       // their JW x 4 patch positions
  int32_t* cols = tl + 4 * JW;
  const int t = threadIdx.x;
  const int jd = (t * JPT) / JC;           // detector of this thread's pairs within the chunk
  const int js = (t * JPT) - jd * JC;      // its first source
  for (int64_t n = blockIdx.x; n < nn; n += gridDim.x) {
    const int32_t r0 = row_ptr[n];
    const int np = row_ptr[n + 1] - r0;
    for (int k = t; k < np; k += JB) cols[k] = col_idx[r0 + k];
    if (Qc) {  // the mesh's patch matrices, built once (k_patch_q)
      const double* q = Qc + qptr[n];
      for (int i = t; i < np * np; i += JB) Q[i] = q[i];
    }
    __syncthreads();
    if (!Qc) build_patch_q(n, el, evol, inc_ptr, inc_list, cols, np, Q, tl, tv);
    for (int s0 = 0; s0 < Ns; s0 += JC) {
      const int ns = min(JC, Ns - s0);
      for (int i = t; i < np * JC; i += JB) {
        const int k = i / JC, j = i - k * JC;
        PS[i] = (j < ns) ? phi[(int64_t)cols[k] * ldphi + scol[s0 + j]] : sc<T>::zero();
      }
      for (int d0 = 0; d0 < Nd; d0 += JC) {
        const int nd = min(JC, Nd - d0);
        for (int i = t; i < np * JC; i += JB) {
          const int k = i / JC, j = i - k * JC;
          PD[i] = (j < nd) ? phi[(int64_t)cols[k] * ldphi + dcol[d0 + j]] : sc<T>::zero();
        }
        __syncthreads();
        // Y = Q PD
        for (int i = t; i < np * JC; i += JB) {
          const int k = i / JC, j = i - k * JC;
          T acc = sc<T>::zero();
          const double* qk = Q + k * np;
          for (int l = 0; l < np; ++l) acc = acc + jscale(qk[l], PD[l * JC + j]);
          Y[i] = acc;
        }
        __syncthreads();
        // pairs (s0 + js + a, d0 + jd): PS^T Y
        T acc[JPT];
#pragma unroll
        for (int a = 0; a < JPT; ++a) acc[a] = sc<T>::zero();
        for (int k = 0; k < np; ++k) {
          const T y = Y[k * JC + jd];
#pragma unroll
          for (int a = 0; a < JPT; ++a) acc[a] = acc[a] + jmul(PS[k * JC + js + a], y);
        }
        if (jd < nd) {
          const int64_t base = (int64_t)(d0 + jd) * Ns + s0 + js;
#pragma unroll
          for (int a = 0; a < JPT; ++a) {
            if (js + a >= ns) continue;
            const int64_t p = pmap ? pmap[base + a] : base + a;
            if (p >= 0) JT[n * ldj + p] = cast_out<TO, T>(jscale(-0.0125, acc[a]));
          }
        }
        __syncthreads();
      }
    }
  }
}

// every node's patch matrix Q_n into Qc[qptr[n] ...] (row-major np x np)
__global__ void __launch_bounds__(JB) k_patch_q(int64_t nn, const int32_t* __restrict__ el,
                                                const double* __restrict__ evol, const int32_t* __restrict__ inc_ptr,
                                                const int32_t* __restrict__ inc_list,
                                                const int32_t* __restrict__ row_ptr,
                                                const int32_t* __restrict__ col_idx, int maxnp,
                                                const int64_t* __restrict__ qptr, double* __restrict__ Qc) {
  extern __shared__ __align__(16) unsigned char smem[];
  double* Q = (double*)smem;  // np x np
  double* tv = Q + maxnp * maxnp;
  int32_t* tl = (int32_t*)(tv + JW);
  int32_t* cols = tl + 4 * JW;
  const int t = threadIdx.x;
  for (int64_t n = blockIdx.x; n < nn; n += gridDim.x) {
    const int32_t r0 = row_ptr[n];
    const int np = row_ptr[n + 1] - r0;
    for (int k = t; k < np; k += JB) cols[k] = col_idx[r0 + k];
    __syncthreads();
    build_patch_q(n, el, evol, inc_ptr, inc_list, cols, np, Q, tl, tv);
    double* q = Qc + qptr[n];
    for (int i = t; i < np * np; i += JB) q[i] = Q[i];
    __syncthreads();
  }
}

// Jelem[e][p] = -0.05 V_e (sum_i ps_i pd_i + S_s S_d), element-major
template <typename T, typename TO>
__global__ void k_jac_elem(int64_t ne, const int32_t* __restrict__ el, const double* __restrict__ evol,
                           const T* __restrict__ phi, int64_t ldphi, int Ns, int Nd, const int32_t* __restrict__ scol,
                           const int32_t* __restrict__ dcol, const int32_t* __restrict__ pmap, TO* __restrict__ JE,
                           int64_t ldj) {
  for (int64_t e = blockIdx.x; e < ne; e += gridDim.x) {
    const int32_t* q = el + 4 * e;
    double V = evol[e];
    int64_t cells = (int64_t)Ns * Nd;
    for (int64_t cell = threadIdx.x; cell < cells; cell += blockDim.x) {
      int d = (int)(cell / Ns), s = (int)(cell - (int64_t)d * Ns);
      int64_t p = pmap ? pmap[cell] : cell;
      if (p < 0) continue;
      T dg = sc<T>::zero(), ss = sc<T>::zero(), sd = sc<T>::zero();
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        T a = phi[(int64_t)q[i] * ldphi + scol[s]];
        T b = phi[(int64_t)q[i] * ldphi + dcol[d]];
        dg = dg + jmul(a, b);
        ss = ss + a;
        sd = sd + b;
      }
      JE[e * ldj + p] = cast_out<TO, T>(jscale(-0.05 * V, dg + jmul(ss, sd)));
    }
  }
}

// dynamic LDS beyond 64 KiB (a node patch of many neighbours) up to the
// device's per-workgroup limit (160 KiB on gfx950), opted in per kernel
static size_t lds_limit() {
  int dev = 0, v = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0)
    v = 64 * 1024;
  return (size_t)v;
}

template <typename K>
static bool allow_lds(K kernel, size_t lds) {
  if (lds <= 64 * 1024) return true;
  return hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)lds) == hipSuccess;
}

static size_t jac_lds(int maxnp, int is_complex) {
  size_t sv = is_complex ? 16 : 8;
  return (size_t)3 * maxnp * JC * sv + (size_t)maxnp * maxnp * 8 + (size_t)JW * (8 + 16) + (size_t)maxnp * 4 + 16;
}

extern "C" int rb_jac_patch_q(int64_t nn, const int32_t* el, const double* evol, const int32_t* inc_ptr,
                              const int32_t* inc_list, const int32_t* row_ptr, const int32_t* col_idx, int maxnp,
                              const int64_t* qptr, double* Qc, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  if (nn <= 0 || maxnp <= 0 || !qptr || !Qc) {
    rb_set_error("rb_jac_patch_q: bad sizes");
    return RB_ERR_ARG;
  }
  size_t lds = (size_t)maxnp * maxnp * 8 + (size_t)JW * (8 + 16) + (size_t)maxnp * 4 + 16;
  if (lds > lds_limit() || !allow_lds(k_patch_q, lds)) {
    rb_set_error("rb_jac_patch_q: node patch of %d nodes needs %zu B of LDS (limit %zu)", maxnp, lds, lds_limit());
    return RB_ERR_SIZE;
  }
  unsigned g = (unsigned)(nn < 262144 ? nn : 262144);
  hipLaunchKernelGGL(k_patch_q, dim3(g), dim3(JB), lds, s, nn, el, evol, inc_ptr, inc_list, row_ptr, col_idx, maxnp,
                     qptr, Qc);
  RB_CHECK_LAUNCH();
  return RB_OK;
}

extern "C" int rb_jac_grid(int64_t nn, int64_t ne, const int32_t* el, const double* evol, const int32_t* inc_ptr,
                           const int32_t* inc_list, const int32_t* row_ptr, const int32_t* col_idx, int maxnp,
                           const void* phi, int64_t ldphi, int is_complex, int Ns, int Nd, const int32_t* scol,
                           const int32_t* dcol, const int32_t* pmap, void* JT, int64_t ldj, int out_dtype,
                           void* Jelem, const int64_t* qptr, const double* Qc, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  if (nn <= 0 || Ns <= 0 || Nd <= 0 || maxnp <= 0 || (!qptr) != (!Qc)) {
    rb_set_error("rb_jac_grid: bad sizes");
    return RB_ERR_ARG;
  }
  size_t lds = jac_lds(maxnp, is_complex);
  if (lds > lds_limit()) {
    rb_set_error("rb_jac_grid: node patch of %d nodes needs %zu B of LDS (limit %zu)", maxnp, lds, lds_limit());
    return RB_ERR_SIZE;
  }
  bool ok = true;
  if (!is_complex)
    ok = out_dtype == RB_F64 ? allow_lds(k_jac_node<double, double>, lds) : allow_lds(k_jac_node<double, float>, lds);
  else
    ok = out_dtype == RB_C128 ? allow_lds(k_jac_node<cplx<double>, cplx<double>>, lds)
                              : allow_lds(k_jac_node<cplx<double>, cplx<float>>, lds);
  if (!ok) {
    rb_set_error("rb_jac_grid: %zu B of dynamic LDS refused", lds);
    return RB_ERR_SIZE;
  }
  unsigned g = (unsigned)(nn < 262144 ? nn : 262144);
  unsigned ge = (unsigned)(ne < 262144 ? ne : 262144);
  if (!is_complex) {
    if (out_dtype == RB_F64) {
      if (JT)
        hipLaunchKernelGGL((k_jac_node<double, double>), dim3(g), dim3(JB), lds, s, nn, el, evol, inc_ptr, inc_list,
                           row_ptr, col_idx, (const double*)phi, ldphi, Ns, Nd, scol, dcol, pmap, (double*)JT, ldj,
                           maxnp, qptr, Qc);
      if (Jelem && ne > 0)
        hipLaunchKernelGGL((k_jac_elem<double, double>), dim3(ge), dim3(256), 0, s, ne, el, evol, (const double*)phi,
                           ldphi, Ns, Nd, scol, dcol, pmap, (double*)Jelem, ldj);
    } else if (out_dtype == RB_F32) {
      if (JT)
        hipLaunchKernelGGL((k_jac_node<double, float>), dim3(g), dim3(JB), lds, s, nn, el, evol, inc_ptr, inc_list,
                           row_ptr, col_idx, (const double*)phi, ldphi, Ns, Nd, scol, dcol, pmap, (float*)JT, ldj,
                           maxnp, qptr, Qc);
      if (Jelem && ne > 0)
        hipLaunchKernelGGL((k_jac_elem<double, float>), dim3(ge), dim3(256), 0, s, ne, el, evol, (const double*)phi,
                           ldphi, Ns, Nd, scol, dcol, pmap, (float*)Jelem, ldj);
    } else {
      rb_set_error("rb_jac_grid: real phi needs RB_F64 or RB_F32 output");
      return RB_ERR_ARG;
    }
  } else {
    typedef cplx<double> C2;
    if (out_dtype == RB_C128) {
      if (JT)
        hipLaunchKernelGGL((k_jac_node<C2, C2>), dim3(g), dim3(JB), lds, s, nn, el, evol, inc_ptr, inc_list, row_ptr,
                           col_idx, (const C2*)phi, ldphi, Ns, Nd, scol, dcol, pmap, (C2*)JT, ldj, maxnp, qptr, Qc);
      if (Jelem && ne > 0)
        hipLaunchKernelGGL((k_jac_elem<C2, C2>), dim3(ge), dim3(256), 0, s, ne, el, evol, (const C2*)phi, ldphi, Ns,
                           Nd, scol, dcol, pmap, (C2*)Jelem, ldj);
    } else if (out_dtype == RB_C64) {
      if (JT)
        hipLaunchKernelGGL((k_jac_node<C2, cplx<float>>), dim3(g), dim3(JB), lds, s, nn, el, evol, inc_ptr, inc_list,
                           row_ptr, col_idx, (const C2*)phi, ldphi, Ns, Nd, scol, dcol, pmap, (cplx<float>*)JT, ldj,
                           maxnp, qptr, Qc);
      if (Jelem && ne > 0)
        hipLaunchKernelGGL((k_jac_elem<C2, cplx<float>>), dim3(ge), dim3(256), 0, s, ne, el, evol, (const C2*)phi,
                           ldphi, Ns, Nd, scol, dcol, pmap, (cplx<float>*)Jelem, ldj);
    } else {
      rb_set_error("rb_jac_grid: complex phi needs RB_C128 or RB_C64 output");
      return RB_ERR_ARG;
    }
  }
  RB_CHECK_LAUNCH();
  return RB_OK;
}

// ---------------------------------------------------------------------------
// detector readout: out[d][s] = sum_j w_j phi[row_j][src0 + s]
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_getdet(int ndet, const int32_t* det_ptr, const int32_t* det_row, const double* det_val,
                         const T* phi, int64_t ldphi, int src0, int nsrc, T* out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)ndet * nsrc) return;
  int d = (int)(i / nsrc), s = (int)(i - (int64_t)d * nsrc);
  T acc = sc<T>::zero();
  for (int32_t j = det_ptr[d]; j < det_ptr[d + 1]; ++j)
    acc = acc + jscale(det_val[j], phi[(int64_t)det_row[j] * ldphi + src0 + s]);
  out[i] = acc;
}

extern "C" int rb_getdet(int ndet, const int32_t* det_ptr, const int32_t* det_row, const double* det_val,
                         const void* phi, int64_t ldphi, int src0, int nsrc, int is_complex, void* out,
                         void* stream) {
  if (ndet <= 0 || nsrc <= 0) return RB_OK;
  hipStream_t s = (hipStream_t)stream;
  unsigned g = (unsigned)ceil_div((int64_t)ndet * nsrc, 256);
  if (is_complex)
    hipLaunchKernelGGL(k_getdet<cplx<double>>, dim3(g), dim3(256), 0, s, ndet, det_ptr, det_row, det_val,
                       (const cplx<double>*)phi, ldphi, src0, nsrc, (cplx<double>*)out);
  else
    hipLaunchKernelGGL(k_getdet<double>, dim3(g), dim3(256), 0, s, ndet, det_ptr, det_row, det_val,
                       (const double*)phi, ldphi, src0, nsrc, (double*)out);
  RB_CHECK_LAUNCH();
  return RB_OK;
}

// ---------------------------------------------------------------------------
// contraction g[n] = sum_{s,d} w[d][s] Jnode[(s,d), n] without forming J:
//   psi_l[s] = sum_d w[d][s] phi_l[d]                       (per node)
//   h_e      = sum_s (sum_{i in e} phi_i[s] psi_i[s] + S_phi[s] S_psi[s])
//   g[n]     = -0.0125 sum_{e ni n} V_e h_e                  (per node, tets ascending)
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_contract_psi(int64_t nn, const T* phi, int64_t ldphi, int Ns, int Nd, const int32_t* dcol,
                               const T* w, T* psi) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nn * Ns) return;
  int64_t n = i / Ns;
  int s = (int)(i - n * Ns);
  T acc = sc<T>::zero();
  for (int d = 0; d < Nd; ++d) acc = acc + jmul(w[(int64_t)d * Ns + s], phi[n * ldphi + dcol[d]]);
  psi[i] = acc;
}

template <typename T>
__global__ void k_contract_elem(int64_t ne, const int32_t* el, const T* phi, int64_t ldphi, int Ns,
                                const int32_t* scol, const T* psi, T* h) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  const int32_t* q = el + 4 * e;
  T acc = sc<T>::zero();
  for (int s = 0; s < Ns; ++s) {
    T dg = sc<T>::zero(), sa = sc<T>::zero(), sb = sc<T>::zero();
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      T a = phi[(int64_t)q[i] * ldphi + scol[s]];
      T b = psi[(int64_t)q[i] * Ns + s];
      dg = dg + jmul(a, b);
      sa = sa + a;
      sb = sb + b;
    }
    acc = acc + dg + jmul(sa, sb);
  }
  h[e] = acc;
}

template <typename T>
__global__ void k_contract_node(int64_t nn, const int32_t* inc_ptr, const int32_t* inc_list, const double* evol,
                                const T* h, T* g) {
  int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nn) return;
  T acc = sc<T>::zero();
  for (int32_t t = inc_ptr[n]; t < inc_ptr[n + 1]; ++t) {
    int32_t e = inc_list[t] >> 2;
    acc = acc + jscale(evol[e], h[e]);
  }
  g[n] = jscale(-0.0125, acc);
}

extern "C" int rb_jac_contract_grid(int64_t nn, int64_t ne, const int32_t* el, const double* evol,
                                    const int32_t* inc_ptr, const int32_t* inc_list, const void* phi, int64_t ldphi,
                                    int is_complex, int Ns, int Nd, const int32_t* scol, const int32_t* dcol,
                                    const void* w, void* psi_work, void* h_work, void* g, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  if (nn <= 0 || ne <= 0 || Ns <= 0 || Nd <= 0) return RB_ERR_ARG;
  unsigned g1 = (unsigned)ceil_div(nn * Ns, 256), g2 = (unsigned)ceil_div(ne, 256), g3 = (unsigned)ceil_div(nn, 256);
  if (is_complex) {
    typedef cplx<double> C2;
    hipLaunchKernelGGL(k_contract_psi<C2>, dim3(g1), dim3(256), 0, s, nn, (const C2*)phi, ldphi, Ns, Nd, dcol,
                       (const C2*)w, (C2*)psi_work);
    hipLaunchKernelGGL(k_contract_elem<C2>, dim3(g2), dim3(256), 0, s, ne, el, (const C2*)phi, ldphi, Ns, scol,
                       (const C2*)psi_work, (C2*)h_work);
    hipLaunchKernelGGL(k_contract_node<C2>, dim3(g3), dim3(256), 0, s, nn, inc_ptr, inc_list, evol,
                       (const C2*)h_work, (C2*)g);
  } else {
    hipLaunchKernelGGL(k_contract_psi<double>, dim3(g1), dim3(256), 0, s, nn, (const double*)phi, ldphi, Ns, Nd,
                       dcol, (const double*)w, (double*)psi_work);
    hipLaunchKernelGGL(k_contract_elem<double>, dim3(g2), dim3(256), 0, s, ne, el, (const double*)phi, ldphi, Ns,
                       scol, (const double*)psi_work, (double*)h_work);
    hipLaunchKernelGGL(k_contract_node<double>, dim3(g3), dim3(256), 0, s, nn, inc_ptr, inc_list, evol,
                       (const double*)h_work, (double*)g);
  }
  RB_CHECK_LAUNCH();
  return RB_OK;
}
