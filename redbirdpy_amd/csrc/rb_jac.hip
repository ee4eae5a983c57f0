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

#define JW 64   // threads per block (one wavefront per node)
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

// Q_n (np x np, row-major in LDS) from the node's incident tets, 64 tets at
// a time: each lane loads one tet (vertices, volume) and finds its vertices'
// positions in the sorted patch (tl, tv in LDS), then one lane per patch row
// adds the tets holding that row's node in ascending tet order (the same
// order for every run, so Q is bit-reproducible).  One wavefront per block.
__device__ inline void build_patch_q(int64_t n, const int32_t* el, const double* evol, const int32_t* inc_ptr,
                                     const int32_t* inc_list, const int32_t* cols, int np, double* Q,
                                     int32_t* tl /* 64 x 4 */, double* tv /* 64 */) {
  const int lane = threadIdx.x;
  for (int i = lane; i < np * np; i += JW) Q[i] = 0.0;
  const int32_t t0 = inc_ptr[n], t1 = inc_ptr[n + 1];
  for (int32_t c0 = t0; c0 < t1; c0 += JW) {
    const int nt = min(JW, t1 - c0);
    if (lane < nt) {
      const int32_t e = inc_list[c0 + lane] >> 2;
      const int32_t* q = el + 4 * (int64_t)e;
      int32_t v[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) v[j] = q[j];
      tv[lane] = evol[e];
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
        tl[lane * 4 + j] = pos;
      }
    }
    __syncthreads();
    for (int k = lane; k < np; k += JW) {
      double* qk = Q + k * np;
      for (int t = 0; t < nt; ++t) {
        const int32_t l0 = tl[t * 4], l1 = tl[t * 4 + 1], l2 = tl[t * 4 + 2], l3 = tl[t * 4 + 3];
        if (l0 != k && l1 != k && l2 != k && l3 != k) continue;
        const double V = tv[t];
        if (l0 >= 0) qk[l0] += (l0 == k) ? 2.0 * V : V;
        if (l1 >= 0) qk[l1] += (l1 == k) ? 2.0 * V : V;
        if (l2 >= 0) qk[l2] += (l2 == k) ? 2.0 * V : V;
        if (l3 >= 0) qk[l3] += (l3 == k) ? 2.0 * V : V;
      }
    }
    __syncthreads();
  }
}

template <typename T, typename TO>
__global__ void __launch_bounds__(JW) k_jac_node(int64_t nn, const int32_t* __restrict__ el,
                                                 const double* __restrict__ evol, const int32_t* __restrict__ inc_ptr,
                                                 const int32_t* __restrict__ inc_list,
                                                 const int32_t* __restrict__ row_ptr,
                                                 const int32_t* __restrict__ col_idx, const T* __restrict__ phi,
                                                 int64_t ldphi, int Ns, int Nd, const int32_t* __restrict__ scol,
                                                 const int32_t* __restrict__ dcol, const int32_t* __restrict__ pmap,
                                                 TO* __restrict__ JT, int64_t ldj, int maxnp) {
  extern __shared__ __align__(16) unsigned char smem[];
  T* PS = (T*)smem;                        // np x JC
  T* PD = PS + maxnp * JC;                 // np x JC
  T* Y = PD + maxnp * JC;                  // np x JC
  double* Q = (double*)(Y + maxnp * JC);   // np x np
  double* tv = Q + maxnp * maxnp;          // 64 staged tet volumes
  int32_t* tl = (int32_t*)(tv + JW);       // their 64 x 4 patch positions
  int32_t* cols = tl + 4 * JW;
  const int lane = threadIdx.x;
  // lane -> 4x4 tile of the 32x32 chunk: 8 tiles along s, 8 along d
  const int ts = (lane & 7) * 4;
  const int td = (lane >> 3) * 4;
  for (int64_t n = blockIdx.x; n < nn; n += gridDim.x) {
    int32_t r0 = row_ptr[n];
    int np = row_ptr[n + 1] - r0;
    for (int k = lane; k < np; k += JW) cols[k] = col_idx[r0 + k];
    __syncthreads();
    build_patch_q(n, el, evol, inc_ptr, inc_list, cols, np, Q, tl, tv);
    __syncthreads();
    for (int s0 = 0; s0 < Ns; s0 += JC) {
      int ns = min(JC, Ns - s0);
      for (int i = lane; i < np * JC; i += JW) {
        int k = i / JC, j = i - k * JC;
        PS[i] = (j < ns) ? phi[(int64_t)cols[k] * ldphi + scol[s0 + j]] : sc<T>::zero();
      }
      for (int d0 = 0; d0 < Nd; d0 += JC) {
        int nd = min(JC, Nd - d0);
        for (int i = lane; i < np * JC; i += JW) {
          int k = i / JC, j = i - k * JC;
          PD[i] = (j < nd) ? phi[(int64_t)cols[k] * ldphi + dcol[d0 + j]] : sc<T>::zero();
        }
        __syncthreads();
        // Y = Q PD
        for (int i = lane; i < np * JC; i += JW) {
          int k = i / JC, j = i - k * JC;
          T acc = sc<T>::zero();
          const double* qk = Q + k * np;
          for (int l = 0; l < np; ++l) acc = acc + jscale(qk[l], PD[l * JC + j]);
          Y[i] = acc;
        }
        __syncthreads();
        T acc[4][4];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
          for (int b = 0; b < 4; ++b) acc[a][b] = sc<T>::zero();
        for (int k = 0; k < np; ++k) {
          T ps[4], y[4];
#pragma unroll
          for (int a = 0; a < 4; ++a) {
            ps[a] = PS[k * JC + ts + a];
            y[a] = Y[k * JC + td + a];
          }
#pragma unroll
          for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] = acc[a][b] + jmul(ps[a], y[b]);
        }
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          int d = d0 + td + b;
          if (td + b >= nd) continue;
#pragma unroll
          for (int a = 0; a < 4; ++a) {
            int s = s0 + ts + a;
            if (ts + a >= ns) continue;
            int64_t cell = (int64_t)d * Ns + s;
            int64_t p = pmap ? pmap[cell] : cell;
            if (p >= 0) JT[n * ldj + p] = cast_out<TO, T>(jscale(-0.0125, acc[a][b]));
          }
        }
        __syncthreads();
      }
    }
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

static size_t jac_lds(int maxnp, int is_complex) {
  size_t sv = is_complex ? 16 : 8;
  return (size_t)3 * maxnp * JC * sv + (size_t)maxnp * maxnp * 8 + (size_t)JW * (8 + 16) + (size_t)maxnp * 4 + 16;
}

extern "C" int rb_jac_grid(int64_t nn, int64_t ne, const int32_t* el, const double* evol, const int32_t* inc_ptr,
                           const int32_t* inc_list, const int32_t* row_ptr, const int32_t* col_idx, int maxnp,
                           const void* phi, int64_t ldphi, int is_complex, int Ns, int Nd, const int32_t* scol,
                           const int32_t* dcol, const int32_t* pmap, void* JT, int64_t ldj, int out_dtype,
                           void* Jelem, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  if (nn <= 0 || Ns <= 0 || Nd <= 0 || maxnp <= 0) {
    rb_set_error("rb_jac_grid: bad sizes");
    return RB_ERR_ARG;
  }
  size_t lds = jac_lds(maxnp, is_complex);
  if (lds > 64 * 1024) {
    rb_set_error("rb_jac_grid: node patch of %d nodes needs %zu B of LDS (> 64 KiB)", maxnp, lds);
    return RB_ERR_SIZE;
  }
  unsigned g = (unsigned)(nn < 262144 ? nn : 262144);
  unsigned ge = (unsigned)(ne < 262144 ? ne : 262144);
  if (!is_complex) {
    if (out_dtype == RB_F64) {
      if (JT)
        hipLaunchKernelGGL((k_jac_node<double, double>), dim3(g), dim3(JW), lds, s, nn, el, evol, inc_ptr, inc_list,
                           row_ptr, col_idx, (const double*)phi, ldphi, Ns, Nd, scol, dcol, pmap, (double*)JT, ldj,
                           maxnp);
      if (Jelem && ne > 0)
        hipLaunchKernelGGL((k_jac_elem<double, double>), dim3(ge), dim3(256), 0, s, ne, el, evol, (const double*)phi,
                           ldphi, Ns, Nd, scol, dcol, pmap, (double*)Jelem, ldj);
    } else if (out_dtype == RB_F32) {
      if (JT)
        hipLaunchKernelGGL((k_jac_node<double, float>), dim3(g), dim3(JW), lds, s, nn, el, evol, inc_ptr, inc_list,
                           row_ptr, col_idx, (const double*)phi, ldphi, Ns, Nd, scol, dcol, pmap, (float*)JT, ldj,
                           maxnp);
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
        hipLaunchKernelGGL((k_jac_node<C2, C2>), dim3(g), dim3(JW), lds, s, nn, el, evol, inc_ptr, inc_list, row_ptr,
                           col_idx, (const C2*)phi, ldphi, Ns, Nd, scol, dcol, pmap, (C2*)JT, ldj, maxnp);
      if (Jelem && ne > 0)
        hipLaunchKernelGGL((k_jac_elem<C2, C2>), dim3(ge), dim3(256), 0, s, ne, el, evol, (const C2*)phi, ldphi, Ns,
                           Nd, scol, dcol, pmap, (C2*)Jelem, ldj);
    } else if (out_dtype == RB_C64) {
      if (JT)
        hipLaunchKernelGGL((k_jac_node<C2, cplx<float>>), dim3(g), dim3(JW), lds, s, nn, el, evol, inc_ptr, inc_list,
                           row_ptr, col_idx, (const C2*)phi, ldphi, Ns, Nd, scol, dcol, pmap, (cplx<float>*)JT, ldj,
                           maxnp);
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
