// K1: FEM system assembly into a precomputed CSR pattern.
//
// One thread per CSR slot walks the slot's contribution list (codes from
// rb_slot_map, ascending) and sums the element/face values in that fixed
// order: a deterministic segmented reduction, no atomics.  Several systems
// that share the pattern (wavelengths) are assembled in the same pass.
//
// Element values follow redbirdpy forward.py:144-246:
//   elem props : D*V grad_i.grad_j + {0.10, 0.05} mua V  (+ i {0.10,0.05} w/c n V)
//   node props : D_e = mean nodal D, mass with exact P1 triple-product weights
//                diag (3 mu_i + sum_{k!=i} mu_k)/60, off (2mu_i+2mu_j+rest)/120
//   Robin faces: area (1-R)/(12(1+R)) on the diagonal, half off-diagonal
#include "rb_common.h"
#include "../../include/rbhip.h"

#define RB_C0 299792458000.0

__device__ __constant__ int c_pairidx[16] = {0, 1, 2, 3, 1, 4, 5, 6, 2, 5, 7, 8, 3, 6, 8, 9};

struct AsmArgs {
  const int32_t* slot_ptr;
  const int32_t* codes;
  int64_t nnz, ne, nf, nn;
  const int32_t* el;
  const int32_t* fc;
  const double* dd;
  const double* evol;
  const double* area;
  int nsys, kind;
  const double* mua;   // nsys x plen
  const double* musp;  // nsys x plen (already mus*(1-g))
  const double* nref;  // nsys x nlen
  int64_t plen, nlen;
  const double* omega;  // nsys
  const double* bc;     // nsys: (1-R)/(12(1+R))
  int is_complex;
  void* out64;
  void* out32;
};

template <int KIND>
__global__ void __launch_bounds__(256) k_assemble(AsmArgs a) {
  int64_t sl = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (sl >= a.nnz) return;
  int32_t c0 = a.slot_ptr[sl], c1 = a.slot_ptr[sl + 1];
  const int64_t face_base = 16 * a.ne;
  for (int s = 0; s < a.nsys; ++s) {
    double re = 0.0, im = 0.0;
    const double om = a.omega[s];
    const double wc = om / RB_C0;
    const double* mua = a.mua + (int64_t)s * a.plen;
    const double* musp = a.musp + (int64_t)s * a.plen;
    const double* nref = a.nref + (int64_t)s * a.nlen;
    for (int32_t c = c0; c < c1; ++c) {
      int32_t code = a.codes[c];
      if (code < face_base) {
        int32_t e = code >> 4;
        int ij = code & 15;
        int i = ij >> 2, j = ij & 3;
        double V = a.evol[e];
        if (KIND == RB_PROP_VOLUME) {
          re += V;
          continue;
        }
        double g = a.dd[10 * (int64_t)e + c_pairidx[ij]];
        if (KIND == RB_PROP_ELEM) {
          double mu = mua[e];
          double D = 1.0 / (3.0 * (mu + musp[e]));
          double w = (i == j) ? 0.10 : 0.05;
          re += g * D + w * mu * V;
          if (om > 0) {
            double n = (a.nlen == 1) ? nref[0] : nref[e];
            im += w * om * (1.0 / RB_C0) * n * V;
          }
        } else {  // RB_PROP_NODE
          const int32_t* q = a.el + 4 * (int64_t)e;
          double mu[4], Ds = 0.0;
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            mu[k] = mua[q[k]];
            Ds += 1.0 / (3.0 * (mu[k] + musp[q[k]]));
          }
          double De = Ds / 4.0;
          double wm = 0.0;
          if (i == j) {
#pragma unroll
            for (int k = 0; k < 4; ++k) wm += mu[k] * ((k == i) ? (3.0 / 60.0) : (1.0 / 60.0));
          } else {
#pragma unroll
            for (int k = 0; k < 4; ++k) wm += mu[k] * ((k == i || k == j) ? (2.0 / 120.0) : (1.0 / 120.0));
          }
          re += g * De + wm * V;
          if (om > 0) {
            double wn = 0.0;
            if (a.nlen == 1) {
              wn = nref[0] * ((i == j) ? 0.10 : 0.05);
            } else if (i == j) {
#pragma unroll
              for (int k = 0; k < 4; ++k) wn += nref[q[k]] * ((k == i) ? (3.0 / 60.0) : (1.0 / 60.0));
            } else {
#pragma unroll
              for (int k = 0; k < 4; ++k) wn += nref[q[k]] * ((k == i || k == j) ? (2.0 / 120.0) : (1.0 / 120.0));
            }
            im += om * (1.0 / RB_C0) * wn * V;
          }
        }
        (void)wc;
      } else {
        if (KIND == RB_PROP_VOLUME) continue;
        int32_t r = code - (int32_t)face_base;
        int32_t f = r / 9;
        int ij = r - 9 * f;
        int i = ij / 3, j = ij - 3 * (ij / 3);
        double ad = a.area[f] * a.bc[s];
        re += (i == j) ? ad : ad * 0.5;
      }
    }
    int64_t o = (int64_t)s * a.nnz + sl;
    if (a.is_complex) {
      double* p = (double*)a.out64;
      p[2 * o] = re;
      p[2 * o + 1] = im;
      if (a.out32) {
        float* q = (float*)a.out32;
        q[2 * o] = (float)re;
        q[2 * o + 1] = (float)im;
      }
    } else {
      ((double*)a.out64)[o] = re;
      if (a.out32) ((float*)a.out32)[o] = (float)re;
    }
  }
}

extern "C" int rb_assemble(const int32_t* slot_ptr, const int32_t* codes, int64_t nnz, const int32_t* el, int64_t ne,
                           const int32_t* fc, int64_t nf, int64_t nn, const double* dd, const double* evol,
                           const double* area, int nsys, int kind, const double* mua, const double* musp,
                           const double* nref, int64_t plen, int64_t nlen, const double* omega, const double* bc,
                           int is_complex, void* out64, void* out32, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  if (nnz <= 0) return RB_OK;
  if (nsys <= 0 || !out64) {
    rb_set_error("rb_assemble: bad arguments");
    return RB_ERR_ARG;
  }
  if (kind != RB_PROP_VOLUME) {
    int64_t need = (kind == RB_PROP_ELEM) ? ne : nn;
    if (plen != need || !(nlen == 1 || nlen == plen) || !mua || !musp || !nref || !omega || !bc) {
      rb_set_error("rb_assemble: property length %lld does not match kind %d (need %lld)", (long long)plen, kind,
                   (long long)need);
      return RB_ERR_ARG;
    }
  }
  AsmArgs a;
  a.slot_ptr = slot_ptr;
  a.codes = codes;
  a.nnz = nnz;
  a.ne = ne;
  a.nf = nf;
  a.nn = nn;
  a.el = el;
  a.fc = fc;
  a.dd = dd;
  a.evol = evol;
  a.area = area;
  a.nsys = nsys;
  a.kind = kind;
  a.mua = mua;
  a.musp = musp;
  a.nref = nref;
  a.plen = plen;
  a.nlen = nlen;
  a.omega = omega;
  a.bc = bc;
  a.is_complex = is_complex;
  a.out64 = out64;
  a.out32 = out32;
  unsigned g = (unsigned)ceil_div(nnz, 256);
  if (kind == RB_PROP_ELEM)
    hipLaunchKernelGGL(k_assemble<RB_PROP_ELEM>, dim3(g), dim3(256), 0, s, a);
  else if (kind == RB_PROP_NODE)
    hipLaunchKernelGGL(k_assemble<RB_PROP_NODE>, dim3(g), dim3(256), 0, s, a);
  else if (kind == RB_PROP_VOLUME)
    hipLaunchKernelGGL(k_assemble<RB_PROP_VOLUME>, dim3(g), dim3(256), 0, s, a);
  else {
    rb_set_error("rb_assemble: unknown kind %d", kind);
    return RB_ERR_ARG;
  }
  RB_CHECK_LAUNCH();
  return RB_OK;
}
