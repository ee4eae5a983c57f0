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

// ---------------------------------------------------------------------------
// Binned assembly (rb_assemble_ws).  A property pass writes per (element, system)
//   element properties: {D = 1/(3(mua+musp)), mua V, (omega/c) n V, V}    (32 bytes)
//   node properties   : {mean nodal D, S = sum of nodal mua}               (16 bytes)
//                       and Sn = sum of nodal n                           (8 bytes)
// (node properties first go through a per-node pass that writes {D, mua, n}
// with the systems of a node side by side, so an element gathers four rows).
// Then each workgroup takes the slots whose first code lies in its run of
// consecutive codes (BM bins of ASM_GRAN): its threads compute the codes'
// contributions (coalesced code loads; the deldotdel and record gathers of
// neighbouring codes share elements) into LDS, several codes per thread with
// their loads in flight together, and one thread per slot adds its slot's
// contributions in the slot map's ascending order.
//
// Every tet code of slot (r, q) has i <-> node r and j <-> node q, so the P1
// mass weight (redbirdpy forward.py:177-228) splits into a per-code term and
// a per-slot one:
//   diagonal     (3 mu_i + sum_{k!=i} mu_k)/60 V = S V/60  + (2 mu_r/60) V
//   off-diagonal (2 mu_i + 2 mu_j + rest)/120 V  = S V/120 + ((mu_r + mu_q)/120) V
// a code contributes g D + S V k (k = 1/60 or 1/120) and V, and the slot adds
// its coefficient times the sum of V (the same for n in the frequency term).
// ---------------------------------------------------------------------------
#define RB_ASM_MAXSYS 4  // systems per slot pass (larger counts run in groups)
#define ASM_GRAN 256     // codes per bin of the bin table; a workgroup (256 threads) takes BM bins
#define ASM_TAIL 256     // LDS room past a workgroup's bins for its last slot's codes

struct AsmTabArgs {
  const int32_t* slot_ptr;
  const int32_t* codes;
  int64_t nnz, ne;
  const int32_t* el;
  const double* dd;
  const double* area;
  const double* mua;   // node kind: nsys x nn
  const double* nref;  // node kind with nlen > 1: nsys x nn
  int64_t plen, nlen;
  const double* omega;
  const double* bc;
  const double4* tab;    // element properties: ne x nsys records
  const double2* ntab;   // node properties: ne x nsys {mean nodal D, sum of nodal mua}
  const double* nsn;     // node properties: ne x nsys sums of nodal n (frequency term, nlen > 1)
  const double* evol;
  const int32_t* bins;   // first slot of each bin (rb_assemble_bins), nbin + 1 entries
  int64_t nbin;
  int nsys, s0, ns;      // systems [s0, s0 + ns) of nsys
  void* out64;
  void* out32;
};

__global__ void __launch_bounds__(256)
    k_asm_props_elem(int64_t ne, int nsys, const double* __restrict__ evol, const double* __restrict__ mua,
                     const double* __restrict__ musp, const double* __restrict__ nref, int64_t plen, int64_t nlen,
                     const double* __restrict__ omega, double4* __restrict__ tab) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  const double V = evol[e];
  for (int s = 0; s < nsys; ++s) {
    const double m = mua[(int64_t)s * plen + e];
    double4 t;
    t.x = 1.0 / (3.0 * (m + musp[(int64_t)s * plen + e]));
    t.y = m * V;
    t.z = omega[s] * (1.0 / RB_C0) * ((nlen == 1) ? nref[(int64_t)s * nlen] : nref[(int64_t)s * nlen + e]) * V;
    t.w = V;
    tab[e * nsys + s] = t;
  }
}

// node properties, per (node, system): rows[n * nsys + s] = {1/(3(mua+musp)), mua, n}, the
// systems of a node side by side (consecutive threads write consecutive rows)
__global__ void __launch_bounds__(256)
    k_asm_node_rows(int64_t nn, int nsys, const double* __restrict__ mua, const double* __restrict__ musp,
                    const double* __restrict__ nref, int64_t nlen, double* __restrict__ rows) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nn * nsys) return;
  const int64_t n = i / nsys, s = i - n * nsys;
  const double m = mua[s * nn + n];
  double* r = rows + 3 * i;
  r[0] = 1.0 / (3.0 * (m + musp[s * nn + n]));
  r[1] = m;
  r[2] = nlen > 1 ? nref[s * nlen + n] : 0.0;
}

// node properties, per (element, system) (consecutive threads write consecutive records):
// the mean of the element's nodal D and the sums of its nodal mua and n
__global__ void __launch_bounds__(256)
    k_asm_props_node(int64_t ne, int nsys, const int32_t* __restrict__ el, const double* __restrict__ rows,
                     double2* __restrict__ ntab, double* __restrict__ nsn) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ne * nsys) return;
  const int64_t e = i / nsys;
  const int s = (int)(i - e * nsys);
  double Ds = 0.0, S = 0.0, Sn = 0.0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const double* r = rows + 3 * ((int64_t)el[4 * e + k] * nsys + s);
    Ds += r[0];
    S += r[1];
    Sn += r[2];
  }
  ntab[i] = make_double2(Ds / 4.0, S);
  nsn[i] = Sn;
}

// first slot whose first code is at or after position v (slot_ptr[0..nnz] ascending)
__device__ inline int64_t slot_lower_bound(const int32_t* __restrict__ slot_ptr, int64_t nnz, int64_t v) {
  int64_t lo = 0, hi = nnz;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if ((int64_t)slot_ptr[mid] < v)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// bins[b] = first slot whose first code is at or after b * ASM_GRAN, b = 0..nbin
__global__ void k_asm_bins(const int32_t* __restrict__ slot_ptr, int64_t nnz, int64_t nbin, int32_t* __restrict__ bins) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b > nbin) return;
  bins[b] = (int32_t)slot_lower_bound(slot_ptr, nnz, b * ASM_GRAN);
}

// one code's contribution to its slot: cr/ci per system, v = V of a tet code under node
// properties (the per-slot mass term multiplies the summed V)
template <int KIND, bool CPLX, int NS>
__device__ __forceinline__ void asm_contrib(const AsmTabArgs& a, int32_t code, double* cr, double* ci, double& v) {
  const int64_t face_base = 16 * a.ne;
  v = 0.0;
  if (code < face_base) {
    const int64_t e = code >> 4;
    const int ij = code & 15;
    const bool dg = (ij >> 2) == (ij & 3);
    const double g = a.dd[10 * e + c_pairidx[ij]];
    if (KIND == RB_PROP_ELEM) {
      const double4* tp = a.tab + e * a.nsys + a.s0;
      const double w = dg ? 0.10 : 0.05;
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        const double4 p = tp[s];
        cr[s] = g * p.x + w * p.y;
        if (CPLX) ci[s] = w * p.z;
      }
    } else {
      const double2* tp = a.ntab + e * a.nsys + a.s0;
      const double V = a.evol[e];
      const double kw = dg ? (1.0 / 60.0) : (1.0 / 120.0);
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        const double2 p = tp[s];
        cr[s] = g * p.x + p.y * V * kw;
        if (CPLX) ci[s] = (a.nlen > 1) ? a.nsn[e * a.nsys + a.s0 + s] * V * kw : 0.0;
      }
      v = V;
    }
  } else {
    const int32_t r = code - (int32_t)face_base;
    const int32_t f = r / 9;
    const int ij = r - 9 * f;
    const double ad = ((ij / 3) == (ij % 3)) ? a.area[f] : a.area[f] * 0.5;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      cr[s] = ad * a.bc[a.s0 + s];
      if (CPLX) ci[s] = 0.0;
    }
  }
}

// the per-slot terms of node properties: the mass and index coefficients of the slot's two
// nodes (from its first code, a tet code: tet codes sort first and every slot has one)
template <bool CPLX, int NS>
__device__ __forceinline__ void asm_slot_coef(const AsmTabArgs& a, int32_t first, double* cm, double* cn) {
  const int ij = first & 15;
  const bool dg = (ij >> 2) == (ij & 3);
  const int64_t e0 = first >> 4;
  const int32_t r = a.el[4 * e0 + (ij >> 2)], q = a.el[4 * e0 + (ij & 3)];
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const int64_t so = (int64_t)(a.s0 + s);
    const double mr = a.mua[so * a.plen + r], mq = a.mua[so * a.plen + q];
    cm[s] = dg ? (2.0 * mr) * (1.0 / 60.0) : (mr + mq) * (1.0 / 120.0);
    if (CPLX) {
      if (a.nlen > 1) {
        const double nr = a.nref[so * a.nlen + r], nq = a.nref[so * a.nlen + q];
        cn[s] = dg ? (2.0 * nr) * (1.0 / 60.0) : (nr + nq) * (1.0 / 120.0);
      } else {
        cn[s] = a.nref[so * a.nlen] * (dg ? 0.10 : 0.05);
      }
    }
  }
}

template <int KIND, bool CPLX, int NS>
__device__ __forceinline__ void asm_write(const AsmTabArgs& a, int64_t sl, const double* re, const double* im,
                                          const double* cm, const double* cn, double sv) {
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    double r = re[s], i = CPLX ? im[s] : 0.0;
    if (KIND == RB_PROP_NODE) {
      r += cm[s] * sv;
      if (CPLX) i = a.omega[a.s0 + s] * (1.0 / RB_C0) * (i + cn[s] * sv);
    }
    const int64_t o = (int64_t)(a.s0 + s) * a.nnz + sl;
    if (CPLX) {
      ((double2*)a.out64)[o] = make_double2(r, i);
      if (a.out32) ((float2*)a.out32)[o] = make_float2((float)r, (float)i);
    } else {
      ((double*)a.out64)[o] = r;
      if (a.out32) ((float*)a.out32)[o] = (float)r;
    }
  }
}

// Workgroup b takes the slots that start in codes [b BM ASM_GRAN, (b+1) BM ASM_GRAN):
// their codes (at most BM ASM_GRAN plus the last slot's tail) are staged in LDS,
// BM + 1 codes per thread with the loads of all of them issued first, then thread t sums
// slots sb + t + 256 j (j < BM) in code order.  A workgroup whose codes exceed the LDS
// room (a slot of more than ASM_TAIL codes) sums its slots straight from memory.
template <int KIND, bool CPLX, int NS, int BM>
__global__ void __launch_bounds__(256) k_assemble_bin(AsmTabArgs a) {
  constexpr int LCAP = BM * ASM_GRAN + ASM_TAIL;
  constexpr int CPT = LCAP / 256;
  constexpr int OWN = BM * ASM_GRAN / 256;
  __shared__ double l_re[NS][LCAP];
  __shared__ double l_im[CPLX ? NS : 1][CPLX ? LCAP : 1];
  __shared__ double l_v[KIND == RB_PROP_NODE ? LCAP : 1];
  const int t = threadIdx.x;
  // workgroups go to the 8 accelerator dies (each with its own L2) round-robin: die x takes
  // one contiguous run of bins, so the element rows that neighbouring bins share are fetched
  // into one L2
  const int64_t G = gridDim.x, q8 = G / 8, r8 = G % 8, x8 = blockIdx.x % 8, k8 = blockIdx.x / 8;
  const int64_t lb = (x8 < r8 ? x8 * (q8 + 1) : r8 * (q8 + 1) + (x8 - r8) * q8) + k8;
  const int64_t b0 = lb * BM, b1 = b0 + BM < a.nbin ? b0 + BM : a.nbin;
  const int64_t sb = a.bins[b0], se = a.bins[b1];
  if (sb >= se) return;  // no slot starts in these bins (uniform over the workgroup)
  const int64_t cb = a.slot_ptr[sb], ce = a.slot_ptr[se];
  if (ce - cb > LCAP) {  // uniform: every slot from memory, one thread per slot
    for (int64_t sl = sb + t; sl < se; sl += 256) {
      const int64_t o0 = a.slot_ptr[sl], o1 = a.slot_ptr[sl + 1];
      double re[NS], im[NS], cm[NS], cn[NS], sv = 0.0;
#pragma unroll
      for (int s = 0; s < NS; ++s) re[s] = im[s] = cm[s] = cn[s] = 0.0;
      if (KIND == RB_PROP_NODE && o1 > o0) asm_slot_coef<CPLX, NS>(a, a.codes[o0], cm, cn);
      for (int64_t c = o0; c < o1; ++c) {
        double cr[NS], ci[NS], v;
        asm_contrib<KIND, CPLX, NS>(a, a.codes[c], cr, ci, v);
#pragma unroll
        for (int s = 0; s < NS; ++s) {
          re[s] += cr[s];
          if (CPLX) im[s] += ci[s];
        }
        sv += v;
      }
      asm_write<KIND, CPLX, NS>(a, sl, re, im, cm, cn, sv);
    }
    return;
  }
  const int wn = (int)(ce - cb);
  // this thread's slots: code ranges and (node properties) their coefficients, loaded
  // ahead of the staging so their latency overlaps it
  int64_t o0[OWN], o1[OWN];
  double cm[OWN][NS], cn[OWN][NS];
#pragma unroll
  for (int j = 0; j < OWN; ++j) {
    const int64_t sl = sb + t + 256 * j;
    o0[j] = o1[j] = 0;
#pragma unroll
    for (int s = 0; s < NS; ++s) cm[j][s] = cn[j][s] = 0.0;
    if (sl < se) {
      o0[j] = a.slot_ptr[sl];
      o1[j] = a.slot_ptr[sl + 1];
      if (KIND == RB_PROP_NODE && o1[j] > o0[j]) asm_slot_coef<CPLX, NS>(a, a.codes[o0[j]], cm[j], cn[j]);
    }
  }
  int32_t cd[CPT];
#pragma unroll
  for (int j = 0; j < CPT; ++j) {
    const int k = t + 256 * j;
    cd[j] = k < wn ? a.codes[cb + k] : 0;
  }
#pragma unroll
  for (int j = 0; j < CPT; ++j) {
    const int k = t + 256 * j;
    if (k < wn) {
      double cr[NS], ci[NS], v;
      asm_contrib<KIND, CPLX, NS>(a, cd[j], cr, ci, v);
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        l_re[s][k] = cr[s];
        if (CPLX) l_im[s][k] = ci[s];
      }
      if (KIND == RB_PROP_NODE) l_v[k] = v;
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < OWN; ++j) {
    const int64_t sl = sb + t + 256 * j;
    if (sl < se) {
      double re[NS], im[NS], sv = 0.0;
#pragma unroll
      for (int s = 0; s < NS; ++s) re[s] = im[s] = 0.0;
#pragma unroll 4
      for (int64_t k = o0[j] - cb; k < o1[j] - cb; ++k) {
#pragma unroll
        for (int s = 0; s < NS; ++s) {
          re[s] += l_re[s][k];
          if (CPLX) im[s] += l_im[s][k];
        }
        if (KIND == RB_PROP_NODE) sv += l_v[k];
      }
      asm_write<KIND, CPLX, NS>(a, sl, re, im, cm[j], cn[j], sv);
    }
  }
}

static int64_t asm_nbin(int64_t ne, int64_t nf) { return ceil_div(16 * ne + 9 * nf, ASM_GRAN) + 1; }

extern "C" int rb_assemble_bins(const int32_t* slot_ptr, int64_t nnz, int64_t ne, int64_t nf, int32_t* bins,
                                void* stream) {
  if (!slot_ptr || !bins || nnz <= 0 || ne <= 0 || nf < 0) {
    rb_set_error("rb_assemble_bins: bad arguments");
    return RB_ERR_ARG;
  }
  const int64_t nbin = asm_nbin(ne, nf);
  hipLaunchKernelGGL(k_asm_bins, dim3((unsigned)ceil_div(nbin + 1, 256)), dim3(256), 0, (hipStream_t)stream, slot_ptr,
                     nnz, nbin, bins);
  RB_CHECK_LAUNCH();
  return RB_OK;
}

extern "C" int rb_assemble_bin_count(int64_t ne, int64_t nf, int64_t* nbins) {
  if (!nbins || ne <= 0 || nf < 0) {
    rb_set_error("rb_assemble_bin_count: bad arguments");
    return RB_ERR_ARG;
  }
  *nbins = asm_nbin(ne, nf) + 1;
  return RB_OK;
}

// workspace: element properties -> ne x nsys double4; node properties -> ne x nsys double2,
// ne x nsys sums of n, nn x nsys node rows of 3 doubles
static size_t asm_ws_bytes(int64_t nn, int64_t ne, int nsys, size_t off[3]) {
  const size_t a = (size_t)(ne > 0 ? ne : 1) * (size_t)nsys;
  const size_t el = a * sizeof(double4);
  off[0] = 0;
  off[1] = a * sizeof(double2);
  off[2] = off[1] + a * sizeof(double);
  const size_t nd = off[2] + (size_t)(nn > 0 ? nn : 1) * (size_t)nsys * 3 * sizeof(double);
  return el > nd ? el : nd;
}

extern "C" int rb_assemble_workspace(int64_t nn, int64_t ne, int nsys, size_t* bytes) {
  if (!bytes || ne < 0 || nn < 0 || nsys <= 0) {
    rb_set_error("rb_assemble_workspace: bad arguments");
    return RB_ERR_ARG;
  }
  size_t off[3];
  *bytes = asm_ws_bytes(nn, ne, nsys, off);
  return RB_OK;
}

extern "C" int rb_assemble_ws(const int32_t* slot_ptr, const int32_t* codes, int64_t nnz, const int32_t* el,
                              int64_t ne, const int32_t* fc, int64_t nf, int64_t nn, const double* dd,
                              const double* evol, const double* area, int nsys, int kind, const double* mua,
                              const double* musp, const double* nref, int64_t plen, int64_t nlen,
                              const double* omega, const double* bc, int is_complex, void* out64, void* out32,
                              const int32_t* bins, void* work, size_t work_bytes, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  if (nnz <= 0) return RB_OK;
  if (kind == RB_PROP_VOLUME)  // no properties: the per-code kernel
    return rb_assemble(slot_ptr, codes, nnz, el, ne, fc, nf, nn, dd, evol, area, nsys, kind, mua, musp, nref, plen,
                       nlen, omega, bc, is_complex, out64, out32, stream);
  if (nsys <= 0 || !out64 || !slot_ptr || !codes || !el || !dd || !evol || !bins || ne <= 0 || nf < 0 ||
      (nf > 0 && !area)) {
    rb_set_error("rb_assemble_ws: bad arguments");
    return RB_ERR_ARG;
  }
  if (kind != RB_PROP_ELEM && kind != RB_PROP_NODE) {
    rb_set_error("rb_assemble_ws: unknown kind %d", kind);
    return RB_ERR_ARG;
  }
  const int64_t need = (kind == RB_PROP_ELEM) ? ne : nn;
  if (plen != need || !(nlen == 1 || nlen == plen) || !mua || !musp || !nref || !omega || !bc) {
    rb_set_error("rb_assemble_ws: property length %lld does not match kind %d (need %lld)", (long long)plen, kind,
                 (long long)need);
    return RB_ERR_ARG;
  }
  size_t off[3];
  const size_t want = asm_ws_bytes(nn, ne, nsys, off);
  if (!work || work_bytes < want) {
    rb_set_error("rb_assemble_ws: workspace of %zu bytes, %zu needed", work_bytes, want);
    return RB_ERR_ARG;
  }
  char* wb = (char*)work;
  const unsigned ge = (unsigned)ceil_div(ne, 256);
  if (kind == RB_PROP_ELEM) {
    hipLaunchKernelGGL(k_asm_props_elem, dim3(ge), dim3(256), 0, s, ne, nsys, evol, mua, musp, nref, plen, nlen,
                       omega, (double4*)work);
  } else {
    double* rows = (double*)(wb + off[2]);
    hipLaunchKernelGGL(k_asm_node_rows, dim3((unsigned)ceil_div(nn * nsys, 256)), dim3(256), 0, s, nn, nsys, mua, musp,
                       nref, nlen, rows);
    hipLaunchKernelGGL(k_asm_props_node, dim3((unsigned)ceil_div(ne * nsys, 256)), dim3(256), 0, s, ne, nsys, el, rows,
                       (double2*)(wb + off[0]), (double*)(wb + off[1]));
  }
  RB_CHECK_LAUNCH();
  AsmTabArgs a;
  a.slot_ptr = slot_ptr;
  a.codes = codes;
  a.nnz = nnz;
  a.ne = ne;
  a.el = el;
  a.dd = dd;
  a.area = area;
  a.mua = mua;
  a.nref = nref;
  a.plen = plen;
  a.nlen = nlen;
  a.omega = omega;
  a.bc = bc;
  a.tab = (const double4*)work;
  a.ntab = (const double2*)(wb + off[0]);
  a.nsn = (const double*)(wb + off[1]);
  a.evol = evol;
  a.bins = bins;
  a.nbin = asm_nbin(ne, nf);
  a.nsys = nsys;
  a.out64 = out64;
  a.out32 = out32;
  // every tet contributes 16 codes and every face 9 (the slot map holds each once); four
  // bins per workgroup, two where the staged contributions of 3-4 systems would cap occupancy
  const unsigned g4 = (unsigned)ceil_div(a.nbin, 4), g2 = (unsigned)ceil_div(a.nbin, 2);
#define RB_ASM_NS(K_, C_)                                                                                 \
  switch (a.ns) {                                                                                       \
    case 1: hipLaunchKernelGGL((k_assemble_bin<K_, C_, 1, 4>), dim3(g4), dim3(256), 0, s, a); break;    \
    case 2: hipLaunchKernelGGL((k_assemble_bin<K_, C_, 2, 4>), dim3(g4), dim3(256), 0, s, a); break;    \
    case 3: hipLaunchKernelGGL((k_assemble_bin<K_, C_, 3, 2>), dim3(g2), dim3(256), 0, s, a); break;    \
    default: hipLaunchKernelGGL((k_assemble_bin<K_, C_, 4, 2>), dim3(g2), dim3(256), 0, s, a); break;   \
  }
  for (int s0 = 0; s0 < nsys; s0 += RB_ASM_MAXSYS) {
    a.s0 = s0;
    a.ns = nsys - s0 < RB_ASM_MAXSYS ? nsys - s0 : RB_ASM_MAXSYS;
    if (kind == RB_PROP_ELEM) {
      if (is_complex) {
        RB_ASM_NS(RB_PROP_ELEM, true);
      } else {
        RB_ASM_NS(RB_PROP_ELEM, false);
      }
    } else {
      if (is_complex) {
        RB_ASM_NS(RB_PROP_NODE, true);
      } else {
        RB_ASM_NS(RB_PROP_NODE, false);
      }
    }
    RB_CHECK_LAUNCH();
  }
#undef RB_ASM_NS
  return RB_OK;
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
