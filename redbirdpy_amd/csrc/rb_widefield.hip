// Wide-field sources and detectors: nodal boundary-flux columns of one
// planar / pattern / Fourier source (redbirdpy utility.py:423-636,
// _process_planar_source + _apply_bc_weighting, before the per-pattern L1
// normalisation, which the caller applies).
//
// A collimated source is sunk to the plane dir . x = dir . (srcpos + dir z0)
// (mode RB_WF_SINK): every tet crossing the plane gives a cut polygon
// (triangle, or a quad split along its first diagonal) whose corners sit on
// the tet's cut edges (n1 < n2) at w p1 + (1 - w) p2.  Otherwise the
// boundary faces themselves are used (mode RB_WF_FACES).  A cut or boundary
// triangle takes part when its centroid, seen along dir, lies in the source
// rectangle (even-odd ray casting in the rectangle's plane) and it faces the
// source (boundary faces: normal . dir < 0; cut triangles: non-degenerate).
// Its flux area * coef * pattern(centroid) goes to its corners, and a cut
// corner splits it between the ends of its edge as w : 1 - w.
//
// The columns are gathered, not scattered: one thread per (node, pattern)
// walks the node's incident tets (or faces) in ascending order and adds
// what each gives to that node, so the sums are deterministic without
// atomics.  A second pass counts the triangles inside the rectangle and
// those also facing the source (the reference's two error conditions).
#include "rb_common.h"
#include "../../include/rbhip.h"
#include <math.h>

struct WfFrame {
  double dir[3];        // source direction
  double e1[3], e2[3];  // in-plane axes (rows 1, 2 of a rotation taking dir to +z)
  double px[5], py[5];  // closed source rectangle in (e1, e2) coordinates
  double ox, oy;        // its first corner
  double ux[2], uy[2];  // unit directions of its two edges from that corner
  double lx, ly;        // edge lengths
  double plane[4];      // sink plane a x + b y + c z + e = 0
  double coef;          // (1 - Reff) / (18 (1 + Reff))
  double M, phi0;       // Fourier modulation depth and phase
  int kind, npat, kx, ky, pdim0, pdim1;
};

__device__ inline void load3(const double* p, int64_t i, double* v) {
  v[0] = p[3 * i];
  v[1] = p[3 * i + 1];
  v[2] = p[3 * i + 2];
}

// even-odd ray casting against the closed 5-point polygon (utility.py:655-669)
__device__ inline bool wf_inside(const WfFrame& f, double x, double y) {
  bool in = false;
  for (int i = 0; i < 4; ++i) {
    const double x1 = f.px[i], y1 = f.py[i], x2 = f.px[i + 1], y2 = f.py[i + 1];
    if ((y1 <= y && y < y2) || (y2 <= y && y < y1)) {
      const double xc = (x2 - x1) * (y - y1) / (y2 - y1 + 1e-30) + x1;
      if (x < xc) in = !in;
    }
  }
  return in;
}

// pattern p at rectangle coordinates of centroid (cx, cy) in plane coordinates
__device__ inline double wf_pattern(const WfFrame& f, const double* img, int p, double x, double y) {
  if (f.kind == RB_WF_UNIFORM) return 1.0;
  const double dx = x - f.ox, dy = y - f.oy;
  double bx = (dx * f.ux[0] + dy * f.ux[1]) / f.lx;
  double by = (dx * f.uy[0] + dy * f.uy[1]) / f.ly;
  bx = fmin(fmax(bx, 0.0), 1.0 - 1e-6);
  by = fmin(fmax(by, 0.0), 1.0 - 1e-6);
  if (f.kind == RB_WF_FOURIER) {
    const int i = p / f.ky, j = p - (p / f.ky) * f.ky;
    return 0.5 * (1.0 + f.M * cos((i * bx + j * by) * 2.0 * M_PI + f.phi0));
  }
  int ix = (int)(bx * f.pdim1), iy = (int)(by * f.pdim0);
  ix = ix < 0 ? 0 : (ix > f.pdim1 - 1 ? f.pdim1 - 1 : ix);
  iy = iy < 0 ? 0 : (iy > f.pdim0 - 1 ? f.pdim0 - 1 : iy);
  return img[((int64_t)iy * f.pdim1 + ix) * f.npat + p];
}

// one triangle: area, centroid in plane coordinates, and whether it lies
// inside the rectangle and faces the source
struct WfTri {
  double area, x, y;
  bool inside, facing;
};

__device__ inline WfTri wf_triangle(const WfFrame& f, const double* a, const double* b, const double* c, bool sink) {
  WfTri t;
  const double ab[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
  const double ac[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
  const double n[3] = {ab[1] * ac[2] - ab[2] * ac[1], ab[2] * ac[0] - ab[0] * ac[2], ab[0] * ac[1] - ab[1] * ac[0]};
  t.area = 0.5 * sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  double cen[3];
  for (int k = 0; k < 3; ++k) cen[k] = (a[k] + b[k] + c[k]) / 3.0;
  t.x = f.e1[0] * cen[0] + f.e1[1] * cen[1] + f.e1[2] * cen[2];
  t.y = f.e2[0] * cen[0] + f.e2[1] * cen[1] + f.e2[2] * cen[2];
  t.inside = wf_inside(f, t.x, t.y);
  double dd = n[0] * f.dir[0] + n[1] * f.dir[1] + n[2] * f.dir[2];
  if (sink) dd = -fabs(dd);  // the sink plane has no outside: every non-degenerate cut faces the source
  t.facing = dd < 0.0;
  return t;
}

// cut polygon of tet e: up to 4 corners (edge ends n1 < n2, weight w of n1,
// position); returns the corner count (0: the tet does not cross the plane)
struct WfCut {
  int n1[4], n2[4];
  double w[4];
  double p[4][3];
};

__device__ inline int wf_cut(const WfFrame& f, const double* node, const int32_t* el, int64_t e, WfCut& cut) {
  int v[4];
  double d[4];
  bool up[4];
  int nup = 0;
  for (int j = 0; j < 4; ++j) {
    v[j] = el[4 * e + j];
    double x[3];
    load3(node, v[j], x);
    d[j] = f.plane[0] * x[0] + f.plane[1] * x[1] + f.plane[2] * x[2] + f.plane[3];
    up[j] = d[j] >= 0.0;
    nup += up[j];
  }
  if (nup == 0 || nup == 4) return 0;
  int ea[4], eb[4], nc;
  if (nup == 1 || nup == 3) {
    int lone = 0;
    for (int j = 0; j < 4; ++j)
      if (up[j] == (nup == 1)) lone = j;
    nc = 0;
    for (int j = 0; j < 4; ++j)
      if (j != lone) {
        ea[nc] = lone;
        eb[nc] = j;
        ++nc;
      }
  } else {
    int pa[2], na[2], ip = 0, in = 0;
    for (int j = 0; j < 4; ++j) {
      if (up[j])
        pa[ip++] = j;
      else
        na[in++] = j;
    }
    // around the polygon: ac, ad, bd, bc
    ea[0] = pa[0], eb[0] = na[0];
    ea[1] = pa[0], eb[1] = na[1];
    ea[2] = pa[1], eb[2] = na[1];
    ea[3] = pa[1], eb[3] = na[0];
    nc = 4;
  }
  for (int k = 0; k < nc; ++k) {
    int la = ea[k], lb = eb[k];
    if (v[lb] < v[la]) {
      int t = la;
      la = lb;
      lb = t;
    }
    cut.n1[k] = v[la];
    cut.n2[k] = v[lb];
    const double w = d[lb] / (d[lb] - d[la]);
    cut.w[k] = w;
    double x1[3], x2[3];
    load3(node, v[la], x1);
    load3(node, v[lb], x2);
    for (int q = 0; q < 3; ++q) cut.p[k][q] = x1[q] * w + x2[q] * (1.0 - w);
  }
  return nc;
}

__global__ void k_wf_sink(int64_t nn, const double* __restrict__ node, const int32_t* __restrict__ el,
                          const int32_t* __restrict__ tinc_ptr, const int32_t* __restrict__ tinc, WfFrame f,
                          const double* __restrict__ img, double* __restrict__ rhs) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int p = blockIdx.y;
  if (i >= nn) return;
  double acc = 0.0;
  for (int32_t k = tinc_ptr[i]; k < tinc_ptr[i + 1]; ++k) {
    const int64_t e = tinc[k] >> 2;
    WfCut cut;
    const int nc = wf_cut(f, node, el, e, cut);
    const int ntri = nc == 4 ? 2 : (nc == 3 ? 1 : 0);
    for (int t = 0; t < ntri; ++t) {
      const int c0 = 0, c1 = t + 1, c2 = t + 2;  // (0,1,2) and, for a quad, (0,2,3)
      const WfTri tr = wf_triangle(f, cut.p[c0], cut.p[c1], cut.p[c2], true);
      if (!(tr.inside && tr.facing)) continue;
      const double wa = tr.area * f.coef * wf_pattern(f, img, p, tr.x, tr.y);
      const int cs[3] = {c0, c1, c2};
      for (int q = 0; q < 3; ++q) {
        const int c = cs[q];
        if (cut.n1[c] == i) acc += wa * cut.w[c];
        if (cut.n2[c] == i) acc += wa * (1.0 - cut.w[c]);
      }
    }
  }
  rhs[i * f.npat + p] = acc;
}

__global__ void k_wf_faces(int64_t nn, const double* __restrict__ node, const int32_t* __restrict__ fc,
                           const double* __restrict__ area, const int32_t* __restrict__ finc_ptr,
                           const int32_t* __restrict__ finc, WfFrame f, const double* __restrict__ img,
                           double* __restrict__ rhs) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int p = blockIdx.y;
  if (i >= nn) return;
  double acc = 0.0;
  for (int32_t k = finc_ptr[i]; k < finc_ptr[i + 1]; ++k) {
    const int64_t fi = finc[k] / 3;
    double a[3], b[3], c[3];
    load3(node, fc[3 * fi], a);
    load3(node, fc[3 * fi + 1], b);
    load3(node, fc[3 * fi + 2], c);
    const WfTri tr = wf_triangle(f, a, b, c, false);
    if (tr.inside && tr.facing) acc += area[fi] * f.coef * wf_pattern(f, img, p, tr.x, tr.y);
  }
  rhs[i * f.npat + p] = acc;
}

// counts[0] += triangles inside the rectangle, counts[1] += those also facing the source
__global__ void k_wf_count(int mode, int64_t nelem, const double* __restrict__ node, const int32_t* __restrict__ el,
                           const int32_t* __restrict__ fc, WfFrame f, unsigned long long* counts) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long in = 0, good = 0;
  if (e < nelem) {
    if (mode == RB_WF_SINK) {
      WfCut cut;
      const int nc = wf_cut(f, node, el, e, cut);
      const int ntri = nc == 4 ? 2 : (nc == 3 ? 1 : 0);
      for (int t = 0; t < ntri; ++t) {
        const WfTri tr = wf_triangle(f, cut.p[0], cut.p[t + 1], cut.p[t + 2], true);
        in += tr.inside;
        good += tr.inside && tr.facing;
      }
    } else {
      double a[3], b[3], c[3];
      load3(node, fc[3 * e], a);
      load3(node, fc[3 * e + 1], b);
      load3(node, fc[3 * e + 2], c);
      const WfTri tr = wf_triangle(f, a, b, c, false);
      in = tr.inside;
      good = tr.inside && tr.facing;
    }
  }
  if (in) atomicAdd(&counts[0], in);
  if (good) atomicAdd(&counts[1], good);
}

extern "C" int rb_widefield(int64_t nn, const double* node, const int32_t* el, int64_t ne, const int32_t* tinc_ptr,
                            const int32_t* tinc, const int32_t* fc, int64_t nf, const double* area,
                            const int32_t* finc_ptr, const int32_t* finc, int mode, const double* frame, int kind,
                            int npat, int kx, int ky, const double* img, int pdim0, int pdim1, double* rhs,
                            unsigned long long* counts, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  if (nn <= 0 || !node || !frame || !rhs || !counts || npat <= 0 || npat > 65535) {
    rb_set_error("rb_widefield: bad arguments (nn %lld, npat %d)", (long long)nn, npat);
    return RB_ERR_ARG;
  }
  if (mode == RB_WF_SINK && (!el || !tinc_ptr || !tinc || ne <= 0)) {
    rb_set_error("rb_widefield: the sink-plane mode needs the tets and the node-to-tet incidence");
    return RB_ERR_ARG;
  }
  if (mode == RB_WF_FACES && (!fc || !area || !finc_ptr || !finc || nf <= 0)) {
    rb_set_error("rb_widefield: the face mode needs the boundary faces, their areas and incidence");
    return RB_ERR_ARG;
  }
  if (mode != RB_WF_SINK && mode != RB_WF_FACES) {
    rb_set_error("rb_widefield: unknown mode %d", mode);
    return RB_ERR_ARG;
  }
  if (kind == RB_WF_FOURIER && (kx <= 0 || ky <= 0 || kx * ky != npat)) {
    rb_set_error("rb_widefield: Fourier patterns need kx * ky == npat (kx %d, ky %d, npat %d)", kx, ky, npat);
    return RB_ERR_ARG;
  }
  if (kind == RB_WF_IMAGE && (!img || pdim0 <= 0 || pdim1 <= 0)) {
    rb_set_error("rb_widefield: image patterns need the image and its size");
    return RB_ERR_ARG;
  }
  if (kind != RB_WF_UNIFORM && kind != RB_WF_FOURIER && kind != RB_WF_IMAGE) {
    rb_set_error("rb_widefield: unknown pattern kind %d", kind);
    return RB_ERR_ARG;
  }
  if (kind == RB_WF_UNIFORM && npat != 1) {
    rb_set_error("rb_widefield: a uniform source has one pattern (npat %d)", npat);
    return RB_ERR_ARG;
  }
  WfFrame f;
  const double* q = frame;
  for (int k = 0; k < 3; ++k) f.dir[k] = q[k], f.e1[k] = q[3 + k], f.e2[k] = q[6 + k];
  for (int k = 0; k < 5; ++k) f.px[k] = q[9 + k], f.py[k] = q[14 + k];
  f.ox = q[19], f.oy = q[20];
  f.ux[0] = q[21], f.ux[1] = q[22], f.uy[0] = q[23], f.uy[1] = q[24];
  f.lx = q[25], f.ly = q[26];
  for (int k = 0; k < 4; ++k) f.plane[k] = q[27 + k];
  f.coef = q[31];
  f.M = q[32];
  f.phi0 = q[33];
  f.kind = kind;
  f.npat = npat;
  f.kx = kx;
  f.ky = ky;
  f.pdim0 = pdim0;
  f.pdim1 = pdim1;
  RB_HIP(hipMemsetAsync(counts, 0, 2 * sizeof(unsigned long long), s));
  const int64_t nel = mode == RB_WF_SINK ? ne : nf;
  hipLaunchKernelGGL(k_wf_count, dim3((unsigned)ceil_div(nel, 256)), dim3(256), 0, s, mode, nel, node, el, fc, f,
                     counts);
  RB_CHECK_LAUNCH();
  const dim3 grid((unsigned)ceil_div(nn, 256), (unsigned)npat);
  if (mode == RB_WF_SINK)
    hipLaunchKernelGGL(k_wf_sink, grid, dim3(256), 0, s, nn, node, el, tinc_ptr, tinc, f, img, rhs);
  else
    hipLaunchKernelGGL(k_wf_faces, grid, dim3(256), 0, s, nn, node, fc, area, finc_ptr, finc, f, img, rhs);
  RB_CHECK_LAUNCH();
  return RB_OK;
}
