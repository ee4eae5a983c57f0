// Mesh preparation on the GPU (replaces redbirdpy utility.py meshprep's
// per-mesh host work: _meshreorient_fallback :1047-1063, _volface_fallback
// :1066-1087, _elemvolume_fallback :1090-1109, _nodevolume :1005-1017 and
// deldotdel :53-98) for one mesh or a batch of meshes concatenated with node
// offsets (faces never cross meshes, so one launch sequence serves all).
#include "rb_common.h"
#include "../../include/rbhip.h"

// local faces of a tet, outward for a positively oriented tet (the reference's order)
__constant__ int c_face_local[4][3] = {{0, 2, 1}, {0, 1, 3}, {0, 3, 2}, {1, 2, 3}};

// per tet: local nodes 0 and 1 swapped when the signed volume is negative;
// evol = |det|/6; deldotdel = V grad(phi_i) . grad(phi_j) over the 10
// upper-triangle pairs (00,01,02,03,11,12,13,22,23,33), gradients from the
// cofactor differences of the reoriented vertex coordinates divided by 6V
__global__ void k_tet_prep(int64_t ne, const double* __restrict__ node, const int32_t* __restrict__ el_in,
                           int32_t* __restrict__ el_out, double* __restrict__ evol, double* __restrict__ dd) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  int32_t v[4];
  double p[4][3];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    v[i] = el_in[4 * e + i];
#pragma unroll
    for (int c = 0; c < 3; ++c) p[i][c] = node[3 * (int64_t)v[i] + c];
  }
  auto det6 = [&]() {
    const double a0 = p[1][0] - p[0][0], a1 = p[1][1] - p[0][1], a2 = p[1][2] - p[0][2];
    const double b0 = p[2][0] - p[0][0], b1 = p[2][1] - p[0][1], b2 = p[2][2] - p[0][2];
    const double c0 = p[3][0] - p[0][0], c1 = p[3][1] - p[0][1], c2 = p[3][2] - p[0][2];
    return (a1 * b2 - a2 * b1) * c0 + (a2 * b0 - a0 * b2) * c1 + (a0 * b1 - a1 * b0) * c2;
  };
  double d = det6();
  if (d < 0) {
    const int32_t t = v[0];
    v[0] = v[1];
    v[1] = t;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double x = p[0][c];
      p[0][c] = p[1][c];
      p[1][c] = x;
    }
    d = det6();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) el_out[4 * e + i] = v[i];
  const double V = fabs(d) / 6.0;
  evol[e] = V;
  // gradient component c of phi_i: cofactor difference over the two other coordinates
  const int cof[4][4] = {{3, 1, 2, 1}, {2, 0, 3, 2}, {1, 3, 0, 3}, {0, 2, 1, 0}};
  double g[3][4];
  const double scale = 1.0 / (V * 6.0);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const int u = c == 0 ? 1 : 0, w = c == 2 ? 1 : 2;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int a = cof[i][0], b = cof[i][1], cc = cof[i][2], dd_ = cof[i][3];
      g[c][i] = ((p[a][u] - p[b][u]) * (p[cc][w] - p[dd_][w]) - (p[cc][u] - p[dd_][u]) * (p[a][w] - p[b][w])) * scale;
    }
  }
  int k = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = i; j < 4; ++j) {
      dd[10 * e + k] = (g[0][i] * g[0][j] + g[1][i] * g[1][j] + g[2][i] * g[2][j]) * V;
      ++k;
    }
}

#define FACE_MAX 32  // boundary faces one node may own (those whose smallest vertex it is)

// boundary faces owned by node a (the faces whose smallest vertex is a that
// lie in exactly one tet), found through a's incident tets; FILL writes them
// sorted by their other two vertices, so the faces of all nodes concatenated
// in node order are in the reference's lexicographic order of sorted triples
template <bool FILL>
__global__ void k_boundary_faces(int64_t nn, const int32_t* __restrict__ el, const int32_t* __restrict__ tinc_ptr,
                                 const int32_t* __restrict__ tinc, int32_t* __restrict__ cnt,
                                 const int32_t* __restrict__ face_ptr, int32_t* __restrict__ faces,
                                 int32_t* __restrict__ bad) {
  const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= nn) return;
  const int32_t t0 = tinc_ptr[a], t1 = tinc_ptr[a + 1];
  int n = 0;
  int32_t kb[FACE_MAX], kc[FACE_MAX], fx[FACE_MAX], fy[FACE_MAX], fz[FACE_MAX];
  for (int32_t t = t0; t < t1; ++t) {
    const int32_t e = tinc[t] >> 2, la = tinc[t] & 3;
    int32_t v[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) v[i] = el[4 * (int64_t)e + i];
    for (int f = 0; f < 4; ++f) {
      const int l0 = c_face_local[f][0], l1 = c_face_local[f][1], l2 = c_face_local[f][2];
      if (l0 != la && l1 != la && l2 != la) continue;
      const int32_t x = v[l0], y = v[l1], z = v[l2];
      const int32_t s0 = min(x, min(y, z)), s2 = max(x, max(y, z)), s1 = x + y + z - s0 - s2;
      if (s0 != a) continue;
      bool shared = false;
      for (int32_t t2 = t0; t2 < t1 && !shared; ++t2) {
        if (t2 == t) continue;
        const int64_t e2 = tinc[t2] >> 2;
        bool h1 = false, h2 = false;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int32_t w = el[4 * e2 + i];
          h1 |= w == s1;
          h2 |= w == s2;
        }
        shared = h1 && h2;
      }
      if (shared) continue;
      if (FILL) {
        if (n < FACE_MAX) {
          kb[n] = s1;
          kc[n] = s2;
          fx[n] = x;
          fy[n] = y;
          fz[n] = z;
        }
      }
      ++n;
    }
  }
  if (!FILL) {
    cnt[a] = n;
    if (n > FACE_MAX) atomicOr(bad, 1);
    return;
  }
  if (n > FACE_MAX) return;  // refused by the count pass
  for (int i = 1; i < n; ++i) {  // insertion sort by (s1, s2)
    const int32_t b = kb[i], c = kc[i], x = fx[i], y = fy[i], z = fz[i];
    int j = i - 1;
    while (j >= 0 && (kb[j] > b || (kb[j] == b && kc[j] > c))) {
      kb[j + 1] = kb[j];
      kc[j + 1] = kc[j];
      fx[j + 1] = fx[j];
      fy[j + 1] = fy[j];
      fz[j + 1] = fz[j];
      --j;
    }
    kb[j + 1] = b;
    kc[j + 1] = c;
    fx[j + 1] = x;
    fy[j + 1] = y;
    fz[j + 1] = z;
  }
  const int64_t o = face_ptr[a];
  for (int i = 0; i < n; ++i) {
    faces[3 * (o + i)] = fx[i];
    faces[3 * (o + i) + 1] = fy[i];
    faces[3 * (o + i) + 2] = fz[i];
  }
}

__global__ void k_face_area(int64_t nf, const double* __restrict__ node, const int32_t* __restrict__ fc,
                            double* __restrict__ area) {
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  const double* p0 = node + 3 * (int64_t)fc[3 * f];
  const double* p1 = node + 3 * (int64_t)fc[3 * f + 1];
  const double* p2 = node + 3 * (int64_t)fc[3 * f + 2];
  const double a0 = p1[0] - p0[0], a1 = p1[1] - p0[1], a2 = p1[2] - p0[2];
  const double b0 = p2[0] - p0[0], b1 = p2[1] - p0[1], b2 = p2[2] - p0[2];
  const double x = a1 * b2 - a2 * b1, y = a2 * b0 - a0 * b2, z = a0 * b1 - a1 * b0;
  area[f] = 0.5 * sqrt(x * x + y * y + z * z);
}

// nvol[n] = a quarter of the volumes of n's tets, summed in ascending tet order
__global__ void k_node_volume(int64_t nn, const double* __restrict__ evol, const int32_t* __restrict__ tinc_ptr,
                              const int32_t* __restrict__ tinc, double* __restrict__ nvol) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nn) return;
  double s = 0.0;
  for (int32_t t = tinc_ptr[n]; t < tinc_ptr[n + 1]; ++t) s += evol[tinc[t] >> 2];
  nvol[n] = 0.25 * s;
}

extern "C" int rb_tet_prep(int64_t nn, int64_t ne, const double* node, const int32_t* el_in, int32_t* el_out,
                           double* evol, double* deldotdel, void* stream) {
  if (nn <= 0 || ne <= 0 || !node || !el_in || !el_out || !evol || !deldotdel) {
    rb_set_error("rb_tet_prep: bad arguments");
    return RB_ERR_ARG;
  }
  hipLaunchKernelGGL(k_tet_prep, dim3((unsigned)ceil_div(ne, 128)), dim3(128), 0, (hipStream_t)stream, ne, node,
                     el_in, el_out, evol, deldotdel);
  RB_CHECK_LAUNCH();
  return RB_OK;
}

extern "C" int rb_boundary_faces(int64_t nn, const int32_t* el, const int32_t* tinc_ptr, const int32_t* tinc,
                                 int32_t* face_ptr, int32_t* faces, int64_t* nf_out, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  if (nn <= 0 || !el || !tinc_ptr || !tinc || !face_ptr || !nf_out) {
    rb_set_error("rb_boundary_faces: bad arguments");
    return RB_ERR_ARG;
  }
  const unsigned g = (unsigned)ceil_div(nn, 128);
  if (!faces) {  // count pass: per-node counts -> face_ptr (exclusive scan, total at face_ptr[nn])
    int32_t* cnt = nullptr;
    int32_t* bad = nullptr;
    RB_HIP(hipMallocAsync((void**)&cnt, (size_t)nn * sizeof(int32_t), s));
    RB_HIP(hipMallocAsync((void**)&bad, sizeof(int32_t), s));
    int rc = rb_fill_i32(bad, 0, 1, s);
    if (rc) return rc;
    hipLaunchKernelGGL((k_boundary_faces<false>), dim3(g), dim3(128), 0, s, nn, el, tinc_ptr, tinc, cnt,
                       (const int32_t*)nullptr, (int32_t*)nullptr, bad);
    RB_CHECK_LAUNCH();
    rc = rb_scan_i32(cnt, face_ptr, nn, s);
    if (rc) return rc;
    int32_t hb = 0, hn = 0;
    RB_HIP(hipMemcpyAsync(&hb, bad, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    RB_HIP(hipMemcpyAsync(&hn, face_ptr + nn, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    RB_HIP(hipStreamSynchronize(s));
    RB_HIP(hipFreeAsync(cnt, s));
    RB_HIP(hipFreeAsync(bad, s));
    if (hb) {
      rb_set_error("rb_boundary_faces: a node owns more than %d boundary faces", FACE_MAX);
      return RB_ERR_SIZE;
    }
    *nf_out = hn;
    return RB_OK;
  }
  hipLaunchKernelGGL((k_boundary_faces<true>), dim3(g), dim3(128), 0, s, nn, el, tinc_ptr, tinc, (int32_t*)nullptr,
                     face_ptr, faces, (int32_t*)nullptr);
  RB_CHECK_LAUNCH();
  return RB_OK;
}

extern "C" int rb_face_area(int64_t nf, const double* node, const int32_t* faces, double* area, void* stream) {
  if (nf <= 0) return RB_OK;
  if (!node || !faces || !area) return RB_ERR_ARG;
  hipLaunchKernelGGL(k_face_area, dim3((unsigned)ceil_div(nf, 256)), dim3(256), 0, (hipStream_t)stream, nf, node,
                     faces, area);
  RB_CHECK_LAUNCH();
  return RB_OK;
}

extern "C" int rb_node_volume(int64_t nn, const double* evol, const int32_t* tinc_ptr, const int32_t* tinc,
                              double* nvol, void* stream) {
  if (nn <= 0 || !evol || !tinc_ptr || !tinc || !nvol) return RB_ERR_ARG;
  hipLaunchKernelGGL(k_node_volume, dim3((unsigned)ceil_div(nn, 256)), dim3(256), 0, (hipStream_t)stream, nn, evol,
                     tinc_ptr, tinc, nvol);
  RB_CHECK_LAUNCH();
  return RB_OK;
}
