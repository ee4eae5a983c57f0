/*
 * rbhip -- C ABI of the MI355X (gfx950) FEM hot path behind redbirdpy's API.
 *
 * Conventions
 *  - Every function returns an int status: 0 ok, < 0 error (message from
 *    rb_last_error(), thread-local).  No C++ exception crosses the ABI.
 *  - Pointers are device pointers unless a parameter is documented as host.
 *    All buffers are owned by the caller; indices are int32 and 0-based.
 *  - `stream` is a hipStream_t passed as void*; work is stream-ordered.
 *  - Real values are float64 / float32; complex values are interleaved
 *    (re, im) pairs of float64 (complex128) or float32 (complex64).
 *
 * The reference (fangq/redbirdpy v0.2.0) is pure Python with no FFI; the
 * functions below replace the compute inside the Python functions cited on
 * each declaration.  The Python host mirror (redbirdpy_amd) keeps the
 * reference's signatures and calls these through ctypes.
 */
#ifndef RBHIP_H
#define RBHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RB_ABI_VERSION 9

/* property layouts for rb_assemble */
#define RB_PROP_ELEM 0   /* one row per element (label tables are expanded by the caller) */
#define RB_PROP_NODE 1   /* one row per node */
#define RB_PROP_VOLUME 2 /* no optics: every tet pair gets V_e (the Jacobian patch weights) */

/* precision for rb_krylov_solve */
#define RB_PREC_FP64 0  /* all arithmetic in float64 / complex128 */
#define RB_PREC_MIXED 1 /* float32 / complex64 inner solves + float64 residual refinement */

/* preconditioner of the CG/COCG solves (rb_krylov_set_precond) */
#define RB_PRECOND_JACOBI 0 /* z = D^-1 r */
#define RB_PRECOND_POLY 1   /* z = p(D^-1 A) D^-1 r, p of degree 1 positive on a Gershgorin bound of the spectrum */
#define RB_PRECOND_POLY2 2  /* p of degree 2 where a fused LDS product applies, else degree 1 */
#define RB_PRECOND_POLY3 3  /* p of degree 3 where a fused LDS product applies, else degree 1 */

/* dtype codes */
#define RB_F64 0
#define RB_C128 1
#define RB_F32 2
#define RB_C64 3

int rb_version(void);
const char* rb_last_error(void);
int rb_device_count(int* n /* host */);

/* Zero `bytes` bytes with a fill kernel (never hipMemset). */
int rb_zero(void* p, size_t bytes, void* stream);

/*
 * Node -> element incidence (replaces the per-node loops of
 * redbirdpy utility.py:1005-1041 `_nodevolume` / `_femnz`).
 * el: ne x nv (0-based); inc_ptr: nn+1; inc_list: ne*nv entries e*nv+local,
 * ascending within each node.
 */
int rb_incidence(const int32_t* el, int64_t ne, int nv, int64_t nn, int32_t* inc_ptr, int32_t* inc_list,
                 void* stream);

/*
 * CSR pattern of the FEM matrix (replaces the COO->CSR conversion of
 * redbirdpy forward.py:249-254 and `_femnz`, utility.py:1020-1041):
 * row n lists, ascending, every node sharing a tet with n.  Call with
 * col_idx == NULL to get *nnz_out (host) only.  diag_slot (nn, optional)
 * receives the slot of each diagonal entry.
 */
int rb_csr_pattern(const int32_t* el, int64_t ne, int64_t nn, const int32_t* inc_ptr, const int32_t* inc_list,
                   int32_t* row_ptr, int32_t* col_idx, int64_t col_cap, int64_t* nnz_out /* host */,
                   int32_t* diag_slot, void* stream);

/*
 * Mesh preparation on the GPU (replaces redbirdpy utility.py meshprep's
 * host steps for one mesh or a batch of meshes concatenated with node
 * offsets).  All indices 0-based.
 * rb_tet_prep: per tet, local nodes 0 and 1 swapped when the signed volume
 *   is negative (_meshreorient_fallback, utility.py:1047-1063), evol = |det|/6
 *   (_elemvolume_fallback, :1090-1109) and deldotdel (ne x 10, :53-98).
 * rb_boundary_faces: faces that lie in exactly one tet, oriented by the local
 *   faces (0,2,1),(0,1,3),(0,3,2),(1,2,3), in lexicographic order of their
 *   sorted vertex triples (_volface_fallback, :1066-1087); tinc from
 *   rb_incidence (nv 4).  faces == NULL: counts into face_ptr (nn + 1,
 *   exclusive scan) and *nf_out (host); then call again with faces (nf x 3).
 *   RB_ERR_SIZE when a node owns more than 32 boundary faces.
 * rb_face_area: half the cross-product norm per triangle (:1090-1109).
 * rb_node_volume: a quarter of the incident tets' volumes (_nodevolume, :1005-1017).
 */
int rb_tet_prep(int64_t nn, int64_t ne, const double* node, const int32_t* el_in, int32_t* el_out, double* evol,
                double* deldotdel, void* stream);
int rb_boundary_faces(int64_t nn, const int32_t* el, const int32_t* tinc_ptr, const int32_t* tinc, int32_t* face_ptr,
                      int32_t* faces, int64_t* nf_out /* host */, void* stream);
int rb_face_area(int64_t nf, const double* node, const int32_t* faces, double* area, void* stream);
int rb_node_volume(int64_t nn, const double* evol, const int32_t* tinc_ptr, const int32_t* tinc, double* nvol,
                   void* stream);

/*
 * Assembly slot map: for each CSR slot, the ascending list of contributing
 * (tet, i, j) / (face, i, j) codes; fixes the summation order of
 * rb_assemble.  tinc/finc are the tet (nv=4) and face (nv=3) incidences.
 */
int rb_slot_map(const int32_t* el, int64_t ne, const int32_t* fc, int64_t nf, int64_t nn, const int32_t* tinc_ptr,
                const int32_t* tinc_list, const int32_t* finc_ptr, const int32_t* finc_list, const int32_t* row_ptr,
                const int32_t* col_idx, int64_t nnz, int32_t* slot_ptr, int32_t* codes, void* stream);

/*
 * FEM system matrix values for nsys systems sharing the pattern
 * (redbirdpy forward.py:90-256 `femlhs`, element branch :156-176, node
 * branch :177-228, Robin boundary :230-246).  mua/musp/nref: nsys x plen
 * (musp already mus*(1-g)); nref rows have length nlen (1 = one value per
 * system).  omega, bc: nsys each, bc = (1-Reff)/(12(1+Reff)).  out64 (and
 * optional out32) receive nsys x nnz values, complex when is_complex.
 */
int rb_assemble(const int32_t* slot_ptr, const int32_t* codes, int64_t nnz, const int32_t* el, int64_t ne,
                const int32_t* fc, int64_t nf, int64_t nn, const double* deldotdel, const double* evol,
                const double* area, int nsys, int kind, const double* mua, const double* musp, const double* nref,
                int64_t plen, int64_t nlen, const double* omega, const double* bc, int is_complex, void* out64,
                void* out32, void* stream);

/*
 * rb_assemble in two passes (the same arguments and results; redbirdpy
 * forward.py:90-256 `femlhs`): one writes the properties of each (element,
 * system) into work (rb_assemble_workspace bytes); then each workgroup takes
 * the slots that start in its bins of 1024 codes, its threads stage the
 * codes' contributions in LDS and one thread per slot sums its slot's
 * codes in the slot map's order (every slot holds at least one code, as in a
 * pattern built from the tets).  RB_PROP_VOLUME runs rb_assemble.  Values
 * agree with rb_assemble to rounding (the P1 mass weights are split into a
 * per-code term and the slot's nodal coefficient times its summed volume).
 */
int rb_assemble_workspace(int64_t nn, int64_t ne, int nsys, size_t* bytes);
/* the workgroup bins of a slot map (once per mesh): rb_assemble_bin_count
 * entries, bins[b] = first slot whose first code is at or after 256 b */
int rb_assemble_bin_count(int64_t ne, int64_t nf, int64_t* nbins);
int rb_assemble_bins(const int32_t* slot_ptr, int64_t nnz, int64_t ne, int64_t nf, int32_t* bins, void* stream);
int rb_assemble_ws(const int32_t* slot_ptr, const int32_t* codes, int64_t nnz, const int32_t* el, int64_t ne,
                   const int32_t* fc, int64_t nf, int64_t nn, const double* deldotdel, const double* evol,
                   const double* area, int nsys, int kind, const double* mua, const double* musp, const double* nref,
                   int64_t plen, int64_t nlen, const double* omega, const double* bc, int is_complex, void* out64,
                   void* out32, const int32_t* bins, void* work, size_t work_bytes, void* stream);

/* wide-field sources (rb_widefield) */
#define RB_WF_SINK 0    /* collimated: the cut of the mesh by the sink plane */
#define RB_WF_FACES 1   /* the boundary faces facing the source */
#define RB_WF_UNIFORM 0 /* planar source: one uniform pattern */
#define RB_WF_FOURIER 1 /* kx * ky Fourier patterns */
#define RB_WF_IMAGE 2   /* patterns from a pdim0 x pdim1 x npat image */
#define RB_WF_FRAME 34  /* doubles in the frame array */

/*
 * Wide-field source or detector: nodal boundary-flux columns of one planar,
 * pattern or Fourier source (replaces redbirdpy utility.py:423-577
 * `_process_planar_source`, with the iso2mesh plane cut `qmeshcut` at
 * :454-479, and :580-634 `_apply_bc_weighting` up to its per-pattern L1
 * normalisation, which the caller applies).  All arrays are device arrays
 * except `frame` (host, RB_WF_FRAME doubles):
 *   [0:3] source direction, [3:6] e1, [6:9] e2 (rows 1, 2 of a rotation
 *   taking the direction to +z), [9:14] / [14:19] x / y of the closed source
 *   rectangle in (e1, e2) coordinates, [19:21] its first corner, [21:23] /
 *   [23:25] unit directions of its two edges, [25] / [26] their lengths,
 *   [27:31] sink plane (a, b, c, e: a x + b y + c z + e = 0),
 *   [31] (1 - Reff) / (18 (1 + Reff)), [32] Fourier depth M, [33] phase.
 * node nn x 3; el ne x 4 with node-to-tet incidence tinc_ptr/tinc (RB_WF_SINK);
 * fc nf x 3, area nf with node-to-face incidence finc_ptr/finc (RB_WF_FACES);
 * img pdim0 x pdim1 x npat (RB_WF_IMAGE, else NULL); rhs nn x npat (written
 * whole); counts[2] = triangles whose centroid lies in the rectangle, and
 * those of them facing the source (the reference raises when either is 0).
 */
int rb_widefield(int64_t nn, const double* node, const int32_t* el, int64_t ne, const int32_t* tinc_ptr,
                 const int32_t* tinc, const int32_t* fc, int64_t nf, const double* area, const int32_t* finc_ptr,
                 const int32_t* finc, int mode, const double* frame, int kind, int npat, int kx, int ky,
                 const double* img, int pdim0, int pdim1, double* rhs, unsigned long long* counts, void* stream);

/*
 * Point location (iso2mesh.tsearchn as called by redbirdpy
 * forward.py:316, :337): lowest-index tet with all barycentric
 * coordinates >= -tol.  node: nn x 3, pts: npts x 3; loc: npts (INT32_MAX
 * when outside), bary: npts x 4 (NaN when outside).
 */
int rb_tsearch(const double* node, const int32_t* el, int64_t ne, const double* pts, int npts, double tol,
               int32_t* loc, double* bary, void* stream);

/*
 * Point-optode RHS columns (redbirdpy forward.py:314-348): B is a zeroed
 * dense n x ld block (row-major); column cols[p] gets bary[p] at the four
 * nodes of tet loc[p].
 */
int rb_rhs_points(const int32_t* el, const int32_t* loc, const double* bary, int npts, const int32_t* cols,
                  int64_t ld, int is_complex, double* B, void* stream);

/*
 * Multi-RHS Krylov solve (replaces redbirdpy solver.py:385-588 `femsolve`,
 * whose default here is SuperLU): Jacobi CG (real) / COCG (complex
 * symmetric) per column over nsys systems x k columns, B and X row-major
 * n x (nsys*k).  precision: RB_PREC_FP64 or RB_PREC_MIXED (needs vals32).
 * Stops column c when ||b_c - A x_c|| <= tol ||b_c||; zero columns give 0.
 * Host outputs: iterations, max relative residual, flag (0 converged,
 * 1 max iterations, 4 breakdown).  work: rb_krylov_workspace() bytes.
 */
int rb_krylov_workspace(int64_t n, int nsys, int k, int is_complex, int precision, size_t* bytes /* host */);
int rb_krylov_solve(int64_t n, int64_t nnz, const int32_t* row_ptr, const int32_t* col_idx,
                    const int32_t* diag_slot, int nsys, int k, int is_complex, const void* vals64,
                    const void* vals32, const void* B, void* X, int precision, double tol, int maxiter, int use_x0,
                    int check_every, void* work, size_t work_bytes, int* iters_out /* host */,
                    double* relres_out /* host */, int* flag_out /* host */, void* stream);

/*
 * rb_krylov_solve without a workspace argument and without any allocation
 * (no hipMalloc/hipFree, no pinned or pageable host buffers, no events).
 * No counterpart in the reference: it is for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Its workspace is a fixed per-device array of
 * RB_NOALLOC_SCRATCH_BYTES bytes inside the library (held by every device
 * the library runs on until the process ends); a solve whose
 * rb_krylov_workspace() size exceeds it fails with RB_ERR_SIZE (-3).  All
 * calls on one device share that array: calls that overlap on one device
 * race.  `stream` must belong to the current device.  Arguments, results and
 * flags are those of rb_krylov_solve (same kernels, same bits).
 */
#define RB_NOALLOC_SCRATCH_BYTES (64u << 20)
int rb_krylov_solve_noalloc(int64_t n, int64_t nnz, const int32_t* row_ptr, const int32_t* col_idx,
                            const int32_t* diag_slot, int nsys, int k, int is_complex, const void* vals64,
                            const void* vals32, const void* B, void* X, int precision, double tol, int maxiter,
                            int use_x0, int check_every, int* iters_out /* host */, double* relres_out /* host */,
                            int* flag_out /* host */, void* stream);

/*
 * LDS-staged SpMM plan (replaces the per-entry row gathers of the
 * femsolve SpMM, solver.py:385-588, for 512-byte rows: 64 complex64 or
 * float64 columns).  Consecutive CSR rows are grouped into blocks of at most
 * RB_LDS_RMAX rows whose neighbour rows (the union of their columns, at most
 * umax) are staged in LDS once per block, with at most recmax records (the
 * kernel's limits: umax = RB_LDS_UMAX, recmax = RB_LDS_RECMAX; others are
 * refused with RB_ERR_ARG); every row's entries become
 * records padded to a multiple of 8 (a record: the LDS byte offset of its
 * column's row and its matrix value; padding records have value 0).  The plan
 * depends on the pattern only; `recv` is scratch the solve fills with the
 * record values of the matrix being solved.
 *
 * rb_spmm_plan_count / rb_spmm_plan_fill run on the HOST over host copies of
 * row_ptr / col_idx (device row order).  The caller allocates (device):
 *   brow, bu: nblk + 1;  ulist: nunion + RB_LDS_UPAD;
 *   rinfo: 2 * (n + 1) + RB_LDS_RPAD;  rmap, loff, recv: nrec + RB_LDS_RECPAD
 * (host buffers of the same sizes for rb_spmm_plan_fill, which also writes
 * the padding).  RB_ERR_SIZE: some row alone exceeds the block limits (the
 * solve then keeps its gather SpMM).
 */
#define RB_LDS_UMAX 128 /* block limits of the kernel */
#define RB_LDS_RECMAX 768
#define RB_LDS_RMAX 64
#define RB_LDS_UPAD 256
#define RB_LDS_RPAD 512
#define RB_LDS_RECPAD 1024
typedef struct rb_spmm_plan {
  int64_t nblk;         /* blocks of consecutive rows */
  int64_t nrec;         /* entry records (padded rows) */
  int32_t umax;         /* the block limits the plan was built with (selects the kernel) */
  int32_t recmax;
  const int32_t* brow;  /* nblk + 1: first row of each block */
  const int32_t* bu;    /* nblk + 1: start of each block's union in ulist */
  const int32_t* ulist; /* union rows of every block, ascending per block */
  const int32_t* rinfo; /* per row (first record, LDS byte offset of the row's own p), row n = (nrec, 0) */
  const int32_t* rmap;  /* record -> CSR slot, -1 for padding */
  const int32_t* loff;  /* record -> LDS byte offset of its column's row within the block */
  void* recv;           /* record values, written by each solve (8 bytes per record) */
  int64_t recv_sets;    /* record value sets recv holds, nrec + RB_LDS_RECPAD records each: for one
                           system, A and (second, used by the polynomial preconditioner when
                           present) A D^-1; one per system pair for real float32 systems */
} rb_spmm_plan;
int rb_spmm_plan_count(int64_t n, const int32_t* row_ptr /* host */, const int32_t* col_idx /* host */, int umax,
                       int recmax, int64_t* nblk /* host */, int64_t* nunion /* host */, int64_t* nrec /* host */);
int rb_spmm_plan_fill(int64_t n, const int32_t* row_ptr /* host */, const int32_t* col_idx /* host */, int umax,
                      int recmax, int32_t* brow, int32_t* bu, int32_t* ulist, int32_t* rinfo, int32_t* rmap,
                      int32_t* loff /* all host */);

/*
 * rb_krylov_solve with an LDS-staged SpMM plan (rb_spmm_plan, device
 * pointers) for the 64-column solves of one system: mixed precision with
 * complex64 or float32 inner solves, or plain float64 real systems; and for
 * mixed solves of an even number of real 64-column systems (one launch per
 * system pair, plan->recv_sets >= nsys / 2; >= nsys for the polynomial
 * preconditioner's pair product).  Other shapes (and plan == NULL) run
 * exactly as rb_krylov_solve.  Same arguments, workspace, results and flags.
 */
int rb_krylov_solve_planned(int64_t n, int64_t nnz, const int32_t* row_ptr, const int32_t* col_idx,
                            const int32_t* diag_slot, int nsys, int k, int is_complex, const void* vals64,
                            const void* vals32, const void* B, void* X, int precision, double tol, int maxiter,
                            int use_x0, int check_every, void* work, size_t work_bytes, const rb_spmm_plan* plan,
                            int* iters_out /* host */, double* relres_out /* host */, int* flag_out /* host */,
                            void* stream);

/*
 * The column sweep's grid (the q = A p launches of rb_krylov_solve and
 * rb_krylov_solve_planned when every 64-column chunk lies in one system):
 * at most cap workgroups (rounded down to a multiple of 8, the accelerator
 * dies), 0 = one resident round of the device (the default).  Process-wide;
 * the autotuner sets it around the solves it has measured.  No counterpart
 * in the reference, whose CPU solvers split right-hand sides with
 * rhsblock / nthread (solver.py:600-639).
 */
int rb_krylov_set_sweep_grid(int64_t cap);
int64_t rb_krylov_get_sweep_grid(void);

/*
 * The preconditioner of every later CG/COCG solve (RB_PRECOND_*), process-wide
 * like the sweep grid.  A solve reads the setting once when it starts;
 * rb_krylov_workspace sizes the workspace for the one set when it is called
 * (the polynomial ones take one more vector of the solve's width, degree 3
 * two).  RB_PRECOND_POLY applies z = p(D^-1 A) D^-1 r with
 * p(x) = (3 / L)(1 - 0.8 x / L), L the Gershgorin bound of D^-1 A per system:
 * one more product per iteration for ~1.8x fewer iterations on 3-D FEM
 * systems.  RB_PRECOND_POLY2 / POLY3: p of degree 2 / 3 in t = x / L with
 * 1 - t p(t) the Chebyshev polynomial of degree 3 / 4 smallest on [0.03, 1]
 * (13.54 - 33.78 t + 21.86 t^2; 20.45 - 91.18 t + 138.51 t^2 - 67.24 t^3,
 * over L), two / three more products for ~2.8x / ~3.7x fewer iterations, on
 * the fused LDS products; degree 1 elsewhere.  The fused products need the
 * plan's A D^-1 records: rb_spmm_plan.recv_sets >= 2 for one system, >= nsys
 * for float32 systems in pairs (without them, and for float64 systems of 64
 * columns each, the pairs run in the one-column layout).  No counterpart in
 * the reference (its iterative solvers take a caller's M, pyamg's smoothed
 * aggregation for 'cg+amg' / 'minres+amg', solver.py:409-414 and :697-700,
 * or none); the autotuner chooses per shape.
 */
int rb_krylov_set_precond(int kind);
int rb_krylov_get_precond(void);

/*
 * Per-kernel timing of the Krylov loop: when enabled, hipEvents are
 * recorded on the launch stream around every launch of the kernels below;
 * rb_prof_read returns the summed milliseconds and the launch count.
 */
#define RB_PROF_SPMM 0    /* q = A p with the p.q reduction */
#define RB_PROF_UPDATE 1  /* x, r update with the r.z and r.r reductions */
#define RB_PROF_PUPDATE 2 /* p = z + beta p */
#define RB_PROF_RESID 3   /* float64 residual of the mixed-precision refinement */
#define RB_PROF_INIT 4    /* initial residual */
#define RB_PROF_POLY 5    /* z = p(D^-1 A) D^-1 r of the polynomial preconditioner, r.z and r.r */
#define RB_PROF_NKERNELS 6
int rb_prof_enable(int on);
int rb_prof_reset(void);
int rb_prof_read(int id, double* total_ms /* host */, int64_t* count /* host */);

/* Y = A_s X for each system s (dtype RB_F64/RB_C128/RB_F32/RB_C64). */
int rb_spmm(int64_t n, int64_t nnz, const int32_t* row_ptr, const int32_t* col_idx, int nsys, int k, int dtype,
            const void* vals, const void* X, void* Y, void* stream);

/*
 * Detector readout (redbirdpy forward.py:365-442 `femgetdet`):
 * out[d][s] = sum over the sparse detector column d (det_ptr/det_row/
 * det_val, CSC, real weights) of w * phi[row][src0 + s].  phi is n x ldphi.
 */
int rb_getdet(int ndet, const int32_t* det_ptr, const int32_t* det_row, const double* det_val, const void* phi,
              int64_t ldphi, int src0, int nsrc, int is_complex, void* out, void* stream);

/*
 * The node Jacobian's patch matrices, once per mesh (redbirdpy forward.py:562-567
 * sums Jelem over each node's tets on every call; Q_n depends on the mesh
 * only): Qc[qptr[n] + i*np + j] = Q_n[i][j] over the CSR row of n
 * (np = row length, qptr = exclusive prefix sum of np^2).
 */
int rb_jac_patch_q(int64_t nn, const int32_t* el, const double* evol, const int32_t* inc_ptr,
                   const int32_t* inc_list, const int32_t* row_ptr, const int32_t* col_idx, int maxnp,
                   const int64_t* qptr, double* Qc, void* stream);

/*
 * mua Jacobian on a source x detector grid (redbirdpy forward.py:509-569
 * `jac`, numpy branch :530-567).  For source column scol[s] and detector
 * column dcol[d] of phi (n x ldphi), cell = d*Ns + s, output row
 * p = pmap ? pmap[cell] : cell (pmap < 0 skips inactive pairs, as the
 * reference drops sd rows with active == 0):
 *   Jnode[n][p] = -0.0125 * Phi_S^T Q_n Phi_D,  Q_n = sum_{e ni n} V_e (I + 1 1^T)_e
 * which equals 0.25 * sum_{e ni n} Jelem[p][e] with the reference's
 * Jelem = -0.05 V_e (sum_i ps_i pd_i + S_s S_d).  JT: nn x ldj node-major;
 * out_dtype RB_F64/RB_F32 (real phi) or RB_C128/RB_C64 (complex phi).
 * Optional Jelem: ne x ldj element-major.  maxnp = longest CSR row.
 * qptr/Qc (both NULL or both given): the patch matrices Q_n from
 * rb_jac_patch_q, read instead of rebuilt from the incident tets.
 */
int rb_jac_grid(int64_t nn, int64_t ne, const int32_t* el, const double* evol, const int32_t* inc_ptr,
                const int32_t* inc_list, const int32_t* row_ptr, const int32_t* col_idx, int maxnp,
                const void* phi, int64_t ldphi, int is_complex, int Ns, int Nd, const int32_t* scol,
                const int32_t* dcol, const int32_t* pmap, void* JT, int64_t ldj, int out_dtype, void* Jelem,
                const int64_t* qptr, const double* Qc, void* stream);

/*
 * Jacobian contraction g[n] = sum_{d,s} w[d*Ns+s] Jnode[n][(s,d)] without
 * forming J: the backward of the forward map (transpose of redbirdpy
 * `jac`, forward.py:509-569).  psi_work: nn x Ns, h_work: ne scalars of the
 * phi type; g: nn (float64 or complex128).
 */
int rb_jac_contract_grid(int64_t nn, int64_t ne, const int32_t* el, const double* evol, const int32_t* inc_ptr,
                         const int32_t* inc_list, const void* phi, int64_t ldphi, int is_complex, int Ns, int Nd,
                         const int32_t* scol, const int32_t* dcol, const void* w, void* psi_work, void* h_work,
                         void* g, void* stream);

/*
 * Gauss-Newton Gram G = diag(d) (J^T)^T J^T diag(d) for JT: K x m (row-major,
 * redbirdpy recon.py:474 `Amat @ Amat.T`), G: m x m float64.
 * mode: 0 float64 FMA, 1 float32 MFMA, 2 bf16 MFMA (split in 3 products:
 * hi*hi + hi*lo + lo*hi), each with float32 accumulation per K-tile and
 * float64 accumulation across tiles.  in_dtype RB_F64 or RB_F32.
 * d (m, optional): row scaling applied before the products.
 */
int rb_gram(int64_t K, int64_t m, const void* JT, int64_t ldj, int in_dtype, const double* d, int mode, double* G,
            void* stream);

/*
 * Gram on the matrix cores: mode 1 = bf16 (hi*hi), mode 3 = bf16x3 split
 * (hi*hi + hi*lo + lo*hi, ~16 mantissa bits per product).  JT is split once
 * into transposed bf16 hi/lo arrays; fp32 MFMA accumulation within K chunks
 * whose partial tiles are summed in fp64 in a fixed order.  colsq (m
 * doubles, may be NULL) receives the exact float64 sum_k JT[k][c]^2 from the
 * same read of JT (undivided by d).  work: rb_gram_workspace() bytes.  Same
 * result as rb_gram (redbirdpy recon.py:474).
 */
int rb_gram_workspace(int64_t K, int64_t m, int mode, size_t* bytes /* host */);
int rb_gram_mfma(int64_t K, int64_t m, const void* JT, int64_t ldj, int in_dtype, const double* d, int mode,
                 double* G, double* colsq, void* work, size_t work_bytes, void* stream);

/*
 * Dual-mesh Jacobian (redbirdpy recon.py:818-876 `_remap_jacobian`, applied
 * at :200-205): Y[r][:M] = sum_k w[k] X[idx[k]][:M], k in [ptr[r], ptr[r+1])
 * taken in order.  X: node-major forward-mesh Jacobian (row stride ldx),
 * Y: nrows recon-mesh rows (row stride ldy); ptr/idx/w: the transpose of the
 * interpolation (mapid, mapweight) as CSR over recon nodes, forward nodes
 * ascending within a row.  Float64.
 */
int rb_jt_remap(int64_t nrows, const int32_t* ptr, const int32_t* idx, const double* w, int64_t M, const double* X,
                int64_t ldx, double* Y, int64_t ldy, void* stream);

/*
 * Multi-RHS BiCGSTAB with right Jacobi preconditioning, float64 or
 * complex128 (redbirdpy solver.py:791-824, method 'bicgstab'): for
 * matrices that are not (complex) symmetric, and as the fallback when a
 * CG/COCG column breaks down.  One system, ncol <= 512 columns of B/X
 * (n x ncol row-major); x0 from X when use_x0.  Conjugated inner products;
 * a column stops at |b - A x| <= tol |b|.  work: rb_bicgstab_workspace()
 * bytes.  iters/relres/flag as rb_krylov_solve (flag 4: breakdown).
 */
int rb_bicgstab_workspace(int64_t n, int ncol, int is_complex, size_t* bytes /* host */);
int rb_bicgstab_solve(int64_t n, int64_t nnz, const int32_t* row_ptr, const int32_t* col_idx, const int32_t* diag,
                      int ncol, int is_complex, const void* vals, const void* B, void* X, double tol, int maxiter,
                      int use_x0, int check_every, void* work, size_t work_bytes, int* iters /* host */,
                      double* relres /* host */, int* flag /* host */, void* stream);

/*
 * Jacobian matvecs with float64 accumulation over a node-major Jacobian
 * JT (K x M, columns in nw blocks of np pairs; redbirdpy recon.py:501
 * `Amat.T @ y` and the J x products of the refinement):
 *   rb_jt_rows: t[k][w] = sum_p JT[k][w*np+p] v[w*np+p]    (t: K x nw)
 *   rb_jt_cols: u[j] = sum_k JT[k][j] z[k][j/np] (mode 0, z: K x nw)
 *               u[j] = sum_k JT[k][j]^2          (mode 1, the Gram diagonal)
 * rb_jt_cols needs rb_jt_cols_workspace() bytes (per-block partials summed in
 * a fixed order).  Both need nw >= 1, np >= 1, nw*np <= M <= ldj; mode 0 of
 * rb_jt_cols needs M == nw*np exactly (RB_ERR_ARG otherwise).
 */
int rb_jt_rows(int64_t K, int64_t M, const void* JT, int64_t ldj, int in_dtype, int nw, int64_t np,
               const double* v, double* t, void* stream);
int rb_jt_cols_workspace(int64_t K, int64_t M, size_t* bytes /* host */);
int rb_jt_cols(int64_t K, int64_t M, const void* JT, int64_t ldj, int in_dtype, int nw, int64_t np,
               const double* z, int mode, double* u, void* work, size_t work_bytes, void* stream);

/*
 * The regularised normal-equation operator applied without forming it:
 *   out = s2 * J W (J^T v) + lam * v,  (J^T v)[k][b] = sum_p JT[k][b*np+p] v[b*np+p]
 * over a float64 node-major JT (K x M, M = nb * np), W an nb x nb float64
 * block-mixing matrix (row-major).  redbirdpy recon.py:429-509 forms
 * J J^T + lam I explicitly and solves it; this product drives iterative
 * refinement of that solve.  One read of JT (rb_jt_rows then rb_jt_cols read
 * it twice).  Needs np a multiple of 64, nb <= 16, M <= 8192 (RB_ERR_ARG
 * otherwise) and rb_jt_hv_workspace() bytes; deterministic (fixed-order sums).
 */
/*
 * In-place Cholesky factorisation of a dense float64 SPD matrix (n x n,
 * row-major, lda >= n, n a multiple of 64): on return the lower triangle
 * holds L (A = L L^T), the upper triangle of each diagonal 64 x 64 block is
 * zero and the blocks above them keep A.  *info (device int) is 0, or the
 * 1-based column at which a non-positive pivot appeared.  Replaces the dense
 * solve of the Gauss-Newton normal equations in redbirdpy recon.py:429-509
 * (reginvunder: `np.linalg.solve(Amat @ Amat.T + lambda I, rhs)`).
 * Needs rb_potrf_workspace() bytes.
 */
int rb_potrf_workspace(int64_t n, size_t* bytes /* host */);
int rb_potrf(int64_t n, double* A, int64_t lda, void* work, size_t work_bytes, int* info, void* stream);

int rb_jt_hv_workspace(int64_t K, int64_t M, size_t* bytes /* host */);
int rb_jt_hv(int64_t K, int64_t M, const double* JT, int64_t ldj, int nb, int64_t np, const double* W,
             const double* v, double s2, double lam, double* out, void* work, size_t work_bytes, void* stream);

#ifdef __cplusplus
}
#endif
#endif
