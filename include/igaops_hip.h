/*
 * libigaops_hip — C ABI of the MI355X (gfx950) implementation of igaops'
 * weighted-quadrature matrix-free operator path.
 *
 * Conventions
 *   - Every function returns IGA_OK (0) or a negative IGA_ERR_* code; the text of
 *     the last failure of the calling thread is available from iga_last_error().
 *     No C++ exception crosses this boundary.
 *   - Pointers named x, y, coef, V, w, ... are device pointers (HBM) unless the
 *     parameter is documented as host.  `stream` is a hipStream_t (NULL = default).
 *   - Tensor-product arrays use direction-1-fastest order: the DOF a = (a1,a2,a3)
 *     lives at a1 + n1*(a2 + n2*a3); with r right-hand sides, element (a, k) lives
 *     at a*r + k (C order of a (N, r) array), as in the reference.
 *   - dtype selects the arithmetic: IGA_F64 (double) or IGA_F32 (float); all
 *     vectors and coefficients of one call share it.
 *
 * Reference citations are into JoaquinCORNEJO/igaops, src/igaops/.
 */
#ifndef IGAOPS_HIP_H
#define IGAOPS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IGA_OK 0
#define IGA_ERR_INVALID -1     /* bad argument / shape (Python: ValueError)        */
#define IGA_ERR_HIP -2         /* HIP runtime failure (Python: RuntimeError)        */
#define IGA_ERR_NOMEM -3       /* allocation failure (Python: MemoryError)          */
#define IGA_ERR_UNSUPPORTED -4 /* configuration not provided (NotImplementedError)  */
#define IGA_ERR_INTERNAL -5    /* unexpected failure (RuntimeError)                 */

#define IGA_F64 0
#define IGA_F32 1
#define IGA_BF16 2 /* FD plans only: float vectors, BF16 matrix-core inputs, FP32 accumulation */

/* ---- library ------------------------------------------------------------- */

/* Library version (major*10000 + minor*100 + patch). */
int iga_version(void);

/* Copies the last error message of this thread into buf (NUL terminated). */
int iga_last_error(char* buf, int buflen);

/* Number of visible HIP devices. */
int iga_device_count(int* count);

/* ---- banded 1-D factors and L-mode products -------------------------------
 * A band holds nmat matrices of shape (nrows x ncols) that share one sparsity
 * band: row i has its possibly-non-zero entries in columns
 * [first[i], first[i] + width).  vals (host, float64) is [nmat][nrows][width].
 * This is the device form of the reference's CSR basis / weight matrices
 * (quadrature/utils/operations.py:52-105, weighted_quad.py:353-363) and of the
 * dense FD eigenvector matrices (width = ncols).
 */
typedef struct iga_band_s* iga_band_t;

int iga_band_create(iga_band_t* out, int nrows, int ncols, int width, int nmat,
                    const int32_t* first, const double* vals);
int iga_band_destroy(iga_band_t band);

/* One n-mode product of a tensor viewed as (outer, ncols, inner):
 *   y_m(o, i, t) = alpha * sum_j V_m[i, j] * x(o, first[i]+j, t)   [+ y_m(o,i,t)]
 * for m < nmat, output viewed as (outer, nrows, inner).  Replaces
 * MatrixFree.tensor_matrix_product (operators/utils/matrix_free.py:39-75); the
 * reference's MatrixFree.apply (matrix_free.py:77-119) is d such calls. */
int iga_band_mode_product(iga_band_t band, int dtype, int64_t outer, int64_t inner,
                          const void* x, void* const* y, int accumulate, double alpha,
                          void* stream);

/* Point-wise coefficient product at the quadrature points with r right-hand
 * sides:  y[q*r + k] = alpha * coef[q] * x[q*r + k]  [+ y[q*r + k]].
 * Replaces the broadcast einsums of operators/bsplines_ops.py:311,341-343,378,416-418. */
int iga_qp_scale(int dtype, int64_t nq, int64_t nrhs, const void* coef, const void* x,
                 void* y, int accumulate, double alpha, void* stream);

/* Dirichlet masks (solvers/linear_solver.py:119-124): zero every entry outside a
 * box (iga_box_mask), or the entries listed in idx (device int64, iga_index_zero). */
int iga_box_mask(int dtype, int ndim, const int64_t* n, const int64_t* lo, const int64_t* len,
                 int64_t nrhs, void* full, void* stream);
int iga_index_zero(int dtype, int64_t nidx, const int64_t* idx, int64_t nrhs, void* x, void* stream);

/* ---- Krylov reductions (FP64 or FP32 vectors, FP64 results) ----------------
 * Deterministic two-stage wavefront-shuffle reductions (fixed chunks, fixed
 * combination order).  `work` is caller-owned device scratch of
 * iga_reduce_workspace_bytes(n, k) bytes.  iga_dot_many takes k <= 16384.  They replace the
 * numpy dot / norm / axpy loops of solvers/linear_solver.py:481-495 and
 * solvers/deflated/helpers.py:44-58,116-133. */
int64_t iga_reduce_workspace_bytes(int64_t n, int k);

/* out[j] = sum_i V[j*ldv + i] * w[i]  for j < k  (out: device, float64[k]) */
int iga_dot_many(int dtype, int64_t n, int k, const void* V, int64_t ldv, const void* w,
                 double* out, void* work, void* stream);

/* w[i] += alpha * sum_j h[j] * V[j*ldv + i]   (h: device, float64[k]) */
int iga_axpy_many(int dtype, int64_t n, int k, const void* V, int64_t ldv, const double* h,
                  double alpha, void* w, void* stream);

/* Complex (interleaved re, im; float64) versions for the recycled solver
 * (solvers/deflated/helpers.py:44-58,116-133; gcrodr.py:266-276,340-345):
 *   out[2j], out[2j+1] = re, im of sum_i conj(V_j[i]) w[i]   (work: bytes for 2k vectors)
 *   w[i] += alpha * sum_j h[j] V_j[i]                          (h: device complex[k]) */
int iga_zdot_many(int64_t n, int k, const double* V, int64_t ldv, const double* w, double* out,
                  void* work, void* stream);
int iga_zaxpy_many(int64_t n, int k, const double* V, int64_t ldv, const double* h, double alpha_re,
                   double alpha_im, double* w, void* stream);

/* One-pass block operations on sets of complex vectors given as device arrays of
 * row addresses (the tall products of gcrodr.py:266-276,288-292,327-363 and
 * helpers.py:240-253):
 *   iga_zgram:    out[a*nb + b] = sum_i conj(X_a[i]) Y_b[i]   (complex, 1 <= na, nb <= 64)
 *   iga_zcombine: Out_c = beta Out_c + alpha sum_j coef[j*nout + c] V_j   (1 <= nout <= 24, alpha real)
 * Rows are complex128 (interleaved re, im), each starting on a 16-byte boundary.
 * iga_zgram runs on the FP64 matrix cores (X^H Y as a real product over the 2n
 * interleaved doubles), or for at most 2 x 2 rows one complex element per thread,
 * with a fixed-order second stage: repeat calls give the same bits.  work for iga_zgram: iga_zgram_workspace_bytes(n, na, nb) bytes. */
int64_t iga_zgram_workspace_bytes(int64_t n, int na, int nb);
int iga_zgram(int64_t n, int na, const double* const* X, int nb, const double* const* Y, double* out,
              void* work, void* stream);
int iga_zcombine(int64_t n, int nv, const double* const* V, int nout, const double* coef, double alpha, double beta,
                 double* const* out, void* stream);

/* y[i] = a * x[i] + b * y[i]; a and b are read from device memory when
 * a_dev / b_dev are non-NULL (then a[0] / b[0] is used, optionally inverted
 * with invert_a), else the host values are used. */
int iga_axpby(int dtype, int64_t n, double a, const double* a_dev, int invert_a,
              const void* x, double b, void* y, void* stream);

/* Casts between float64 and float32 vectors: y = (dtype_y) x. */
int iga_cast(int dtype_x, int dtype_y, int64_t n, const void* x, void* y, void* stream);

/* ---- device-side GMRES Arnoldi step ------------------------------------------
 * Replaces the inner loop of _GMRES (solvers/linear_solver.py:473-508): the
 * orthogonalisation of the new Krylov vector (481-490, two classical Gram-Schmidt
 * passes there), the Hessenberg column and its least-squares residual, with the
 * least-squares problem kept in Givens-rotated form instead of lstsq (501; the
 * residual estimate |g| equals its residual norm).  The reorthogonalisation is
 * delayed by one step (DCGS2), so a step makes two passes over the basis and step
 * k finishes basis vector k-1 and the Hessenberg column k-2.
 * `ls` is a float64 device state of iga_gmres_state_doubles(m) values;
 * iga_gmres_layout gives the offsets of R (m x m column-major, the rotated
 * Hessenberg), g (m+1, the rotated beta e_1) and hist (2(m+1): g[d] and the
 * subdiagonal H[d, d-1] of Krylov dimension d, for the host to poll).
 * iga_gmres_start: beta = ||r||, g[0] = beta, V0 = r / beta (state cleared first).
 * iga_gmres_step (k = 1..m+1, m <= 2000): V[0..k-1) hold the final basis, V[k-1]
 * the tentative vector u and w = A' u; writes the final V[k-1], the next tentative
 * V[k], and for k >= 2 the column k-2 of R, g[k-1] and hist entry k-1.  V holds
 * k+1 rows of stride ldv.  `work` holds iga_reduce_workspace_bytes(n, m+1) bytes;
 * vectors are `dtype`, the state float64. */
int64_t iga_gmres_state_doubles(int m);
int iga_gmres_layout(int m, int64_t* offsets);  /* offsets[3] = R, g, hist */
int iga_gmres_start(int dtype, int64_t n, int m, const void* r, void* V0, double* ls, void* work, void* stream);
int iga_gmres_step(int dtype, int64_t n, int k, int m, void* V, int64_t ldv, void* w, double* ls, void* work,
                   void* stream);

/* ---- Sum-factorised sparse assembly ------------------------------------------
 * Replaces Operations.sumfact_assembling (operators/utils/operations.py:59-115) and
 * the sums of it in BsplineOperations / NurbsOperations.assemble_scalar_*
 * (operators/bsplines_ops.py:138-246, operators/nurbs_ops.py:182-287):
 *   A[i, j] = sum_t scale[t] sum_q prod_d V_{t,d}(i_d, j_d, q_d) coef[t][q]   (* w[i] w[j] if w)
 * with V_{t,d}(i, j, q) = W_{t,d}[i, q] B_{t,d}[q, j], on the Kronecker product of
 * the 1-D patterns.  Host tables per direction d: pat_ptr[d] (nr[d]+1) and pat_col[d]
 * (npairs[d]) list the pairs (i_d, j_d); lst_ptr[t*ndim+d] (npairs[d]+1) indexes,
 * for term t, each pair's entries of lst_q[d] (quadrature point q_d) and lst_v[d]
 * (value), nlist[d] entries each.  Device: coef[t] (prod nq values), w (optional,
 * square matrices), and the outputs row_ptr (N_r+1, int64), cols (int64) and vals
 * (float64) of iga_sumfact_nnz entries; rows in C order (direction 1 fastest),
 * columns ascending.  ndim <= 4, nterms <= 32; every table index is checked on the
 * host; the call returns once the matrix is written. */
int64_t iga_sumfact_nnz(int ndim, const int64_t* nr, const int32_t* const* pat_ptr);
int iga_sumfact_assemble(int ndim, int nterms, const int64_t* nr, const int64_t* nq, const int64_t* nc,
                         const int32_t* const* pat_ptr, const int32_t* const* pat_col, const int64_t* npairs,
                         const int32_t* const* lst_ptr, const int32_t* const* lst_q, const double* const* lst_v,
                         const int64_t* nlist, const double* const* coef, const double* scale,
                         int64_t* row_ptr, const double* w, int64_t* cols, double* vals, void* stream);

/* Gather / scatter of a box of a 3-D (or lower) tensor-product array (the free
 * nodes of the FD preconditioner, fastdiag/utils/space_precond.py:104-122,296,315):
 *   gather:  sub[s] = full[offset(s)]        (box lo[d]..lo[d]+len[d]-1, dims n[d])
 *   scatter: full[offset(s)] = sub[s]; entries of full outside the box are set to 0
 *            when zero_outside != 0.  ndim <= 4, r right-hand sides. */
int iga_box_gather(int dtype, int ndim, const int64_t* n, const int64_t* lo, const int64_t* len,
                   int64_t nrhs, const void* full, void* sub, void* stream);
int iga_box_scatter(int dtype, int ndim, const int64_t* n, const int64_t* lo, const int64_t* len,
                    int64_t nrhs, const void* sub, void* full, int zero_outside, void* stream);

/* ---- fused 3-D sum-factorised operator -------------------------------------
 * One plan per (three 1-D quadrature rules, coefficient form, dtype).  It replaces
 * the chain of MatrixFree.apply + einsum calls of
 * BsplineOperations.compute_mf_scalar_{u_v,gradu_gradv} (operators/bsplines_ops.py:290-355)
 * and NurbsOperations.compute_mf_scalar_{u_v,gradu_gradv} (operators/nurbs_ops.py:320-390)
 * for 3-D rules whose test weights depend only on the test derivative
 * (Gauss rules and WQ type "2": weights = [W0, W0, W1, W1]).
 *
 * Direction d is described by its band start first[d][q] (non-decreasing,
 * first + P <= nc[d]) and tabs[d] = float64 [4][nq[d]][P] holding, for each
 * quadrature point q and band slot j (DOF first[d][q] + j):
 *   B0[q][j] = basis[0][q, c], B1[q][j] = basis[1][q, c],
 *   W0[q][j] = weights[0][c, q], W1[q][j] = weights[2][c, q].
 * form: 0 mass (1 coefficient/point), 1 stiffness with symmetric C (6),
 *       2 stiffness + mass (7), 3 symmetric 4x4 (10: NURBS stiffness [+ mass]),
 *       4 general 4x4 (16), 5 the fields of 2 with the mass term on a second input
 *       (iga_fused_apply_sliced2), 6 first order with the gradient on the test
 *       function (4: compute_mf_scalar_gradu_v, operators/bsplines_ops.py:357-390),
 *       7 first order with the gradient on the trial function (4:
 *       compute_mf_scalar_u_gradv, bsplines_ops.py:392-429); for forms 6 and 7 the C of
 *       iga_fused_pack is the (3, nq) coefficient array c and m is unused (the NURBS
 *       form adds its value-value correction, operators/nurbs_ops.py:392-430).
 * nslab: number of q3 slabs (0 = automatic). */
typedef struct iga_fused_s* iga_fused_t;

int iga_fused_create(iga_fused_t* out, const int32_t* nc, const int32_t* nq, int P,
                     const int32_t* const* first, const double* const* tabs, int form, int dtype,
                     int nslab);
int iga_fused_destroy(iga_fused_t plan);
/* Right-hand sides per work-group of the sweep when several share one coefficient set
 * (an even nrhs; forms 1, 2 and 3): 2 reads each coefficient and runs each plane-phase
 * barrier once for two of them (twice the registers per thread), 1 (default) processes
 * them one by one.  The results agree to rounding (the compiler may contract the
 * per-point sums differently in the two variants). */
int iga_fused_set_rhs_block(iga_fused_t plan, int nr);

/* Sizes: packed coefficient length (elements), partial-block workspace per
 * right-hand side (elements), tile and slab counts. */
int iga_fused_info(iga_fused_t plan, int64_t* coef_elems, int64_t* part_elems_per_rhs,
                   int* ntiles1, int* ntiles2, int* nslabs);
int64_t iga_fused_workspace_bytes(iga_fused_t plan, int nrhs);

/* Packs coefficients given in the reference layout (float64 device arrays):
 * C (3,3,nq) [C[i][j] multiplies trial d_j and test d_i], m (nq), and for NURBS
 * the weight function W (nq) and its parametric gradient dW (3,nq) at the
 * quadrature points (both NULL for B-splines).  NURBS terms are folded once here
 * instead of on every call (operators/nurbs_ops.py:355-387). */
int iga_fused_pack(iga_fused_t plan, const double* C, const double* m, const double* W,
                   const double* dW, void* packed, void* stream);

/* y = alpha * A x [+ y] for nrhs right-hand sides ((N, nrhs) C order).  For NURBS
 * pass nurbs_w (per-DOF weights, plan dtype): input and output are scaled by it
 * as in operators/nurbs_ops.py:359-361,388-390.  work: device scratch of
 * iga_fused_workspace_bytes(plan, nrhs) bytes. */
int iga_fused_apply(iga_fused_t plan, const void* packed, const void* x, void* y, int nrhs,
                    const void* nurbs_w, double alpha, int accumulate, void* work, void* stream);

/* iga_fused_apply without the workspace argument, for callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops and harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  The partial blocks go to a built-in
 * device array of 64 MiB (67,108,864 bytes) that each GPU running the library holds
 * until the process ends; the call allocates nothing (no hipMalloc / hipFree, no
 * allocating library routine) and rejects with IGA_ERR_INVALID any call whose
 * iga_fused_workspace_bytes(plan, nrhs) exceeds 64 MiB.  Every call on one GPU shares
 * that GPU's array and leaves its last contents there: calls that overlap on one GPU
 * (other streams or host threads) race, so serialise them.  Results are those of
 * iga_fused_apply with the same arguments (same kernels, same summation order). */
int iga_fused_apply_noalloc(iga_fused_t plan, const void* packed, const void* x, void* y, int nrhs,
                            const void* nurbs_w, double alpha, int accumulate, void* stream);

/* The operators of nmember solves that share one plan (same rules and form, hence the
 * same shape; coefficients of their own, or the same ones for several right-hand
 * sides of one operator) in one launch: member k's coefficients at packed[k] (a host
 * array of nmember device pointers, 1 <= nmember <= 16), entry a of its input at
 * x[a * x_dstride + k * x_mstride] ((N, nmember) C order: nmember, 1; member-major:
 * 1, N), its output at y[k]
 * (a host array of nmember device pointers, N values each), nurbs_w shared (the
 * shape's).  With ndim = 3, y is zero outside the DOF box
 * [box_lo, box_lo + box_len) (the Dirichlet mask of linear_solver.py:119-124); ndim = 0,
 * no mask.  Each column is the member's iga_fused_apply result bit for bit (same
 * kernels and summation order).  The reference batches only right-hand sides of one
 * operator (operators/utils/matrix_free.py:36,117-119); this entry batches the
 * same-shape operator work of different requests of a solve server.  work:
 * iga_fused_workspace_bytes(plan, nmember) bytes. */
int iga_fused_apply_group(iga_fused_t plan, int nmember, const void* const* packed, const void* x,
                          int64_t x_dstride, int64_t x_mstride, void* const* y, const void* nurbs_w, void* work,
                          int ndim, const int64_t* box_lo, const int64_t* box_len, void* stream);

/* The two launches of iga_fused_apply, separately (for timing): stage 1 runs the
 * coefficient sweep into the partial blocks of `work`, stage 2 adds them into y. */
int iga_fused_apply_stage(iga_fused_t plan, int stage, const void* packed, const void* x, void* y,
                          int nrhs, const void* nurbs_w, double alpha, int accumulate, void* work,
                          void* stream);

/* Space-time form of the fused operator: nslices independent 3-D applies, slice s
 * reading x[dof * x_dstride + s * x_sstride] and its own packed coefficients at
 * packed + s * coef_slice_stride (elements; 0 = shared), writing
 * y[dof * y_dstride + s * y_sstride].  With time the slowest direction, the slices
 * are the time quadrature points of the space-time terms of
 * BsplineOperations.compute_mf_scalar_{u_v,gradu_gradv}(..., time_ders=...)
 * (operators/bsplines_ops.py:290-355): the time trial basis is applied before and
 * the time test weights after (iga_band_mode_product).  work: bytes of
 * iga_fused_workspace_bytes(plan, nslices). */
int iga_fused_apply_sliced(iga_fused_t plan, const void* packed, int64_t coef_slice_stride, const void* x,
                           int64_t x_dstride, int64_t x_sstride, void* y, int64_t y_dstride, int64_t y_sstride,
                           int nslices, const void* nurbs_w, double alpha, int accumulate, void* work,
                           void* stream);

/* The space-time operator's two sliced sweeps in one (form 5, the KM7 fields): the
 * stiffness term on x and the mass term on x1 (same strides), i.e.
 * gradu_gradv(time_ders=(0,0)) + u_v(time_ders=(0,1)) of operators/bsplines_ops.py
 * on the time-interpolated and time-differentiated vectors; the plan must have
 * form 5 and the other apply entries reject it. */
int iga_fused_apply_sliced2(iga_fused_t plan, const void* packed, int64_t coef_slice_stride, const void* x,
                            const void* x1, int64_t x_dstride, int64_t x_sstride, void* y, int64_t y_dstride,
                            int64_t y_sstride, int nslices, const void* nurbs_w, double alpha, int accumulate,
                            void* work, void* stream);

/* iga_fused_pack with the components of C (and dW) comp_stride elements apart
 * (>= the plan's number of points): packs one time slice of space-time coefficient
 * arrays (3, 3, N_q_space * n_q_time) without a copy. */
int iga_fused_pack_strided(iga_fused_t plan, const double* C, const double* m, const double* W,
                           const double* dW, int64_t comp_stride, void* packed, void* stream);

/* Packs the coefficients of a cross-section map extruded along xi_3 without the
 * (3,3,N_q) array: the same values as conductivity_coefficients of the extruded
 * patch times a lognormal field, followed by iga_fused_pack, generated point by point
 * (q = plane + n_plane q3, plane = q1 + nq1 q2):
 *   det = dz[q3] det2[plane],  kappa = exp(sigma Re sum_k Ez[q3][k] H[plane][k]),
 *   C_ab = kappa det inv_ab[plane] (a, b in {1,2}, symmetric),  C_33 = kappa det / dz^2,
 *   m = det when with_mass, else 0.
 * plane (device float64, 4 x n_plane): det2, inv_11, inv_22, inv_12 of the cross-section
 * (inverse of J_2^T J_2); dz (n3): dz/dxi_3.  H (n_plane x modes) and Ez (n3 x modes)
 * are complex (re, im interleaved); H = NULL gives kappa = 1.  NURBS (W, dW not NULL):
 * the control weights are those of the cross-section repeated along xi_3, so the
 * weight function is a cross-section field: W (n_plane) and dW (2 x n_plane, its
 * xi_1 and xi_2 derivatives; the xi_3 derivative is zero), folded as in
 * iga_fused_pack. */
int iga_fused_pack_extruded(iga_fused_t plan, const double* plane, const double* dz, const double* H,
                            const double* Ez, int modes, double sigma, int with_mass, const double* W,
                            const double* dW, void* packed, void* stream);

/* Cross-section factor of the separable Fourier field of the benchmark's lognormal
 * conductivity (no counterpart in the reference, whose callers pass coefficients):
 *   H[p][c] = sum_{a,b} coef[a][b][c] exp(i pi (a+1) x_p) exp(i pi (b+1) y_p),
 * xy (device, 2 x n_plane): the scaled cross-section coordinates; coef (device,
 * complex modes^3, c fastest); modes 1..16. */
int iga_lognormal_plane_modes(int64_t nplane, int modes, const double* xy, const double* coef, double* H,
                              void* stream);

/* ---- fast-diagonalization apply --------------------------------------------
 * Replaces SpaceFD.apply (fastdiag/utils/space_precond.py:273-320) for one DOF per
 * node: y[free] = U (Lambda^{-1} o U^T x[free]), y = 0 outside the free box, where
 * U = U_ndim (x) ... (x) U_1.  The free box of direction d is [lo[d], lo[d]+len[d])
 * of n[d] DOFs (the Dirichlet trimming of space_precond.py:160-171); U[d] is the
 * host float64 row-major len[d] x len[d] eigenvector matrix of eigh(K_d, M_d)
 * (space_precond.py:175).  Mode products run on the matrix cores in the plan dtype:
 * for three directions in three launches (gather and directions 1-2; direction 3
 * with the Lambda^{-1} scaling; directions 1-2 back and the scatter) when the box
 * lengths are at most 192 and a plane fits the 160 KiB of LDS (FP64: up to about
 * 132), otherwise one GEMM launch per mode product.  dtype IGA_BF16 takes float
 * vectors and rounds every matrix-core input to bf16, accumulating in FP32 (three
 * directions only). */
typedef struct iga_fd_s* iga_fd_t;

int iga_fd_create(iga_fd_t* out, int ndim, const int64_t* n, const int64_t* lo, const int64_t* len,
                  const double* const* U, int dtype);
/* Lambda over the free box (host float64, direction 1 fastest), e.g. the output of
 * SpaceFD.update_space_eigenvalues (space_precond.py:195-236); stored inverted. */
int iga_fd_set_eigenvalues(iga_fd_t plan, const double* eigenvalues);
int64_t iga_fd_workspace_bytes(iga_fd_t plan, int nrhs);
int iga_fd_apply(iga_fd_t plan, const void* x, void* y, int nrhs, void* work, void* stream);
/* iga_fd_apply on right-hand-side-major arrays: right-hand side k of x and y at k * N
 * (N values each, N the full grid) instead of interleaved; every kernel then reads
 * and writes each right-hand side contiguously.  Same results per right-hand side. */
int iga_fd_apply_rhs_major(iga_fd_t plan, const void* x, void* y, int nrhs, void* work, void* stream);
int iga_fd_destroy(iga_fd_t plan);

/* ---- one right-preconditioned Arnoldi step per call ---------------------------
 * The loop body of GMRES (solvers/linear_solver.py:438-525: z = P v_{k-1},
 * w = A z, the Dirichlet mask of w, linear_solver.py:119-124, then the
 * orthogonalisation and Givens update of iga_gmres_step) queued by one call.  A
 * context fixes the fused operator (plan, packed coefficients, NURBS weights or
 * NULL, its partial-block workspace), the FD preconditioner (NULL: z = v_{k-1}; the
 * FD writes zeros outside its free box, which must be the mask's box), the mask box
 * (ndim 3, or 0 for none; the operator's gather writes the zeros), the run's vectors
 * z and w (n values of dtype), its least-squares state ls
 * (iga_gmres_state_doubles(m)) and reduction workspace, and the host record.
 * Host record: host_rec, host-visible memory from iga_host_alloc of at least
 * 2 (m + 2) doubles, and tag, a multiple of 2^32 that identifies the solve.  Hist
 * entry d (g[d] and H[d, d-1] of Krylov dimension d; for d = 0 both are beta) is
 * written to host_rec[2 + 2d], host_rec[3 + 2d], and then tag + d to the int64 at
 * host_rec[0] with release ordering: a host that reads tag + d there may read
 * every entry up to d.  Entry 0 comes with iga_arnoldi_start, entry k - 1 with
 * step k >= 2.  A buffer may serve solves one after the other (new tag each): the
 * steps of a solve are queued on one stream, so they have finished once a later
 * operation on that stream has.
 * iga_arnoldi_start: iga_gmres_start(r -> V row 0).
 * iga_arnoldi_step (k = 1..m+1): step k of iga_gmres_step on the basis V (rows of
 * stride ldv, at least k + 1 of them); preconditioned = 1 takes z as already holding
 * P v_{k-1} (a caller that batches the FD of several solves), preconditioned = 2 takes
 * w as already holding the masked A P v_{k-1} (a caller that also batches the operator,
 * iga_fused_apply_group) and queues only the orthogonalisation.
 * iga_host_alloc: coherent pinned host memory (zeroed), device-writable;
 * iga_host_free releases it (no device work may still write to it). */
typedef struct iga_arnoldi_s* iga_arnoldi_t;

int iga_arnoldi_create(iga_arnoldi_t* out, int dtype, int64_t n, int m, iga_fused_t op, const void* packed,
                       const void* nurbs_w, void* op_work, iga_fd_t fd, void* fd_work, int ndim, const int64_t* box_n,
                       const int64_t* box_lo, const int64_t* box_len, void* z, void* w, double* ls, void* work,
                       double* host_rec, int64_t tag);
int iga_arnoldi_start(iga_arnoldi_t ctx, const void* r, void* V, void* stream);
int iga_arnoldi_step(iga_arnoldi_t ctx, int k, void* V, int64_t ldv, int preconditioned, void* stream);
int iga_arnoldi_destroy(iga_arnoldi_t ctx);

/* The reference's third stopping criterion, the relative increment
 * ||x_d - x_{d-1}|| / (||x_d|| + 1e-14) (solvers/linear_solver.py:107-117,503-515,
 * solvers/utils/convergence_manager.py:281-327), for one-call steps: each step keeps
 * z_{k-1} = P v_{k-1} as row k - 1 of Z (row stride ldz, at least m + 1 rows; Z may be
 * set again, with its rows copied, when the caller grows it), forms the Gram row
 * z_{k-1} . z_0..z_{k-1} into gram (device, m + 1 doubles) and the small dense kernel
 * writes the increment record entry j = k - 1 to inc_rec (host-visible, from
 * iga_host_alloc, iga_arnoldi_increment_doubles(m) doubles): at 1 + j (2m + 3), R[0..j-1][j-1]
 * of the rotated Hessenberg (m slots), g[j - 1] (slot m) and the Gram row of z_j (slots
 * m + 1 on); the int64 at inc_rec[0] then holds tag + j (release ordering), before the
 * residual estimate of entry j is published.  The host forms y_d = R_d^-1 g_d and the
 * increment from the Gram matrix, so x is never formed per step.  Z == NULL switches
 * the criterion off.  Only the step's own FD and operator (preconditioned = 0). */
int64_t iga_arnoldi_increment_doubles(int m);
int iga_arnoldi_set_increment(iga_arnoldi_t ctx, void* Z, int64_t ldz, double* gram, double* inc_rec);

/* One Arnoldi step of g solves of one shape (1 <= g <= 16; contexts sharing the FD
 * plan, the fused-operator plan, the mask box, the NURBS weights, the vector type and
 * length; each with its own coefficients, basis and state) as one call: step ks[i] of
 * solve i on basis Vs[i] (row stride ldvs[i]).  The vectors V_i[k_i - 1] go, one
 * after the other, through one multi-right-hand-side FD apply (iga_fd_apply_rhs_major)
 * and the results through one
 * iga_fused_apply_group launch into each context's w (masked); then the solves'
 * orthogonalisations (iga_gmres_step's results bit for bit, residual estimates
 * published as by iga_arnoldi_step), one launch per kernel of the step for all of them
 * (per-member launches on streams[i] when the members' basis layouts differ).  All of it
 * runs on batch_stream, after every member stream's earlier work, and each member
 * stream is ordered after it (events; a member stream equal to batch_stream needs
 * none).  xy: scratch of
 * iga_arnoldi_group_workspace_bytes(ctx, g) bytes; fd_work: iga_fd_workspace_bytes(fd,
 * g); op_work: iga_fused_workspace_bytes(op, g).  The reference has no counterpart: it
 * runs one solve at a time (solvers/linear_solver.py:438-525); this entry point
 * batches the same-shape steps of a solve server or of several right-hand sides. */
int64_t iga_arnoldi_group_workspace_bytes(iga_arnoldi_t ctx, int g);
int iga_arnoldi_step_group(int g, const iga_arnoldi_t* ctxs, const int* ks, void* const* Vs, const int64_t* ldvs,
                           void* const* streams, void* xy, void* fd_work, void* op_work, void* batch_stream);
int iga_host_alloc(int64_t bytes, void** out);
int iga_host_free(void* p);

/* ---- space-time arrowhead fast-diagonalization apply ----------------------------
 * Replaces SpTimeFD.apply / apply_transpose (fastdiag/utils/time_arrow.py:111-192)
 * for one DOF per node, time the slowest direction:
 *   y[free] = Re[(U_s (x) U_t) A^{-1} (U_s (x) conj(U_t))^T x[free]],  y = 0 elsewhere,
 * A^{-1} being the arrowhead solve of solve_special_arrowhead
 * (fastdiag/utils/operations.py:35-93) with H = adv D (+) Lambda_s and b = adv g
 * (time_arrow.py:82-97); the transposed apply swaps U_t and conj(U_t) and uses
 * b = -adv conj(g).  n/lo/len describe the 4-D free box (three space directions, then
 * time, whose first DOF is constrained: lo[3] = 1, time_template.py:52-56).  U[d]
 * are the real space eigenvectors (row-major len[d] x len[d]); ut_re/ut_im the time
 * eigenvectors U_t of compute_time_schurdecomposition (time_arrow.py:44-80),
 * row-major len[3] x len[3]. */
typedef struct iga_stfd_s* iga_stfd_t;

int iga_stfd_create(iga_stfd_t* out, const int64_t* n, const int64_t* lo, const int64_t* len,
                    const double* const* U, const double* ut_re, const double* ut_im, int dtype);
/* lam_space: the spatial Lambda over the space box (direction 1 fastest); d_re/d_im:
 * the diagonal of D_t (len[3]); g_re/g_im: its border g (len[3] - 1); adv: the
 * advection corrector (time_arrow.py:99-109). */
int iga_stfd_set_blocks(iga_stfd_t plan, const double* lam_space, const double* d_re, const double* d_im,
                        const double* g_re, const double* g_im, double adv);
int64_t iga_stfd_workspace_bytes(iga_stfd_t plan, int nrhs);
int iga_stfd_apply(iga_stfd_t plan, const void* x, void* y, int nrhs, int transpose, void* work, void* stream);
int iga_stfd_destroy(iga_stfd_t plan);

/* The legacy QZ form of the same apply (SpTimeFD of fastdiag/utils/time_legacy.py:16-171,
 * chosen by SingleFD.compute_time_schurdecomposition(..., use_arrowhead=False),
 * fastdiag/preconditioner.py:65-88): the complex QZ decomposition W_t = Q S Z^H,
 * M_t = Q T Z^H of the time pencil; apply = (U_s^T (x) Q^H) transform, per spatial
 * eigenvalue the upper-triangular solve with adv S + lambda T, (U_s (x) Z)
 * back-transform, real part; transpose = Z^T, the transposed (lower) solve, conj(Q).
 * q_*, z_* (and s_*, t_*) row-major len[3] x len[3]. */
int iga_stfd_create_qz(iga_stfd_t* out, const int64_t* n, const int64_t* lo, const int64_t* len,
                       const double* const* U, const double* q_re, const double* q_im, const double* z_re,
                       const double* z_im, int dtype);
int iga_stfd_set_qz(iga_stfd_t plan, const double* lam_space, const double* s_re, const double* s_im,
                    const double* t_re, const double* t_im, double adv);

#ifdef __cplusplus
}
#endif

#endif /* IGAOPS_HIP_H */
