/*
 * sie_capi.h — C ABI of libsie_amd (MI355X / gfx950), the SIE-MoM hot path:
 * RWG impedance-matrix assembly (T-EFIE, T-PMCHWT, SIE-GSTC), dense complex LU
 * with triangular solves, FP64 iterative refinement, GMRES and the adjoint
 * gradient of the GSTC susceptibilities.
 *
 * Conventions
 *  - Complex numbers are interleaved (re, im) doubles or floats.
 *  - Matrices are column-major with a leading dimension, as in the reference
 *    (columnMajorIndex, include/sie_mom/cuda/sie_kernels_assembly.cuh:12-14).
 *  - Unknowns are ordered block-major by RWG edge id: [J1|M1|J2|M2] for GSTC
 *    (src/sie_mom/solvers/gstc_sie.cu:122), [J|M] for PMCHWT, [J] for EFIE.
 *  - Pointers named *_dev are device pointers owned by the caller; everything
 *    else is host memory. `stream` is a hipStream_t (NULL = default stream);
 *    every call is stream-ordered and returns once the work is enqueued unless
 *    it documents a host synchronisation.
 *  - Every function returns a status: SIE_OK or one of the SIE_ERR_* codes,
 *    with a message readable through sie_last_error() on the calling thread.
 *    The codes mirror the reference's exception classes: invalid_argument,
 *    overflow_error, and runtime_error for singular / non-finite / device
 *    failures (gstc_sie.cu:71-179, dense_lu_solver.cuh:102-111).
 */
#ifndef SIE_CAPI_H
#define SIE_CAPI_H

#ifdef __cplusplus
extern "C" {
#endif

#define SIE_OK 0
#define SIE_ERR_INVALID 1
#define SIE_ERR_OVERFLOW 2
#define SIE_ERR_SINGULAR 3
#define SIE_ERR_NONFINITE 4
#define SIE_ERR_DEVICE 5

/* Message of the last failed call on this thread ("" after a success). */
const char* sie_last_error(void);

/* ------------------------------------------------------------------ mesh */
/* A closed, outward-oriented triangle mesh uploaded to the device once.
 * Replaces mesh::MeshData + copy_problem_to_device (include/sie_mom/mesh/
 * spherical_surface.hpp:10-15, src/sie_mom/solvers/gstc_sie.cu:181-242).
 * vertices: nv x 3 doubles; faces, face_edges: nf x 3 ints (local edge j runs
 * vertex j -> j+1); ne: number of RWG edges. Every edge must be shared by
 * exactly two faces. */
typedef struct sie_mesh* sie_mesh_t;
int sie_mesh_create(const double* vertices, int nv, const int* faces, int nf,
                    const int* face_edges, int ne, sie_mesh_t* out);
int sie_mesh_destroy(sie_mesh_t mesh);

/* Mesh generators. Call once with NULL arrays to get counts = {nv, nf, ne},
 * then again with arrays of those sizes (vertices nv*3, faces nf*3, edges ne*2,
 * face_edges nf*3). Edge ids follow the reference's first-encounter order
 * (src/sie_mom/mesh/spherical_surface.cpp:100-135).
 * sie_icosphere replaces mesh::SphericalSurface(radius, level)
 * (spherical_surface.cpp:19-98). sie_spindle is the circular-arc spindle of
 * the reference's gmsh CircularOgive (src/sie_mom/mesh/circular_ogive.py:13-80):
 * the revolution of an arc of radius arc_radius about its chord (length chord). */
int sie_icosphere(float radius, int level, double* vertices, int* faces, int* edges,
                  int* face_edges, int* counts);
int sie_spindle(double chord, double arc_radius, int n_rings, double* vertices, int* faces,
                int* edges, int* face_edges, int* counts);

/* Strict OBJ loader (the reference's ObjMeshLoader, src/sie_mom/mesh/
 * obj_mesh_loader.cpp:83-386): `v` and triangular `f` records (1-based,
 * negative relative indices and a/t/n forms), vertex and face order and
 * winding preserved, closed-surface checks (the reference's messages, as
 * SIE_ERR_INVALID), RWG edges in first-encounter order. Same two-call
 * protocol and array layouts as the generators above. */
int sie_obj_load(const char* path, double* vertices, int* faces, int* edges, int* face_edges,
                 int* counts);

/* Launch parameters of the reference assembly kernels, already rounded to
 * float exactly as the reference casts them (gstc_sie.cu:281-290). */
typedef struct {
    int rule;         /* Dunavant rule 1..20 */
    float k1, k2;     /* interior / exterior wavenumber */
    float mu_r1, mu_r2, eps_r1, eps_r2;
    float mu0, eps0, omega;
    float threshold;  /* near-pair centroid distance (0.1 lambda0 by default) */
} sie_calderon_params;

/* Raw Calderon quadrature of every face pair, no epilogue: out_dev holds
 * nf*nf*72 doubles, pair (m,n) at (m*nf+n)*72 as 9 complex TL(k1), TL(k2),
 * TK(k1), TK(k2). The device counterpart of calculateDoubleQuadrature +
 * calculateSingularPartQuadrature per pair (include/sie_mom/cuda/
 * dunavant_quadrature.cuh:207-324, 769-875); for validation on small meshes. */
int sie_calderon_pair_blocks(sie_mesh_t mesh, const sie_calderon_params* p,
                             double* out_dev, void* stream);

#define SIE_ASM_CALDERON 1
#define SIE_ASM_MASS 2
#define SIE_ASM_UNSCALED 4 /* eta0 -> 1: assemble_Z_calderon_kernel (sie_kernels_assembly.cuh:24-148)
                              and assemble_Z_mass_kernel (:282-452), the operator of the
                              reference's matrix-free matvec */
/* eta0-scaled SIE-GSTC matrix D A D (dim 4E), overwriting Z_dev.
 * Replaces assemble_system_matrix (gstc_sie.cu:244-317): the Calderon kernel
 * assemble_Z_calderon_eta0_scaled_kernel (sie_kernels_assembly.cuh:156-272)
 * and the mass kernel assemble_Z_mass_eta0_scaled_kernel (:460-594).
 * tensors_dev: 32 complex per face (or 32 if uniform): the 4x4 row-major
 * fields-to-currents map followed by its inverse (gstc_sie.hpp:205-213).
 * basis: 0 = face frame, 1 = spherical frame. flags: SIE_ASM_* (both = full). */
int sie_assemble_gstc(sie_mesh_t mesh, const sie_calderon_params* p, const double* tensors_dev,
                      int uniform, int basis, int flags, double* Z_dev, long long ldz,
                      void* stream);

/* PEC T-EFIE matrix jw mu2 TL(k2), dim E (assemble_pec_efie_operator_kernel,
 * sie_kernels_assembly.cuh:610-705, launched with k2 twice by pec_solver.cu:352-359). */
int sie_assemble_pec(sie_mesh_t mesh, const sie_calderon_params* p, double* Z_dev,
                     long long ldz, void* stream);

/* T-PMCHWT matrix, dim 2E, eta0 scaled (t_pmchwt_kernel,
 * src/sie_mom/solvers/dielectric/dielectric_kernels.cu:23-150). */
int sie_assemble_pmchwt(sie_mesh_t mesh, const sie_calderon_params* p, double* Z_dev,
                        long long ldz, void* stream);

/* Matrix-free products y = Z x with the operators above, Z never stored:
 * the blocks of each chunk of source faces are computed and applied at once
 * (x and y: device vectors of the operator's dimension, distinct).
 * sie_gstc_matvec replaces matvec_Z_calderon_kernel + matvec_Z_mass_kernel
 * (sie_kernels_matvec.cuh:9-327, launched by PhysicsMatVec,
 * test_gmres_solver.cu:247-287); pass SIE_ASM_UNSCALED for that operator.
 * Synchronises the stream. */
int sie_gstc_matvec(sie_mesh_t mesh, const sie_calderon_params* p, const double* tensors_dev,
                    int uniform, int basis, int flags, const double* x_dev, double* y_dev,
                    void* stream);
int sie_pec_matvec(sie_mesh_t mesh, const sie_calderon_params* p, const double* x_dev,
                   double* y_dev, void* stream);
int sie_pmchwt_matvec(sie_mesh_t mesh, const sie_calderon_params* p, const double* x_dev,
                      double* y_dev, void* stream);

/* Plane wave bound to its medium (cuda::PlaneWave, include/sie_mom/excitation/
 * source_functors.cuh:14-42; region 1 interior, 2 exterior). */
typedef struct {
    double E0[3];
    double khat[3];
    double k, eta;
    int region;
} sie_plane_wave;

#define SIE_RHS_GSTC 0
#define SIE_RHS_PMCHWT 1
#define SIE_RHS_EFIE 2
/* Tested incident fields of nwaves plane waves, one RHS column each, written
 * (not accumulated) to rhs_dev (leading dimension ldb, complex).
 * GSTC: assemble_V_vector_eta0_scaled_kernel (sie_kernels_utils.cuh:185-232),
 * using eta0 = sqrt(double(mu0)/double(eps0)) on the M rows;
 * PMCHWT: assemble_dielectric_rhs_kernel with sign +1 exterior / -1 interior
 * and sign*sqrt(vac_mu/vac_eps) on M rows (dielectric_solver.cu:400-531);
 * EFIE: assemble_pec_rhs_kernel, +<f, E_inc> (pec_solver.cu:243-264). */
int sie_rhs_plane_waves(sie_mesh_t mesh, int kind, int rule, const sie_plane_wave* waves,
                        int nwaves, float mu0, float eps0, double vac_mu, double vac_eps,
                        double* rhs_dev, long long ldb, void* stream);

/* E and H radiated at nobs observation points (obs_dev: nobs x 3 doubles) by
 * physical RWG currents J_dev, M_dev (one complex per edge) in a medium
 * (epsilon, mu absolute; k = omega sqrt(mu epsilon)): E_dev, H_dev receive
 * nobs x 3 complex. Replaces evaluate_surface_integral_fields
 * (src/sie_mom/solvers/surface_current_fields.cu:170-299) and its kernel
 * compute_scattered_fields_kernel (include/sie_mom/cuda/sie_kernels_utils.cuh:
 * 21-120): E = jw mu L(J) + K(M), H = -K(J) + jw eps L(M), faces within 0.1
 * lambda of a point integrated with the Wilton singular part. Synchronises
 * the stream. */
int sie_surface_fields(sie_mesh_t mesh, int rule, double omega, double epsilon, double mu,
                       const double* J_dev, const double* M_dev, const double* obs_dev, int nobs,
                       double* E_dev, double* H_dev, void* stream);

/* ------------------------------------------------------------ dense algebra */
/* precision: 64 = complex double (interleaved doubles), 32 = complex float.
 * op: 0 = N, 1 = T (transpose), 2 = C (conjugate transpose); the numbering
 * follows cublasOperation_t (CUBLAS_OP_N = 0, CUBLAS_OP_T = 1, CUBLAS_OP_C = 2),
 * which the reference's DenseLuSolver::solve takes (dense_lu_solver.cuh:122-166).
 * sie_lu_solve accepts all three; GEMV and iterative refinement take N and C. */

/* C = beta C + alpha A B, complex column-major (m x k times k x n) on the
 * matrix cores. variant (FP32 only): 0 = 4M real decomposition, 1 = 3M,
 * 2 = FP32-accurate split-FP16 (each FP32 operand split into FP16 hi + lo,
 * three FP16 MFMAs per real product, FP32 accumulation). */
int sie_gemm(int precision, int m, int n, int k, double alpha, const void* A_dev, long long lda,
             const void* B_dev, long long ldb, double beta, void* C_dev, long long ldc, int variant,
             void* stream);

/* In-place LU with partial pivoting, A = P L U (unit L), column-major n x n.
 * ipiv_dev[k] (0-based) is the row swapped with row k. Replaces
 * DenseLuSolver::factorize / cusolverDnZgetrf (dense_lu_solver.cuh:42-112);
 * returns SIE_ERR_SINGULAR with "dense matrix is singular at U diagonal k"
 * (k 1-based, also stored in *info) when a pivot column is exactly zero, and
 * SIE_ERR_DEVICE when a panel's grid-wide exchange gave up (sie_set_wait_bound).
 * Synchronises the stream. variant: as sie_gemm, for the trailing update. */
int sie_lu_factor(int precision, int n, void* A_dev, long long lda, int* ipiv_dev, int* info,
                  int variant, void* stream);

/* Solve op(A) X = B in place with the factors of sie_lu_factor; B is n x nrhs
 * (ldb). Replaces DenseLuSolver::solve / cusolverDnZgetrs
 * (dense_lu_solver.cuh:122-166); op = 2 is the adjoint solve with A^H and
 * op = 1 (A^T) is solved as A^H conj(X) = conj(B). With nrhs <= 4 (blocked
 * sweeps that wait on published blocks) the call synchronises the stream and
 * returns SIE_ERR_DEVICE if such a wait gave up. With more columns and op = 0
 * the rows beyond each 256-row block are updated by one product in the
 * factors' precision (each column accurate relative to its own scale). */
int sie_lu_solve(int precision, int op, int n, int nrhs, const void* LU_dev, long long lda,
                 const int* ipiv_dev, void* B_dev, long long ldb, void* stream);

/* y = beta y + alpha op(A) x, complex double, A m x n (lda). op 2 uses A^H
 * (y has n entries). Replaces the cublasZgemv residual (gstc_sie.cu:398-457). */
int sie_zgemv(int op, int m, int n, double alpha, const double* A_dev, long long lda,
              const double* x_dev, double beta, double* y_dev, void* stream);

/* sie_zgemv without any allocation: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight refinement or serving loops, harnesses that forbid hipMalloc). The
 * partial sums live in a static device array of 64 MiB per GPU (held until
 * the process ends), so the call allocates nothing and returns
 * SIE_ERR_INVALID when they would exceed it: op N needs
 * 16 m ceil(n / 2048) bytes, op C 16 n ceil(m / 8192) (square operators up to
 * n ~ 92,000 with op N). Same kernels, same operation order and results as
 * sie_zgemv. Every call on one GPU shares that GPU's array: calls that
 * overlap in time on one GPU (other streams or threads) race. Stream-ordered;
 * no synchronisation. */
int sie_zgemv_noalloc(int op, int m, int n, double alpha, const double* A_dev, long long lda,
                      const double* x_dev, double beta, double* y_dev, void* stream);

/* B = scale * A with precision conversion (64->32, 32->64 or 64->64).
 * fro_norm (may be NULL; complex double sources only): receives the Frobenius
 * norm of A, formed in the same pass (deterministic order) -- the matrix norm
 * of the refinement's backward-error test below; then the call synchronises
 * the stream. */
int sie_convert(int from_precision, int to_precision, long long rows, long long cols,
                const void* A_dev, long long lda, void* B_dev, long long ldb, double scale,
                double* fro_norm, void* stream);

/* FP64 iterative refinement: solve op(Z) X = B with Z in complex double and
 * an FP32 LU of Z (sie_lu_factor(32, ...)) as the inner solver:
 *   x = 0, r = b;  repeat: x += |r| LU32^{-1}(r/|r|);  r = b - op(Z) x (FP64)
 * until the normwise backward error reaches tol,
 *   |r| <= tol (znorm |x| + |b|)          (2-norms of the vectors),
 * or maxit sweeps. znorm is a norm of Z, normally its Frobenius norm from
 * sie_convert (the minimal perturbation dZ with (Z + dZ) x = b is of rank one,
 * so its Frobenius and 2-norms agree); znorm = 0 turns the test into the
 * relative residual |r| <= tol |b|. Per column: rel_residual[c] = |r|/|b|,
 * backward_error[c] = |r| / (znorm |x| + |b|) (either may be NULL),
 * iterations[c]. A column that ends above tol is reported, not an error: the
 * caller compares backward_error (or rel_residual) with tol. No counterpart in
 * the reference (its solves are plain FP64 cuSOLVER). Synchronises the stream
 * every sweep (host norm). */
int sie_ir_solve(int op, int n, int nrhs, const double* Z_dev, long long ldz, const float* LU_dev,
                 long long ldl, const int* ipiv_dev, const double* B_dev, long long ldb, double* X_dev,
                 long long ldx, double tol, double znorm, int maxit, double* rel_residual,
                 double* backward_error, int* iterations, void* stream);

/* The same solve by GMRES(restart) right-preconditioned with the FP32 LU
 * (GMRES-IR): z_k = LU32^{-1} v_k with FP64 arithmetic on the FP32 factors,
 * w = op(Z) z_k and its CGS2 orthogonalisation in FP64, x += [z_0..z_{k-1}] y;
 * each cycle ends with the true FP64 residual. Same stopping rule and outputs
 * as sie_ir_solve, and one more: a restart cycle that does not halve the true
 * residual ends the solve (the residual's FP64 floor), with the backward error
 * reached reported; iterations[c] counts inner steps (one preconditioner solve
 * and one FP64 GEMV each). No counterpart in the reference. Synchronises the
 * stream every step. */
int sie_fgmres_ir_solve(int op, int n, int nrhs, const double* Z_dev, long long ldz,
                        const float* LU_dev, long long ldl, const int* ipiv_dev,
                        const double* B_dev, long long ldb, double* X_dev, long long ldx,
                        double tol, double znorm, int maxit, int restart, double* rel_residual,
                        double* backward_error, int* iterations, void* stream);

/* Releases the library's device state outside the mesh handles (workspaces,
 * captured sweep graphs and their stream, the wait-error word, profiling
 * events) after draining the device; later calls re-create what they need.
 * Call it before the process exits while the HIP runtime is still up (the
 * Python binding registers it with atexit): the library runs no device call
 * from static destructors. No counterpart in the reference (its RAII buffers
 * live per solve). */
int sie_release(void);

/* Bound of the library's device-side waits (the LU panels' grid-wide pivot
 * exchanges and the triangular sweeps' published blocks), default 5 s: a
 * wait that reaches it ends its kernel, and the call that synchronised
 * returns SIE_ERR_DEVICE instead of leaving the device hung (e.g. when the
 * panel's workgroups cannot all be resident because another process holds
 * the GPU). force_timeout (testing only): the panel exchanges behave as if a
 * workgroup never arrived, so the next factorisation gives up at the bound.
 * No counterpart in the reference. */
int sie_set_wait_bound(double seconds, int force_timeout);

/* Sustained rate of the FP16 matrix cores on the current device under the
 * split-FP16 GEMM's instruction mix (register operands, no memory traffic):
 * one 512-thread workgroup per CU runs `iters` slices of 48 32x32x16 FP16
 * MFMAs per wave, `reps` times; *tflops receives the best rate (FP16 MFMA
 * flops, 2 M N K per instruction) and *clock_mhz the shader clock held during
 * that launch (s_memtime cycles over 100 MHz s_memrealtime ticks). A
 * measurement for rooflines; no counterpart in the reference. Synchronises. */
int sie_mfma_f16_rate(int iters, int reps, double* tflops, double* clock_mhz, void* stream);

/* Launch form of the LU panel kernels, whose workgroups wait for one another
 * and so must all be resident at once: 0 (default) plain dispatches, which
 * the dispatcher makes resident on an otherwise idle device (at most one
 * workgroup per CU); 1 cooperative launches, where the runtime guarantees it
 * (~21 us more per panel). With plain dispatches, kernels of other streams
 * running during a factorisation can keep a panel's workgroups out; its
 * bounded waits then give up and sie_lu_factor returns SIE_ERR_DEVICE. No
 * counterpart in the reference (cuSOLVER's panel). */
int sie_set_panel_launch(int cooperative);

/* Euclidean norm of n complex doubles (deterministic two-stage reduction,
 * synchronises the stream; returns -1 on error). */
double sie_norm2(long long n, const double* x_dev, void* stream);

/* ------------------------------------------------------------- GMRES */
/* y_dev = A x_dev for an operator supplied by the caller, enqueued on stream;
 * returns SIE_OK or an SIE_ERR_* code (the solve stops with that code). */
typedef int (*sie_matvec_fn)(void* user, const double* x_dev, double* y_dev, void* stream);

/* Restarted GMRES(m) for A x = b, complex double, replacing gmres_restarted /
 * GMRESSolver::solve (include/sie_mom/cuda/gmres/gmres_solver.cuh:258-537) with
 * its contract: x_dev holds the initial guess on entry; tol is relative to |b|;
 * |b| < 1e-20 gives x = 0; *out_res is the absolute residual estimate of the
 * last Arnoldi step; *out_iters counts inner steps over all restarts; *info is
 * 0 when converged, 1 when maxnit restarts did not reach tol; history (may be
 * NULL) receives the relative residual after each inner step, up to
 * history_cap values, *history_len the number written. Arnoldi by classical
 * Gram-Schmidt with one reorthogonalisation (two GEMV passes over the basis);
 * device workspace n x (m+1) complex. Synchronises the stream. */
int sie_gmres(int n, int m, int maxnit, double tol, sie_matvec_fn matvec, void* user, const double* b_dev,
              double* x_dev, double* out_res, int* out_iters, int* info, double* history, int history_cap,
              int* history_len, void* stream);

/* sie_gmres with the stored-matrix operator y = Z x (Z column-major, ldz):
 * the stored-Z counterpart of the reference's matvec_Z_* kernels. */
int sie_gmres_dense(int n, const double* Z_dev, long long ldz, int m, int maxnit, double tol,
                    const double* b_dev, double* x_dev, double* out_res, int* out_iters, int* info,
                    double* history, int history_cap, int* history_len, void* stream);

/* ----------------------------------------------------------- adjoint */
/* GSTC surface maps from real per-face parameters (kind = parameters per
 * face): Lambda = [[jB, K], [-K^T, jX]] (the reciprocal lossless form of
 * src/sie_mom/gstc/surface_parameters.py:280-306) with
 *   kind 3  (omega, omega_biisotropic.py:196-219): B = b p0 I, X = x p1 I,
 *            K = k p2 [[0, 1], [-1, 0]];
 *   kind 10 (general 2x2): B = b [[p0, p2], [p2, p1]], X = x [[p3, p5], [p5, p4]],
 *            K = k [[p6, p7], [p8, p9]],
 * where b, x, k are b_scale, x_scale, k_scale (e.g. w eps0, w mu0, k0 for
 * susceptibility parameters). Writes the 32-complex records per face
 * (Lambda row-major, then Lambda^-1 inverted in impedance-balanced form with
 * eta_ref, as src/sie_mom/gstc/tensor_io.py:17-130 does); *n_singular counts
 * faces whose map could not be inverted (then SIE_ERR_SINGULAR). */
int sie_gstc_records(int nf, int kind, const double* params_dev, double b_scale, double x_scale,
                     double k_scale, double eta_ref, double* records_dev, int* n_singular,
                     void* stream);

/* Far field of the exterior currents of a scaled GSTC solution xs (dim 4E,
 * physical J2 = xs[2E:3E], M2 = eta0 xs[3E:4E]) in ndir directions. dirs is a
 * host array of ndir x 9 doubles (r_hat, theta_hat, phi_hat); F_dev receives
 * ndir x 2 complex (F_theta, F_phi) with the Schelkunoff convention of
 * applications/anisotropic_gstc_sphere/plot_far_field.py:150-219. */
int sie_farfield(sie_mesh_t mesh, int rule, double k0, double eta0, const double* x_dev, int ndir,
                 const double* dirs, double* F_dev, void* stream);

/* Adjoint of sie_farfield: g = A^H u for F = A xs, with u_dev ndir x 2
 * complex. g_dev (dim 4E complex) is the gradient dJ/d conj(xs) of
 * J = sum_d (u_d . conj F_d) when u_d = w_d F_d / (2 eta0). */
int sie_farfield_adjoint(sie_mesh_t mesh, int rule, double k0, double eta0, int ndir, const double* dirs,
                         const double* u_dev, double* g_dev, void* stream);

/* Per-face derivative of S = lambda^H Zs_mass x with respect to the 32 record
 * entries (Lambda then Lambda^-1) of that face: G_dev nf x 32 complex. Zs_mass
 * is the eta0-scaled GSTC mass operator of sie_assemble_gstc. */
int sie_gstc_mass_sensitivity(sie_mesh_t mesh, int rule, int basis, float mu0, float eps0,
                              const double* lambda_dev, const double* x_dev, double* G_dev, void* stream);

/* Chain rule to the real parameters: grad_dev (nf x kind) = -2 Re dS/dtheta
 * with dLambda^-1 = -Lambda^-1 dLambda Lambda^-1. */
int sie_gstc_param_grad(int nf, int kind, const double* records_dev, const double* G_dev,
                        double b_scale, double x_scale, double k_scale, double* grad_dev, void* stream);

/* ------------------------------------------------------------ profiling */
/* Per-kernel timing with HIP events on the launch stream (no counterpart in
 * the reference, which times phases with std::chrono). kernel: 0 FP32 GEMM,
 * 1 FP64 GEMM, 2 far-pair assembly, 3 LU panel, 4 near-pair assembly,
 * 5 per-edge gather, 6 GEMV (products of at least 2^24 entries), 7 split-FP16
 * GEMM (128 x 128 tiles), 8 split-FP16 GEMM (256 x 128 tiles, K >= 512 and
 * N > 1024), 9 single-RHS triangular solve (both blocked sweeps), 10 adjoint
 * mass-block contraction, 11 matrix-free operator product (whole call).
 * enable(k, 1) clears and starts recording; read() synchronises on the
 * recorded events and returns the launch count, summed duration and summed
 * algorithmic flops / bytes. */
int sie_profile_enable(int kernel, int on);
int sie_profile_read(int kernel, int* launches, double* total_ms, double* flops, double* bytes);

#ifdef __cplusplus
}
#endif
#endif /* SIE_CAPI_H */
