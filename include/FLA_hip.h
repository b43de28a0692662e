/*
 * FLA_hip.h -- thin C ABI between libflame_amd's C host layer and its
 * hand-written gfx950 HIP kernels.  Plain pointers, sizes and an int status
 * (0 = success, otherwise a hipError_t value or one of FLA_HIP_E*).
 *
 * Every matrix argument is a DEVICE pointer to column-major storage with
 * leading dimension ld (element (i,j) at A[i + j*ld]).  All calls are
 * asynchronous on the library's streams unless they return host data
 * (info / pivot indices), in which case they synchronize once at the end.
 *
 * Each entry point replaces one seam of the reference's HIP/SuperMatrix path
 * (paths relative to the libflame source tree):
 *   fla_hip_dgemm        FLA_Gemm_external_hip   src/base/flamec/wrappers/blas/3/hip/FLA_Gemm_external_hip.c:19 (rocblas_dgemm :142)
 *   fla_hip_dsyrk        FLA_Herk_external_hip / FLA_Syrk_external_hip
 *                        src/base/flamec/wrappers/blas/3/hip/FLA_Herk_external_hip.c (rocblas_dsyrk :107)
 *   fla_hip_dtrsm        FLA_Trsm_external_hip   src/base/flamec/wrappers/blas/3/hip/FLA_Trsm_external_hip.c:19 (rocblas_dtrsm :102)
 *   fla_hip_dpotrf       FLA_Chol_blk_external_hip src/base/flamec/wrappers/lapack/hip/FLA_Chol_blk_external_hip.c:30 (rocsolver_dpotrf :75)
 *   fla_hip_dgetrf       FLA_LU_piv_blk_external_hip src/base/flamec/wrappers/lapack/hip/FLA_LU_piv_blk_external_hip.c:20 (rocsolver_dgetrf :75)
 *   fla_hip_dlaswp       FLA_Apply_pivots_unb_external_hip src/base/flamec/wrappers/lapack/hip/FLA_Apply_pivots_unb_external_hip.c:20 (rocsolver_dlaswp :92)
 * and the SuperMatrix HIP scheduler (FLASH_Queue_exec_hip,
 * src/base/flamec/supermatrix/main/FLASH_Queue_exec.c:2302) is replaced by the
 * one-GPU look-ahead schedule inside fla_hip_dpotrf / fla_hip_dgetrf.
 */
#ifndef FLA_HIP_H
#define FLA_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FLA_HIP_OK 0
#define FLA_HIP_EARG 9001   /* invalid argument */
#define FLA_HIP_ENODEV 9002 /* no gfx950 device available */
#define FLA_HIP_ESPIN 9003  /* a bounded inter-workgroup wait expired */
#define FLA_HIP_ERANGE 9004 /* input does not fit the reduced precision */
#define FLA_HIP_ENOINIT 9005   /* the library is not initialised (no-allocation entry points) */
#define FLA_HIP_ESCRATCH 9006  /* problem exceeds a no-allocation entry point's scratch */

/* sides / uplo / trans / diag use the FLAME constants (FLAME.h). */

/* ---- runtime ---- */
int fla_hip_init(void);     /* selects device 0, creates the two streams */
int fla_hip_finalize(void); /* releases streams, events and workspaces */
int fla_hip_synchronize(void);
/* The caller's stream (a hipStream_t; NULL: none, the default).  Replaces
   the per-call rocblas_handle, and the stream it carries, that the
   reference's FLA_*_external_hip wrappers take
   (src/base/flamec/supermatrix/hip/main/FLASH_Queue_hip.c:297-318).  While
   one is set, every entry point that queues device work orders that work
   after the work already queued on the caller's stream.  It also orders the
   caller's later work on that stream after its own, so asynchronous entry
   points (dgemm, dlacpy, ...) compose with the caller's kernels without a
   host synchronisation.  The library keeps its own streams; the ordering
   uses events created at initialisation.  fla_hip_set_stream initialises
   the library if needed; FLA_Finalize / fla_hip_finalize clear the setting. */
int fla_hip_set_stream(void* stream);
void* fla_hip_get_stream(void);
int fla_hip_malloc(void** ptr, size_t bytes);
int fla_hip_free(void* ptr);
int fla_hip_memcpy_h2d(void* dst, const void* src, size_t bytes);
int fla_hip_memcpy_d2h(void* dst, const void* src, size_t bytes);
int fla_hip_memcpy_d2d(void* dst, const void* src, size_t bytes);
/* 2-D strided copies of an m x n column-major double matrix (sync). */
int fla_hip_dmatrix_h2d(int m, int n, const double* h, int ldh, double* d, int ldd);
int fla_hip_dmatrix_d2h(int m, int n, const double* d, int ldd, double* h, int ldh);
int fla_hip_dmatrix_d2d(int m, int n, const double* s, int lds, double* d, int ldd);
/* Pitched copy of `height` rows of `width` bytes; kind 0 = host->device,
   1 = device->host, 2 = device->device (synchronous). */
int fla_hip_memcpy2d(void* dst, size_t dpitch, const void* src, size_t spitch, size_t width,
                     size_t height, int kind);
/* 1 if ptr is device memory the library can launch kernels on, else 0. */
int fla_hip_is_device_ptr(const void* ptr);
const char* fla_hip_error_string(int status);

/* ---- level-3 kernels (MFMA FP64) ---- */
/* C := alpha*op(A)*op(B) + beta*C; trans = FLA_NO_TRANSPOSE / FLA_TRANSPOSE. */
int fla_hip_dgemm(int transa, int transb, int m, int n, int k, double alpha,
                  const double* A, int lda, const double* B, int ldb, double beta,
                  double* C, int ldc);
/* C := alpha*op(A)*op(A)' + beta*C on the uplo triangle only. */
int fla_hip_dsyrk(int uplo, int trans, int n, int k, double alpha, const double* A,
                  int lda, double beta, double* C, int ldc);
/* B := alpha*inv(op(A))*B (left) or alpha*B*inv(op(A)) (right). */
int fla_hip_dtrsm(int side, int uplo, int trans, int diag, int m, int n, double alpha,
                  const double* A, int lda, double* B, int ldb);
/* fla_hip_dtrsm without any device allocation.  No counterpart in the
   reference: it is for callers that cannot allocate (code under HIP graph
   capture, tight serving loops).  Same arguments, kernels and results as
   fla_hip_dtrsm; the workspace (inverted 128-blocks, their condition flags
   and two wait words) is a file-scope device array sized for a triangle of
   order <= 8192 (8 MB per GPU, held until the process ends).  Requires FLA_Init / fla_hip_init
   first (FLA_HIP_ENOINIT otherwise, no implicit initialization); a larger
   triangle returns FLA_HIP_ESCRATCH.  Calls that overlap on one GPU (from
   several host threads) share the array and race. */
int fla_hip_dtrsm_noalloc(int side, int uplo, int trans, int diag, int m, int n, double alpha,
                          const double* A, int lda, double* B, int ldb);

/* ---- factorizations ---- */
/* Blocked Cholesky, in place on the uplo triangle.  *info = FLA_SUCCESS (-1)
   or the 0-based index of the first column whose pivot is <= 0 or NaN. */
int fla_hip_dpotrf(int uplo, int n, double* A, int lda, int* info);
/* LU without pivoting of an m x n matrix in place (unit L below, U on and
   above the diagonal).  *info = FLA_SUCCESS or the 0-based index of the
   first exact zero on U's diagonal.  Replaces FLA_LU_nopiv
   (src/lapack/dec/lu/nopiv/front/flamec/FLA_LU_nopiv.c:16). */
int fla_hip_dgetrf_nopiv(int m, int n, double* A, int lda, int* info);
/* Triangular inverse in place on the uplo triangle, diag FLA_UNIT_DIAG or
   FLA_NONUNIT_DIAG (the other triangle, and a unit diagonal, are not
   touched).  Replaces FLA_Trinv (src/lapack/inv/tri/front/flamec/FLA_Trinv.c:16). */
int fla_hip_dtrtri(int uplo, int diag, int n, double* A, int lda);
/* A := L' L (lower) or U U' (upper) on the uplo triangle.  Replaces FLA_Ttmm
   (src/lapack/misc/ttmm/front/flamec/FLA_Ttmm.c). */
int fla_hip_dlauum(int uplo, int n, double* A, int lda);
/* A := inv(A) for SPD A on the uplo triangle: Cholesky, triangular inverse,
   then the triangle times its transpose.  *info = FLA_SUCCESS or the
   0-based failing column of the Cholesky (A then holds the partial factor).
   Replaces FLA_SPDinv (src/lapack/inv/spd/front/flamec/FLA_SPDinv_internal.c:22). */
int fla_hip_dspdinv(int uplo, int n, double* A, int lda, int* info);
/* Blocked LU with partial pivoting.  p (device, int[min(m,n)]) receives
   FLAME-native pivots (0-based, relative to the row).  *info = FLA_SUCCESS or
   the 0-based index of the first null pivot. */
int fla_hip_dgetrf(int m, int n, double* A, int lda, int* p, int* info);
/* Row interchanges i <-> i + p[i * |incp|] for i = k1..k2 (in that order, or
   reversed when incp < 0) on the ncols columns of A; p is a device
   pointer to FLAME-native (0-based, relative) pivots.  The final source of
   every touched row is resolved on the device and only touched rows move;
   asynchronous on the library's stream.  Row indices are bounded by lda; an
   interchange whose target row lies outside [0, lda) is skipped. */
int fla_hip_dlaswp(int ncols, double* A, int lda, int k1, int k2, const int* p, int incp);
/* The same with the rows of the m x ncols matrix: targets outside [0, m) are
   skipped (FLA_Apply_pivots on an m-row object). */
int fla_hip_dlaswp_rows(int m, int ncols, double* A, int lda, int k1, int k2, const int* p,
                        int incp);
/* ---- hierarchical (FLASH) storage ----
   Copies ntiles pieces between tiles and the flat column-major matrix
   (flat, ldf): desc holds 6 host int64 per piece -- the tile's device
   address, its leading dimension, rows, cols, and the row and column of the
   piece in the flat matrix.  to_tiles != 0 writes the tiles, else the flat
   matrix.  elem_bytes 8 (double) or 4 (int).  Synchronous.  Replaces the
   host walk of FLASH_Copy_flat_to_hier / FLASH_Copy_hier_to_flat
   (src/base/flamec/hierarchy/main/FLASH_Copy_other.c:81-138). */
int fla_hip_tile_copy(int ntiles, const long long* desc, void* flat, long ldf, int elem_bytes,
                      int to_tiles);

/* ---- LU with incremental pivoting (FLASH_LU_incpiv) ----
   On the flat m x n matrix A whose b x b tiles are the hierarchical object's
   blocks (m and n multiples of b), inner block nb_alg.  p: the flattened
   pivot object, m x (n/b) ints, tile (i,k) at rows i*b, column k; L: the
   flattened L object, m x (n/b)*nb_alg, tile (i,k) at rows i*b, columns
   k*nb_alg.  Synchronous.  Replaces FLASH_LU_incpiv_var2
   (src/lapack/dec/lu/incpiv/zaux/flamec/flash/FLASH_LU_incpiv_var2.c:13-114)
   and FLASH_FS_incpiv_aux1 (FLASH_FS_incpiv_aux1.c:13-118). */
int fla_hip_dgetrf_incpiv(int m, int n, double* A, int lda, int b, int nb_alg, int* p, int ldp,
                          double* L, int ldl);
/* The forward substitution of those transformations on the m x nrhs X. */
int fla_hip_dfs_incpiv(int m, int n, const double* A, int lda, int b, int nb_alg, const int* p,
                       int ldp, const double* L, int ldl, int nrhs, double* X, int ldx);

/* ---- LAPACK auxiliaries on device matrices ----
   norm: '1' / 'O' (max column sum), 'I' (max row sum), 'M' (max |a_ij|),
   'F' / 'E' (Frobenius); *value receives the norm (synchronizes).  NaN
   propagates.  Replace dlange_ / dlansy_ (src/map/common/lapacksrc/f2c/
   dlange.c, dlansy.c) for device buffers; dlansy reads only the uplo
   triangle of the symmetric matrix. */
int fla_hip_dlange(int norm, int m, int n, const double* A, int lda, double* value);
int fla_hip_dlansy(int norm, int uplo, int n, const double* A, int lda, double* value);
/* B := A on the uplo trapezoid (FLA_LOWER_TRIANGULAR: i >= j,
   FLA_UPPER_TRIANGULAR: i <= j, anything else: all of A); dlacpy_
   (src/map/common/lapacksrc/f2c/dlacpy.c).  Asynchronous. */
int fla_hip_dlacpy(int uplo, int m, int n, const double* A, int lda, double* B, int ldb);
/* Off-diagonal entries of the uplo part (strictly lower / strictly upper /
   all) := alpha, diagonal := beta; dlaset_ (src/map/common/lapacksrc/f2c/
   dlaset.c).  Asynchronous. */
int fla_hip_dlaset(int uplo, int m, int n, double alpha, double beta, double* A, int lda);
/* p[i] += (i+1) (to LAPACK) or -= (i+1) (to FLAME) on a device int vector. */
int fla_hip_shift_pivots(int ptype, int m, int* p);

/* ---- tuning ---- */
/* Block sizes; 0 keeps current, -1 restores the default.  chol_nb: outer
   Cholesky block (default: chosen from n; else a multiple of
   128, <= 1024).  lu_nb: 32..128 (multiple of 32) = one panel of that width
   per step; 256 (default) = blocks of two 128-column panels, whose trailing
   update contracts over 256. */
int fla_hip_set_blocksizes(int chol_nb, int lu_nb);
int fla_hip_get_blocksizes(int* chol_nb, int* lu_nb);
/* Outer Cholesky block used for order n (chol_nb 0 = automatic: n/32
   rounded to a multiple of 128 within [256, 1024]). */
int fla_hip_chol_blocksize(int n);
/* Tile configuration of the MFMA update kernel (0 = default, 1..19 the
   measured alternatives; anything else is an argument error). */
int fla_hip_set_gemm_config(int cfg);
/* Width G (tile columns) of the groups in which update tiles are ordered, so
   that concurrently running tiles share operand panels in L2 (default 4). */
int fla_hip_set_gemm_group(int G);
/* CUs left out of the update stream's CU mask for the panel stream
   (default 8; 0 = no mask). */
int fla_hip_set_reserved_cus(int reserved);

/* ---- measurement ---- */
/* When enabled, the factorization drivers bracket every trailing-update
   launch with HIP events on its stream and accumulate, per kind, the
   algorithmic flops and the elapsed milliseconds.
   kind 0 = Cholesky trailing update (lower SYRK), 1 = LU trailing update
   (GEMM), 2 = LU panel, 3 = Cholesky panel chain, 4 = reduced-precision
   Cholesky trailing update (FP32 / BF16 MFMA), 5 = reduced-precision LU
   trailing update, 6 = LU row interchanges (time and launches; flops 0). */
int fla_hip_stats_enable(int on);
int fla_hip_stats_reset(void);
int fla_hip_stats_get(int kind, double* flops, double* ms, long* launches);

/* ---- mixed precision (dsposv_ semantics, src/map/common/lapacksrc/f2c/dsposv.c:192) ---- */
#define FLA_MP_FP32 0     /* FP32 trailing matrix, FP32 MFMA updates */
#define FLA_MP_BF16 1     /* FP32 trailing matrix, BF16 MFMA updates, FP32 accumulation */
#define FLA_IR_CLASSIC 0  /* x += (L L')^-1 r, as dsposv_ */
#define FLA_IR_GMRES 1    /* GMRES-IR: GMRES on A d = r, preconditioned by (L L')^-1 */
typedef struct {
  int precision;     /* FLA_MP_FP32 (default) or FLA_MP_BF16 */
  int refinement;    /* FLA_IR_CLASSIC (default) or FLA_IR_GMRES */
  int max_iter;      /* refinement steps before the FP64 fall-back (default 30) */
  int gmres_restart; /* Krylov dimension per GMRES-IR step (default 20) */
  double gmres_tol;  /* relative preconditioned residual per GMRES-IR step (default 1e-4) */
  double stop_factor; /* stop when max|r_j| <= max|x_j| * ||A||_inf * eps * stop_factor for
                         every column; 0 (default) = sqrt(n), dsposv_'s rule */
} fla_hip_mp_opts;
/* Reduced-precision Cholesky of the symmetric A (uplo triangle): trailing
   updates in FP32 or BF16 on MFMA, each panel factored in FP64 from the
   reduced-precision trailing values.  Writes the lower factor (FP64 values)
   into L; L's strictly upper part is not referenced.  *info = FLA_SUCCESS or
   the 0-based failing column.  FLA_HIP_ERANGE if A overflows FP32. */
int fla_hip_mp_chol(int uplo, int n, const double* A, int lda, int precision, double* L,
                    int ldl, int* info);
/* Solves A X = B for SPD A (uplo triangle) by the reduced-precision factor
   and refinement with an FP64 residual, with dsposv_'s stopping rule,
   iteration limit and FP64 fall-back.  *iter: number of refinement steps,
   or -2 (overflow), -3 (factorization failed), -31 (no convergence), in
   which case A holds the FP64 factor.  *info: 0, or k > 0 when the FP64
   fall-back found the leading minor of order k not positive definite.
   opts may be NULL (FP32, classic refinement, 30 steps). */
int fla_hip_dsposv(int uplo, int n, int nrhs, double* A, int lda, const double* B, int ldb,
                   double* X, int ldx, const fla_hip_mp_opts* opts, int* iter, int* info);
/* Reduced-precision LU with partial pivoting of the general n x n A
   (dsgesv_'s factorization, src/map/common/lapacksrc/f2c/dsgesv.c:314):
   FP32 trailing matrix, FP64 panels (the LU panel kernel) and
   U12 := inv(L11) A12 in FP64, A22 -= L21 U12 in FP32 or BF16 on MFMA.
   F receives the factors (FP64 values, unit L below the diagonal), p the
   FLAME-native pivots (device, n ints).  *info = FLA_SUCCESS or the 0-based
   column of the first null pivot.  FLA_HIP_ERANGE if A overflows FP32. */
int fla_hip_mp_lu(int n, const double* A, int lda, int precision, double* F, int ldf, int* p,
                  int* info);
/* Solves A X = B for general A by the reduced-precision LU and refinement
   with an FP64 residual: dsgesv_ (src/map/common/lapacksrc/f2c/dsgesv.c:189)
   with the iteration codes, limit and FP64 fall-back of fla_hip_dsposv.  p
   (device, n ints) receives the FLAME-native pivots of the factorization
   used: the reduced-precision one, or after a fall-back the FP64 dgetrf's,
   with A then overwritten by the FP64 factors.  *info: 0, or k > 0 when
   U(k-1, k-1) of the FP64 fall-back is exactly zero. */
int fla_hip_dsgesv(int n, int nrhs, double* A, int lda, int* p, const double* B, int ldb,
                   double* X, int ldx, const fla_hip_mp_opts* opts, int* iter, int* info);
/* Timings and counters of the last fla_hip_dsposv / fla_hip_dsgesv: [0] factor ms (incl.
   conversion and norm), [1] refinement ms, [2] preconditioner solves,
   [3] FP64 products with A, [4] ||A||_inf, [5] GMRES iterations, [6] the
   reduced-precision factorization's info. */
int fla_hip_mp_last_stats(double* out, int nout);
/* Panel solve of the reduced-precision factor: 0 blocked substitution,
   1 (default) multiplication by the inverse of each 128 x 128 diagonal
   factor. */
int fla_hip_mp_set_panel(int inverse);
/* CUs left out of the reduced-precision factor's update stream for its panel
   chain (default 64; 0 = no mask). */
int fla_hip_mp_set_reserved_cus(int reserved);
/* Outer block of the reduced-precision factor (multiple of 128; 0 = the
   FP64 Cholesky's choice for the order). */
int fla_hip_mp_set_blocksize(int nb);

/* ---- batched small factorizations and solves (no counterpart in the
   reference: thousands of independent matrices of uneven sizes in one
   persistent launch with a device work queue) ----
   Host arrays n[], lda[], ldb[], nrhs[] and host arrays of DEVICE pointers
   A[], B[], ipiv[]; info is a DEVICE int array[count] (FLA_SUCCESS or the
   0-based failing column / first null pivot).  n[i] <= 512, nrhs[i] <= 256.
   B may be NULL (factor only); with B, B[i] is overwritten by the solution
   of A_i X = B_i.  Cholesky: lower triangle, FLA_Chol semantics per matrix;
   LU: partial pivoting with FLAME-native pivots (FLA_LU_piv semantics).
   Asynchronous on the library's stream; repeated calls with the same count
   and configuration replay a cached hipGraph. */
#define FLA_HIP_BATCH_MAXN 512
int fla_hip_dposv_batched(int count, const int* n, double* const* A, const int* lda,
                          double* const* B, const int* ldb, const int* nrhs, int* info);
int fla_hip_sposv_batched(int count, const int* n, float* const* A, const int* lda,
                          float* const* B, const int* ldb, const int* nrhs, int* info);
int fla_hip_dgesv_batched(int count, const int* n, double* const* A, const int* lda,
                          int* const* ipiv, double* const* B, const int* ldb, const int* nrhs,
                          int* info);
/* Autotuner: kind 0 = dposv, 1 = sposv, 2 = dgesv.  Times the candidate
   configurations (panel width, workgroups per CU) on synthetic matrices of
   these sizes and keeps the fastest for batches with the same size
   histogram.  Outputs may be NULL. */
int fla_hip_batched_tune(int kind, int count, const int* n, const int* nrhs, int* best_nb,
                         int* best_wg_per_cu, double* best_ms);
/* Forces one configuration for every batch of that kind (nb: 16 or 32 for
   kinds 0 and 1, 16 for kind 2; 1 <= wg_per_cu <= 4). */
int fla_hip_batched_set_config(int kind, int nb, int wg_per_cu);
int fla_hip_batched_set_graph(int on);

/* ---- adjoints (no counterpart in the reference) ----
   Reverse mode of A = L L' (L lower, n x n): given Lbar (lower), Abar :=
   (G + G')/2 with G = L^-T Phi(L' Lbar) L^-1, Phi = lower triangle with the
   diagonal halved, so df = <Abar, dA> for symmetric dA.  The strictly upper
   parts of L and Lbar are not referenced; Abar is written in full. */
int fla_hip_dchol_adjoint(int n, const double* L, int ldl, const double* Lbar, int ldlb,
                          double* Abar, int ldab);
/* Reverse mode of X = A^-1 B with A = L L': Bbar := A^-1 Xbar and
   Abar -= (Bbar X' + X Bbar')/2 (Abar accumulated). */
int fla_hip_dchol_solve_adjoint(int n, int nrhs, const double* L, int ldl, const double* X,
                                int ldx, const double* Xbar, int ldxb, double* Bbar, int ldbb,
                                double* Abar, int ldab);
/* Gaussian process with K = s^2 exp(-|x-x'|^2 / (2 l^2)) + noise*I on n
   points x (n x d, column-major): the kernel matrix; the contraction
   g = (<Kbar, dK/dl>, <Kbar, dK/ds>) into host g[2]; and the log marginal
   likelihood of y (device vector) with its gradient in (l, s) (host lml,
   grad[2]) through the Cholesky and Cholesky-solve adjoints.  info is the
   Cholesky's (FLA_SUCCESS or the failing column). */
int fla_hip_gp_rbf_kernel(int n, int d, const double* x, int ldx, double l, double s,
                          double noise, double* K, int ldk);
int fla_hip_gp_rbf_grad(int n, int d, const double* x, int ldx, double l, double s,
                        const double* Kbar, int ldkb, double* g);
int fla_hip_gp_rbf_lml_grad(int n, int d, const double* x, int ldx, const double* y, double l,
                            double s, double noise, double* lml, double* grad, int* info);
/* Batched FP32 Cholesky adjoint (no counterpart in the reference): for
   count problems of order n <= 256 stored as consecutive column-major n x n
   blocks (ld n), Abar_i = Phi-sym(L_i^-T Phi(L_i^T Lbar_i) L_i^-1); only the
   lower triangles of L_i and Lbar_i are read.  One workgroup per problem. */
int fla_hip_schol_adjoint_batched(int count, int n, const float* L, const float* Lbar,
                                  float* Abar);
/* Batched FP32 GP log marginal likelihood and its gradient in (l_i, s) for
   count problems with n <= 256 points in d <= 8 dimensions: x = count
   consecutive n x d column-major blocks (ld n), y = count vectors of n,
   l = count length scales; K_i = s^2 exp(-|x-x'|^2/(2 l_i^2)) + noise I is
   factored by fla_hip_sposv_batched, Kbar_i assembled by the batched
   Cholesky adjoint (Lbar = -diag(1/L_jj)) plus the Cholesky-solve adjoint.
   Device outputs: lml[count], grad[2*count] = (d/dl_i, d/ds) pairs, and
   info[count] (FLA_SUCCESS, or the failing column with NaN outputs). */
int fla_hip_sgp_rbf_lml_grad_batched(int count, int n, int d, const float* x, const float* y,
                                     const float* l, float s, float noise, float* lml,
                                     float* grad, int* info);

/* ---- synthetic inputs and checks (device side) ---- */
/* Counter-based generators: identical values to the CPU generators. */
int fla_hip_gen_uniform(unsigned long long seed, int m, int n, double* A, int lda);
int fla_hip_gen_normal(unsigned long long seed, int m, int n, double* A, int lda);
/* A := scale*B*B' + shift*I (lower triangle and diagonal written). */
int fla_hip_gram_lower(int n, int k, const double* B, int ldb, double scale,
                       double shift, double* A, int lda);
/* SPD A = Q diag(s) Q' with s log-spaced on [1, kappa] and Q a product of
   nrefl Householder reflectors with N(0,1) vectors (full matrix written). */
int fla_hip_gen_spd_reflectors(unsigned long long seed, int n, int nrefl, double kappa,
                               double* A, int lda);
/* General A = Q1 diag(s) Q2' with the same s, Q1 a product of nrefl
   reflectors drawn from seed and Q2 an independent product drawn from seed2. */
int fla_hip_gen_ge_reflectors(unsigned long long seed, unsigned long long seed2, int n, int nrefl,
                              double kappa, double* A, int lda);
/* S := float(A), m x n (no overlap). */
int fla_hip_convert_d2s(int m, int n, const double* A, int lda, float* S, int lds);
/* A(i,:) *= 10^(-i/max(m,n)) */
int fla_hip_scale_rows_logspace(int m, int n, double* A, int lda);
/* Y := op(M)*X + beta*Y with M given by its uplo triangle as a symmetric
   matrix (sym=1), a triangular factor (sym=0, unit diag per `diag`), or a
   full matrix (uplo = FLA_FULL_MATRIX).  FP64, used by residual checks. */
int fla_hip_dmatmul_check(int kind, int uplo, int trans, int diag, int n, int nrhs,
                          const double* M, int ldm, const double* X, int ldx, double beta,
                          double* Y, int ldy);

#ifdef __cplusplus
}
#endif
#endif
