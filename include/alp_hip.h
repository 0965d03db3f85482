/*
 * alp_hip.h -- C ABI of the MI355X-native `hip` backend for ALP/GraphBLAS.
 *
 * This is the thin boundary between host code (the grb:: C++ layer in
 * alp_amd/include/graphblas/hip/, the sparse_cg_* transition ABI in
 * alp_hip_solver.h, or any FFI caller) and the hand-written gfx950 kernels in
 * libalp_hip.so. Signatures carry plain pointers and sizes only.
 *
 * Conventions
 *   - Every function returns an alp_rc_t whose values equal grb::RC
 *     (include/graphblas/rc.hpp:47-155 in ALP): 0 = SUCCESS.
 *   - Pointers named d_* are device pointers (from alp_hip_malloc or any HIP
 *     allocation on the current device); h_* are host pointers.
 *   - Matrices are opaque handles owning device-resident compressed storage in
 *     ALP's layout (reference/compressed_storage.hpp:48-77): uint64 offsets,
 *     uint32 indices, values; CRS always, CCS when requested.
 *   - All work is enqueued on the library's stream (alp_hip_set_stream); a call
 *     that returns a host scalar synchronises that stream. Calls are blocking
 *     with respect to host-visible results, as ALP's reference backends are
 *     (reference/properties.hpp: isBlockingExecution).
 *   - Calls that share containers are not re-entrant, as in ALP.
 */
#ifndef ALP_HIP_H
#define ALP_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* grb::RC (include/graphblas/rc.hpp:47-155) */
typedef enum {
	ALP_SUCCESS = 0,
	ALP_PANIC = 1,
	ALP_OUTOFMEM = 2,
	ALP_MISMATCH = 3,
	ALP_OVERLAP = 4,
	ALP_OVERFLW = 5,
	ALP_UNSUPPORTED = 6,
	ALP_ILLEGAL = 7,
	ALP_FAILED = 8
} alp_rc_t;

/* semirings, each the device form of a grb::Semiring (include/graphblas/semiring.hpp:190) */
typedef enum {
	ALP_PLUS_TIMES_F64 = 0,     /* add<double>, mul<double>, zero, one */
	ALP_PLUS_TIMES_F32 = 1,     /* add<float>, mul<float>, zero, one */
	ALP_LOR_LAND_BOOL = 2,      /* logical_or<bool>, logical_and<bool>, false, true */
	ALP_PLUS_PATTERN_F64 = 3,   /* add<double>, left_assign_if<double,bool> over a pattern
	                               matrix (simple_pagerank.hpp:224-237; vxm over void) */
	ALP_MIN_PLUS_F64 = 4        /* min<double>, add<double>, infinity, zero */
} alp_semiring_t;

/* descriptor bits, same values as grb::descriptors (include/graphblas/descriptors.hpp:63-267) */
#define ALP_DESC_INVERT_MASK 1u
#define ALP_DESC_TRANSPOSE_MATRIX 2u
#define ALP_DESC_STRUCTURAL 8u
#define ALP_DESC_DENSE 16u
#define ALP_DESC_ADD_IDENTITY 32u
#define ALP_DESC_SAFE_OVERLAP 1024u
#define ALP_DESC_TRANSPOSE_LEFT 2048u
#define ALP_DESC_TRANSPOSE_RIGHT 4096u
/* library-only bit: the output is known to be all zero (a grb::set(y,0) folded into
 * the product), so y is written without being read */
#define ALP_DESC_OVERWRITE (1u << 20)

typedef struct alp_hip_matrix_s *alp_hip_matrix_t;

/* ---- runtime: grb::init / grb::finalize (src/graphblas/reference/init.cpp:46,81) ---- */
alp_rc_t alp_hip_init(int device);
alp_rc_t alp_hip_finalize(void);
alp_rc_t alp_hip_set_stream(void *hip_stream); /* NULL = the library's own stream */
alp_rc_t alp_hip_sync(void);
const char *alp_hip_version(void);
/* number of gfx950 devices; 0 without a GPU (never initialises the runtime's device) */
int alp_hip_device_count(void);

/* ---- timing on the library stream (the role of grb::utils::Timer,
 *      include/graphblas/utils/timer.hpp, for device work) ----
 * alp_hip_timer_start records a HIP event on the current stream; alp_hip_timer_stop
 * records a second one, waits for it and returns the elapsed milliseconds. */
alp_rc_t alp_hip_timer_start(void);
alp_rc_t alp_hip_timer_stop(double *ms);

/* ---- device memory ---- */
alp_rc_t alp_hip_malloc(void **d_ptr, size_t bytes);
alp_rc_t alp_hip_free(void *d_ptr);
alp_rc_t alp_hip_memcpy_h2d(void *d_dst, const void *h_src, size_t bytes);
alp_rc_t alp_hip_memcpy_d2h(void *h_dst, const void *d_src, size_t bytes);
alp_rc_t alp_hip_memcpy_d2d(void *d_dst, const void *d_src, size_t bytes);

/* ---- matrices (Matrix<D,reference>, reference/matrix.hpp:1018) ---- */
/* value_type: 0 = f64, 1 = f32, 2 = pattern (Matrix<void>) */
#define ALP_VAL_F64 0
#define ALP_VAL_F32 1
#define ALP_VAL_PATTERN 2
/* flags */
#define ALP_MAT_WITH_CCS 1u      /* also build the CCS (transpose / vxm gathers) */
#define ALP_MAT_WITH_F32 2u      /* also keep an f32 copy of f64 values (mixed precision) */

/* from host CRS arrays (copied; the caller keeps ownership of its arrays) */
alp_rc_t alp_hip_matrix_create_crs(alp_hip_matrix_t *A, size_t nrows, size_t ncols, size_t nnz,
                                   const uint64_t *h_off, const uint32_t *h_idx,
                                   const void *h_val, int value_type, unsigned flags);
/* from host coordinates, deduplicated by the caller: buildMatrixUnique
 * (reference/io.hpp:2304 -> matrix.hpp:1854-1951); the counting sort runs on the
 * device and reproduces the reference's within-row order (reverse input order) */
alp_rc_t alp_hip_matrix_create_coo(alp_hip_matrix_t *A, size_t nrows, size_t ncols, size_t nnz,
                                   const uint32_t *h_I, const uint32_t *h_J, const void *h_V,
                                   int value_type, unsigned flags);
/* MatrixMarket coordinate files (utils/parser/matrixFileReader.hpp,
 * matrixFileIterator.hpp): real / double / integer / pattern fields, general or
 * symmetric storage, 1-based indices. A symmetric file's off-diagonal entry is
 * followed by its mirror (matrixFileIterator.hpp:411-436). Complex, hermitian,
 * skew-symmetric and dense "array" files give ALP_UNSUPPORTED; a malformed file
 * (bad banner, index out of range, too few or too many entries) ALP_ILLEGAL.
 * alp_hip_mtx_read parses on the host: with h_I == NULL it reports the sizes (nnz
 * after expansion, the file's entry count; entries and kind may be NULL) and the
 * kind (bits ALP_MTX_PATTERN, ALP_MTX_SYMMETRIC); with arrays of *nnz entries it
 * fills them in that order (h_V may be NULL; pattern files give 1.0). */
#define ALP_MTX_PATTERN 1
#define ALP_MTX_SYMMETRIC 2
alp_rc_t alp_hip_mtx_read(const char *path, size_t *nrows, size_t *ncols, size_t *nnz, size_t *entries,
                          uint32_t *h_I, uint32_t *h_J, double *h_V, int *kind);
/* the file's matrix built as alp_hip_matrix_create_coo from those entries, in
 * that order (buildMatrixUnique from the file iterator, reference/io.hpp:2304);
 * value_type ALP_VAL_PATTERN keeps no values, F64 / F32 take the file's values
 * (1 for a pattern file) */
alp_rc_t alp_hip_matrix_read_mtx(alp_hip_matrix_t *A, const char *path, int value_type, unsigned flags);
/* the 27-point stencil of build_ndims_system_matrix (algorithms/hpcg/
 * matrix_building_utils.hpp:60-73), generated on the device */
alp_rc_t alp_hip_matrix_stencil27(alp_hip_matrix_t *A, size_t nx, size_t ny, size_t nz,
                                  double diag, double offdiag, unsigned flags);
alp_rc_t alp_hip_matrix_destroy(alp_hip_matrix_t A);
alp_rc_t alp_hip_matrix_info(alp_hip_matrix_t A, size_t *nrows, size_t *ncols, size_t *nnz);
/* download CRS (which = 0) or CCS (which = 1); h_val may be NULL */
alp_rc_t alp_hip_matrix_download(alp_hip_matrix_t A, int which, uint64_t *h_off,
                                 uint32_t *h_idx, void *h_val);
/* select the SpMV kernel: 0 = auto, 1 = CSR-vector (`lanes` per row: 2..64, `rows`
 * per lane group: 1, 2 or 4; 0 = default), 2 = CSR-stream (`rows` picks the launch:
 * 1 = persistent grid 4 blocks/CU, 2 = persistent 8 blocks/CU, 4 = a block per tile,
 * 8 = a block per tile in plain block order, 16 / 32 = software-pipelined persistent
 * grid with 4 / 2 blocks per CU, 64 = the tile's indices and values staged in LDS and
 * the rows gathered lane by lane; `lanes` 16 = 4096-nonzero tiles),
 * 3 = CSR-wave (a 512-nonzero tile per wavefront), 4 = sliced ELL (64-row slices,
 * one lane per row; `lanes` 0/1 = plain / XCD-contiguous slice order, `rows` 4, 8
 * or 16 entries in flight), 5 = the sliced ELL with its index-free diagonal slices
 * (same arguments). 4 and 5 exist only for matrices with regular rows, 5 only when
 * most slices share one column-offset pattern per slice (stencils). */
alp_rc_t alp_hip_matrix_set_kernel(alp_hip_matrix_t A, int kernel, int lanes, int rows);
alp_rc_t alp_hip_matrix_get_kernel(alp_hip_matrix_t A, int *kernel, int *lanes, int *rows);
/* time each candidate SpMV kernel on this matrix and keep the fastest */
alp_rc_t alp_hip_matrix_autotune(alp_hip_matrix_t A);
/* the sliced layouts built for A's CRS (which = 0) or CCS (which = 1): *slices
 * 64-row slices holding *entries padded entries (0, 0: none built), of which
 * *dia_slices are in index-free diagonal form (no reference counterpart: the
 * reference stores CRS/CCS only; this reports what the kernels stream) */
alp_rc_t alp_hip_matrix_layout(alp_hip_matrix_t A, int which, size_t *slices, size_t *entries,
                               size_t *dia_slices);

/* ---- device storage, for callers that run kernels of their own on it ----
 * (the grb:: header layer instantiates its generic semiring kernels this way) */
/* the library's stream (work enqueued there is ordered with every library call) */
alp_rc_t alp_hip_get_stream(void **hip_stream);
/* device arrays of A's CRS (which = 0) or CCS (which = 1, built on first request):
 * internal::getCRS / getCCS (reference/matrix.hpp:72-90); *d_val is NULL for a pattern
 * matrix. The pointers stay valid until A is destroyed. */
alp_rc_t alp_hip_matrix_storage(alp_hip_matrix_t A, int which, const uint64_t **d_off,
                                const uint32_t **d_idx, const void **d_val);
/* bytes of device memory A holds: CRS, CCS and every derived layout (sliced ELL,
 * diagonal slices, FP32 copies, tile metadata) */
alp_rc_t alp_hip_matrix_device_bytes(alp_hip_matrix_t A, size_t *bytes);
/* buildMatrixUnique of a pattern matrix from host coordinates (as
 * alp_hip_matrix_create_coo, reference/matrix.hpp:1854-1951) that also reports where
 * each input entry went: d_crs_src[p] = the input index stored at CRS position p, and,
 * with ALP_MAT_WITH_CCS, d_ccs_src[p] likewise for the CCS (device arrays of nnz
 * entries; either may be NULL). A caller keeps values of any type in that order. */
alp_rc_t alp_hip_matrix_create_coo_map(alp_hip_matrix_t *A, size_t nrows, size_t ncols, size_t nnz,
                                       const uint32_t *h_I, const uint32_t *h_J, unsigned flags,
                                       uint64_t *d_crs_src, uint64_t *d_ccs_src);
/* a matrix holding a copy of device CRS arrays (d_val NULL for a pattern); with
 * ALP_MAT_WITH_CCS its CCS lists each column's rows in ascending order, as
 * mxm_generic fills an output (reference/blas3.hpp:1622-1636), and d_ccs_src (may be
 * NULL) gets each CCS entry's CRS position */
alp_rc_t alp_hip_matrix_create_device(alp_hip_matrix_t *A, size_t nrows, size_t ncols, size_t nnz,
                                      const uint64_t *d_off, const uint32_t *d_idx, const void *d_val,
                                      int value_type, unsigned flags, uint64_t *d_ccs_src);

/* ---- level-3: mxm (reference/blas3.hpp:1148 mxm_generic, :1744 mxm) ----
 * *C = *C (+) op(A) (x) op(B), op = transpose with ALP_DESC_TRANSPOSE_LEFT / _RIGHT.
 * When *C is NULL or holds no entries the product is written to a new matrix; when
 * it holds entries they are updated in place (their positions come first in each
 * row, products fold into them), and *C is replaced by the result. Rows list their
 * columns in the order the reference's SPA first reaches them; each entry folds its
 * products in the reference's order, so results are the reference's bit for bit.
 * Semirings: ALP_PLUS_TIMES_F64 and ALP_MIN_PLUS_F64 (double or pattern operands,
 * a pattern entry being the multiplicative identity), ALP_PLUS_TIMES_F32 (float or
 * pattern), ALP_LOR_LAND_BOOL (pattern operands, pattern result). The result has
 * a CCS when A or B had one (ALP_MAT_WITH_CCS). */
alp_rc_t alp_hip_mxm(alp_hip_matrix_t *C, alp_hip_matrix_t A, alp_hip_matrix_t B, alp_semiring_t ring,
                     unsigned descr);

/* ---- level-2: mxv / vxm (reference/blas2.hpp:870-2200) ----
 * y (+)= A (x) x  (mxv; ALP_DESC_TRANSPOSE_MATRIX gathers over the CCS instead)
 * d_mask (nullable) is an output mask of bytes (non-zero = true); with
 * ALP_DESC_INVERT_MASK rows with a true mask are skipped. Element types follow the
 * semiring: f64, f32, or bytes (0/1) for ALP_LOR_LAND_BOOL. */
alp_rc_t alp_hip_mxv(alp_hip_matrix_t A, alp_semiring_t ring, unsigned descr, void *d_y,
                     const void *d_x, const uint8_t *d_mask);
/* y (+)= x (x) A   (vxm: the CCS gather; ALP_DESC_TRANSPOSE_MATRIX gathers over CRS) */
alp_rc_t alp_hip_vxm(alp_hip_matrix_t A, alp_semiring_t ring, unsigned descr, void *d_y,
                     const void *d_x, const uint8_t *d_mask);

/* ---- level-1 (reference/blas1.hpp), f64 ---- */
/* z = z + dot(x, y) with a fixed, deterministic reduction tree */
alp_rc_t alp_hip_dot_f64(size_t n, const double *d_x, const double *d_y, double *h_z);
alp_rc_t alp_hip_dot_f32(size_t n, const float *d_x, const float *d_y, double *h_z);
alp_rc_t alp_hip_set_f64(size_t n, double *d_x, double v);
alp_rc_t alp_hip_set_f32(size_t n, float *d_x, float v);
alp_rc_t alp_hip_set_u8(size_t n, uint8_t *d_x, uint8_t v);
/* y = alpha * x + y   (eWiseMul(y, alpha, x, ring) / eWiseMulAdd, blas1.hpp:8684,6566) */
alp_rc_t alp_hip_axpy_f64(size_t n, double alpha, const double *d_x, double *d_y);
/* z = alpha * x + y   (eWiseMulAdd into a cleared z, hpcg.hpp:179-180) */
alp_rc_t alp_hip_xpay_f64(size_t n, double alpha, const double *d_x, const double *d_y,
                          double *d_z);
/* z = x op y, op: 0 add, 1 subtract, 2 multiply, 3 divide (eWiseApply, blas1.hpp) */
alp_rc_t alp_hip_ewise_f64(size_t n, int op, const double *d_x, const double *d_y, double *d_z);
/* x = alpha * x  (foldr(alpha, x, mulMonoid), blas1.hpp:1685) */
alp_rc_t alp_hip_scale_f64(size_t n, double alpha, double *d_x);
/* h_z = z + sum(x)  (foldl(z, x, addMonoid), blas1.hpp:2765) */
alp_rc_t alp_hip_sum_f64(size_t n, const double *d_x, double *h_z);
alp_rc_t alp_hip_convert_f64_to_f32(size_t n, const double *d_x, float *d_y);
alp_rc_t alp_hip_convert_f32_to_f64(size_t n, const float *d_x, double *d_y);

/* ---- CG (algorithms/conjugate_gradient.hpp:191-484), device resident ----
 * Solves A x = b, stopping when ||r|| < tol*||b|| or after max_iter iterations;
 * *residual = ||r||^2 (the reference's convention), *iters = iterations done.
 * Returns ALP_FAILED when not converged, as the reference does.
 * precision: 0 = FP64 throughout; 1 = CG iterations on A's FP32 copy with FP32
 * r, p, Ap (each product row summed in FP64, rounded once) and the solution
 * update x += alpha p done in FP64 on the FP64 x every iteration, with reliable
 * updates: whenever the FP32 recurrence residual has fallen by 100x since the
 * last update, r = b - A x is formed in FP64 and CG continues from it with its
 * search direction kept; convergence is decided on the FP64 residual, and
 * *iters counts the FP32 iterations;
 * 2 = restarted FP32 inner CG solves (each to 1e-5 relative) inside FP64
 * iterative refinement;
 * 3 = as 1, but a segment's increments of x accumulate in FP32 and are added to
 * the FP64 x before each FP64 residual (a group-wise update: 8 fewer bytes per
 * row and iteration). Modes 1-3 use A's FP32 copy (made on first use). */
alp_rc_t alp_hip_cg(alp_hip_matrix_t A, const double *d_b, double *d_x, size_t max_iter,
                    double tol, int precision, size_t *iters, double *residual);

/* CG on A^T (gathering over the CCS): the adjoint solve A^T lambda = g; same
 * arguments and conventions as alp_hip_cg */
alp_rc_t alp_hip_cg_transposed(alp_hip_matrix_t A, const double *d_b, double *d_x, size_t max_iter,
                               double tol, int precision, size_t *iters, double *residual);

/* FP64 CG that allocates nothing: no counterpart in the reference. It is for
 * callers that cannot allocate device memory (code under stream capture, tight
 * serving loops, harnesses that forbid hipMalloc). Same arguments, results and
 * return codes as alp_hip_cg with precision 0 (bit-identical x and iteration
 * count), except:
 *   - its work vectors, state and partial sums live in one file-scope device
 *     array that the library holds on each device until the process ends
 *     (24 MiB + 33 KiB); A may have at most 2^20 rows (1,048,576), larger
 *     systems return ALP_OUTOFMEM;
 *   - the library must be initialised first (alp_hip_init allocates; without
 *     it the call returns ALP_PANIC);
 *   - every call on a device shares that array, so calls that overlap on one
 *     device (other host threads, other streams) race: serialise them.
 * The matrix and the vectors d_b, d_x are the caller's. */
alp_rc_t alp_hip_cg_noalloc(alp_hip_matrix_t A, const double *d_b, double *d_x, size_t max_iter,
                            double tol, size_t *iters, double *residual);

/* ---- differentiable solve (no counterpart in the reference) ----
 * A(theta): the 27-point stencil graph's weighted Laplacian plus `shift` on the
 * diagonal (A_ij = -theta_e for the edge e = (i, j), A_ii = sum of theta at i + shift).
 * theta holds 13 conductances per vertex, slot d for the d-th positive neighbour
 * offset in the order dz in {0,1}, dy in {-1,0,1}, dx in {-1,0,1} (positive: dz > 0,
 * or dz = 0 and dy > 0, or dz = dy = 0 and dx > 0); out-of-grid slots are unused. */
alp_rc_t alp_hip_matrix_conductance27(alp_hip_matrix_t *A, size_t nx, size_t ny, size_t nz,
                                      const double *d_theta, double shift, unsigned flags);
/* dL/dtheta (13 per vertex, 0 in unused slots) = -(lambda_i - lambda_j)(x_i - x_j) */
alp_rc_t alp_hip_grad_conductance27(size_t nx, size_t ny, size_t nz, const double *d_lambda,
                                    const double *d_x, double *d_dtheta);
/* dL/dA = -lambda x^T on A's pattern, one value per stored CRS entry */
alp_rc_t alp_hip_grad_matrix(alp_hip_matrix_t A, const double *d_lambda, const double *d_x,
                             double *d_dA);

/* ---- HPCG (algorithms/hpcg/hpcg.hpp:97-223, multigrid_v_cycle.hpp, red_black_gauss_seidel.hpp,
 *      system_building_utils.hpp:84-150) ----
 * A handle owns the whole multigrid hierarchy of the nx*ny*nz 27-point problem
 * (diagonal 26, off-diagonals -1), built on the device as build_hpcg_system does:
 * coarsening step 2, minimum side 4, at most max_levels coarse levels, injection
 * coarseners. colours = k > 1 colours rows by index mod k (build_static_color_masks,
 * matrix_building_utils.hpp:139-165); colours = 0 uses the 8-colour parity colouring
 * (x, y, z mod 2). presmooth/postsmooth: symmetric Gauss-Seidel sweeps per level.
 * flags: ALP_MAT_WITH_F32 also keeps FP32 copies for precision = 1. */
typedef struct alp_hip_hpcg_s *alp_hip_hpcg_t;
alp_rc_t alp_hip_hpcg_create(alp_hip_hpcg_t *H, size_t nx, size_t ny, size_t nz, size_t max_levels,
                             size_t colours, size_t presmooth, size_t postsmooth, unsigned flags);
alp_rc_t alp_hip_hpcg_destroy(alp_hip_hpcg_t H);
/* n, nnz of the finest level, number of grids (finest included), and whether every
 * level's colouring is proper (then colour steps run fused) */
alp_rc_t alp_hip_hpcg_info(alp_hip_hpcg_t H, size_t *n, size_t *nnz, size_t *levels, int *proper);
/* z = M r: one V-cycle of the preconditioner (FP64, original row numbering) */
alp_rc_t alp_hip_hpcg_vcycle(alp_hip_hpcg_t H, const double *d_r, double *d_z);
/* Solve A x = b from the given x: stop when ||r||/||r0|| <= tol or after max_iter
 * iterations; *residual = ||r||. precision 0 = FP64 throughout; 1 = PCG iterations
 * in FP32 (V-cycle, r, z, p, A p with FP64 row sums) on an FP64 x, with reliable
 * updates: whenever the FP32 recurrence residual has fallen by 100x, r = b - A x is
 * formed in FP64 and replaces it (*outer = updates; the stopping test and
 * *residual are FP64); 2 = FP64 CG (SpMV, dots, residual and solution updates in
 * FP64) whose preconditioner, the whole V-cycle (smoother, restriction SpMV,
 * transfers), runs in FP32; 3 = FP32 SpMV, smoother and transfers in an inner FP32
 * PCG restarted inside FP64 iterative refinement (*iters = inner iterations,
 * *outer = refinement steps). outer may be NULL.
 * with_precond = 0 runs hpcg's unpreconditioned branch. */
alp_rc_t alp_hip_hpcg_solve(alp_hip_hpcg_t H, const double *d_b, double *d_x, size_t max_iter,
                            double tol, int precision, int with_precond, size_t *iters,
                            double *residual, size_t *outer);

/* ---- graphs (pattern matrices, Matrix<void>) ---- */
#define ALP_GRAPH_SYMMETRIZE 1u       /* add (j, i) for every edge (i, j) */
#define ALP_GRAPH_DROP_SELF_LOOPS 2u  /* drop edges (i, i) */
/* Graph500-style R-MAT edges (no counterpart in ALP): edge e picks, at each of
 * `scale` levels, a quadrant with probabilities a, b, c, 1-a-b-c from a counter-based
 * generator seeded by `seed`; deterministic for any launch shape. */
alp_rc_t alp_hip_rmat_edges(unsigned scale, size_t nedges, double a, double b, double c,
                            uint64_t seed, uint32_t *d_src, uint32_t *d_dst);
/* an n x n pattern matrix (CRS and CCS) from device edge lists; duplicates are
 * removed (buildMatrixUnique's uniqueness, reference/io.hpp:2304, established on the
 * device by sort + unique); rows list columns in ascending order */
alp_rc_t alp_hip_graph_from_edges(alp_hip_matrix_t *A, size_t n, size_t nedges,
                                  const uint32_t *d_src, const uint32_t *d_dst, unsigned flags,
                                  size_t *nnz);
/* the first vertex of maximum out-degree */
alp_rc_t alp_hip_max_degree_vertex(alp_hip_matrix_t A, uint32_t *vertex, size_t *degree);
/* breadth-first levels from `source` over the lor-land semiring: d_level[v] = hops
 * from source, -1 if unreachable; push/pull switched per level like vxm_generic's
 * scatter/gather choice (reference/blas2.hpp:1108-1132, 1311-1334) */
alp_rc_t alp_hip_bfs(alp_hip_matrix_t A, uint32_t source, int32_t *d_level, size_t *nlevels,
                     size_t *reached);
/* simple_pagerank (algorithms/simple_pagerank.hpp:146-533): d_pr (n doubles) gets the
 * ranks; returns ALP_FAILED when `conv` was not reached within max_iter */
alp_rc_t alp_hip_pagerank(alp_hip_matrix_t L, double *d_pr, double alpha, double conv,
                          size_t max_iter, size_t *iters, double *quality);
/* knn (algorithms/knn.hpp:82-140): d_out[v] = 1 iff v is within k hops of source.
 * A matrix built by alp_hip_graph_from_edges with ALP_GRAPH_SYMMETRIZE answers by a
 * breadth-first search stopped at depth k; any other matrix by k lor-land products,
 * as the reference does (same result on a symmetric pattern) */
alp_rc_t alp_hip_khop(alp_hip_matrix_t A, uint32_t source, size_t k, uint8_t *d_out);

/* ---- batched multi-right-hand-side CG (no counterpart in the reference; each
 *      instance follows conjugate_gradient.hpp:191-484) ----
 * nsys anisotropic 7-point Laplacians on nx*ny*nz grids with per-axis coefficients
 * (diagonal 2(cx+cy+cz), off-diagonals -cx, -cy, -cz; Dirichlet boundaries), stacked
 * block-diagonally; nrhs in {1, 2, 4, 8, 16} right-hand sides per system. Blocks of
 * vectors are row-major: entry (row, k) at row * nrhs + k. */
typedef struct alp_hip_batch_s *alp_hip_batch_t;
alp_rc_t alp_hip_batch_create_laplace7(alp_hip_batch_t *B, size_t nsys, const uint32_t *h_nx,
                                       const uint32_t *h_ny, const uint32_t *h_nz,
                                       const double *h_cx, const double *h_cy, const double *h_cz,
                                       size_t nrhs);
/* nsys caller-supplied systems as one stacked CRS (host arrays): system s owns
 * rows [h_row0[s], h_row0[s+1]) (h_row0[0] = 0, every system non-empty);
 * h_off (h_row0[nsys] + 1 offsets) indexes the stacked rows, and h_idx[k] is the
 * column of entry k inside its own system (< the system's row count), h_val its
 * value. Each system must be symmetric positive definite (CG's requirement; not
 * checked). ALP_MISMATCH for a column outside its system, ALP_ILLEGAL for
 * malformed offsets. Each system is what one grb::Matrix<double> would hold for
 * a conjugate_gradient call (conjugate_gradient.hpp:191-484). */
alp_rc_t alp_hip_batch_create_crs(alp_hip_batch_t *B, size_t nsys, const uint64_t *h_row0,
                                  const uint64_t *h_off, const uint32_t *h_idx, const double *h_val,
                                  size_t nrhs);
alp_rc_t alp_hip_batch_destroy(alp_hip_batch_t B);
/* total rows, nonzeros, and (h_row0, nsys + 1 entries, may be NULL) each system's first row */
alp_rc_t alp_hip_batch_info(alp_hip_batch_t B, size_t *nrows, size_t *nnz, uint64_t *h_row0);
/* slices of the batch's diagonal form (at most 64 rows each, inside the SpMM tiles)
   and how many qualified; *ndiagonal is 0 when the form was not built. No
   counterpart in the reference (the batch itself has none). */
alp_rc_t alp_hip_batch_layout(alp_hip_batch_t B, size_t *nslices, size_t *ndiagonal);
/* Y = A X over all systems (FP64). Block vectors (d_X, d_Y here, d_B and d_X of
   alp_hip_batch_solve) must be 16-byte aligned, as every hipMalloc pointer is:
   the products move each row's right-hand sides in 16-byte chunks; others give
   ALP_ILLEGAL. */
alp_rc_t alp_hip_batch_spmm(alp_hip_batch_t B, const double *d_X, double *d_Y);
/* time the SpMM for each right-hand-side tile width (lanes per row) and keep the fastest,
   separately for the FP64 product and the FP32-data product of the mixed iterations;
   *width (may be NULL) gets the FP64 width */
alp_rc_t alp_hip_batch_autotune(alp_hip_batch_t B, int *width);
/* all systems from X = 0 to ||b - A x|| <= tol ||b|| per right-hand side; precision
 * 0 = FP64 CG, 1 = FP32 SpMM (FP64 row sums) and CG vectors on an FP64 X (a
 * segment's increments accumulated in FP32 and added at each update) with
 * reliable FP64 residual updates per instance, 2 = restarted FP32 SpMM/CG inside
 * FP64 iterative refinement. h_iters and
 * h_relres (nsys * nrhs entries each, may be NULL): CG iterations and final true
 * relative residual per instance. ALP_FAILED if some instance missed tol. */
alp_rc_t alp_hip_batch_solve(alp_hip_batch_t B, const double *d_B, double *d_X, double tol,
                             size_t max_iter, int precision, size_t *h_iters, double *h_relres,
                             size_t *outer);

#ifdef __cplusplus
}
#endif

#endif /* ALP_HIP_H */
