/*
 * alp_hip_solver.h -- the sparse CG solver C ABI of ALP's transition path,
 * exported by libalp_hip.so with the hip backend underneath.
 *
 * Names, argument meanings, the error enum and its values are those of ALP's
 * include/transition/solver.h, so a C program written against that header
 * links against libalp_hip.so unchanged. What differs is where the data lives:
 *   - init copies the CRS arrays (a, ja, ia) to the device; the reference keeps
 *     a non-owning view of them (src/transition/solver.cpp:153-170). Changing
 *     the caller's arrays after init therefore has no effect here.
 *   - solve copies x and b to the device, runs the device CG, and copies x back
 *     (on NO_ERROR and FAILED).
 *   - a preconditioner is a host callback: each CG iteration copies r to the
 *     host, calls it, and copies z back (two PCIe transfers of n values).
 *   - init validates the structure (ia non-decreasing, 0 <= ja[k] < n, n below
 *     2^32) and returns ILLEGAL_ARGUMENT otherwise, where the reference would
 *     read out of bounds; a handle of another variant is rejected in the same
 *     way.
 * The variant suffix is <value><ia type><ja type>... as in the reference:
 * s/d = float/double values; i/z = int/size_t, the first for ja (column
 * indices) and the second for ia (row offsets): sii, siz, szz, dii, diz, dzz.
 * Arithmetic is in the value type; dot products accumulate in FP64 with a
 * fixed reduction tree.
 */
#ifndef ALP_HIP_SOLVER_H
#define ALP_HIP_SOLVER_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* include/transition/solver.h:120-155 (same order, same values) */
typedef enum {
	NO_ERROR,         /* success */
	NULL_ARGUMENT,    /* a required pointer argument was NULL */
	ILLEGAL_ARGUMENT, /* an argument had an illegal value */
	OUT_OF_MEMORY,    /* host or device allocation failed */
	FAILED,           /* the solve did not reach the tolerance within the iteration limit,
	                     or the preconditioner returned non-zero */
	UNKNOWN           /* any other failure; the library state is undefined */
} sparse_err_t;

/* include/transition/solver.h:159 */
typedef void *sparse_cg_handle_t;

/* include/transition/solver.h:182-212: out = M^-1 in; non-zero return = failure */
typedef int (*sparse_cg_preconditioner_sxx_t)(float *const, const float *const, void *const);
typedef int (*sparse_cg_preconditioner_dxx_t)(double *const, const double *const, void *const);

/* include/transition/solver.h:248-330 (src/transition/solver.cpp:277-338):
 * n == 0 -> ILLEGAL_ARGUMENT; NULL handle/a/ja/ia -> NULL_ARGUMENT. Defaults:
 * tolerance 1e-5 (relative: stop when ||r|| < tol ||b||), 1000 iterations, no
 * preconditioner, residual = +inf, iteration count 0. */
sparse_err_t sparse_cg_init_sii(sparse_cg_handle_t *const handle, const size_t n, const float *const a,
                                const int *const ja, const int *const ia);
sparse_err_t sparse_cg_init_dii(sparse_cg_handle_t *const handle, const size_t n, const double *const a,
                                const int *const ja, const int *const ia);
sparse_err_t sparse_cg_init_siz(sparse_cg_handle_t *const handle, const size_t n, const float *const a,
                                const int *const ja, const size_t *const ia);
sparse_err_t sparse_cg_init_diz(sparse_cg_handle_t *const handle, const size_t n, const double *const a,
                                const int *const ja, const size_t *const ia);
sparse_err_t sparse_cg_init_szz(sparse_cg_handle_t *const handle, const size_t n, const float *const a,
                                const size_t *const ja, const size_t *const ia);
sparse_err_t sparse_cg_init_dzz(sparse_cg_handle_t *const handle, const size_t n, const double *const a,
                                const size_t *const ja, const size_t *const ia);

/* include/transition/solver.h:347-480: the system size */
sparse_err_t sparse_cg_get_size_sii(sparse_cg_handle_t const handle, size_t *const size);
sparse_err_t sparse_cg_get_size_siz(sparse_cg_handle_t const handle, size_t *const size);
sparse_err_t sparse_cg_get_size_szz(sparse_cg_handle_t const handle, size_t *const size);
sparse_err_t sparse_cg_get_size_dii(sparse_cg_handle_t const handle, size_t *const size);
sparse_err_t sparse_cg_get_size_diz(sparse_cg_handle_t const handle, size_t *const size);
sparse_err_t sparse_cg_get_size_dzz(sparse_cg_handle_t const handle, size_t *const size);

/* include/transition/solver.h:503-569: the relative tolerance */
sparse_err_t sparse_cg_get_tolerance_sii(const sparse_cg_handle_t handle, float *const tol);
sparse_err_t sparse_cg_get_tolerance_siz(const sparse_cg_handle_t handle, float *const tol);
sparse_err_t sparse_cg_get_tolerance_szz(const sparse_cg_handle_t handle, float *const tol);
sparse_err_t sparse_cg_get_tolerance_dii(const sparse_cg_handle_t handle, double *const tol);
sparse_err_t sparse_cg_get_tolerance_diz(const sparse_cg_handle_t handle, double *const tol);
sparse_err_t sparse_cg_get_tolerance_dzz(const sparse_cg_handle_t handle, double *const tol);

/* include/transition/solver.h:591-657; a non-positive tolerance makes the next
 * solve return UNKNOWN, as the reference's CG rejects it (conjugate_gradient.hpp:300-306) */
sparse_err_t sparse_cg_set_tolerance_sii(sparse_cg_handle_t handle, const float tol);
sparse_err_t sparse_cg_set_tolerance_siz(sparse_cg_handle_t handle, const float tol);
sparse_err_t sparse_cg_set_tolerance_szz(sparse_cg_handle_t handle, const float tol);
sparse_err_t sparse_cg_set_tolerance_dii(sparse_cg_handle_t handle, const double tol);
sparse_err_t sparse_cg_set_tolerance_diz(sparse_cg_handle_t handle, const double tol);
sparse_err_t sparse_cg_set_tolerance_dzz(sparse_cg_handle_t handle, const double tol);

/* include/transition/solver.h:678-744 */
sparse_err_t sparse_cg_set_max_iter_count_sii(sparse_cg_handle_t handle, const size_t max_iters);
sparse_err_t sparse_cg_set_max_iter_count_siz(sparse_cg_handle_t handle, const size_t max_iters);
sparse_err_t sparse_cg_set_max_iter_count_szz(sparse_cg_handle_t handle, const size_t max_iters);
sparse_err_t sparse_cg_set_max_iter_count_dii(sparse_cg_handle_t handle, const size_t max_iters);
sparse_err_t sparse_cg_set_max_iter_count_diz(sparse_cg_handle_t handle, const size_t max_iters);
sparse_err_t sparse_cg_set_max_iter_count_dzz(sparse_cg_handle_t handle, const size_t max_iters);

/* include/transition/solver.h:778-874: the last solve's squared residual
 * norm r.r (conjugate_gradient.hpp:424-431); +inf before the first solve */
sparse_err_t sparse_cg_get_residual_sii(const sparse_cg_handle_t handle, float *const tol);
sparse_err_t sparse_cg_get_residual_siz(const sparse_cg_handle_t handle, float *const tol);
sparse_err_t sparse_cg_get_residual_szz(const sparse_cg_handle_t handle, float *const tol);
sparse_err_t sparse_cg_get_residual_dii(const sparse_cg_handle_t handle, double *const tol);
sparse_err_t sparse_cg_get_residual_diz(const sparse_cg_handle_t handle, double *const tol);
sparse_err_t sparse_cg_get_residual_dzz(const sparse_cg_handle_t handle, double *const tol);

/* include/transition/solver.h:905-1001: the last solve's iteration count */
sparse_err_t sparse_cg_get_iter_count_sii(const sparse_cg_handle_t handle, size_t *const iters);
sparse_err_t sparse_cg_get_iter_count_siz(const sparse_cg_handle_t handle, size_t *const iters);
sparse_err_t sparse_cg_get_iter_count_szz(const sparse_cg_handle_t handle, size_t *const iters);
sparse_err_t sparse_cg_get_iter_count_dii(const sparse_cg_handle_t handle, size_t *const iters);
sparse_err_t sparse_cg_get_iter_count_diz(const sparse_cg_handle_t handle, size_t *const iters);
sparse_err_t sparse_cg_get_iter_count_dzz(const sparse_cg_handle_t handle, size_t *const iters);

/* include/transition/solver.h:1034-1205 */
sparse_err_t sparse_cg_get_preconditioner_sii(const sparse_cg_handle_t handle,
                                              sparse_cg_preconditioner_sxx_t *const preconditioner,
                                              void **const data);
sparse_err_t sparse_cg_get_preconditioner_siz(const sparse_cg_handle_t handle,
                                              sparse_cg_preconditioner_sxx_t *const preconditioner,
                                              void **const data);
sparse_err_t sparse_cg_get_preconditioner_szz(const sparse_cg_handle_t handle,
                                              sparse_cg_preconditioner_sxx_t *const preconditioner,
                                              void **const data);
sparse_err_t sparse_cg_get_preconditioner_dii(const sparse_cg_handle_t handle,
                                              sparse_cg_preconditioner_dxx_t *const preconditioner,
                                              void **const data);
sparse_err_t sparse_cg_get_preconditioner_diz(const sparse_cg_handle_t handle,
                                              sparse_cg_preconditioner_dxx_t *const preconditioner,
                                              void **const data);
sparse_err_t sparse_cg_get_preconditioner_dzz(const sparse_cg_handle_t handle,
                                              sparse_cg_preconditioner_dxx_t *const preconditioner,
                                              void **const data);

/* include/transition/solver.h:1236-1430 (src/transition/solver.cpp:639-654): a NULL
 * preconditioner removes it; a NULL preconditioner with non-NULL data is ILLEGAL_ARGUMENT */
sparse_err_t sparse_cg_set_preconditioner_sii(const sparse_cg_handle_t handle,
                                              const sparse_cg_preconditioner_sxx_t preconditioner,
                                              void *const data);
sparse_err_t sparse_cg_set_preconditioner_dii(const sparse_cg_handle_t handle,
                                              const sparse_cg_preconditioner_dxx_t preconditioner,
                                              void *const data);
sparse_err_t sparse_cg_set_preconditioner_siz(const sparse_cg_handle_t handle,
                                              const sparse_cg_preconditioner_sxx_t preconditioner,
                                              void *const data);
sparse_err_t sparse_cg_set_preconditioner_diz(const sparse_cg_handle_t handle,
                                              const sparse_cg_preconditioner_dxx_t preconditioner,
                                              void *const data);
sparse_err_t sparse_cg_set_preconditioner_szz(const sparse_cg_handle_t handle,
                                              const sparse_cg_preconditioner_sxx_t preconditioner,
                                              void *const data);
sparse_err_t sparse_cg_set_preconditioner_dzz(const sparse_cg_handle_t handle,
                                              const sparse_cg_preconditioner_dxx_t preconditioner,
                                              void *const data);

/* include/transition/solver.h:1443-1580 (src/transition/solver.cpp:764-790): x is the
 * initial guess on input, the last iterate on NO_ERROR and FAILED; NULL_ARGUMENT
 * leaves it untouched; UNKNOWN leaves it undefined */
sparse_err_t sparse_cg_solve_sii(sparse_cg_handle_t handle, float *const x, const float *const b);
sparse_err_t sparse_cg_solve_siz(sparse_cg_handle_t handle, float *const x, const float *const b);
sparse_err_t sparse_cg_solve_szz(sparse_cg_handle_t handle, float *const x, const float *const b);
sparse_err_t sparse_cg_solve_dii(sparse_cg_handle_t handle, double *const x, const double *const b);
sparse_err_t sparse_cg_solve_diz(sparse_cg_handle_t handle, double *const x, const double *const b);
sparse_err_t sparse_cg_solve_dzz(sparse_cg_handle_t handle, double *const x, const double *const b);

/* include/transition/solver.h:1593-1653 */
sparse_err_t sparse_cg_destroy_sii(sparse_cg_handle_t handle);
sparse_err_t sparse_cg_destroy_siz(sparse_cg_handle_t handle);
sparse_err_t sparse_cg_destroy_szz(sparse_cg_handle_t handle);
sparse_err_t sparse_cg_destroy_dii(sparse_cg_handle_t handle);
sparse_err_t sparse_cg_destroy_diz(sparse_cg_handle_t handle);
sparse_err_t sparse_cg_destroy_dzz(sparse_cg_handle_t handle);

#ifdef __cplusplus
}
#endif

#endif /* ALP_HIP_SOLVER_H */
