// The transition path's sparse CG C ABI (include/transition/solver.h) on the
// hip backend. The reference implementation (src/transition/solver.cpp:42-276)
// keeps a CG_Data per handle around a non-owning CRS view and calls
// grb::algorithms::(preconditioned_)conjugate_gradient with host vectors; here a
// handle owns a device matrix (copied at init) and two device vectors, and a
// solve is one device CG (cg_native) bracketed by two host<->device copies.
#include "alp_internal.hpp"

#include "../../include/alp_hip_solver.h"

#include <cmath>
#include <limits>
#include <new>
#include <type_traits>
#include <vector>

namespace alp {
namespace {

// one tag per variant, so a handle passed to another variant's function is rejected
template <typename T, typename RSI, typename NZI>
constexpr uint32_t variant_tag() {
	return 0x41500000u | (std::is_same<T, double>::value ? 4u : 0u) | (std::is_same<RSI, size_t>::value ? 2u : 0u) |
	       (std::is_same<NZI, size_t>::value ? 1u : 0u);
}

template <typename T, typename RSI, typename NZI>
struct CgHandle {
	using precond_t = int (*)(T *const, const T *const, void *const);
	uint32_t tag = variant_tag<T, RSI, NZI>();
	size_t n = 0;
	// defaults of CG_Data (src/transition/solver.cpp:157-163)
	T tolerance = T(1e-5);
	size_t max_iter = 1000;
	T residual = std::numeric_limits<T>::infinity();
	size_t iters = 0;
	precond_t precond = nullptr;
	void *precond_data = nullptr;
	alp_hip_matrix_t A = nullptr;
	T *d_x = nullptr;
	T *d_b = nullptr;

	~CgHandle() {
		if (A) alp_hip_matrix_destroy(A);
		if (d_x) hipFree(d_x);
		if (d_b) hipFree(d_b);
	}
};

template <typename T, typename RSI, typename NZI>
CgHandle<T, RSI, NZI> *as_handle(void *h) {
	auto *p = static_cast<CgHandle<T, RSI, NZI> *>(h);
	return p->tag == variant_tag<T, RSI, NZI>() ? p : nullptr;
}

template <typename I>
bool negative(I v) {
	if constexpr (std::is_signed<I>::value)
		return v < 0;
	else
		return false;
}

sparse_err_t from_rc(alp_rc_t rc) {
	switch (rc) {
		case ALP_SUCCESS: return NO_ERROR;
		case ALP_OUTOFMEM: return OUT_OF_MEMORY;
		case ALP_FAILED: return FAILED;
		default: return UNKNOWN;
	}
}

// solver.cpp:277-297 plus a structural check of the CRS arrays before they reach
// the device: ia non-decreasing from its first entry, 0 <= ja < n
template <typename T, typename RSI, typename NZI>
sparse_err_t cg_init(sparse_cg_handle_t *const handle, const size_t n, const T *const a, const RSI *const ja,
                     const NZI *const ia) {
	if (n == 0) return ILLEGAL_ARGUMENT;
	if (!handle || !a || !ja || !ia) return NULL_ARGUMENT;
	if (n > (size_t)std::numeric_limits<uint32_t>::max()) return ILLEGAL_ARGUMENT;
	std::vector<uint64_t> off;
	std::vector<uint32_t> idx;
	try {
		off.resize(n + 1);
		const NZI base = ia[0];
		if (negative(base)) return ILLEGAL_ARGUMENT;
		for (size_t i = 0; i <= n; ++i) {
			if (ia[i] < base || (i > 0 && ia[i] < ia[i - 1])) return ILLEGAL_ARGUMENT;
			off[i] = (uint64_t)(ia[i] - base);
		}
		const size_t nnz = off[n];
		idx.resize(nnz ? nnz : 1);
		const RSI *jb = ja + base;
		for (size_t k = 0; k < nnz; ++k) {
			if (negative(jb[k]) || (size_t)jb[k] >= n) return ILLEGAL_ARGUMENT;
			idx[k] = (uint32_t)jb[k];
		}
	} catch (const std::bad_alloc &) {
		return OUT_OF_MEMORY;
	}
	auto *h = new (std::nothrow) CgHandle<T, RSI, NZI>();
	if (!h) return OUT_OF_MEMORY;
	h->n = n;
	const int vt = std::is_same<T, double>::value ? ALP_VAL_F64 : ALP_VAL_F32;
	alp_rc_t rc = alp_hip_matrix_create_crs(&h->A, n, n, off[n], off.data(), idx.data(), a + ia[0], vt, 0);
	if (!rc && hipMalloc(&h->d_x, sizeof(T) * n) != hipSuccess) rc = ALP_OUTOFMEM;
	if (!rc && hipMalloc(&h->d_b, sizeof(T) * n) != hipSuccess) rc = ALP_OUTOFMEM;
	if (rc) {
		delete h;
		return rc == ALP_OUTOFMEM ? OUT_OF_MEMORY : rc == ALP_ILLEGAL || rc == ALP_MISMATCH ? ILLEGAL_ARGUMENT : UNKNOWN;
	}
	*handle = h;
	return NO_ERROR;
}

// solver.cpp:764-790
template <typename T, typename RSI, typename NZI>
sparse_err_t cg_solve(sparse_cg_handle_t handle, T *const x, const T *const b) {
	if (!handle || !x || !b) return NULL_ARGUMENT;
	auto *h = as_handle<T, RSI, NZI>(handle);
	if (!h) return ILLEGAL_ARGUMENT;
	// the CG's own argument checks (conjugate_gradient.hpp:300-312) surface as UNKNOWN
	if (!(h->tolerance > T(0)) || h->max_iter == 0) return UNKNOWN;
	const size_t bytes = sizeof(T) * h->n;
	hipStream_t s = stream();
	if (hipMemcpyAsync(h->d_x, x, bytes, hipMemcpyHostToDevice, s) != hipSuccess ||
	    hipMemcpyAsync(h->d_b, b, bytes, hipMemcpyHostToDevice, s) != hipSuccess)
		return UNKNOWN;
	HostPrecond<T> pc;
	pc.fn = h->precond;
	pc.data = h->precond_data;
	size_t iters = 0;
	double res = std::numeric_limits<double>::infinity();
	const alp_rc_t rc = cg_native<T>(h->A, h->d_b, h->d_x, h->max_iter, (double)h->tolerance, pc, &iters, &res);
	h->iters = iters;
	h->residual = (T)res;
	if (rc != ALP_SUCCESS && rc != ALP_FAILED) return UNKNOWN;
	if (hipMemcpyAsync(x, h->d_x, bytes, hipMemcpyDeviceToHost, s) != hipSuccess ||
	    hipStreamSynchronize(s) != hipSuccess)
		return UNKNOWN;
	return from_rc(rc);
}

template <typename T, typename RSI, typename NZI>
sparse_err_t cg_destroy(sparse_cg_handle_t handle) {
	if (!handle) return NULL_ARGUMENT;
	auto *h = as_handle<T, RSI, NZI>(handle);
	if (!h) return ILLEGAL_ARGUMENT;
	delete h;
	return NO_ERROR;
}

// the accessors (solver.cpp:340-762): NULL handle or output -> NULL_ARGUMENT
#define ALP_CG_HANDLE_OR_FAIL(h, handle)                \
	if (!(handle)) return NULL_ARGUMENT;                \
	auto *h = as_handle<T, RSI, NZI>(handle);           \
	if (!h) return ILLEGAL_ARGUMENT

template <typename T, typename RSI, typename NZI>
sparse_err_t cg_get_size(sparse_cg_handle_t handle, size_t *const size) {
	if (!size) return NULL_ARGUMENT;
	ALP_CG_HANDLE_OR_FAIL(h, handle);
	*size = h->n;
	return NO_ERROR;
}

template <typename T, typename RSI, typename NZI>
sparse_err_t cg_get_tolerance(sparse_cg_handle_t handle, T *const tol) {
	if (!tol) return NULL_ARGUMENT;
	ALP_CG_HANDLE_OR_FAIL(h, handle);
	*tol = h->tolerance;
	return NO_ERROR;
}

template <typename T, typename RSI, typename NZI>
sparse_err_t cg_set_tolerance(sparse_cg_handle_t handle, const T tol) {
	ALP_CG_HANDLE_OR_FAIL(h, handle);
	h->tolerance = tol;
	return NO_ERROR;
}

template <typename T, typename RSI, typename NZI>
sparse_err_t cg_set_max_iter(sparse_cg_handle_t handle, const size_t max_iters) {
	ALP_CG_HANDLE_OR_FAIL(h, handle);
	h->max_iter = max_iters;
	return NO_ERROR;
}

template <typename T, typename RSI, typename NZI>
sparse_err_t cg_get_residual(sparse_cg_handle_t handle, T *const res) {
	if (!res) return NULL_ARGUMENT;
	ALP_CG_HANDLE_OR_FAIL(h, handle);
	*res = h->residual;
	return NO_ERROR;
}

template <typename T, typename RSI, typename NZI>
sparse_err_t cg_get_iters(sparse_cg_handle_t handle, size_t *const iters) {
	if (!iters) return NULL_ARGUMENT;
	ALP_CG_HANDLE_OR_FAIL(h, handle);
	*iters = h->iters;
	return NO_ERROR;
}

template <typename T, typename RSI, typename NZI>
sparse_err_t cg_get_precond(sparse_cg_handle_t handle, int (**const fn)(T *const, const T *const, void *const),
                            void **const data) {
	if (!fn || !data) return NULL_ARGUMENT;
	ALP_CG_HANDLE_OR_FAIL(h, handle);
	*fn = h->precond;
	*data = h->precond_data;
	return NO_ERROR;
}

template <typename T, typename RSI, typename NZI>
sparse_err_t cg_set_precond(sparse_cg_handle_t handle, int (*fn)(T *const, const T *const, void *const),
                            void *const data) {
	ALP_CG_HANDLE_OR_FAIL(h, handle);
	if (!fn && data) return ILLEGAL_ARGUMENT;
	h->precond = fn;
	h->precond_data = data;
	return NO_ERROR;
}

#undef ALP_CG_HANDLE_OR_FAIL

}  // namespace
}  // namespace alp

using namespace alp;

// One block of exported functions per variant: value type T, ja type RSI, ia type NZI.
#define ALP_SPARSE_CG_VARIANT(SFX, T, RSI, NZI, PRECOND_T)                                                       \
	ALP_EXPORT sparse_err_t sparse_cg_init_##SFX(sparse_cg_handle_t *const handle, const size_t n,               \
	                                             const T *const a, const RSI *const ja, const NZI *const ia) {   \
		return cg_init<T, RSI, NZI>(handle, n, a, ja, ia);                                                       \
	}                                                                                                            \
	ALP_EXPORT sparse_err_t sparse_cg_get_size_##SFX(sparse_cg_handle_t const handle, size_t *const size) {      \
		return cg_get_size<T, RSI, NZI>(handle, size);                                                           \
	}                                                                                                            \
	ALP_EXPORT sparse_err_t sparse_cg_get_tolerance_##SFX(const sparse_cg_handle_t handle, T *const tol) {       \
		return cg_get_tolerance<T, RSI, NZI>(handle, tol);                                                       \
	}                                                                                                            \
	ALP_EXPORT sparse_err_t sparse_cg_set_tolerance_##SFX(sparse_cg_handle_t handle, const T tol) {              \
		return cg_set_tolerance<T, RSI, NZI>(handle, tol);                                                       \
	}                                                                                                            \
	ALP_EXPORT sparse_err_t sparse_cg_set_max_iter_count_##SFX(sparse_cg_handle_t handle, const size_t m) {      \
		return cg_set_max_iter<T, RSI, NZI>(handle, m);                                                          \
	}                                                                                                            \
	ALP_EXPORT sparse_err_t sparse_cg_get_residual_##SFX(const sparse_cg_handle_t handle, T *const tol) {        \
		return cg_get_residual<T, RSI, NZI>(handle, tol);                                                        \
	}                                                                                                            \
	ALP_EXPORT sparse_err_t sparse_cg_get_iter_count_##SFX(const sparse_cg_handle_t handle, size_t *const it) {  \
		return cg_get_iters<T, RSI, NZI>(handle, it);                                                            \
	}                                                                                                            \
	ALP_EXPORT sparse_err_t sparse_cg_get_preconditioner_##SFX(const sparse_cg_handle_t handle,                  \
	                                                           PRECOND_T *const fn, void **const data) {         \
		return cg_get_precond<T, RSI, NZI>(handle, fn, data);                                                    \
	}                                                                                                            \
	ALP_EXPORT sparse_err_t sparse_cg_set_preconditioner_##SFX(const sparse_cg_handle_t handle,                  \
	                                                           const PRECOND_T fn, void *const data) {           \
		return cg_set_precond<T, RSI, NZI>(handle, fn, data);                                                    \
	}                                                                                                            \
	ALP_EXPORT sparse_err_t sparse_cg_solve_##SFX(sparse_cg_handle_t handle, T *const x, const T *const b) {     \
		return cg_solve<T, RSI, NZI>(handle, x, b);                                                              \
	}                                                                                                            \
	ALP_EXPORT sparse_err_t sparse_cg_destroy_##SFX(sparse_cg_handle_t handle) {                                 \
		return cg_destroy<T, RSI, NZI>(handle);                                                                  \
	}

ALP_SPARSE_CG_VARIANT(sii, float, int, int, sparse_cg_preconditioner_sxx_t)
ALP_SPARSE_CG_VARIANT(siz, float, int, size_t, sparse_cg_preconditioner_sxx_t)
ALP_SPARSE_CG_VARIANT(szz, float, size_t, size_t, sparse_cg_preconditioner_sxx_t)
ALP_SPARSE_CG_VARIANT(dii, double, int, int, sparse_cg_preconditioner_dxx_t)
ALP_SPARSE_CG_VARIANT(diz, double, int, size_t, sparse_cg_preconditioner_dxx_t)
ALP_SPARSE_CG_VARIANT(dzz, double, size_t, size_t, sparse_cg_preconditioner_dxx_t)

#undef ALP_SPARSE_CG_VARIANT
