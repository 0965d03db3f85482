// grb::mxm for the hip backend's built-in semirings (reference/blas3.hpp:1148
// mxm_generic, :1744 mxm): C (+)= op(A) (x) op(B). The product itself is the
// expand-sort-compress kernel set of graphblas/hip/spgemm.hpp, which the grb::
// header layer instantiates for any other semiring; it repeats mxm_generic's
// evaluation order, so results and the storage order of C equal the
// reference's.
#include "alp_internal.hpp"

#include "graphblas/hip/spgemm.hpp"

#include <cstdlib>

namespace alp {

struct MulTimesF64 {
	__device__ double operator()(double a, double b) const { return a * b; }
};
struct AddPlusF64 {
	__device__ double operator()(double a, double b) const { return a + b; }
};
struct MulTimesF32 {
	__device__ float operator()(float a, float b) const { return a * b; }
};
struct AddPlusF32 {
	__device__ float operator()(float a, float b) const { return a + b; }
};
struct MulPlusF64 {  // min-plus: the multiplication is +
	__device__ double operator()(double a, double b) const { return a + b; }
};
struct AddMinF64 {
	__device__ double operator()(double a, double b) const { return b < a ? b : a; }
};
struct PatternOp {  // lor-land over patterns: structure only
	__device__ uint8_t operator()(uint8_t, uint8_t) const { return 1; }
};

struct Operand {
	const uint64_t *off = nullptr;
	const uint32_t *idx = nullptr;
	const void *val = nullptr;
	size_t rows = 0, cols = 0;
};

static alp_rc_t operand(alp_hip_matrix_t M, bool transpose, Operand &o) {
	if (transpose && !M->ccs.off) ALP_TRY(build_ccs_from_crs(M));
	const Storage &S = transpose ? M->ccs : M->crs;
	o.off = S.off;
	o.idx = S.idx;
	o.val = S.val;
	o.rows = transpose ? M->ncols : M->nrows;
	o.cols = transpose ? M->nrows : M->ncols;
	return ALP_SUCCESS;
}

static alp_rc_t from_hip(hipError_t e) {
	if (e == hipSuccess) return ALP_SUCCESS;
	(void)hipGetLastError();
	return e == hipErrorOutOfMemory ? ALP_OUTOFMEM : ALP_PANIC;
}

// products per block of rows; ALP_HIP_MXM_BUDGET overrides it (tests force many
// small blocks with it)
static size_t block_budget() {
	size_t b = size_t(1) << 26;
	if (const char *e = std::getenv("ALP_HIP_MXM_BUDGET")) {
		const long long v = std::atoll(e);
		if (v > 0) b = (size_t)v;
	}
	return b;
}

template <bool VALUES, typename T, class Mul, class Add>
static alp_rc_t run(const Operand &a, const Operand &b, alp_hip_matrix_t old, T one, Mul mul, Add add,
                    alp_spgemm::Result &res) {
	const uint64_t *coff = nullptr;
	const uint32_t *cidx = nullptr;
	const T *cval = nullptr;
	if (old && old->nnz) {
		coff = old->crs.off;
		cidx = old->crs.idx;
		cval = (const T *)old->crs.val;
	}
	return from_hip(alp_spgemm::multiply<VALUES, T, T, T>(stream(), a.rows, a.off, a.idx, (const T *)a.val, one, b.off,
	                                                      b.idx, (const T *)b.val, one, coff, cidx, cval, mul, add,
	                                                      res, block_budget()));
}

}  // namespace alp

using namespace alp;

ALP_EXPORT alp_rc_t alp_hip_mxm(alp_hip_matrix_t *C, alp_hip_matrix_t A, alp_hip_matrix_t B, alp_semiring_t ring,
                                unsigned descr) {
	ALP_TRY(ensure_init());
	if (!C || !A || !B) return ALP_ILLEGAL;
	Operand a, b;
	ALP_TRY(operand(A, (descr & ALP_DESC_TRANSPOSE_LEFT) != 0, a));
	ALP_TRY(operand(B, (descr & ALP_DESC_TRANSPOSE_RIGHT) != 0, b));
	if (a.cols != b.rows) return ALP_MISMATCH;
	int out_type;
	switch (ring) {
		case ALP_PLUS_TIMES_F64:
		case ALP_MIN_PLUS_F64:
			if ((A->value_type != ALP_VAL_F64 && A->value_type != ALP_VAL_PATTERN) ||
			    (B->value_type != ALP_VAL_F64 && B->value_type != ALP_VAL_PATTERN))
				return ALP_MISMATCH;
			out_type = ALP_VAL_F64;
			break;
		case ALP_PLUS_TIMES_F32:
			if ((A->value_type != ALP_VAL_F32 && A->value_type != ALP_VAL_PATTERN) ||
			    (B->value_type != ALP_VAL_F32 && B->value_type != ALP_VAL_PATTERN))
				return ALP_MISMATCH;
			out_type = ALP_VAL_F32;
			break;
		case ALP_LOR_LAND_BOOL:
			// a pattern product: every stored entry of A and B is true
			if (A->value_type != ALP_VAL_PATTERN || B->value_type != ALP_VAL_PATTERN) return ALP_MISMATCH;
			out_type = ALP_VAL_PATTERN;
			break;
		default:
			return ALP_ILLEGAL;
	}
	alp_hip_matrix_t old = *C;
	if (old) {
		if (old->nrows != a.rows || old->ncols != b.cols) return ALP_MISMATCH;
		if (old->nnz && old->value_type != out_type) return ALP_MISMATCH;
	}
	alp_spgemm::Result res;
	alp_rc_t rc = ALP_SUCCESS;
	// a pattern operand contributes the multiplicative identity (the value a void
	// matrix's getValue returns, compressed_storage.hpp:1311)
	if (ring == ALP_PLUS_TIMES_F64)
		rc = run<true, double>(a, b, old, 1.0, MulTimesF64(), AddPlusF64(), res);
	else if (ring == ALP_MIN_PLUS_F64)
		rc = run<true, double>(a, b, old, 0.0, MulPlusF64(), AddMinF64(), res);
	else if (ring == ALP_PLUS_TIMES_F32)
		rc = run<true, float>(a, b, old, 1.0f, MulTimesF32(), AddPlusF32(), res);
	else
		rc = run<false, uint8_t>(a, b, old, (uint8_t)1, PatternOp(), PatternOp(), res);
	if (rc) return rc;
	alp_hip_matrix_t out = nullptr;
	const unsigned flags = (A->flags | B->flags) & ALP_MAT_WITH_CCS;
	rc = alp_hip_matrix_create_device(&out, a.rows, b.cols, res.nnz, res.off, res.idx, res.val, out_type, flags,
	                                  nullptr);
	hipFree(res.off);
	hipFree(res.idx);
	if (res.val) hipFree(res.val);
	if (rc) return rc;
	if (old) alp_hip_matrix_destroy(old);
	*C = out;
	return ALP_SUCCESS;
}
