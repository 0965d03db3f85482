/*
 * blas3.hpp -- grb::mxm over any semiring, or any additive monoid with a
 * multiplicative operator (reference/blas3.hpp:1148 mxm_generic, :1744, :1798).
 *
 * The product is the expand-sort-compress SpGEMM of spgemm.hpp, instantiated
 * here for the program's algebra; libalp_hip.so runs the same kernels for its
 * built-in semirings (alp_hip_mxm), which this layer uses when the algebra and
 * domains are one of those. Either way C's rows list their columns in the order
 * the reference's SPA first reaches them, each entry folds its products in the
 * reference's order, and an output that holds entries is updated in place.
 * RESIZE does nothing: EXECUTE sizes the output itself.
 */
#ifndef ALP_HIP_GRB_BLAS3_HPP
#define ALP_HIP_GRB_BLAS3_HPP

#include "blas2.hpp"
#include "spgemm.hpp"

namespace grb {

namespace internal {
namespace hip {

template <class MUL, class ID0, typename OUT, typename TA, typename TB>
struct SpgemmMul {
	__device__ OUT operator()(const TA &a, const TB &b) const {
		typename MUL::D3 p = static_cast<typename MUL::D3>(ID0::template value<OUT>());
		apply_op<MUL>(p, a, b);
		return static_cast<OUT>(p);
	}
};
template <class ADD, typename OUT>
struct SpgemmAdd {
	__device__ OUT operator()(OUT x, const OUT &y) const {
		foldl_op<ADD>(x, y);
		return x;
	}
};

// the library's alp_hip_mxm covers: plus-times / min-plus over double (double or
// pattern operands), plus-times over float, lor-land over patterns
template <class ADD, class MUL, typename OutT, typename T1, typename T2> struct lib_mxm { static constexpr int value = -1; };
template <typename T1, typename T2>
struct lib_mxm_f64 {
	static constexpr bool ok = (std::is_same<T1, double>::value || std::is_void<T1>::value) &&
	                           (std::is_same<T2, double>::value || std::is_void<T2>::value);
};
template <Backend b1, Backend b2, typename T1, typename T2>
struct lib_mxm<operators::add<double, double, double, b1>, operators::mul<double, double, double, b2>, double, T1, T2> {
	static constexpr int value = lib_mxm_f64<T1, T2>::ok ? ALP_PLUS_TIMES_F64 : -1;
};
template <Backend b1, Backend b2, typename T1, typename T2>
struct lib_mxm<operators::min<double, double, double, b1>, operators::add<double, double, double, b2>, double, T1, T2> {
	static constexpr int value = lib_mxm_f64<T1, T2>::ok ? ALP_MIN_PLUS_F64 : -1;
};
template <Backend b1, Backend b2>
struct lib_mxm<operators::add<float, float, float, b1>, operators::mul<float, float, float, b2>, float, float, float> {
	static constexpr int value = ALP_PLUS_TIMES_F32;
};
template <Backend b1, Backend b2>
struct lib_mxm<operators::logical_or<bool, bool, bool, b1>, operators::logical_and<bool, bool, bool, b2>, void, void,
               void> {
	static constexpr int value = ALP_LOR_LAND_BOOL;
};

template <int RING, typename OutT, typename T1, typename T2, Backend b, typename R, typename C, typename N>
RC mxm_library(Matrix<OutT, b, R, C, N> &Cm, const Matrix<T1, b, R, C, N> &A, const Matrix<T2, b, R, C, N> &B,
               unsigned d, std::true_type) {
	alp_hip_matrix_t h = Cm.handle();
	Cm.release_handle();
	RC r = rc(alp_hip_mxm(&h, A.handle(), B.handle(), (alp_semiring_t)RING, d));
	Cm.reset(h);
	return r;
}
template <int RING, typename OutT, typename T1, typename T2, Backend b, typename R, typename C, typename N>
RC mxm_library(Matrix<OutT, b, R, C, N> &, const Matrix<T1, b, R, C, N> &, const Matrix<T2, b, R, C, N> &, unsigned,
               std::false_type) {
	return PANIC;
}

template <Descriptor descr, class ADDM, class MUL, template <typename> class One, typename OutT, typename T1,
          typename T2, Backend b, typename R, typename C, typename N>
RC mxm(Matrix<OutT, b, R, C, N> &Cm, const Matrix<T1, b, R, C, N> &A, const Matrix<T2, b, R, C, N> &B,
       const Phase phase) {
	constexpr bool tl = (descr & descriptors::transpose_left) != 0;
	constexpr bool trr = (descr & descriptors::transpose_right) != 0;
	const size_t m = tl ? A.cols() : A.rows(), k = tl ? A.rows() : A.cols();
	const size_t kb = trr ? B.cols() : B.rows(), n = trr ? B.rows() : B.cols();
	if (Cm.rows() != m || Cm.cols() != n || k != kb) return MISMATCH;
	if (phase == RESIZE) return SUCCESS;
	typedef typename ADDM::Operator ADD;
	if (!A.handle() || !B.handle() || A.nonzeroes() == 0 || B.nonzeroes() == 0) return SUCCESS;  // C unchanged
	// a forcing point: the product's size is decided on the host (the reference's
	// nonblocking backend also runs level-3 primitives blocking)
	Forcing f;
	if (f.rc) return f.rc;
	constexpr int ring = lib_mxm<ADD, MUL, OutT, T1, T2>::value;
	const unsigned d = (tl ? ALP_DESC_TRANSPOSE_LEFT : 0u) | (trr ? ALP_DESC_TRANSPOSE_RIGHT : 0u);
	if (ring >= 0) return mxm_library<ring>(Cm, A, B, d, std::integral_constant<bool, (ring >= 0)>());

	typedef typename std::conditional<std::is_void<OutT>::value, char, OutT>::type OUT;
	typedef typename std::conditional<std::is_void<T1>::value, typename MUL::D1, T1>::type TA;
	typedef typename std::conditional<std::is_void<T2>::value, typename MUL::D2, T2>::type TB;
	typedef IdentityOf<ADDM::template Identity> ID0;
	typename Matrix<T1, b, R, C, N>::Storage SA;
	typename Matrix<T2, b, R, C, N>::Storage SB;
	RC r = A.storage(tl ? 1 : 0, SA);
	r = r ? r : B.storage(trr ? 1 : 0, SB);
	if (r) return r;
	typename Matrix<OutT, b, R, C, N>::Storage SC;
	const bool inplace = Cm.handle() && Cm.nonzeroes() > 0;
	if (inplace && (r = Cm.storage(0, SC))) return r;
	constexpr bool VALUES = !std::is_void<OutT>::value;
	alp_spgemm::Result res;
	r = from_hip(alp_spgemm::multiply<VALUES, OUT, TA, TB>(
	    stream(), m, SA.off, SA.idx, (const TA *)SA.val, One<TA>::value(), SB.off, SB.idx, (const TB *)SB.val,
	    One<TB>::value(), inplace ? SC.off : nullptr, inplace ? SC.idx : nullptr, inplace ? (const OUT *)SC.val : nullptr,
	    SpgemmMul<MUL, ID0, OUT, TA, TB>(), SpgemmAdd<ADD, OUT>(), res));
	if (r) return r;
	struct Free {
		alp_spgemm::Result &x;
		~Free() {
			if (x.off) (void)hipFree(x.off);
			if (x.idx) (void)hipFree(x.idx);
			if (x.val) (void)hipFree(x.val);
		}
	} guard{res};
	alp_hip_matrix_t h = nullptr;
	constexpr int vt = std::is_void<OutT>::value ? ALP_VAL_PATTERN : lib_type<OUT>::value;
	if (std::is_void<OutT>::value || lib_holds_values<OUT>::value) {
		r = rc(alp_hip_matrix_create_device(&h, m, n, res.nnz, res.off, res.idx, res.val, vt, ALP_MAT_WITH_CCS,
		                                    nullptr));
		if (r) return r;
		Cm.reset(h);
		return SUCCESS;
	}
	// values of another type stay here, in CRS order and, through the position
	// map, in CCS order
	DevArray<uint64_t> ccs_src(res.nnz);
	r = rc(alp_hip_matrix_create_device(&h, m, n, res.nnz, res.off, res.idx, nullptr, ALP_VAL_PATTERN,
	                                    ALP_MAT_WITH_CCS, ccs_src.p));
	if (r) return r;
	DevArray<OUT> crs(res.nnz), ccs(res.nnz);
	if (res.nnz) {
		r = from_hip(hipMemcpyAsync(crs.p, res.val, sizeof(OUT) * res.nnz, hipMemcpyDeviceToDevice, stream()));
		hipLaunchKernelGGL(k_gather_values<OUT>, dim3(grid(res.nnz)), dim3(kBlock), 0, stream(), res.nnz,
		                   (const uint64_t *)ccs_src.p, (const OUT *)res.val, ccs.p);
	}
	r = r ? r : sync();
	if (r) {
		alp_hip_matrix_destroy(h);
		return r;
	}
	Cm.reset(h, std::move(crs), std::move(ccs));
	return SUCCESS;
}

}  // namespace hip
}  // namespace internal

// C (+)= op(A) (x) op(B) over a semiring (blas3.hpp:1744)
template <Descriptor descr = descriptors::no_operation, typename OutT, typename T1, typename T2, class Ring,
          enum Backend b, typename R, typename C, typename N>
RC mxm(Matrix<OutT, b, R, C, N> &Cm, const Matrix<T1, b, R, C, N> &A, const Matrix<T2, b, R, C, N> &B,
       const Ring & = Ring(), const Phase phase = EXECUTE,
       typename std::enable_if<is_semiring<Ring>::value>::type * = nullptr) {
	return internal::hip::mxm<descr, typename Ring::AdditiveMonoid, typename Ring::MultiplicativeOperator,
	                          Ring::template One>(Cm, A, B, phase);
}
// the (additive monoid, multiplicative operator) form (blas3.hpp:1798)
template <Descriptor descr = descriptors::no_operation, typename OutT, typename T1, typename T2, class AddMonoid,
          class MulOp, enum Backend b, typename R, typename C, typename N>
RC mxm(Matrix<OutT, b, R, C, N> &Cm, const Matrix<T1, b, R, C, N> &A, const Matrix<T2, b, R, C, N> &B,
       const AddMonoid &, const MulOp &, const Phase phase = EXECUTE,
       typename std::enable_if<is_monoid<AddMonoid>::value && is_operator<MulOp>::value>::type * = nullptr) {
	static_assert(!std::is_void<T1>::value && !std::is_void<T2>::value,
	              "grb::mxm: the monoid-operator form takes no pattern (void) operands (reference/blas3.hpp:1844)");
	return internal::hip::mxm<descr, AddMonoid, MulOp, identities::one>(Cm, A, B, phase);
}

}  // namespace grb

#endif
