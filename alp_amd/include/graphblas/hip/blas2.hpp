/*
 * blas2.hpp -- grb::mxv and grb::vxm over any semiring, or any additive monoid
 * with a multiplicative operator (ALP's base/blas2.hpp:60-243; the reference's
 * vxm_generic, reference/blas2.hpp:870-1528, and its inner kernels :295-769).
 *
 * Dispatch, in the reference's order of checks: dimensions (MISMATCH), the dense
 * descriptor against sparse inputs or masks (ILLEGAL), mask sizes (MISMATCH), the
 * RESIZE phase, the trivial cases (an empty input or matrix, an empty mask: the
 * output keeps its entries, explicit_zero adds the identity where it had none),
 * then overlapping output and input or output and mask without safe_overlap
 * (OVERLAP).
 *
 * Execution:
 *   - dense vectors under a semiring the library has tuned kernels for (plus-
 *     times over double or float, lor-land, min-plus), unmasked or with a dense
 *     boolean value mask: the library's SpMV (alp_hip_mxv / alp_hip_vxm);
 *   - otherwise a gather per output element over the CRS (mxv) or CCS (vxm) --
 *     the reference's vxm_inner_kernel_gather (:295-568): the row is folded in
 *     storage order, each product on the left of the running sum, then folded
 *     into an assigned output or assigned to an unassigned one; add_identity,
 *     input and output masks (structural, inverted), sparse inputs and the
 *     immutable-nonzero exit of lor monoids included. Results equal the
 *     reference's for any semiring, in its evaluation order;
 *   - a sparse input (at most 1/16 of its entries) under an exact commutative
 *     monoid over an integral or boolean domain: a push from the input's
 *     nonzeroes over the transposed storage with atomic folds -- the reference's
 *     vxm_inner_kernel_scatter (:646-769). Exact arithmetic makes the fold order
 *     immaterial, so results equal the gather's.
 */
#ifndef ALP_HIP_GRB_BLAS2_HPP
#define ALP_HIP_GRB_BLAS2_HPP

#include "blas1.hpp"

namespace grb {

namespace internal {

// the library semiring a (monoid, operator, domains) combination runs as, or -1
template <class ADD, class MUL, typename IO, typename X, typename T> struct lib_ring { static constexpr int value = -1; };
template <Backend b1, Backend b2>
struct lib_ring<operators::add<double, double, double, b1>, operators::mul<double, double, double, b2>, double, double,
                double> {
	static constexpr int value = ALP_PLUS_TIMES_F64;
};
template <Backend b1, Backend b2>
struct lib_ring<operators::add<float, float, float, b1>, operators::mul<float, float, float, b2>, float, float, float> {
	static constexpr int value = ALP_PLUS_TIMES_F32;
};
template <Backend b1, Backend b2>
struct lib_ring<operators::min<double, double, double, b1>, operators::add<double, double, double, b2>, double, double,
                double> {
	static constexpr int value = ALP_MIN_PLUS_F64;
};
template <Backend b1, Backend b2, typename T>
struct lib_ring<operators::logical_or<bool, bool, bool, b1>, operators::logical_and<bool, bool, bool, b2>, bool, bool,
                T> {
	static constexpr int value = ALP_LOR_LAND_BOOL;
};

namespace hip {

// an operator declared associative and commutative (ALP's operators carry these
// flags; an operator without them counts as neither)
template <class OP, class = void> struct exact_fold_op : std::false_type {};
template <class OP>
struct exact_fold_op<OP, decltype((void)OP::is_commutative, (void)OP::is_associative)>
    : std::integral_constant<bool, OP::is_commutative && OP::is_associative> {};

// whether the library's SpMV covers ring RING on a matrix of T (values it holds;
// lor-land only on patterns, where every stored entry is true)
template <int RING, typename T> struct lib_matrix_ok {
	static constexpr bool value =
	    (std::is_same<T, double>::value && (RING == ALP_PLUS_TIMES_F64 || RING == ALP_MIN_PLUS_F64)) ||
	    (std::is_same<T, float>::value && RING == ALP_PLUS_TIMES_F32) ||
	    (std::is_void<T>::value && (RING == ALP_PLUS_TIMES_F64 || RING == ALP_LOR_LAND_BOOL));
};

template <class ADD>
__device__ __forceinline__ bool is_identity(const typename ADD::D3 &v, const typename ADD::D3 &id) {
	return v == id;
}

// the gather for output element i (vxm_inner_kernel_gather, reference/blas2.hpp:295-568)
template <Descriptor descr, bool LEFT, bool MASKED, bool IN_MASKED, bool IMMUTABLE, class ADD, class MUL, class ID0,
          class ONE, typename IO, typename X, typename V, typename M, typename VM>
__global__ void k_mv_gather(size_t m, size_t nsrc, const uint64_t *off, const uint32_t *idx, const V *aval, Arg<X> x,
                            IO *y, uint8_t *ya, const M *mval, const uint8_t *masg, const VM *vmval,
                            const uint8_t *vmasg, unsigned long long *count) {
	typedef typename ADD::D3 OUT;
	typedef typename MUL::D3 PROD;
	typedef typename std::conditional<LEFT, typename MUL::D2, typename MUL::D1>::type NZ;
	typedef typename std::conditional<LEFT, typename MUL::D1, typename MUL::D2>::type SRC;
	unsigned long long c = 0;
	for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (size_t)gridDim.x * blockDim.x) {
		const bool yin = ya[i] != 0;
		bool yout = yin;
		const OUT id = ID0::template value<OUT>();
		bool skip = MASKED && !mask_at<descr>(mval, masg, i);
		if (!skip && IMMUTABLE && yin && !is_identity<ADD>(static_cast<OUT>(y[i]), id)) skip = true;
		if (!skip) {
			OUT out = id;
			bool set = false;
			if ((descr & descriptors::add_identity) && i < nsrc &&
			    (!IN_MASKED || mask_at<descr>(vmval, vmasg, i)) &&
			    ((descr & descriptors::dense) || x.has(i))) {
				typename ADD::D1 t;
				if (LEFT) {
					t = static_cast<typename ADD::D1>(x.get(i));
				} else {
					PROD p = static_cast<PROD>(id);
					apply_op<MUL>(p, ONE::template value<NZ>(), static_cast<SRC>(x.get(i)));
					t = static_cast<typename ADD::D1>(p);
				}
				out = static_cast<OUT>(t);
				set = true;
			}
			for (uint64_t k = off[i]; k < off[i + 1]; ++k) {
				const uint32_t j = idx[k];
				if (IN_MASKED && !mask_at<descr>(vmval, vmasg, j)) continue;
				if (!(descr & descriptors::dense) && !x.has(j)) continue;
				const NZ a = aval ? static_cast<NZ>(aval[k]) : ONE::template value<NZ>();
				const SRC s = static_cast<SRC>(x.get(j));
				PROD p = static_cast<PROD>(id);
				if (LEFT) apply_op<MUL>(p, s, a); else apply_op<MUL>(p, a, s);
				foldr_op<ADD>(p, out);  // out = p (+) out
				set = true;
			}
			if (set || (descr & descriptors::explicit_zero)) {
				if ((descr & descriptors::dense) || yin) foldl_op<ADD>(y[i], out);
				else y[i] = static_cast<IO>(out);
				yout = true;
			}
		}
		ya[i] = yout ? 1 : 0;
		c += yout ? 1 : 0;
	}
	if (!count) return;
	count_add(c, count);
}

// the push: for each input nonzero j, fold its products into the outputs i of
// row j of the transposed storage (vxm_inner_kernel_scatter, :646-769)
template <typename OUT>
__device__ __forceinline__ unsigned long long pack(const OUT &v) {
	unsigned long long u = 0;
	memcpy(&u, &v, sizeof(OUT));
	return u;
}
template <typename OUT>
__device__ __forceinline__ OUT unpack(unsigned long long u) {
	OUT v;
	memcpy(&v, &u, sizeof(OUT));
	return v;
}

template <Descriptor descr, bool LEFT, bool MASKED, bool IN_MASKED, class ADD, class MUL, class ID0, class ONE,
          typename X, typename V, typename M, typename VM>
__global__ void k_mv_push(size_t nsrc, const uint64_t *off, const uint32_t *idx, const V *aval, Arg<X> x,
                          unsigned long long *acc, uint8_t *touched, const M *mval, const uint8_t *masg,
                          const VM *vmval, const uint8_t *vmasg) {
	typedef typename ADD::D3 OUT;
	typedef typename MUL::D3 PROD;
	typedef typename std::conditional<LEFT, typename MUL::D2, typename MUL::D1>::type NZ;
	typedef typename std::conditional<LEFT, typename MUL::D1, typename MUL::D2>::type SRC;
	for (size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x; j < nsrc; j += (size_t)gridDim.x * blockDim.x) {
		if (!x.has(j)) continue;
		if (IN_MASKED && !mask_at<descr>(vmval, vmasg, j)) continue;
		const SRC s = static_cast<SRC>(x.get(j));
		for (uint64_t k = off[j]; k < off[j + 1]; ++k) {
			const uint32_t i = idx[k];
			if (MASKED && !mask_at<descr>(mval, masg, i)) continue;
			const NZ a = aval ? static_cast<NZ>(aval[k]) : ONE::template value<NZ>();
			PROD p = static_cast<PROD>(ID0::template value<OUT>());
			if (LEFT) apply_op<MUL>(p, s, a); else apply_op<MUL>(p, a, s);
			const OUT pv = static_cast<OUT>(p);
			unsigned long long old = acc[i], assumed;
			do {
				assumed = old;
				OUT cur = unpack<OUT>(assumed);
				foldl_op<ADD>(cur, pv);
				old = atomicCAS(acc + i, assumed, pack<OUT>(cur));
			} while (old != assumed);
			touched[i] = 1;
		}
	}
}

template <Descriptor descr, bool MASKED, class ADD, class ID0, typename IO, typename X, typename M>
__global__ void k_mv_push_finish(size_t m, size_t nsrc, Arg<X> x, const unsigned long long *acc,
                                 const uint8_t *touched, IO *y, uint8_t *ya, const M *mval, const uint8_t *masg,
                                 unsigned long long *count) {
	typedef typename ADD::D3 OUT;
	unsigned long long c = 0;
	for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (size_t)gridDim.x * blockDim.x) {
		const bool yin = ya[i] != 0;
		bool yout = yin;
		bool allowed = !MASKED || mask_at<descr>(mval, masg, i);
		OUT out = ID0::template value<OUT>();
		bool set = false;
		if (allowed && touched[i]) {
			out = unpack<OUT>(acc[i]);
			set = true;
		}
		if (allowed && (descr & descriptors::add_identity) && i < nsrc && x.has(i)) {
			OUT t = static_cast<OUT>(x.get(i));
			if (set) foldl_op<ADD>(out, t); else out = t;
			set = true;
		}
		if (allowed && (set || (descr & descriptors::explicit_zero))) {
			if (yin) foldl_op<ADD>(y[i], out); else y[i] = static_cast<IO>(out);
			yout = true;
		}
		ya[i] = yout ? 1 : 0;
		c += yout ? 1 : 0;
	}
	if (!count) return;
	count_add(c, count);
}

template <typename T>
__global__ void k_fill_identity(size_t n, unsigned long long *acc, uint8_t *touched, T id) {
	unsigned long long u = 0;
	memcpy(&u, &id, sizeof(T));
	for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
		acc[i] = u;
		touched[i] = 0;
	}
}

// explicit_zero in the trivial case: unassigned outputs get the identity
template <typename IO>
__global__ void k_explicit_zero(size_t n, IO *y, uint8_t *ya, IO id) {
	for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
		if (!ya[i]) {
			y[i] = id;
			ya[i] = 1;
		}
}

template <template <typename> class I> struct IdentityOf {
	template <typename D> static __host__ __device__ D value() { return I<D>::value(); }
};

// the library's kernels: dense vectors, no mask or a dense boolean value mask
template <Descriptor descr, int RING, bool MXV, typename IO, typename X, typename M, typename T, Backend b,
          typename C, typename R, typename Cc, typename N>
bool mv_library(Vector<IO, b, C> &u, const Vector<M, b, C> *mask, const Matrix<T, b, R, Cc, N> &A,
                const Vector<X, b, C> &v, RC &rc, std::true_type) {
	constexpr Descriptor supported = descriptors::invert_mask | descriptors::transpose_matrix | descriptors::dense |
	                                 descriptors::add_identity | descriptors::safe_overlap | descriptors::no_casting;
	if ((descr & ~supported) != 0 || !u.dense() || !v.dense() || !A.handle()) return false;
	const uint8_t *mp = nullptr;
	if (mask && mask->size()) {
		if (!std::is_same<M, bool>::value || !mask->dense()) return false;
		mp = reinterpret_cast<const uint8_t *>(mask->raw());
	}
	const unsigned d = descr & (descriptors::invert_mask | descriptors::transpose_matrix | descriptors::dense |
	                            descriptors::add_identity);
	(void)stream();  // enqueued: inside a pipeline this joins the open segment
	rc = MXV ? grb::internal::rc(alp_hip_mxv(A.handle(), (alp_semiring_t)RING, d, u.raw(), v.raw(), mp))
	         : grb::internal::rc(alp_hip_vxm(A.handle(), (alp_semiring_t)RING, d, u.raw(), v.raw(), mp));
	return true;
}
template <Descriptor descr, int RING, bool MXV, typename IO, typename X, typename M, typename T, Backend b,
          typename C, typename R, typename Cc, typename N>
bool mv_library(Vector<IO, b, C> &, const Vector<M, b, C> *, const Matrix<T, b, R, Cc, N> &, const Vector<X, b, C> &,
                RC &, std::false_type) {
	return false;
}

template <Descriptor descr, bool MXV, class ADDM, class MUL, template <typename> class One, typename IO, typename M,
          typename X, typename VM, typename T, Backend b, typename C, typename R, typename Cc, typename N>
RC mv(Vector<IO, b, C> &u, const Vector<M, b, C> *mask, const Matrix<T, b, R, Cc, N> &A, const Vector<X, b, C> &v,
      const Vector<VM, b, C> *v_mask, const Phase phase) {
	typedef typename ADDM::Operator ADD;
	typedef typename ADD::D3 OUT;
	constexpr bool tr = (descr & descriptors::transpose_matrix) != 0;
	// mxv gathers over the CRS, vxm over the CCS; transpose_matrix swaps them
	const bool use_crs = MXV != tr;
	const size_t out_n = use_crs ? A.rows() : A.cols();
	const size_t in_n = use_crs ? A.cols() : A.rows();
	if (u.size() != out_n || v.size() != in_n) return MISMATCH;
	const bool masked = mask && mask->size() > 0;
	const bool in_masked = v_mask && v_mask->size() > 0;
	if (descr & descriptors::dense) {
		if (!v.dense()) return ILLEGAL;
		if (masked && !mask->dense()) return ILLEGAL;
		if (in_masked && !v_mask->dense()) return ILLEGAL;
	}
	if (masked && mask->size() != out_n) return MISMATCH;
	if (in_masked && v_mask->size() != in_n) return MISMATCH;
	if (phase == RESIZE) return SUCCESS;
	// trivial cases (reference/blas2.hpp:1004-1031)
	if (v.nnz() == 0 || A.rows() == 0 || A.cols() == 0 || A.nonzeroes() == 0 ||
	    (masked && mask->nnz() == 0 && !(descr & descriptors::invert_mask)) ||
	    (in_masked && v_mask->nnz() == 0 && !(descr & descriptors::invert_mask))) {
		if ((descr & descriptors::explicit_zero) && u.size()) {
			hipLaunchKernelGGL(k_explicit_zero<IO>, dim3(grid(u.size())), dim3(kBlock), 0, stream(), u.size(), u.raw(),
			                   u.assigned(), static_cast<IO>(ADDM::template getIdentity<OUT>()));
			return u.recount();
		}
		return SUCCESS;
	}
	if (!(descr & descriptors::safe_overlap) && (const void *)u.raw() == (const void *)v.raw()) return OVERLAP;
	if (masked && (const void *)u.raw() == (const void *)mask->raw()) return OVERLAP;
	// the library's tuned SpMV
	constexpr int ring = lib_ring<ADD, MUL, IO, X, T>::value;
	if (!in_masked) {
		RC r = SUCCESS;
		if (mv_library<descr, ring, MXV>(u, mask, A, v, r,
		                                 std::integral_constant<bool, (ring >= 0) && lib_matrix_ok<ring, T>::value>())) {
			return r;  // u was dense and stays so
		}
	}
	typename Matrix<T, b, R, Cc, N>::Storage S;
	RC r = A.storage(use_crs ? 0 : 1, S);
	if (r) return r;
	// an output that was dense stays dense (entries are only added); otherwise the
	// kernel counts into u's device slot
	const bool dense_out = u.known_dense();
	unsigned long long *count = dense_out ? nullptr : u.count_slot();
	const M *mval = masked ? mask->raw() : nullptr;
	const uint8_t *masg = masked && !mask->known_dense() ? mask->assigned() : nullptr;
	const VM *vmval = in_masked ? v_mask->raw() : nullptr;
	const uint8_t *vmasg = in_masked && !v_mask->known_dense() ? v_mask->assigned() : nullptr;
	typedef IdentityOf<ADDM::template Identity> ID0;
	typedef IdentityOf<One> ONE;
	constexpr bool exact = std::is_integral<OUT>::value && exact_fold_op<ADD>::value && sizeof(OUT) <= 8 &&
	                       std::is_trivially_copyable<OUT>::value;
	const bool push = exact && v.nnz() * 16 <= v.size() && !(descr & descriptors::dense);
	if (push) {
		// the transposed storage lists, for input index j, the outputs it reaches
		typename Matrix<T, b, R, Cc, N>::Storage ST;
		r = A.storage(use_crs ? 1 : 0, ST);
		if (r) return r;
		// per-output accumulators and touched flags in the reused workspace
		unsigned long long *acc = static_cast<unsigned long long *>(workspace(out_n * (sizeof(unsigned long long) + 1)));
		uint8_t *touched = reinterpret_cast<uint8_t *>(acc + out_n);
		hipLaunchKernelGGL(k_fill_identity<OUT>, dim3(grid(out_n)), dim3(kBlock), 0, stream(), out_n, acc,
		                   touched, ADDM::template getIdentity<OUT>());
#define ALP_HIP_PUSH(MK, IM)                                                                                     \
	hipLaunchKernelGGL((k_mv_push<descr, !MXV, MK, IM, ADD, MUL, ID0, ONE, X, typename std::remove_cv<         \
	                                 typename std::remove_pointer<decltype(ST.val)>::type>::type, M, VM>),      \
	                   dim3(grid(in_n)), dim3(kBlock), 0, stream(), in_n, ST.off, ST.idx, ST.val, arg(v), acc, \
	                   touched, mval, masg, vmval, vmasg)
		if (masked) {
			if (in_masked) ALP_HIP_PUSH(true, true); else ALP_HIP_PUSH(true, false);
		} else {
			if (in_masked) ALP_HIP_PUSH(false, true); else ALP_HIP_PUSH(false, false);
		}
#undef ALP_HIP_PUSH
		if (masked)
			hipLaunchKernelGGL((k_mv_push_finish<descr, true, ADD, ID0, IO, X, M>), dim3(grid(out_n)), dim3(kBlock), 0,
			                   stream(), out_n, in_n, arg(v), (const unsigned long long *)acc,
			                   (const uint8_t *)touched, u.raw(), u.assigned(), mval, masg, count);
		else
			hipLaunchKernelGGL((k_mv_push_finish<descr, false, ADD, ID0, IO, X, M>), dim3(grid(out_n)), dim3(kBlock),
			                   0, stream(), out_n, in_n, arg(v), (const unsigned long long *)acc,
			                   (const uint8_t *)touched, u.raw(), u.assigned(), mval, masg, count);
	} else {
		typedef typename std::remove_cv<typename std::remove_pointer<decltype(S.val)>::type>::type VT;
		constexpr bool imm = has_immutable_nonzeroes<ADDM>::value;
#define ALP_HIP_GATHER(MK, IM)                                                                                    \
	hipLaunchKernelGGL((k_mv_gather<descr, !MXV, MK, IM, imm, ADD, MUL, ID0, ONE, IO, X, VT, M, VM>),            \
	                   dim3(grid(out_n)), dim3(kBlock), 0, stream(), out_n, in_n, S.off, S.idx, S.val, arg(v),     \
	                   u.raw(), u.assigned(), mval, masg, vmval, vmasg, count)
		if (masked) {
			if (in_masked) ALP_HIP_GATHER(true, true); else ALP_HIP_GATHER(true, false);
		} else {
			if (in_masked) ALP_HIP_GATHER(false, true); else ALP_HIP_GATHER(false, false);
		}
#undef ALP_HIP_GATHER
	}
	if (dense_out) u.set_nnz(out_n); else u.mark_counted();
	return from_hip(hipGetLastError());
}

}  // namespace hip
}  // namespace internal

// ---- mxv: u (+)= A (x) v (reference/blas2.hpp:1836-1946) -------------------------------

template <Descriptor descr = descriptors::no_operation, class Ring, typename IO, typename X, typename T,
          enum Backend b, typename C, typename R, typename Cc, typename N>
RC mxv(Vector<IO, b, C> &u, const Matrix<T, b, R, Cc, N> &A, const Vector<X, b, C> &v, const Ring & = Ring(),
       const Phase phase = EXECUTE, typename std::enable_if<is_semiring<Ring>::value>::type * = nullptr) {
	return internal::hip::mv<descr, true, typename Ring::AdditiveMonoid, typename Ring::MultiplicativeOperator,
	                         Ring::template One>(u, (const Vector<bool, b, C> *)nullptr, A, v,
	                                             (const Vector<bool, b, C> *)nullptr, phase);
}
template <Descriptor descr = descriptors::no_operation, class Ring, typename IO, typename M, typename X, typename T,
          enum Backend b, typename C, typename R, typename Cc, typename N>
RC mxv(Vector<IO, b, C> &u, const Vector<M, b, C> &mask, const Matrix<T, b, R, Cc, N> &A, const Vector<X, b, C> &v,
       const Ring & = Ring(), const Phase phase = EXECUTE,
       typename std::enable_if<is_semiring<Ring>::value>::type * = nullptr) {
	return internal::hip::mv<descr, true, typename Ring::AdditiveMonoid, typename Ring::MultiplicativeOperator,
	                         Ring::template One>(u, &mask, A, v, (const Vector<bool, b, C> *)nullptr, phase);
}
template <Descriptor descr = descriptors::no_operation, class Ring, typename IO, typename M, typename X, typename VM,
          typename T, enum Backend b, typename C, typename R, typename Cc, typename N>
RC mxv(Vector<IO, b, C> &u, const Vector<M, b, C> &mask, const Matrix<T, b, R, Cc, N> &A, const Vector<X, b, C> &v,
       const Vector<VM, b, C> &v_mask, const Ring & = Ring(), const Phase phase = EXECUTE,
       typename std::enable_if<is_semiring<Ring>::value>::type * = nullptr) {
	return internal::hip::mv<descr, true, typename Ring::AdditiveMonoid, typename Ring::MultiplicativeOperator,
	                         Ring::template One>(u, &mask, A, v, &v_mask, phase);
}
// (additive monoid, multiplicative operator) forms
template <Descriptor descr = descriptors::no_operation, class AddMonoid, class MulOp, typename IO, typename X,
          typename T, enum Backend b, typename C, typename R, typename Cc, typename N>
RC mxv(Vector<IO, b, C> &u, const Matrix<T, b, R, Cc, N> &A, const Vector<X, b, C> &v, const AddMonoid &,
       const MulOp &, const Phase phase = EXECUTE,
       typename std::enable_if<is_monoid<AddMonoid>::value && is_operator<MulOp>::value>::type * = nullptr) {
	return internal::hip::mv<descr, true, AddMonoid, MulOp, identities::one>(
	    u, (const Vector<bool, b, C> *)nullptr, A, v, (const Vector<bool, b, C> *)nullptr, phase);
}
template <Descriptor descr = descriptors::no_operation, class AddMonoid, class MulOp, typename IO, typename M,
          typename X, typename T, enum Backend b, typename C, typename R, typename Cc, typename N>
RC mxv(Vector<IO, b, C> &u, const Vector<M, b, C> &mask, const Matrix<T, b, R, Cc, N> &A, const Vector<X, b, C> &v,
       const AddMonoid &, const MulOp &, const Phase phase = EXECUTE,
       typename std::enable_if<is_monoid<AddMonoid>::value && is_operator<MulOp>::value>::type * = nullptr) {
	return internal::hip::mv<descr, true, AddMonoid, MulOp, identities::one>(u, &mask, A, v,
	                                                                         (const Vector<bool, b, C> *)nullptr, phase);
}

// ---- vxm: u (+)= v (x) A (reference/blas2.hpp:1540-1694) -------------------------------

template <Descriptor descr = descriptors::no_operation, class Ring, typename IO, typename X, typename T,
          enum Backend b, typename C, typename R, typename Cc, typename N>
RC vxm(Vector<IO, b, C> &u, const Vector<X, b, C> &v, const Matrix<T, b, R, Cc, N> &A, const Ring & = Ring(),
       const Phase phase = EXECUTE, typename std::enable_if<is_semiring<Ring>::value>::type * = nullptr) {
	return internal::hip::mv<descr, false, typename Ring::AdditiveMonoid, typename Ring::MultiplicativeOperator,
	                         Ring::template One>(u, (const Vector<bool, b, C> *)nullptr, A, v,
	                                             (const Vector<bool, b, C> *)nullptr, phase);
}
template <Descriptor descr = descriptors::no_operation, class Ring, typename IO, typename M, typename X, typename T,
          enum Backend b, typename C, typename R, typename Cc, typename N>
RC vxm(Vector<IO, b, C> &u, const Vector<M, b, C> &mask, const Vector<X, b, C> &v, const Matrix<T, b, R, Cc, N> &A,
       const Ring & = Ring(), const Phase phase = EXECUTE,
       typename std::enable_if<is_semiring<Ring>::value>::type * = nullptr) {
	return internal::hip::mv<descr, false, typename Ring::AdditiveMonoid, typename Ring::MultiplicativeOperator,
	                         Ring::template One>(u, &mask, A, v, (const Vector<bool, b, C> *)nullptr, phase);
}
template <Descriptor descr = descriptors::no_operation, class Ring, typename IO, typename M, typename X, typename VM,
          typename T, enum Backend b, typename C, typename R, typename Cc, typename N>
RC vxm(Vector<IO, b, C> &u, const Vector<M, b, C> &mask, const Vector<X, b, C> &v, const Vector<VM, b, C> &v_mask,
       const Matrix<T, b, R, Cc, N> &A, const Ring & = Ring(), const Phase phase = EXECUTE,
       typename std::enable_if<is_semiring<Ring>::value>::type * = nullptr) {
	return internal::hip::mv<descr, false, typename Ring::AdditiveMonoid, typename Ring::MultiplicativeOperator,
	                         Ring::template One>(u, &mask, A, v, &v_mask, phase);
}
template <Descriptor descr = descriptors::no_operation, class AddMonoid, class MulOp, typename IO, typename X,
          typename T, enum Backend b, typename C, typename R, typename Cc, typename N>
RC vxm(Vector<IO, b, C> &u, const Vector<X, b, C> &v, const Matrix<T, b, R, Cc, N> &A, const AddMonoid &,
       const MulOp &, const Phase phase = EXECUTE,
       typename std::enable_if<is_monoid<AddMonoid>::value && is_operator<MulOp>::value>::type * = nullptr) {
	return internal::hip::mv<descr, false, AddMonoid, MulOp, identities::one>(
	    u, (const Vector<bool, b, C> *)nullptr, A, v, (const Vector<bool, b, C> *)nullptr, phase);
}
template <Descriptor descr = descriptors::no_operation, class AddMonoid, class MulOp, typename IO, typename M,
          typename X, typename T, enum Backend b, typename C, typename R, typename Cc, typename N>
RC vxm(Vector<IO, b, C> &u, const Vector<M, b, C> &mask, const Vector<X, b, C> &v, const Matrix<T, b, R, Cc, N> &A,
       const AddMonoid &, const MulOp &, const Phase phase = EXECUTE,
       typename std::enable_if<is_monoid<AddMonoid>::value && is_operator<MulOp>::value>::type * = nullptr) {
	return internal::hip::mv<descr, false, AddMonoid, MulOp, identities::one>(u, &mask, A, v,
	                                                                          (const Vector<bool, b, C> *)nullptr, phase);
}

}  // namespace grb

#endif
