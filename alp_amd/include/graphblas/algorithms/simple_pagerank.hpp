/*
 * grb::algorithms::simple_pagerank on the hip backend, with the signature, checks
 * and loop of ALP's algorithms/simple_pagerank.hpp:146-533 written in grb::
 * primitives (the branch for backends whose lambdas cannot reduce into captured
 * scalars, :299-316 and :407-446): out-degrees by a pattern vxm, the dangling
 * mask, row scaling by alpha / degree, then per iteration the dangling mass (a
 * masked fold), pr_next = pr .* row_sum, pr_nextnext = pr_next L plus the
 * uniform share, the L1 residual (a dot under abs_diff), and a swap of pr and
 * pr_nextnext; FAILED when `conv` is not reached within `max` iterations.
 *
 * Without a descriptor, for double ranks that start from the reference's uniform
 * vector (pr not dense on entry) on a matrix the library holds, the iteration runs
 * on the device in alp_hip_pagerank instead (the same recurrences, fused); the
 * workspaces are then left as they were.
 */
#ifndef ALP_HIP_GRAPHBLAS_ALGORITHMS_PAGERANK_HPP
#define ALP_HIP_GRAPHBLAS_ALGORITHMS_PAGERANK_HPP

#include <utility>

#include "../../graphblas.hpp"

namespace grb {
namespace algorithms {

namespace internal {

template <enum Backend b, typename C, typename NonzeroT, typename R, typename Cc, typename N>
RC pagerank_device(Vector<double, b, C> &pr, const Matrix<NonzeroT, b, R, Cc, N> &L, const double alpha,
                   const double conv, const size_t max, size_t *const iterations, double *const quality, bool &done) {
	done = false;
	if (!L.handle() || nnz(pr) == size(pr)) return SUCCESS;
	done = true;
	RC r = set(pr, 0.0);  // dense output; the device writes every rank
	if (r) return r;
	size_t it = 0;
	double q = 0;
	grb::internal::hip::Forcing force;  // the library's solver is blocking
	if (force.rc) return force.rc;
	r = grb::internal::rc(alp_hip_pagerank(L.handle(), pr.raw(), alpha, conv, max, &it, &q));
	if (iterations) *iterations = it;
	if (quality) *quality = q;
	return r;
}
template <typename V, typename M, typename T>
RC pagerank_device(V &, const M &, const T, const T, const size_t, size_t *const, double *const, bool &done) {
	done = false;
	return SUCCESS;
}

}  // namespace internal

template <Descriptor descr = descriptors::no_operation, typename IOType, typename NonzeroT, enum Backend b,
          typename C, typename R, typename Cc, typename N>
RC simple_pagerank(Vector<IOType, b, C> &pr, const Matrix<NonzeroT, b, R, Cc, N> &L, Vector<IOType, b, C> &pr_next,
                   Vector<IOType, b, C> &pr_nextnext, Vector<IOType, b, C> &row_sum, Vector<bool, b, C> &is_dangling,
                   const IOType alpha = 0.85, const IOType conv = 0.0000001, const size_t max = 1000,
                   size_t *const iterations = nullptr, double *const quality = nullptr) {
	Monoid<operators::add<IOType>, identities::zero> addM;
	Semiring<operators::add<IOType>, operators::mul<IOType>, identities::zero, identities::one> realRing;
	const size_t n = nrows(L);
	const IOType zero = realRing.template getZero<IOType>();
	// the reference's checks (simple_pagerank.hpp:178-203)
	if (n != ncols(L)) return ILLEGAL;
	if (size(pr) != n) return MISMATCH;
	if (size(pr_next) != n || size(pr_nextnext) != n || size(row_sum) != n) return MISMATCH;
	if (alpha <= 0 || alpha >= 1) return ILLEGAL;
	if (max <= 0) return ILLEGAL;
	if (descr == descriptors::no_operation) {
		bool done = false;
		const RC r = internal::pagerank_device(pr, L, alpha, conv, max, iterations, quality, done);
		if (done) return r;
	}
	// prelude (:205-297): start vector, out-degrees, dangling vertices, row scaling
	RC ret = SUCCESS;
	if (nnz(pr) != n) ret = set(pr, static_cast<IOType>(1) / static_cast<IOType>(n));
	ret = ret ? ret : set(pr_nextnext, zero);
	{
		Semiring<operators::add<IOType>, operators::left_assign_if<IOType, bool, IOType>, identities::zero,
		         identities::logical_true>
		    pattern_ring;
		ret = ret ? ret : set(pr_next, 1);
		ret = ret ? ret : set(row_sum, 0);
		ret = ret ? ret
		          : vxm<descr | descriptors::dense | descriptors::transpose_matrix>(row_sum, pr_next, L,
		                                                                            pattern_ring);
	}
	{
		// pr_next marks the dangling vertices (out-degree zero)
		const DeviceView<IOType> rs = view(row_sum), pn = view(pr_next);
		ret = ret ? ret
		          : eWiseLambda([=] __device__(const size_t i) { pn[i] = rs[i] == zero ? 1 : 0; }, row_sum, pr_next);
		IOType dangling_count = zero;
		ret = ret ? ret : foldl<descriptors::dense>(dangling_count, pr_next, addM);
		if (ret == SUCCESS && (double)size(is_dangling) < (double)dangling_count) return ILLEGAL;
		ret = ret ? ret : clear(is_dangling);
		ret = ret ? ret : set(is_dangling, pr_next, true);
	}
	{
		const DeviceView<IOType> rs = view(row_sum);
		const IOType a = alpha;
		ret = ret ? ret : eWiseLambda([=] __device__(const size_t i) {
			if (rs[i] > zero) rs[i] = a / rs[i];
		}, row_sum);
	}
	// iterations (:299-498)
	size_t iter = 0;
	IOType dangling = zero, residual = zero;
	do {
		residual = dangling = 0;
		ret = ret ? ret : foldl<descriptors::structural>(dangling, pr, is_dangling, addM);
		ret = ret ? ret : eWiseApply<descriptors::dense>(pr_next, pr, row_sum, operators::mul<IOType>());
		if (ret == SUCCESS) dangling = (alpha * dangling + 1 - alpha) / static_cast<IOType>(n);
		ret = ret ? ret : set<descriptors::dense>(pr_nextnext, 0);
		ret = ret ? ret : vxm<descriptors::dense | descr>(pr_nextnext, pr_next, L, realRing);
		if (ret == SUCCESS) {
			ret = foldl<descriptors::dense>(pr_nextnext, dangling, addM);
			if (ret == SUCCESS) {
				residual = zero;
				ret = dot<descriptors::dense>(residual, pr, pr_nextnext, addM, operators::abs_diff<IOType>());
			}
			if (ret == SUCCESS) std::swap(pr, pr_nextnext);
		}
		++iter;
		if (conv != zero && residual <= conv) break;
	} while (ret == SUCCESS && iter < max);
	if (iterations) *iterations = iter;
	if (quality) *quality = residual;
	if (ret != SUCCESS) return ret;
	return residual <= conv ? SUCCESS : FAILED;
}

}  // namespace algorithms
}  // namespace grb

#endif
