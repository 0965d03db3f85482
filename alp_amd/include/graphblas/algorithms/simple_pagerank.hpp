/*
 * grb::algorithms::simple_pagerank on the hip backend: the signature of ALP's
 * algorithms/simple_pagerank.hpp:146-533. The iteration (dangling mass, damping,
 * L1 residual, the `conv` test, at most `max` iterations; FAILED when not
 * converged) runs on the device in alp_hip_pagerank over the matrix's pattern;
 * the workspace vectors are checked for size and left as they were.
 */
#ifndef ALP_HIP_GRAPHBLAS_ALGORITHMS_PAGERANK_HPP
#define ALP_HIP_GRAPHBLAS_ALGORITHMS_PAGERANK_HPP

#include "../../graphblas.hpp"

namespace grb {
namespace algorithms {

template <Descriptor descr = descriptors::no_operation, typename IOType, typename NonzeroT, enum Backend b,
          typename C, typename R, typename Cc, typename N>
RC simple_pagerank(Vector<IOType, b, C> &pr, const Matrix<NonzeroT, b, R, Cc, N> &L, Vector<IOType, b, C> &pr_next,
                   Vector<IOType, b, C> &pr_nextnext, Vector<IOType, b, C> &row_sum, Vector<bool, b, C> &is_dangling,
                   const IOType alpha = 0.85, const IOType conv = 0.0000001, const size_t max = 1000,
                   size_t *const iterations = nullptr, double *const quality = nullptr) {
	static_assert(std::is_same<IOType, double>::value, "simple_pagerank on the hip backend: double ranks");
	const size_t n = nrows(L);
	// the reference's argument checks (simple_pagerank.hpp:178-203)
	if (ncols(L) != n) return ILLEGAL;
	if (size(pr) != n || size(pr_next) != n || size(pr_nextnext) != n || size(row_sum) != n ||
	    size(is_dangling) != n)
		return MISMATCH;
	if (!(alpha > 0.0 && alpha < 1.0) || max == 0) return ILLEGAL;
	if (n == 0) return SUCCESS;
	RC r = set(pr, 0.0);  // dense output; the device writes every rank
	if (r) return r;
	if (!L.handle()) {  // no edges: every vertex dangling, uniform ranks
		if (iterations) *iterations = 1;
		if (quality) *quality = 0.0;
		return set(pr, 1.0 / (double)n);
	}
	size_t it = 0;
	double q = 0;
	grb::internal::hip::Forcing force;  // the library's solver is blocking
	if (force.rc) return force.rc;
	r = grb::internal::rc(alp_hip_pagerank(L.handle(), pr.raw(), alpha, conv, max, &it, &q));
	if (iterations) *iterations = it;
	if (quality) *quality = q;
	return r;
}

}  // namespace algorithms
}  // namespace grb

#endif
