/*
 * grb::algorithms::simple_pagerank on the hip backend: the signature of ALP's
 * algorithms/simple_pagerank.hpp:146-533 (the workspace vectors are accepted and
 * left untouched); the iteration is alp_hip_pagerank on the device.
 */
#ifndef ALP_HIP_GRAPHBLAS_ALGORITHMS_PAGERANK_HPP
#define ALP_HIP_GRAPHBLAS_ALGORITHMS_PAGERANK_HPP

#include "../../graphblas.hpp"

namespace grb {
namespace algorithms {

template <Descriptor descr = descriptors::no_operation, typename IOType, typename NonzeroT, enum Backend b,
          typename R, typename C, typename N>
RC simple_pagerank(Vector<IOType, b> &pr, const Matrix<NonzeroT, b, R, C, N> &L, Vector<IOType, b> &pr_next,
                   Vector<IOType, b> &pr_nextnext, Vector<IOType, b> &row_sum, Vector<bool, b> &is_dangling,
                   const IOType alpha = 0.85, const IOType conv = 0.0000001, const size_t max = 1000,
                   size_t *const iterations = nullptr, double *const quality = nullptr) {
	static_assert(std::is_same<IOType, double>::value, "simple_pagerank on the hip backend: double ranks");
	const size_t n = nrows(L);
	if (ncols(L) != n) return ILLEGAL;
	if (size(pr) != n || size(pr_next) != n || size(pr_nextnext) != n || size(row_sum) != n ||
	    size(is_dangling) != n)
		return MISMATCH;
	if (n == 0) return SUCCESS;
	if (!L.handle()) return ILLEGAL;
	size_t it = 0;
	double q = 0;
	const RC rc = internal::rc(alp_hip_pagerank(L.handle(), pr.raw(), alpha, conv, max, &it, &q));
	if (iterations) *iterations = it;
	if (quality) *quality = q;
	pr.mark_dense();
	return rc;
}

}  // namespace algorithms
}  // namespace grb

#endif
