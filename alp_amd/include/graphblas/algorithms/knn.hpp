/*
 * grb::algorithms::knn on the hip backend: the signature of ALP's
 * algorithms/knn.hpp:82-140; u(v) = true iff v is within k hops of source
 * (lor-land products with add_identity, on the device: alp_hip_khop).
 */
#ifndef ALP_HIP_GRAPHBLAS_ALGORITHMS_KNN_HPP
#define ALP_HIP_GRAPHBLAS_ALGORITHMS_KNN_HPP

#include "../../graphblas.hpp"

namespace grb {
namespace algorithms {

template <Descriptor descr = descriptors::no_operation, typename InputType, enum Backend b, typename R,
          typename C, typename N>
RC knn(Vector<bool, b> &u, const Matrix<InputType, b, R, C, N> &A, const size_t source, const size_t k,
       Vector<bool, b> &buf1) {
	const size_t n = nrows(A);
	if (n != ncols(A)) return MISMATCH;
	if (size(buf1) != n || size(u) != n) return MISMATCH;
	if (source >= n) return ILLEGAL;
	if (n == 0) return SUCCESS;
	const RC rc = internal::rc(alp_hip_khop(A.handle(), (uint32_t)source, k, u.raw()));
	u.mark_dense();
	return rc;
}

}  // namespace algorithms
}  // namespace grb

#endif
