/*
 * grb::algorithms::knn on the hip backend, with the signature and steps of ALP's
 * algorithms/knn.hpp:82-140: u holds true exactly at the vertices within k hops
 * of source, by k lor-land products with add_identity (mpv.hpp:94-170) seeded
 * with source alone in buf1; transpose_matrix in descr flips the direction.
 *
 * Without a descriptor, for a boolean u on a matrix the library holds, the hops
 * run on the device in alp_hip_khop instead (a breadth-first search cut at depth
 * k on a graph built symmetrised, k boolean products otherwise); u is the same
 * set, and buf1 is left as it was.
 */
#ifndef ALP_HIP_GRAPHBLAS_ALGORITHMS_KNN_HPP
#define ALP_HIP_GRAPHBLAS_ALGORITHMS_KNN_HPP

#include "../../graphblas.hpp"
#include "mpv.hpp"

namespace grb {
namespace algorithms {

namespace internal {
template <typename F>
__global__ void k_khop_assigned(size_t n, const F *val, uint8_t *asg) {
	for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
		asg[i] = val[i] ? 1 : 0;
}

template <enum Backend b, typename C, typename InputType, typename R, typename Cc, typename N>
RC knn_device(Vector<bool, b, C> &u, const Matrix<InputType, b, R, Cc, N> &A, const size_t source, const size_t k,
              bool &done) {
	done = false;
	const size_t n = nrows(A);
	if (!A.handle() || source >= n) return SUCCESS;
	done = true;
	grb::internal::hip::Forcing force;  // the library's search is blocking
	if (force.rc) return force.rc;
	RC r = grb::internal::rc(alp_hip_khop(A.handle(), (uint32_t)source, k, reinterpret_cast<uint8_t *>(u.raw())));
	if (r) return r;
	hipLaunchKernelGGL(k_khop_assigned<bool>, dim3(grb::internal::hip::grid(n)), dim3(grb::internal::hip::kBlock), 0,
	                   grb::internal::hip::stream(), n, (const bool *)u.raw(), u.assigned());
	return u.recount();
}
template <typename U, typename A_>
RC knn_device(U &, const A_ &, const size_t, const size_t, bool &done) {
	done = false;
	return SUCCESS;
}
}  // namespace internal

template <Descriptor descr = descriptors::no_operation, typename OutputType, typename InputType, enum Backend b,
          typename C, typename R, typename Cc, typename N>
RC knn(Vector<OutputType, b, C> &u, const Matrix<InputType, b, R, Cc, N> &A, const size_t source, const size_t k,
       Vector<bool, b, C> &buf1) {
	Semiring<operators::logical_or<bool>, operators::logical_and<bool>, identities::logical_false,
	         identities::logical_true>
	    ring;
	// the reference's checks (knn.hpp:93-110); a source out of range is setElement's
	// MISMATCH below
	const size_t n = nrows(A);
	if (n != ncols(A)) return MISMATCH;
	if (size(buf1) != n || size(u) != n) return MISMATCH;
	if (descr == descriptors::no_operation) {
		bool done = false;
		const RC r = internal::knn_device(u, A, source, k, done);
		if (done) return r;
	}
	RC ret = SUCCESS;
	if (nnz(u) != 0) ret = clear(u);
	if (nnz(buf1) != 0) ret = ret ? ret : clear(buf1);
	ret = ret ? ret : setElement(buf1, true, source);
	if (ret == SUCCESS) {
		if (descr & descriptors::transpose_matrix)
			ret = mpv<(descr | descriptors::add_identity) & ~(descriptors::transpose_matrix)>(u, A, k, buf1, buf1,
			                                                                                   ring);
		else
			ret = mpv<descr | descriptors::add_identity | descriptors::transpose_matrix>(u, A, k, buf1, buf1, ring);
	}
	return ret;
}

}  // namespace algorithms
}  // namespace grb

#endif
