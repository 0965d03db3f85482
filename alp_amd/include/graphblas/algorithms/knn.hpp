/*
 * grb::algorithms::knn on the hip backend: the signature of ALP's
 * algorithms/knn.hpp:82-140. u holds true exactly at the vertices within k hops
 * of source (lor-land products with add_identity, knn.hpp:128-133 over
 * mpv.hpp:94-170), computed on the device by alp_hip_khop (a depth-limited BFS on
 * a symmetrised graph, k lor-land products otherwise).
 */
#ifndef ALP_HIP_GRAPHBLAS_ALGORITHMS_KNN_HPP
#define ALP_HIP_GRAPHBLAS_ALGORITHMS_KNN_HPP

#include "../../graphblas.hpp"

namespace grb {
namespace algorithms {

namespace internal {
template <typename F>
__global__ void k_khop_assigned(size_t n, const F *val, uint8_t *asg) {
	for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
		asg[i] = val[i] ? 1 : 0;
}
}  // namespace internal

template <Descriptor descr = descriptors::no_operation, typename InputType, enum Backend b, typename C, typename R,
          typename Cc, typename N>
RC knn(Vector<bool, b, C> &u, const Matrix<InputType, b, R, Cc, N> &A, const size_t source, const size_t k,
       Vector<bool, b, C> &buf1) {
	const size_t n = nrows(A);
	if (n != ncols(A)) return MISMATCH;
	if (size(buf1) != n || size(u) != n) return MISMATCH;
	if (source >= n) return ILLEGAL;
	if (n == 0) return SUCCESS;
	if (!A.handle()) {  // no edges: the source alone
		RC r = clear(u);
		return r ? r : setElement(u, true, source);
	}
	grb::internal::hip::Forcing force;  // the library's solver is blocking
	if (force.rc) return force.rc;
	RC r = grb::internal::rc(alp_hip_khop(A.handle(), (uint32_t)source, k, reinterpret_cast<uint8_t *>(u.raw())));
	if (r) return r;
	hipLaunchKernelGGL(internal::k_khop_assigned<bool>, dim3(grb::internal::hip::grid(n)),
	                   dim3(grb::internal::hip::kBlock), 0, grb::internal::hip::stream(), n,
	                   (const bool *)u.raw(), u.assigned());
	return u.recount();
}

}  // namespace algorithms
}  // namespace grb

#endif
