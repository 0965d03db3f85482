/*
 * grb::algorithms::mpv on the hip backend: k products with A over any semiring,
 * with the signature and loop of ALP's algorithms/mpv.hpp:94-170. The products
 * run as grb::mxv calls, alternately into temp and u, and u and temp are swapped
 * at the end when the last product went into temp. As in the reference, each
 * product is folded into what its output already holds (mxv is in place) and
 * only temp is cleared, once: over plus-times k = 3 yields A v + A^3 v; over
 * lor-land with add_identity (knn) the k-hop set. k = 0 copies v.
 */
#ifndef ALP_HIP_GRAPHBLAS_ALGORITHMS_MPV_HPP
#define ALP_HIP_GRAPHBLAS_ALGORITHMS_MPV_HPP

#include <utility>

#include "../../graphblas.hpp"

namespace grb {
namespace algorithms {

template <Descriptor descr, class Ring, typename IOType, typename InputType, enum Backend b, typename C,
          typename R, typename Cc, typename N>
RC mpv(Vector<IOType, b, C> &u, const Matrix<InputType, b, R, Cc, N> &A, const size_t k,
       const Vector<IOType, b, C> &v, Vector<IOType, b, C> &temp, const Ring &ring) {
	static_assert(!(descr & descriptors::no_casting) ||
	                  (std::is_same<IOType, typename Ring::D4>::value &&
	                   std::is_same<InputType, typename Ring::D2>::value &&
	                   std::is_same<IOType, typename Ring::D1>::value &&
	                   std::is_same<IOType, typename Ring::D3>::value),
	              "grb::mpv: container element types differ from the semiring's domains (no_casting)");
	const size_t n = nrows(A);
	// the reference's checks (mpv.hpp:117-132)
	if (n != ncols(A)) return ILLEGAL;
	if (size(u) != n || size(v) != n || size(temp) != n) return MISMATCH;
	if (k == 0) return set<descr>(u, v);
	RC ret = mxv<descr>(u, A, v, ring);
	if (k == 1) return ret;
	// v may be temp itself (knn passes one buffer for both): it is consumed now
	ret = ret ? ret : clear(temp);
	bool into_temp = false;
	for (size_t it = 1; ret == SUCCESS && it < k; it += 2) {
		into_temp = true;
		ret = mxv<descr>(temp, A, u, ring);
		if (it + 1 == k || ret != SUCCESS) break;
		into_temp = false;
		ret = mxv<descr>(u, A, temp, ring);
	}
	if (ret == SUCCESS && into_temp) std::swap(u, temp);
	return ret;
}

}  // namespace algorithms
}  // namespace grb

#endif
