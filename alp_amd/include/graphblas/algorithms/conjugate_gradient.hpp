/*
 * grb::algorithms::conjugate_gradient on the hip backend: the signature of ALP's
 * algorithms/conjugate_gradient.hpp:493-583; the solve is the device CG of
 * alp_hip_cg (three fused kernels per iteration, replayed from a captured graph),
 * with the reference's recurrences, stopping rule (sqrt(r.r) < tol ||b||) and
 * outputs (iterations, residual = r.r, FAILED when not converged).
 */
#ifndef ALP_HIP_GRAPHBLAS_ALGORITHMS_CG_HPP
#define ALP_HIP_GRAPHBLAS_ALGORITHMS_CG_HPP

#include "../../graphblas.hpp"

namespace grb {
namespace algorithms {

template <Descriptor descr = descriptors::no_operation, typename IOType, typename ResidualType,
          typename NonzeroType, typename InputType,
          class Ring = Semiring<operators::add<IOType>, operators::mul<IOType>, identities::zero, identities::one>,
          class Minus = operators::subtract<IOType>, class Divide = operators::divide<IOType>, enum Backend b,
          typename R, typename C, typename N>
RC conjugate_gradient(Vector<IOType, b> &x, const Matrix<NonzeroType, b, R, C, N> &A,
                      const Vector<InputType, b> &b_, const size_t max_iterations, ResidualType tol,
                      size_t &iterations, ResidualType &residual, Vector<IOType, b> &r, Vector<IOType, b> &u,
                      Vector<IOType, b> &temp, const Ring & = Ring(), const Minus & = Minus(),
                      const Divide & = Divide()) {
	static_assert(std::is_same<IOType, double>::value && std::is_same<NonzeroType, double>::value &&
	                  std::is_same<InputType, double>::value,
	              "conjugate_gradient on the hip backend: double vectors and matrix");
	static_assert(internal::ring_id<Ring>::value == ALP_PLUS_TIMES_F64, "plus-times over double");
	const size_t n = nrows(A);
	if (ncols(A) != n || size(x) != n || size(b_) != n || size(r) != n || size(u) != n || size(temp) != n)
		return MISMATCH;
	// the workspace vectors r, u, temp are the reference's; the device solver keeps its own
	(void)r;
	(void)u;
	(void)temp;
	if (!(tol > ResidualType(0))) return ILLEGAL;
	if (max_iterations == 0) return ILLEGAL;
	iterations = 0;
	residual = std::numeric_limits<ResidualType>::infinity();
	if (n == 0) return SUCCESS;
	if (!A.handle()) return ILLEGAL;
	size_t it = 0;
	double res = 0;
	const RC rc = internal::rc(alp_hip_cg(A.handle(), b_.raw(), x.raw(), max_iterations, (double)tol, 0, &it, &res));
	iterations = it;
	residual = static_cast<ResidualType>(res);
	x.mark_dense();
	return rc;
}

}  // namespace algorithms
}  // namespace grb

#endif
