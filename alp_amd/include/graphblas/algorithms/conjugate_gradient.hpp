/*
 * grb::algorithms::conjugate_gradient and preconditioned_conjugate_gradient on
 * the hip backend, with the signatures of ALP's algorithms/conjugate_gradient.hpp
 * (:191-484 the preconditioned loop, :493-583 the plain entry point).
 *
 * The loop is the reference's, written in grb:: primitives in its order (so it
 * runs over any ring and domain the primitives support): r = b - A x, u = z,
 * sigma = r.z; per iteration temp = A u, alpha = sigma / (temp.u), x += alpha u,
 * r -= alpha temp, the stopping test sqrt(r.r) < tol ||b||, z = M^-1 r,
 * u = z + (r.z / sigma) u. residual is r.r; FAILED when the tolerance was not
 * reached. A sparse x is first reset to zero, as the reference's prelude does.
 *
 * For FP64 plus-times on a matrix whose values the library holds, the plain
 * entry point runs the library's device CG instead (alp_hip_cg: the same
 * recurrences and stopping rule, three fused kernels per iteration replayed
 * from a captured graph); afterwards r holds b - A x of the returned x and
 * temp holds A x.
 */
#ifndef ALP_HIP_GRAPHBLAS_ALGORITHMS_CG_HPP
#define ALP_HIP_GRAPHBLAS_ALGORITHMS_CG_HPP

#include <cmath>
#include <functional>

#include "../../graphblas.hpp"

namespace grb {
namespace algorithms {

namespace internal {

template <Descriptor descr, bool PRECONDITIONED, typename IOType, typename ResidualType, typename NonzeroType,
          typename InputType, class Ring, class Minus, class Divide, enum Backend b, typename C, typename R,
          typename Cc, typename N, class Precond>
RC cg_loop(Vector<IOType, b, C> &x, const Matrix<NonzeroType, b, R, Cc, N> &A, const Vector<InputType, b, C> &bv,
           const size_t max_iterations, ResidualType tol, size_t &iterations, ResidualType &residual,
           Vector<IOType, b, C> &r, Vector<IOType, b, C> &u, Vector<IOType, b, C> &temp, Vector<IOType, b, C> &z,
           const Ring &ring, const Minus &minus, const Divide &, const Precond &Minv) {
	constexpr Descriptor dd = descr | descriptors::dense;
	typedef typename Divide::D3 DT;
	const IOType zero = ring.template getZero<IOType>();
	const size_t n = nrows(A);
	RC ret = SUCCESS;
	if (nnz(x) != n) ret = set<descriptors::invert_mask | descriptors::structural>(x, x, zero);
	if (ret) return ret;
	IOType sigma, bnorm, alpha, beta;
	ret = set(temp, 0);
	ret = ret ? ret : mxv<dd>(temp, A, x, ring);
	ret = ret ? ret : set(r, zero);
	ret = ret ? ret : foldl(r, bv, ring.getAdditiveMonoid());
	ret = ret ? ret : foldl<dd>(r, temp, minus);
	if (PRECONDITIONED) {
		ret = ret ? ret : set(z, 0);
		ret = ret ? ret : Minv(z, r);
	}
	ret = ret ? ret : set(u, z);
	sigma = zero;
	ret = ret ? ret : dot<dd>(sigma, r, z, ring);
	bnorm = zero;
	ret = ret ? ret : dot<dd>(bnorm, bv, bv, ring);
	if (ret) return ret;
	tol *= std::sqrt(bnorm);
	size_t iter = 0;
	do {
		++iter;
		ret = ret ? ret : set<dd>(temp, 0);
		ret = ret ? ret : mxv<dd>(temp, A, u, ring);
		beta = zero;
		ret = ret ? ret : dot<dd>(beta, temp, u, ring);
		{
			DT a = DT();
			grb::internal::apply_op<Divide>(a, sigma, beta);
			alpha = static_cast<IOType>(a);
		}
		ret = ret ? ret : eWiseMul<dd>(x, alpha, u, ring);
		ret = ret ? ret : foldr<dd>(alpha, temp, ring.getMultiplicativeMonoid());
		ret = ret ? ret : foldl<dd>(r, temp, minus);
		alpha = zero;
		ret = ret ? ret : dot<dd>(alpha, r, r, ring);
		residual = static_cast<ResidualType>(alpha);
		if (ret == SUCCESS && (std::sqrt(residual) < tol || iter >= max_iterations)) break;
		if (PRECONDITIONED) {
			beta = zero;
			ret = ret ? ret : Minv(z, r);
			ret = ret ? ret : dot<dd>(beta, r, z, ring);
		} else {
			beta = alpha;
		}
		{
			DT a = DT();
			grb::internal::apply_op<Divide>(a, beta, sigma);
			alpha = static_cast<IOType>(a);
		}
		ret = ret ? ret : foldr<dd>(alpha, u, ring.getMultiplicativeMonoid());
		ret = ret ? ret : foldr<dd>(z, u, ring.getAdditiveMonoid());
		sigma = beta;
	} while (ret == SUCCESS);
	iterations = iter;
	if (ret == SUCCESS && std::sqrt(residual) >= tol) return FAILED;
	return ret;
}

struct NoPrecond {
	template <typename V> RC operator()(V &, const V &) const { return SUCCESS; }
};

// the library's device CG: FP64 plus-times, values held by the library
template <typename IOType, typename NonzeroType, typename InputType, class Ring, class Minus>
struct device_cg {
	static constexpr bool value =
	    std::is_same<IOType, double>::value && std::is_same<NonzeroType, double>::value &&
	    std::is_same<InputType, double>::value &&
	    std::is_same<typename Ring::AdditiveOperator, operators::add<double>>::value &&
	    std::is_same<typename Ring::MultiplicativeOperator, operators::mul<double>>::value &&
	    std::is_same<Minus, operators::subtract<double>>::value;
};

template <typename ResidualType, enum Backend b, typename C, typename R, typename Cc, typename N>
RC cg_device(Vector<double, b, C> &x, const Matrix<double, b, R, Cc, N> &A, const Vector<double, b, C> &bv,
             const size_t max_iterations, ResidualType tol, size_t &iterations, ResidualType &residual,
             Vector<double, b, C> &r, Vector<double, b, C> &temp, bool &done, std::true_type) {
	done = false;
	if (!A.handle() || !bv.dense()) return SUCCESS;
	done = true;
	RC ret = SUCCESS;
	if (!x.dense()) ret = set<descriptors::invert_mask | descriptors::structural>(x, x, 0.0);
	if (ret) return ret;
	size_t it = 0;
	double res = 0.0;
	grb::internal::hip::Forcing force;  // the library's solver is blocking
	if (force.rc) return force.rc;
	const RC rc = grb::internal::rc(
	    alp_hip_cg(A.handle(), bv.raw(), x.raw(), max_iterations, (double)tol, 0, &it, &res));
	iterations = it;
	residual = static_cast<ResidualType>(res);
	if (rc != SUCCESS && rc != FAILED) return rc;
	// r = b - A x, temp = A x of the returned x
	Semiring<operators::add<double>, operators::mul<double>, identities::zero, identities::one> ring;
	ret = set(temp, 0.0);
	ret = ret ? ret : mxv(temp, A, x, ring);
	ret = ret ? ret : set(r, bv);
	ret = ret ? ret : foldl(r, temp, operators::subtract<double>());
	return ret ? ret : rc;
}
template <typename ResidualType, typename X, typename A_, typename B_, typename V>
RC cg_device(X &, const A_ &, const B_ &, const size_t, ResidualType, size_t &, ResidualType &, V &, V &, bool &done,
             std::false_type) {
	done = false;
	return SUCCESS;
}

template <typename IOType, typename ResidualType, typename NonzeroType, typename InputType, enum Backend b,
          typename C, typename R, typename Cc, typename N>
RC cg_checks(const Vector<IOType, b, C> &x, const Matrix<NonzeroType, b, R, Cc, N> &A,
             const Vector<InputType, b, C> &bv, const size_t max_iterations, const ResidualType tol,
             const Vector<IOType, b, C> &r, const Vector<IOType, b, C> &u, const Vector<IOType, b, C> &temp) {
	const size_t n = nrows(A);
	if (ncols(A) != n || size(x) != n || size(bv) != n || size(r) != n || size(u) != n || size(temp) != n)
		return MISMATCH;
	// the reference's argument checks (conjugate_gradient.hpp:300-312)
	if (!(tol > ResidualType(0))) return ILLEGAL;
	if (max_iterations == 0) return ILLEGAL;
	return SUCCESS;
}

}  // namespace internal

// plain CG (conjugate_gradient.hpp:493-583)
template <Descriptor descr = descriptors::no_operation, typename IOType, typename ResidualType,
          typename NonzeroType, typename InputType,
          class Ring = Semiring<operators::add<IOType>, operators::mul<IOType>, identities::zero, identities::one>,
          class Minus = operators::subtract<IOType>, class Divide = operators::divide<IOType>, enum Backend b,
          typename C, typename R, typename Cc, typename N>
RC conjugate_gradient(Vector<IOType, b, C> &x, const Matrix<NonzeroType, b, R, Cc, N> &A,
                      const Vector<InputType, b, C> &bv, const size_t max_iterations, ResidualType tol,
                      size_t &iterations, ResidualType &residual, Vector<IOType, b, C> &r, Vector<IOType, b, C> &u,
                      Vector<IOType, b, C> &temp, const Ring &ring = Ring(), const Minus &minus = Minus(),
                      const Divide &divide = Divide()) {
	RC ret = internal::cg_checks(x, A, bv, max_iterations, tol, r, u, temp);
	if (ret) return ret;
	iterations = 0;
	residual = std::numeric_limits<ResidualType>::infinity();
	if (nrows(A) == 0) return SUCCESS;
	bool done = false;
	ret = internal::cg_device(
	    x, A, bv, max_iterations, tol, iterations, residual, r, temp, done,
	    std::integral_constant<bool, internal::device_cg<IOType, NonzeroType, InputType, Ring, Minus>::value>());
	if (done) return ret;
	// without a preconditioner z is r (by reference)
	return internal::cg_loop<descr, false>(x, A, bv, max_iterations, tol, iterations, residual, r, u, temp, r, ring,
	                                       minus, divide, internal::NoPrecond());
}

// CG with a preconditioner z = M^-1 r given as Minv(z, r) -> RC
// (conjugate_gradient.hpp:191-484)
template <Descriptor descr = descriptors::no_operation, typename IOType, typename ResidualType,
          typename NonzeroType, typename InputType,
          class Ring = Semiring<operators::add<IOType>, operators::mul<IOType>, identities::zero, identities::one>,
          class Minus = operators::subtract<IOType>, class Divide = operators::divide<IOType>, enum Backend b,
          typename C, typename R, typename Cc, typename N>
RC preconditioned_conjugate_gradient(
    Vector<IOType, b, C> &x, const Matrix<NonzeroType, b, R, Cc, N> &A, const Vector<InputType, b, C> &bv,
    const std::function<RC(Vector<IOType, b, C> &, const Vector<IOType, b, C> &)> &Minv, const size_t max_iterations,
    ResidualType tol, size_t &iterations, ResidualType &residual, Vector<IOType, b, C> &r, Vector<IOType, b, C> &u,
    Vector<IOType, b, C> &temp, Vector<IOType, b, C> &temp_precond, const Ring &ring = Ring(),
    const Minus &minus = Minus(), const Divide &divide = Divide()) {
	RC ret = internal::cg_checks(x, A, bv, max_iterations, tol, r, u, temp);
	if (ret) return ret;
	if (size(temp_precond) != nrows(A)) return MISMATCH;
	iterations = 0;
	residual = std::numeric_limits<ResidualType>::infinity();
	if (nrows(A) == 0) return SUCCESS;
	return internal::cg_loop<descr, true>(x, A, bv, max_iterations, tol, iterations, residual, r, u, temp,
	                                      temp_precond, ring, minus, divide, Minv);
}

}  // namespace algorithms
}  // namespace grb

#endif
