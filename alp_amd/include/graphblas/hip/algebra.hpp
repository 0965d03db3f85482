/*
 * algebra.hpp -- identities, binary operators, monoids and semirings of the
 * grb:: header layer (ALP's identities.hpp, ops.hpp, base/internalops.hpp,
 * monoid.hpp:139-147, semiring.hpp:190).
 *
 * Every operator is a class with the domains D1 (left), D2 (right), D3 (result)
 * and a static `__host__ __device__ void apply(const D1 *a, const D2 *b, D3 *c)`,
 * the form of ALP's internal operators (base/internalops.hpp). The device kernels
 * of this layer are templates on these classes, so a program may pass any of
 * them -- or an operator of its own with the same members -- in a Monoid or
 * Semiring and the kernels for that algebra are compiled into the program.
 */
#ifndef ALP_HIP_GRB_ALGEBRA_HPP
#define ALP_HIP_GRB_ALGEBRA_HPP

#include "../utils/iscomplex.hpp"
#include "base.hpp"

namespace grb {

#define ALP_HD __host__ __device__ __forceinline__

// ---- identities (identities.hpp) ---------------------------------------------------

namespace identities {
template <typename D> struct zero { static ALP_HD constexpr D value() { return static_cast<D>(0); } };
template <typename D> struct one { static ALP_HD constexpr D value() { return static_cast<D>(1); } };
template <typename D> struct infinity {
	static ALP_HD constexpr D value() {
		return std::numeric_limits<D>::has_infinity ? std::numeric_limits<D>::infinity() : std::numeric_limits<D>::max();
	}
};
template <typename D> struct negative_infinity {
	static ALP_HD constexpr D value() {
		return std::numeric_limits<D>::has_infinity ? -std::numeric_limits<D>::infinity()
		                                            : std::numeric_limits<D>::lowest();
	}
};
template <typename D> struct logical_false { static ALP_HD constexpr D value() { return static_cast<D>(false); } };
template <typename D> struct logical_true { static ALP_HD constexpr D value() { return static_cast<D>(true); } };
}  // namespace identities

// ---- operators (ops.hpp; semantics of base/internalops.hpp) ----------------------

namespace operators {

#define ALP_HIP_OPERATOR(NAME, ASSOC, COMM, BODY)                                           \
	template <typename D1_, typename D2_ = D1_, typename D3_ = D2_, enum Backend impl = hip> \
	struct NAME {                                                                           \
		typedef D1_ D1;                                                                     \
		typedef D2_ D2;                                                                     \
		typedef D3_ D3;                                                                     \
		typedef D1_ left_type;                                                              \
		typedef D2_ right_type;                                                             \
		typedef D3_ result_type;                                                            \
		static constexpr bool is_associative = ASSOC;                                       \
		static constexpr bool is_commutative = COMM;                                        \
		static ALP_HD void apply(const D1 *a, const D2 *b, D3 *c) { BODY }                  \
	};

// c = a (the right operand is ignored)
ALP_HIP_OPERATOR(left_assign, true, false, (void)b; *c = static_cast<D3>(*a);)
// c = b
ALP_HIP_OPERATOR(right_assign, true, false, (void)a; *c = static_cast<D3>(*b);)
// c = a when b holds, else c keeps its value
ALP_HIP_OPERATOR(left_assign_if, false, false, if (static_cast<bool>(*b)) *c = static_cast<D3>(*a);)
// c = b when a holds, else c keeps its value
ALP_HIP_OPERATOR(right_assign_if, false, false, if (static_cast<bool>(*a)) *c = static_cast<D3>(*b);)
ALP_HIP_OPERATOR(add, true, true, *c = static_cast<D3>(*a + *b);)
ALP_HIP_OPERATOR(mul, true, true, *c = static_cast<D3>(*a * *b);)
ALP_HIP_OPERATOR(max, true, true, if (*a < *b) *c = static_cast<D3>(*b); else *c = static_cast<D3>(*a);)
ALP_HIP_OPERATOR(min, true, true, if (*a > *b) *c = static_cast<D3>(*b); else *c = static_cast<D3>(*a);)
ALP_HIP_OPERATOR(subtract, false, false, *c = static_cast<D3>(*a - *b);)
ALP_HIP_OPERATOR(divide, false, false, *c = static_cast<D3>(*a / *b);)
ALP_HIP_OPERATOR(divide_reverse, false, false, *c = static_cast<D3>(*b / *a);)
ALP_HIP_OPERATOR(equal, false, true, *c = static_cast<D3>(*a == *b);)
ALP_HIP_OPERATOR(not_equal, false, true, *c = static_cast<D3>(*a != *b);)
// the first operand that evaluates true, else the left one
ALP_HIP_OPERATOR(any_or, true, true,
                 if (*a) *c = static_cast<D3>(*a); else if (*b) *c = static_cast<D3>(*b); else *c = static_cast<D3>(*a);)
ALP_HIP_OPERATOR(logical_or, true, true, *c = static_cast<D3>(*a || *b);)
ALP_HIP_OPERATOR(logical_and, true, true, *c = static_cast<D3>(*a && *b);)
ALP_HIP_OPERATOR(relu, false, false, if (*a < *b) *c = static_cast<D3>(*b); else *c = static_cast<D3>(*a);)
ALP_HIP_OPERATOR(abs_diff, false, true,
                 if (*a < *b) *c = static_cast<D3>(*b - *a); else *c = static_cast<D3>(*a - *b);)
ALP_HIP_OPERATOR(square_diff, false, true, *c = static_cast<D3>((*a - *b) * (*a - *b));)
ALP_HIP_OPERATOR(less_than, false, false, *c = static_cast<D3>(*a < *b);)
ALP_HIP_OPERATOR(leq, false, false, *c = static_cast<D3>(*a <= *b);)
ALP_HIP_OPERATOR(greater_than, false, false, *c = static_cast<D3>(*a > *b);)
ALP_HIP_OPERATOR(geq, false, false, *c = static_cast<D3>(*a >= *b);)

#undef ALP_HIP_OPERATOR

// c = conj(a) b and c = a conj(b) (ops.hpp:900-990 over internalops.hpp's
// conjugate_mul); for real domains conjugation is the identity, and the
// operators are associative and commutative unless a domain is complex
#define ALP_HIP_CONJ_MUL(NAME, BODY)                                                        \
	template <typename D1_, typename D2_ = D1_, typename D3_ = D2_, enum Backend impl = hip> \
	struct NAME {                                                                           \
		typedef D1_ D1;                                                                     \
		typedef D2_ D2;                                                                     \
		typedef D3_ D3;                                                                     \
		typedef D1_ left_type;                                                              \
		typedef D2_ right_type;                                                             \
		typedef D3_ result_type;                                                            \
		static constexpr bool is_associative =                                              \
		    !utils::is_complex<D1>::value && !utils::is_complex<D2>::value;                 \
		static constexpr bool is_commutative = is_associative;                              \
		static ALP_HD void apply(const D1 *a, const D2 *b, D3 *c) { BODY }                  \
	};
ALP_HIP_CONJ_MUL(conjugate_left_mul, *c = static_cast<D3>(utils::is_complex<D1>::conjugate(*a) * *b);)
ALP_HIP_CONJ_MUL(conjugate_right_mul, *c = static_cast<D3>(*a * utils::is_complex<D2>::conjugate(*b));)
#undef ALP_HIP_CONJ_MUL

}  // namespace operators

// ---- algebraic structures -----------------------------------------------------

// Monoid<OP, ID> (monoid.hpp:139-147)
template <class OP, template <typename> class ID>
class Monoid {
public:
	typedef OP Operator;
	typedef typename OP::D1 D1;
	typedef typename OP::D2 D2;
	typedef typename OP::D3 D3;
	template <typename D> using Identity = ID<D>;
	template <typename D> static ALP_HD constexpr D getIdentity() { return ID<D>::value(); }
	OP getOperator() const { return OP(); }
};

// Semiring<ADD, MUL, ID0, ID1> (semiring.hpp:190): D1 x D2 -> D3 under MUL, D3 (+) D4
// -> D4 under ADD
template <class ADD, class MUL, template <typename> class ID0, template <typename> class ID1>
class Semiring {
public:
	typedef ADD AdditiveOperator;
	typedef MUL MultiplicativeOperator;
	typedef typename MUL::D1 D1;
	typedef typename MUL::D2 D2;
	typedef typename MUL::D3 D3;
	typedef typename ADD::D3 D4;
	typedef Monoid<ADD, ID0> AdditiveMonoid;
	typedef Monoid<MUL, ID1> MultiplicativeMonoid;
	template <typename D> using Zero = ID0<D>;
	template <typename D> using One = ID1<D>;
	template <typename D> static ALP_HD constexpr D getZero() { return ID0<D>::value(); }
	template <typename D> static ALP_HD constexpr D getOne() { return ID1<D>::value(); }
	AdditiveMonoid getAdditiveMonoid() const { return AdditiveMonoid(); }
	MultiplicativeMonoid getMultiplicativeMonoid() const { return MultiplicativeMonoid(); }
	ADD getAdditiveOperator() const { return ADD(); }
	MUL getMultiplicativeOperator() const { return MUL(); }
};

template <typename T> struct is_semiring : std::false_type {};
template <class A, class M, template <typename> class I0, template <typename> class I1>
struct is_semiring<Semiring<A, M, I0, I1>> : std::true_type {};
template <typename T> struct is_monoid : std::false_type {};
template <class O, template <typename> class I> struct is_monoid<Monoid<O, I>> : std::true_type {};
template <typename T> struct is_operator {
private:
	template <typename U> static auto test(int) -> decltype(&U::apply, std::true_type());
	template <typename> static std::false_type test(...);

public:
	static constexpr bool value = !is_semiring<T>::value && !is_monoid<T>::value && decltype(test<T>(0))::value;
};

// Monoids whose nonzeroes cannot change once set (a true under logical_or): a
// gather may stop at the first one (reference/blas2.hpp:354-360)
template <typename M> struct has_immutable_nonzeroes : std::false_type {};
template <typename D1, typename D2, typename D3, Backend b, template <typename> class I>
struct has_immutable_nonzeroes<Monoid<operators::logical_or<D1, D2, D3, b>, I>> : std::true_type {};
template <typename D1, typename D2, typename D3, Backend b, template <typename> class I>
struct has_immutable_nonzeroes<Monoid<operators::any_or<D1, D2, D3, b>, I>> : std::true_type {};

namespace internal {

// c = a op b on the given domains
template <class OP, typename A, typename B, typename C>
ALP_HD void apply_op(C &c, const A &a, const B &b) {
	const typename OP::D1 x = static_cast<typename OP::D1>(a);
	const typename OP::D2 y = static_cast<typename OP::D2>(b);
	typename OP::D3 z = static_cast<typename OP::D3>(c);
	OP::apply(&x, &y, &z);
	c = static_cast<C>(z);
}
// c = c op b (foldl)
template <class OP, typename C, typename B>
ALP_HD void foldl_op(C &c, const B &b) { apply_op<OP>(c, c, b); }
// c = a op c (foldr)
template <class OP, typename A, typename C>
ALP_HD void foldr_op(const A &a, C &c) { apply_op<OP>(c, a, c); }

template <class M> struct op_of { typedef M type; };
template <class O, template <typename> class I> struct op_of<Monoid<O, I>> { typedef O type; };

}  // namespace internal

}  // namespace grb

#endif
