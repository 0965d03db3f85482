/*
 * blas0.hpp -- level-0 primitives of the grb:: header layer: an operator applied
 * to scalars on the host (ALP's blas0.hpp:179-330: apply, foldl, foldr), with the
 * operand conversions of the reference (each operand cast to the operator's
 * domain, the result cast back).
 */
#ifndef ALP_HIP_GRB_BLAS0_HPP
#define ALP_HIP_GRB_BLAS0_HPP

#include "containers.hpp"

namespace grb {

// containers and algebraic structures, which the level-0 forms do not take
// (type_traits.hpp is_object)
template <typename T> struct is_object : std::integral_constant<bool, is_monoid<T>::value || is_semiring<T>::value> {};
template <typename T, enum Backend b, typename C> struct is_object<Vector<T, b, C>> : std::true_type {};
template <typename T, enum Backend b, typename R, typename C, typename N>
struct is_object<Matrix<T, b, R, C, N>> : std::true_type {};

// out = x op y
template <Descriptor descr = descriptors::no_operation, class OP, typename InputType1, typename InputType2,
          typename OutputType>
RC apply(OutputType &out, const InputType1 &x, const InputType2 &y, const OP & = OP(),
         typename std::enable_if<is_operator<OP>::value && !is_object<InputType1>::value &&
                                 !is_object<InputType2>::value && !is_object<OutputType>::value>::type * = nullptr) {
	static_assert(!(descr & descriptors::no_casting) ||
	                  (std::is_same<InputType1, typename OP::D1>::value &&
	                   std::is_same<InputType2, typename OP::D2>::value &&
	                   std::is_same<OutputType, typename OP::D3>::value),
	              "grb::apply (level 0): operand types differ from the operator's domains (no_casting)");
	internal::apply_op<OP>(out, x, y);
	return SUCCESS;
}

// x = x op y
template <Descriptor descr = descriptors::no_operation, class OP, typename IOType, typename InputType>
RC foldl(IOType &x, const InputType &y, const OP & = OP(),
         typename std::enable_if<is_operator<OP>::value && !is_object<InputType>::value &&
                                 !is_object<IOType>::value>::type * = nullptr) {
	static_assert(!(descr & descriptors::no_casting) ||
	                  (std::is_same<IOType, typename OP::D1>::value &&
	                   std::is_same<InputType, typename OP::D2>::value &&
	                   std::is_same<IOType, typename OP::D3>::value),
	              "grb::foldl (level 0): operand types differ from the operator's domains (no_casting)");
	internal::foldl_op<OP>(x, y);
	return SUCCESS;
}

// y = x op y
template <Descriptor descr = descriptors::no_operation, class OP, typename InputType, typename IOType>
RC foldr(const InputType &x, IOType &y, const OP & = OP(),
         typename std::enable_if<is_operator<OP>::value && !is_object<InputType>::value &&
                                 !is_object<IOType>::value>::type * = nullptr) {
	static_assert(!(descr & descriptors::no_casting) ||
	                  (std::is_same<InputType, typename OP::D1>::value &&
	                   std::is_same<IOType, typename OP::D2>::value &&
	                   std::is_same<IOType, typename OP::D3>::value),
	              "grb::foldr (level 0): operand types differ from the operator's domains (no_casting)");
	internal::foldr_op<OP>(x, y);
	return SUCCESS;
}

}  // namespace grb

#endif
