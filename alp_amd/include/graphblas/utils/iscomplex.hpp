/*
 * utils/iscomplex.hpp -- grb::utils::is_complex<T>: whether T is std::complex,
 * and the scalar helpers the algorithms use on either kind (conjugate, modulus,
 * norm, inverse), as ALP's utils/iscomplex.hpp provides them. The real forms
 * are host and device functions; the complex forms are host only.
 */
#ifndef ALP_HIP_GRB_UTILS_ISCOMPLEX_HPP
#define ALP_HIP_GRB_UTILS_ISCOMPLEX_HPP

#include <hip/hip_runtime.h>

#include <complex>
#include <type_traits>

namespace grb {
namespace utils {

template <typename C>
struct is_complex {
	static_assert(std::is_floating_point<C>::value, "is_complex: C is not a floating point type");
	typedef C type;
	static constexpr bool value = false;
	static __host__ __device__ C conjugate(const C &x) noexcept { return x; }
	static __host__ __device__ C modulus(const C &x) noexcept { return x > 0 ? x : -x; }
	static __host__ __device__ C norm(const C &x) noexcept { return x * x; }
	static __host__ __device__ C inverse(const C &x) noexcept { return static_cast<C>(1) / x; }
};

template <typename T>
struct is_complex<std::complex<T>> {
	static_assert(std::is_floating_point<T>::value, "is_complex: T is not a floating point type");
	typedef T type;
	static constexpr bool value = true;
	static std::complex<T> conjugate(const std::complex<T> &x) noexcept { return std::conj(x); }
	static T modulus(const std::complex<T> &x) noexcept { return std::abs(x); }
	static T norm(const std::complex<T> &x) noexcept { return std::norm(x); }
	static std::complex<T> inverse(const std::complex<T> &x) noexcept { return static_cast<T>(1) / x; }
};

}  // namespace utils
}  // namespace grb

#endif
