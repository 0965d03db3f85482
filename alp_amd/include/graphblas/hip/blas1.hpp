/*
 * blas1.hpp -- level-1 primitives of the grb:: header layer over any operator,
 * monoid or semiring (ALP's base/blas1.hpp; reference/blas1.hpp): dot, eWiseApply,
 * eWiseMul, eWiseAdd, eWiseMulAdd, foldl / foldr and eWiseLambda.
 *
 * Sparsity follows ALP's rules: an operator annihilates (an entry exists where
 * both inputs have one), a monoid is neutral (where either has one; a missing
 * side acts as the identity); eWiseApply is out of place (the output holds the
 * result's entries only), eWiseMul / eWiseAdd / eWiseMulAdd / foldl / foldr
 * update their output in place. Scalars count as dense operands.
 *
 * Reductions (dot, foldl / foldr into a scalar) combine elements in index order:
 * each thread folds one contiguous chunk, the chunks of a block are folded in
 * order, and the blocks' results in order; the result is deterministic and valid
 * for any associative monoid, commutative or not.
 */
#ifndef ALP_HIP_GRB_BLAS1_HPP
#define ALP_HIP_GRB_BLAS1_HPP

#include "io.hpp"

namespace grb {

namespace internal {
namespace hip {

// an operand of an element-wise kernel: a vector (values, assigned flags; null
// flags = dense) or a scalar
template <typename T>
struct Arg {
	const T *v = nullptr;
	const uint8_t *a = nullptr;
	T s = T();
	bool scalar = false;
	__device__ __forceinline__ bool has(size_t i) const { return scalar || !a || a[i]; }
	__device__ __forceinline__ T get(size_t i) const { return scalar ? s : v[i]; }
};
// a vector operand: without its flags when the host knows it is dense (a pending
// count is not waited for: the flags are passed instead)
template <typename T, enum Backend b, typename C>
Arg<T> arg(const Vector<T, b, C> &x) {
	Arg<T> r;
	r.v = x.raw();
	r.a = x.known_dense() ? nullptr : x.assigned();
	return r;
}
template <typename T>
Arg<T> arg_scalar(const T &s) {
	Arg<T> r;
	r.s = s;
	r.scalar = true;
	return r;
}

enum EwMode {
	kApplyOp,      // z = x op y where both exist (out of place)
	kApplyMonoid,  // z = x op y where either exists, the missing side neutral (out of place)
	kMulIn,        // z = z (+) (x (x) y) where x and y exist (in place)
	kAddIn,        // z = z (+) (x (+) y) where either exists (in place)
	kMulAddIn,     // z = z (+) ((x (x) y) (+) w): the product where x, y exist, w where it exists
	kFoldl,        // z = z op y where both exist (operator) / where y exists (monoid)
	kFoldr         // z = x op z likewise
};

// element i of an element-wise pass: masked-out entries of an out-of-place
// result are removed, those of an in-place one left alone; returns whether z
// holds element i afterwards
template <int MODE, bool MONOID, Descriptor descr, bool MASKED, class ADD, class MUL, typename Z, typename X,
          typename Y, typename W, typename M>
__device__ __forceinline__ bool ew_element(size_t i, Z *z, uint8_t *za, const Arg<X> &x, const Arg<Y> &y,
                                           const Arg<W> &w, const M *mval, const uint8_t *masg) {
	const bool allowed = !MASKED || mask_at<descr>(mval, masg, i);
	const bool zin = za[i] != 0;
	bool zout = zin;
	if (MODE == kApplyOp || MODE == kApplyMonoid) {
		zout = false;
		if (allowed) {
			const bool hx = x.has(i), hy = y.has(i);
			if (hx && hy) {
				apply_op<ADD>(z[i], x.get(i), y.get(i));
				zout = true;
			} else if (MODE == kApplyMonoid && hx) {
				z[i] = static_cast<Z>(x.get(i));
				zout = true;
			} else if (MODE == kApplyMonoid && hy) {
				z[i] = static_cast<Z>(y.get(i));
				zout = true;
			}
		}
	} else if (allowed) {
		if (MODE == kMulIn) {
			if (x.has(i) && y.has(i)) {
				typename MUL::D3 t = typename MUL::D3();
				apply_op<MUL>(t, x.get(i), y.get(i));
				if (zin) foldl_op<ADD>(z[i], t); else z[i] = static_cast<Z>(t);
				zout = true;
			}
		} else if (MODE == kAddIn) {
			const bool hx = x.has(i), hy = y.has(i);
			if (hx || hy) {
				typename ADD::D3 t;
				if (hx && hy) apply_op<ADD>(t, x.get(i), y.get(i));
				else if (hx) t = static_cast<typename ADD::D3>(x.get(i));
				else t = static_cast<typename ADD::D3>(y.get(i));
				if (zin) foldl_op<ADD>(z[i], t); else z[i] = static_cast<Z>(t);
				zout = true;
			}
		} else if (MODE == kMulAddIn) {
			const bool hp = x.has(i) && y.has(i), hw = w.has(i);
			if (hp || hw) {
				typename ADD::D3 t;
				if (hp) {
					typename MUL::D3 p = typename MUL::D3();
					apply_op<MUL>(p, x.get(i), y.get(i));
					if (hw) apply_op<ADD>(t, p, w.get(i)); else t = static_cast<typename ADD::D3>(p);
				} else {
					t = static_cast<typename ADD::D3>(w.get(i));
				}
				if (zin) foldl_op<ADD>(z[i], t); else z[i] = static_cast<Z>(t);
				zout = true;
			}
		} else if (MODE == kFoldl) {
			if (y.has(i)) {
				if (zin) {
					foldl_op<ADD>(z[i], y.get(i));
				} else if (MONOID) {
					z[i] = static_cast<Z>(y.get(i));
					zout = true;
				}
			}
		} else if (MODE == kFoldr) {
			if (x.has(i)) {
				if (zin) {
					foldr_op<ADD>(x.get(i), z[i]);
				} else if (MONOID) {
					z[i] = static_cast<Z>(x.get(i));
					zout = true;
				}
			}
		}
	}
	za[i] = zout ? 1 : 0;
	return zout;
}

// one element-wise pass over n entries, counting z's entries (count may be null)
template <int MODE, bool MONOID, Descriptor descr, bool MASKED, class ADD, class MUL, typename Z, typename X,
          typename Y, typename W, typename M>
__global__ void k_ewise(size_t n, Z *z, uint8_t *za, Arg<X> x, Arg<Y> y, Arg<W> w, const M *mval, const uint8_t *masg,
                        unsigned long long *count) {
	unsigned long long c = 0;
	for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
		c += ew_element<MODE, MONOID, descr, MASKED, ADD, MUL, Z, X, Y, W, M>(i, z, za, x, y, w, mval, masg) ? 1 : 0;
	if (!count) return;
	count_add(c, count);
}

// the same pass as a stage of a fused kernel (its result count is known on the host)
template <typename Z, typename X, typename Y, typename W, typename M>
struct EwArgs {
	Z *z;
	uint8_t *za;
	Arg<X> x;
	Arg<Y> y;
	Arg<W> w;
	const M *mval;
	const uint8_t *masg;
};
template <int MODE, bool MONOID, Descriptor descr, bool MASKED, class ADD, class MUL, typename Z, typename X,
          typename Y, typename W, typename M>
__device__ void ew_stage(size_t b, size_t e, const void *args) {
	const EwArgs<Z, X, Y, W, M> a = *static_cast<const EwArgs<Z, X, Y, W, M> *>(args);
	for (size_t i = b + threadIdx.x; i < e; i += blockDim.x)
		(void)ew_element<MODE, MONOID, descr, MASKED, ADD, MUL, Z, X, Y, W, M>(i, a.z, a.za, a.x, a.y, a.w, a.mval,
		                                                                          a.masg);
}

// ---- ordered reductions ------------------------------------------------------------

constexpr unsigned kReduceBlocks = 512;

// per block: the fold, in index order, of the terms of its contiguous share of
// [0, n); term(i) gives (present, value). part / part_has hold one per block.
template <class OP, typename T, class Term>
__global__ void k_reduce_blocks(size_t n, Term term, T *part, uint8_t *part_has) {
	__shared__ T lv[kBlock];
	__shared__ uint8_t lh[kBlock];
	const size_t nthreads = (size_t)gridDim.x * blockDim.x;
	const size_t chunk = (n + nthreads - 1) / nthreads;
	const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	const size_t b = t * chunk, e = b + chunk < n ? b + chunk : n;
	T acc = T();
	bool has = false;
	for (size_t i = b; i < e; ++i) {
		T v;
		if (term(i, v)) {
			if (has) foldl_op<OP>(acc, v); else acc = v;
			has = true;
		}
	}
	lv[threadIdx.x] = acc;
	lh[threadIdx.x] = has ? 1 : 0;
	__syncthreads();
	if (threadIdx.x == 0) {
		T r = T();
		bool rh = false;
		for (unsigned k = 0; k < blockDim.x; ++k) {
			if (!lh[k]) continue;
			if (rh) foldl_op<OP>(r, lv[k]); else r = lv[k];
			rh = true;
		}
		part[blockIdx.x] = r;
		part_has[blockIdx.x] = rh ? 1 : 0;
	}
}

// the fold of the block results in order (has = any present), into the host
// report, then seq
template <class OP, typename T>
__global__ void k_reduce_final(unsigned nb, const T *part, const uint8_t *part_has, HostReport *rep,
                               unsigned long long seq) {
	if (blockIdx.x != 0 || threadIdx.x != 0) return;
	T r = T();
	bool rh = false;
	for (unsigned k = 0; k < nb; ++k) {
		if (!part_has[k]) continue;
		if (rh) foldl_op<OP>(r, part[k]); else r = part[k];
		rh = true;
	}
	memcpy(rep->value, &r, sizeof(T));
	rep->has = rh ? 1u : 0u;
	__hip_atomic_store(&rep->seq, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// a scalar output: a forcing point (the host waits for the value)
template <class OP, typename T, class Term>
RC reduce(size_t n, Term term, T &result, bool &present) {
	present = false;
	Forcing f;
	if (f.rc) return f.rc;
	if (n == 0) return SUCCESS;
	unsigned nb = (unsigned)((n + kBlock - 1) / kBlock);
	if (nb > kReduceBlocks) nb = kReduceBlocks;
	static_assert(sizeof(T) <= sizeof(HostReport::value), "reduce: result type too large for the host report");
	HostReportSlot *rs = nullptr;
	RC r = host_report(rs);
	if (r) return r;
	// workspace: nb block results, then nb flags
	const size_t vbytes = (sizeof(T) * nb + 15) / 16 * 16;
	char *ws = static_cast<char *>(workspace(vbytes + nb));
	T *part = reinterpret_cast<T *>(ws);
	uint8_t *part_has = reinterpret_cast<uint8_t *>(ws + vbytes);
	const unsigned long long seq = ++rs->seq;
	hipLaunchKernelGGL((k_reduce_blocks<OP, T, Term>), dim3(nb), dim3(kBlock), 0, library_stream(), n, term, part,
	                   part_has);
	hipLaunchKernelGGL((k_reduce_final<OP, T>), dim3(1), dim3(64), 0, library_stream(), nb, (const T *)part,
	                   (const uint8_t *)part_has, rs->d, seq);
	r = from_hip(hipGetLastError());
	if (r) return r;
	r = host_report_wait(*rs, seq);
	if (r) return r;
	std::memcpy(&result, rs->h->value, sizeof(T));
	present = rs->h->has != 0;
	return SUCCESS;
}

// term of a dot product: x_i (x) y_i where both exist
template <class MUL, typename T, typename X, typename Y>
struct DotTerm {
	Arg<X> x;
	Arg<Y> y;
	__device__ bool operator()(size_t i, T &v) const {
		if (!x.has(i) || !y.has(i)) return false;
		v = T();
		apply_op<MUL>(v, x.get(i), y.get(i));
		return true;
	}
};
// term of a fold into a scalar: x_i where it exists and the mask allows
template <Descriptor descr, bool MASKED, typename T, typename X, typename M>
struct FoldTerm {
	Arg<X> x;
	const M *mval;
	const uint8_t *masg;
	__device__ bool operator()(size_t i, T &v) const {
		if (!x.has(i)) return false;
		if (MASKED && !mask_at<descr>(mval, masg, i)) return false;
		v = static_cast<T>(x.get(i));
		return true;
	}
};

// whether z comes out dense whatever the values are, from what the host knows:
// dense (flag-free) operands, a dense z for the in-place modes, and no mask or a
// structural mask that holds everywhere
template <int MODE, bool MONOID, typename X, typename Y, typename W>
bool ewise_dense_result(bool z_dense, const Arg<X> &x, const Arg<Y> &y, const Arg<W> &w, bool mask_all) {
	const bool xd = x.scalar || !x.a, yd = y.scalar || !y.a, wd = w.scalar || !w.a;
	if (MODE == kApplyOp) return mask_all && xd && yd;
	if (MODE == kApplyMonoid) return mask_all && (xd || yd);
	if (z_dense) return true;  // an in-place result keeps every entry it had
	if (!mask_all) return false;
	if (MODE == kMulIn) return xd && yd;
	if (MODE == kAddIn) return xd || yd;
	if (MODE == kMulAddIn) return (xd && yd) || wd;
	if (MODE == kFoldl) return MONOID && yd;
	return MONOID && xd;  // kFoldr
}

template <int MODE, bool MONOID, Descriptor descr, class ADD, class MUL, typename Z, typename X, typename Y,
          typename W, typename M, enum Backend b, typename C>
RC ewise(Vector<Z, b, C> &z, const Arg<X> &x, const Arg<Y> &y, const Arg<W> &w, const Vector<M, b, C> *mask) {
	const size_t n = z.size();
	const bool masked = mask && mask->size() > 0;
	const bool mask_all = !masked || ((descr & descriptors::structural) && !(descr & descriptors::invert_mask) &&
	                                  mask->known_dense());
	const bool dense_out = ewise_dense_result<MODE, MONOID>(z.known_dense(), x, y, w, mask_all);
	if (dense_out) {
		// the count is known: a stage of the next fused pass
		z.set_nnz(n);
		const M *mv = masked ? mask->raw() : nullptr;
		const uint8_t *ma = masked && !mask->known_dense() ? mask->assigned() : nullptr;
		const EwArgs<Z, X, Y, W, M> a{z.raw(), z.assigned(), x, y, w, mv, ma};
		return masked ? enqueue_stage<ew_stage<MODE, MONOID, descr, true, ADD, MUL, Z, X, Y, W, M>>(n, a)
		              : enqueue_stage<ew_stage<MODE, MONOID, descr, false, ADD, MUL, Z, X, Y, W, M>>(n, a);
	}
	unsigned long long *count = z.count_slot();
	if (n) {
		if (masked)
			hipLaunchKernelGGL((k_ewise<MODE, MONOID, descr, true, ADD, MUL, Z, X, Y, W, M>), dim3(grid(n)),
			                   dim3(kBlock), 0, stream(), n, z.raw(), z.assigned(), x, y, w, mask->raw(),
			                   mask->known_dense() ? nullptr : mask->assigned(), count);
		else
			hipLaunchKernelGGL((k_ewise<MODE, MONOID, descr, false, ADD, MUL, Z, X, Y, W, M>), dim3(grid(n)),
			                   dim3(kBlock), 0, stream(), n, z.raw(), z.assigned(), x, y, w, (const M *)nullptr,
			                   (const uint8_t *)nullptr, count);
	}
	z.mark_counted();
	return from_hip(hipGetLastError());
}

template <typename V, typename T, enum Backend b, typename C>
bool sizes_match(const Vector<T, b, C> &z, const V &v) {
	return v.size() == z.size();
}

// descriptors::dense: every vector argument must be dense (reference/blas1.hpp)
template <Descriptor descr, typename T, enum Backend b, typename C>
bool dense_ok(const Vector<T, b, C> &v) {
	return !(descr & descriptors::dense) || v.dense();
}

template <typename A> struct is_vec : std::false_type {};
template <typename T, enum Backend b, typename C> struct is_vec<Vector<T, b, C>> : std::true_type {};

template <typename T, enum Backend b, typename C> Arg<T> as_arg(const Vector<T, b, C> &v) { return arg(v); }
template <typename T> Arg<T> as_arg(const T &s, typename std::enable_if<!is_vec<T>::value>::type * = nullptr) {
	return arg_scalar(s);
}
template <typename T> struct val_of { typedef T type; };
template <typename T, enum Backend b, typename C> struct val_of<Vector<T, b, C>> { typedef T type; };

template <typename A> size_t arg_size(const A &, size_t n, std::false_type) { return n; }
template <typename A> size_t arg_size(const A &a, size_t, std::true_type) { return a.size(); }
template <Descriptor descr, typename A> bool arg_dense(const A &, std::false_type) { return true; }
template <Descriptor descr, typename A> bool arg_dense(const A &a, std::true_type) { return dense_ok<descr>(a); }

template <Descriptor descr, typename Z, typename... As>
RC check_args(const Z &z, const As &...as) {
	const size_t n = z.size();
	const bool sizes[] = {true, (arg_size(as, n, is_vec<As>()) == n)...};
	const bool dense[] = {true, arg_dense<descr>(as, is_vec<As>())...};
	for (bool s : sizes)
		if (!s) return MISMATCH;
	if (!dense_ok<descr>(z)) return ILLEGAL;
	for (bool d : dense)
		if (!d) return ILLEGAL;
	return SUCCESS;
}

template <typename T, enum Backend b, typename C> const void *addr(const Vector<T, b, C> &v) { return v.raw(); }
template <typename T> const void *addr(const T &, typename std::enable_if<!is_vec<T>::value>::type * = nullptr) {
	return nullptr;
}

}  // namespace hip
}  // namespace internal

// ---- dot (reference/blas1.hpp:9520-9639) ----------------------------------------------

// z = z (+) sum over i in x and y of x_i (x) y_i, with the semiring's monoid
template <Descriptor descr = descriptors::no_operation, class Ring, typename Z, typename X, typename Y,
          enum Backend b, typename C>
RC dot(Z &z, const Vector<X, b, C> &x, const Vector<Y, b, C> &y, const Ring &ring = Ring(),
       const Phase phase = EXECUTE, typename std::enable_if<is_semiring<Ring>::value>::type * = nullptr) {
	(void)ring;
	if (x.size() != y.size()) return MISMATCH;
	if (!internal::hip::dense_ok<descr>(x) || !internal::hip::dense_ok<descr>(y)) return ILLEGAL;
	if (phase == RESIZE) return SUCCESS;
	typedef typename Ring::AdditiveOperator ADD;
	typedef typename Ring::MultiplicativeOperator MUL;
	typedef typename ADD::D3 T;
	// the library's deterministic FP64 dot for dense plus-times vectors of double
	if (std::is_same<ADD, operators::add<double>>::value && std::is_same<MUL, operators::mul<double>>::value &&
	    std::is_same<X, double>::value && std::is_same<Y, double>::value && x.dense() && y.dense()) {
		double v = 0.0;
		internal::hip::Forcing f;
		if (f.rc) return f.rc;
		RC r = internal::rc(alp_hip_dot_f64(x.size(), (const double *)x.raw(), (const double *)y.raw(), &v));
		if (r) return r;
		internal::foldl_op<ADD>(z, v);
		return SUCCESS;
	}
	T v = T();
	bool present = false;
	internal::hip::DotTerm<MUL, T, X, Y> term{internal::hip::arg(x), internal::hip::arg(y)};
	RC r = internal::hip::reduce<ADD>(x.size(), term, v, present);
	if (r) return r;
	if (present) internal::foldl_op<ADD>(z, v);
	return SUCCESS;
}
// the (additive monoid, multiplicative operator) form
template <Descriptor descr = descriptors::no_operation, class AddMonoid, class AnyOp, typename Z, typename X,
          typename Y, enum Backend b, typename C>
RC dot(Z &z, const Vector<X, b, C> &x, const Vector<Y, b, C> &y, const AddMonoid &, const AnyOp &,
       const Phase phase = EXECUTE,
       typename std::enable_if<is_monoid<AddMonoid>::value && is_operator<AnyOp>::value>::type * = nullptr) {
	Semiring<typename AddMonoid::Operator, AnyOp, AddMonoid::template Identity, identities::one> ring;
	return dot<descr>(z, x, y, ring, phase);
}

// ---- eWiseApply (out of place) ---------------------------------------------------------

// z = x op y: with an operator where both exist, with a monoid where either does
template <Descriptor descr = descriptors::no_operation, class OP, typename Z, typename X, typename Y,
          enum Backend b, typename C>
RC eWiseApply(Vector<Z, b, C> &z, const X &x, const Y &y, const OP & = OP(), const Phase phase = EXECUTE,
              typename std::enable_if<(is_operator<OP>::value || is_monoid<OP>::value) &&
                                      (internal::hip::is_vec<X>::value || internal::hip::is_vec<Y>::value)>::type * =
                  nullptr) {
	RC r = internal::hip::check_args<descr>(z, x, y);
	if (r || phase == RESIZE) return r;
	typedef typename internal::op_of<OP>::type O;
	return internal::hip::ewise<is_monoid<OP>::value ? internal::hip::kApplyMonoid : internal::hip::kApplyOp,
	                            is_monoid<OP>::value, descr, O, O, Z, typename internal::hip::val_of<X>::type,
	                            typename internal::hip::val_of<Y>::type, char, bool>(
	    z, internal::hip::as_arg(x), internal::hip::as_arg(y), internal::hip::Arg<char>(),
	    (const Vector<bool, b, C> *)nullptr);
}
// masked
template <Descriptor descr = descriptors::no_operation, class OP, typename Z, typename M, typename X, typename Y,
          enum Backend b, typename C>
RC eWiseApply(Vector<Z, b, C> &z, const Vector<M, b, C> &mask, const X &x, const Y &y, const OP & = OP(),
              const Phase phase = EXECUTE,
              typename std::enable_if<(is_operator<OP>::value || is_monoid<OP>::value)>::type * = nullptr) {
	RC r = internal::hip::check_args<descr>(z, x, y);
	if (r || phase == RESIZE) return r;
	if (mask.size() && mask.size() != z.size()) return MISMATCH;
	typedef typename internal::op_of<OP>::type O;
	return internal::hip::ewise<is_monoid<OP>::value ? internal::hip::kApplyMonoid : internal::hip::kApplyOp,
	                            is_monoid<OP>::value, descr, O, O, Z, typename internal::hip::val_of<X>::type,
	                            typename internal::hip::val_of<Y>::type, char, M>(
	    z, internal::hip::as_arg(x), internal::hip::as_arg(y), internal::hip::Arg<char>(), &mask);
}

// ---- eWiseMul / eWiseAdd / eWiseMulAdd (in place) --------------------------------------

// z = z (+) x (x) y where x and y exist (blas1.hpp:8684 for the scalar-vector form)
template <Descriptor descr = descriptors::no_operation, class Ring, typename Z, typename X, typename Y,
          enum Backend b, typename C>
RC eWiseMul(Vector<Z, b, C> &z, const X &x, const Y &y, const Ring & = Ring(), const Phase phase = EXECUTE,
            typename std::enable_if<is_semiring<Ring>::value>::type * = nullptr) {
	RC r = internal::hip::check_args<descr>(z, x, y);
	if (r || phase == RESIZE) return r;
	typedef typename internal::hip::val_of<X>::type XT;
	typedef typename internal::hip::val_of<Y>::type YT;
	return internal::hip::ewise<internal::hip::kMulIn, false, descr, typename Ring::AdditiveOperator,
	                            typename Ring::MultiplicativeOperator, Z, XT, YT, char, bool>(
	    z, internal::hip::as_arg(x), internal::hip::as_arg(y), internal::hip::Arg<char>(),
	    (const Vector<bool, b, C> *)nullptr);
}
// masked
template <Descriptor descr = descriptors::no_operation, class Ring, typename Z, typename M, typename X, typename Y,
          enum Backend b, typename C>
RC eWiseMul(Vector<Z, b, C> &z, const Vector<M, b, C> &mask, const X &x, const Y &y, const Ring & = Ring(),
            const Phase phase = EXECUTE, typename std::enable_if<is_semiring<Ring>::value>::type * = nullptr) {
	RC r = internal::hip::check_args<descr>(z, x, y);
	if (r || phase == RESIZE) return r;
	if (mask.size() && mask.size() != z.size()) return MISMATCH;
	return internal::hip::ewise<internal::hip::kMulIn, false, descr, typename Ring::AdditiveOperator,
	                            typename Ring::MultiplicativeOperator, Z, typename internal::hip::val_of<X>::type,
	                            typename internal::hip::val_of<Y>::type, char, M>(
	    z, internal::hip::as_arg(x), internal::hip::as_arg(y), internal::hip::Arg<char>(), &mask);
}

// z = z (+) (x (+) y) where either exists
template <Descriptor descr = descriptors::no_operation, class Ring, typename Z, typename X, typename Y,
          enum Backend b, typename C>
RC eWiseAdd(Vector<Z, b, C> &z, const X &x, const Y &y, const Ring & = Ring(), const Phase phase = EXECUTE,
            typename std::enable_if<is_semiring<Ring>::value>::type * = nullptr) {
	RC r = internal::hip::check_args<descr>(z, x, y);
	if (r || phase == RESIZE) return r;
	return internal::hip::ewise<internal::hip::kAddIn, true, descr, typename Ring::AdditiveOperator,
	                            typename Ring::MultiplicativeOperator, Z, typename internal::hip::val_of<X>::type,
	                            typename internal::hip::val_of<Y>::type, char, bool>(
	    z, internal::hip::as_arg(x), internal::hip::as_arg(y), internal::hip::Arg<char>(),
	    (const Vector<bool, b, C> *)nullptr);
}

// z = z (+) ((a (x) x) (+) y) (blas1.hpp eWiseMulAdd): the product where a and x
// exist, y where it exists
template <Descriptor descr = descriptors::no_operation, class Ring, typename Z, typename A, typename X, typename Y,
          enum Backend b, typename C>
RC eWiseMulAdd(Vector<Z, b, C> &z, const A &a, const X &x, const Y &y, const Ring & = Ring(),
               const Phase phase = EXECUTE, typename std::enable_if<is_semiring<Ring>::value>::type * = nullptr) {
	RC r = internal::hip::check_args<descr>(z, a, x, y);
	if (r || phase == RESIZE) return r;
	return internal::hip::ewise<internal::hip::kMulAddIn, false, descr, typename Ring::AdditiveOperator,
	                            typename Ring::MultiplicativeOperator, Z, typename internal::hip::val_of<A>::type,
	                            typename internal::hip::val_of<X>::type, typename internal::hip::val_of<Y>::type,
	                            bool>(z, internal::hip::as_arg(a), internal::hip::as_arg(x), internal::hip::as_arg(y),
	                                  (const Vector<bool, b, C> *)nullptr);
}

// ---- foldl / foldr --------------------------------------------------------------------

// x = x op y, element-wise (blas1.hpp:2639): where both exist with an operator,
// y's entries join x with a monoid; y may be a scalar
template <Descriptor descr = descriptors::no_operation, class OP, typename Z, typename Y, enum Backend b,
          typename C>
RC foldl(Vector<Z, b, C> &x, const Y &y, const OP & = OP(), const Phase phase = EXECUTE,
         typename std::enable_if<is_operator<OP>::value || is_monoid<OP>::value>::type * = nullptr) {
	RC r = internal::hip::check_args<descr>(x, y);
	if (r || phase == RESIZE) return r;
	typedef typename internal::op_of<OP>::type O;
	return internal::hip::ewise<internal::hip::kFoldl, is_monoid<OP>::value, descr, O, O, Z, char,
	                            typename internal::hip::val_of<Y>::type, char, bool>(
	    x, internal::hip::Arg<char>(), internal::hip::as_arg(y), internal::hip::Arg<char>(),
	    (const Vector<bool, b, C> *)nullptr);
}
// masked
template <Descriptor descr = descriptors::no_operation, class OP, typename Z, typename M, typename Y,
          enum Backend b, typename C>
RC foldl(Vector<Z, b, C> &x, const Vector<M, b, C> &mask, const Y &y, const OP & = OP(),
         const Phase phase = EXECUTE,
         typename std::enable_if<is_operator<OP>::value || is_monoid<OP>::value>::type * = nullptr) {
	RC r = internal::hip::check_args<descr>(x, y);
	if (r || phase == RESIZE) return r;
	if (mask.size() && mask.size() != x.size()) return MISMATCH;
	typedef typename internal::op_of<OP>::type O;
	return internal::hip::ewise<internal::hip::kFoldl, is_monoid<OP>::value, descr, O, O, Z, char,
	                            typename internal::hip::val_of<Y>::type, char, M>(
	    x, internal::hip::Arg<char>(), internal::hip::as_arg(y), internal::hip::Arg<char>(), &mask);
}
// y = x op y, element-wise (blas1.hpp:2049; x may be a scalar, blas1.hpp:1685)
template <Descriptor descr = descriptors::no_operation, class OP, typename X, typename Z, enum Backend b,
          typename C>
RC foldr(const X &x, Vector<Z, b, C> &y, const OP & = OP(), const Phase phase = EXECUTE,
         typename std::enable_if<is_operator<OP>::value || is_monoid<OP>::value>::type * = nullptr) {
	RC r = internal::hip::check_args<descr>(y, x);
	if (r || phase == RESIZE) return r;
	typedef typename internal::op_of<OP>::type O;
	return internal::hip::ewise<internal::hip::kFoldr, is_monoid<OP>::value, descr, O, O, Z,
	                            typename internal::hip::val_of<X>::type, char, char, bool>(
	    y, internal::hip::as_arg(x), internal::hip::Arg<char>(), internal::hip::Arg<char>(),
	    (const Vector<bool, b, C> *)nullptr);
}

// z = z (+) the fold of x's entries (blas1.hpp:2765), in index order
template <Descriptor descr = descriptors::no_operation, class Monoid_, typename Z, typename X, enum Backend b,
          typename C>
RC foldl(Z &z, const Vector<X, b, C> &x, const Monoid_ & = Monoid_(), const Phase phase = EXECUTE,
         typename std::enable_if<is_monoid<Monoid_>::value && !internal::hip::is_vec<Z>::value>::type * = nullptr) {
	if (!internal::hip::dense_ok<descr>(x)) return ILLEGAL;
	if (phase == RESIZE) return SUCCESS;
	typedef typename Monoid_::Operator O;
	typedef typename O::D3 T;
	T v = T();
	bool present = false;
	internal::hip::FoldTerm<descr, false, T, X, bool> term{internal::hip::arg(x), nullptr, nullptr};
	RC r = internal::hip::reduce<O>(x.size(), term, v, present);
	if (r) return r;
	if (present) internal::foldl_op<O>(z, v);
	return SUCCESS;
}
// masked
template <Descriptor descr = descriptors::no_operation, class Monoid_, typename Z, typename X, typename M,
          enum Backend b, typename C>
RC foldl(Z &z, const Vector<X, b, C> &x, const Vector<M, b, C> &mask, const Monoid_ & = Monoid_(),
         const Phase phase = EXECUTE,
         typename std::enable_if<is_monoid<Monoid_>::value && !internal::hip::is_vec<Z>::value>::type * = nullptr) {
	if (mask.size() == 0) return foldl<descr>(z, x, Monoid_(), phase);
	if (mask.size() != x.size()) return MISMATCH;
	if (phase == RESIZE) return SUCCESS;
	typedef typename Monoid_::Operator O;
	typedef typename O::D3 T;
	T v = T();
	bool present = false;
	internal::hip::FoldTerm<descr, true, T, X, M> term{internal::hip::arg(x), mask.raw(),
	                                                   mask.dense() ? nullptr : mask.assigned()};
	RC r = internal::hip::reduce<O>(x.size(), term, v, present);
	if (r) return r;
	if (present) internal::foldl_op<O>(z, v);
	return SUCCESS;
}
// z = (the fold of x's entries) (+) z
template <Descriptor descr = descriptors::no_operation, class Monoid_, typename X, typename Z, enum Backend b,
          typename C>
RC foldr(const Vector<X, b, C> &x, Z &z, const Monoid_ & = Monoid_(), const Phase phase = EXECUTE,
         typename std::enable_if<is_monoid<Monoid_>::value && !internal::hip::is_vec<Z>::value>::type * = nullptr) {
	if (phase == RESIZE) return SUCCESS;
	typedef typename Monoid_::Operator O;
	typedef typename O::D3 T;
	T v = T();
	bool present = false;
	internal::hip::FoldTerm<descr, false, T, X, bool> term{internal::hip::arg(x), nullptr, nullptr};
	RC r = internal::hip::reduce<O>(x.size(), term, v, present);
	if (r) return r;
	if (present) internal::foldr_op<O>(v, z);
	return SUCCESS;
}

// ---- eWiseLambda (reference/blas1.hpp:9743-9860) ---------------------------------------

namespace internal {
namespace hip {
template <class F>
__global__ void k_lambda(size_t n, const uint8_t *asg, F f) {
	for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
		if (!asg || asg[i]) f(i);
}
inline RC lambda_sizes(size_t) { return SUCCESS; }
template <typename T, enum Backend b, typename C, typename... Vs>
RC lambda_sizes(size_t n, const Vector<T, b, C> &v, const Vs &...vs) {
	return v.size() != n ? MISMATCH : lambda_sizes(n, vs...);
}
}  // namespace hip
}  // namespace internal

// Calls f(i) on the device for every index i that x holds, all in parallel; f may
// read and write elements of x and of the other vectors (of x's size) through
// grb::view(...) captured by value, but not change any vector's sparsity. f is a
// device callable, e.g. [=] __device__ (const size_t i) { ... }.
template <typename F, typename T, enum Backend b, typename C, typename... Vs>
RC eWiseLambda(const F f, const Vector<T, b, C> &x, const Vs &...others) {
	RC r = internal::hip::lambda_sizes(x.size(), others...);
	if (r) return r;
	const size_t n = x.size();
	if (n == 0 || (!x.count_pending() && x.nnz() == 0)) return SUCCESS;
	hipLaunchKernelGGL((internal::hip::k_lambda<F>), dim3(internal::hip::grid(n)), dim3(internal::hip::kBlock), 0,
	                   internal::hip::stream(), n, x.known_dense() ? nullptr : x.assigned(), f);
	return internal::hip::from_hip(hipGetLastError());
}

}  // namespace grb

#endif
