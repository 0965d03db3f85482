// An ALP/GraphBLAS program exercising the grb:: layer's generic, sparse
// semantics on the device, in the style of the reference's unit tests
// (tests/unit/{mxv,vxm,dot,eWiseApply,eWiseMul,mxm,masked_mxv,sparse_mxv}.cpp):
// sparse inputs and outputs, masks, descriptors and their error codes, and
// semirings the library has no fixed kernel for (max-times over double,
// plus-min over int, a program-defined operator). Inputs come from a seeded
// generator and are printed with the results as one JSON object; the Python
// test recomputes every result by brute force from the printed inputs.
//
//   grb_semantics_program <seed>
#include <graphblas.hpp>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

using namespace grb;

namespace {

uint64_t g_state = 0;
uint64_t next_u64() {
	uint64_t z = (g_state += 0x9E3779B97F4A7C15ull);
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
	return z ^ (z >> 31);
}
double uniform() { return (double)(next_u64() >> 11) * (1.0 / 9007199254740992.0); }

std::string out;
void key(const char *k) {
	if (out.size() > 1) out += ", ";
	out += "\"";
	out += k;
	out += "\": ";
}
template <typename T> void emit_list(const char *k, const std::vector<T> &v) {
	key(k);
	out += "[";
	char buf[64];
	for (size_t i = 0; i < v.size(); ++i) {
		std::snprintf(buf, sizeof buf, "%s%.17g", i ? ", " : "", (double)v[i]);
		out += buf;
	}
	out += "]";
}
void emit_num(const char *k, double v) {
	key(k);
	char buf[64];
	std::snprintf(buf, sizeof buf, "%.17g", v);
	out += buf;
}
// a sparse vector as index and value lists
template <typename T> void emit_vec(const std::string &name, const Vector<T> &v) {
	std::vector<double> idx, val;
	for (const auto &p : v) {
		idx.push_back((double)p.first);
		val.push_back((double)p.second);
	}
	emit_list((name + "_idx").c_str(), idx);
	emit_list((name + "_val").c_str(), val);
	emit_num((name + "_nnz").c_str(), (double)nnz(v));
}
template <typename T> void emit_mat(const std::string &name, const Matrix<T> &A) {
	std::vector<double> I, J, V;
	for (const auto &e : A) {
		I.push_back((double)e.first.first);
		J.push_back((double)e.first.second);
		V.push_back((double)e.second);
	}
	emit_list((name + "_I").c_str(), I);
	emit_list((name + "_J").c_str(), J);
	emit_list((name + "_V").c_str(), V);
}

// a program-defined operator: addition modulo 7 over int (associative, commutative)
template <typename L, typename R = L, typename O = R>
struct add_mod7 {
	typedef L D1;
	typedef R D2;
	typedef O D3;
	static constexpr bool is_associative = true;
	static constexpr bool is_commutative = true;
	static __host__ __device__ void apply(const D1 *a, const D2 *b, D3 *c) { *c = (D3)(((*a + *b) % 7 + 7) % 7); }
};

}  // namespace

int main(int argc, char **argv) {
	g_state = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 1;
	if (init() != SUCCESS) return 1;
	const size_t m = 37, n = 29;
	// a random m x n matrix in random coordinate order (some rows empty)
	std::vector<size_t> I, J;
	std::vector<double> V;
	std::vector<int> Vi;
	for (size_t i = 0; i < m; ++i) {
		if (uniform() < 0.15) continue;
		for (size_t j = 0; j < n; ++j)
			if (uniform() < 0.18) {
				I.push_back(i);
				J.push_back(j);
				V.push_back(std::floor(uniform() * 19.0) - 9.0 + 0.25 * std::floor(uniform() * 4.0));
				Vi.push_back((int)std::floor(uniform() * 21.0) - 10);
			}
	}
	for (size_t k = I.size(); k > 1; --k) {  // shuffle the input order
		const size_t r = next_u64() % k;
		std::swap(I[k - 1], I[r]);
		std::swap(J[k - 1], J[r]);
		std::swap(V[k - 1], V[r]);
		std::swap(Vi[k - 1], Vi[r]);
	}
	Matrix<double> A(m, n);
	Matrix<int> Ai(m, n);
	Matrix<void> P(m, n);
	RC rc = buildMatrixUnique(A, I.data(), J.data(), V.data(), I.size(), SEQUENTIAL);
	rc = rc ? rc : buildMatrixUnique(Ai, I.data(), J.data(), Vi.data(), I.size(), SEQUENTIAL);
	rc = rc ? rc : buildMatrixUnique(P, I.data(), J.data(), I.size(), SEQUENTIAL);
	out = "{";
	emit_list("I", I);
	emit_list("J", J);
	emit_list("V", V);
	emit_list("Vi", Vi);
	emit_num("m", (double)m);
	emit_num("n", (double)n);
	emit_mat("A", A);  // storage order as the library stores it

	// sparse inputs: about a third of the entries of x (double) and xi (int)
	Vector<double> x(n);
	Vector<int> xi(n);
	Vector<bool> xb(n);
	for (size_t j = 0; j < n; ++j)
		if (uniform() < 0.35) {
			rc = rc ? rc : setElement(x, std::floor(uniform() * 9.0) - 4.0, j);
			rc = rc ? rc : setElement(xi, (int)(next_u64() % 13) - 6, j);
			rc = rc ? rc : setElement(xb, true, j);
		}
	emit_vec("x", x);
	emit_vec("xi", xi);
	emit_vec("xb", xb);

	// 1. plus-times over double, sparse input, fresh output: entries only where a
	//    row meets the input
	Semiring<operators::add<double>, operators::mul<double>, identities::zero, identities::one> plus_times;
	Vector<double> y1(m);
	rc = rc ? rc : mxv(y1, A, x, plus_times);
	emit_vec("y_plus_times", y1);

	// 2. min-plus over double into a fresh output (its entries take the product's
	//    value, not min(0, ...)), then accumulated into an existing one
	Semiring<operators::min<double>, operators::add<double>, identities::infinity, identities::zero> min_plus;
	Vector<double> y2(m);
	rc = rc ? rc : mxv(y2, A, x, min_plus);
	emit_vec("y_min_plus", y2);
	Vector<double> y2d(m);
	rc = rc ? rc : set(y2d, 1.5);
	rc = rc ? rc : mxv(y2d, A, x, min_plus);
	emit_vec("y_min_plus_acc", y2d);

	// 3. max-times over double (no library kernel: the generic gather)
	Semiring<operators::max<double>, operators::mul<double>, identities::negative_infinity, identities::one> max_times;
	Vector<double> y3(m);
	rc = rc ? rc : mxv(y3, A, x, max_times);
	emit_vec("y_max_times", y3);

	// 4. plus-min over int on an int matrix (values held by the layer)
	Semiring<operators::add<int>, operators::min<int>, identities::zero, identities::infinity> plus_min;
	Vector<int> y4(m);
	rc = rc ? rc : mxv(y4, Ai, xi, plus_min);
	emit_vec("y_plus_min_int", y4);

	// 5. a program-defined additive operator: (add mod 7, times) over int, vxm
	Semiring<add_mod7<int>, operators::mul<int>, identities::zero, identities::one> mod7;
	Vector<int> xm(m), y5(n);
	for (size_t i = 0; i < m; i += 2) rc = rc ? rc : setElement(xm, (int)(i % 11) - 5, i);
	emit_vec("xm", xm);
	rc = rc ? rc : vxm(y5, xm, Ai, mod7);
	emit_vec("y_mod7_vxm", y5);

	// 6. lor-land over the pattern: boolean output nonzeroes as the reference's
	//    (only rows that meet the sparse input get an entry), a pushed sparse
	//    input, a masked product (structural complement) and add_identity
	Semiring<operators::logical_or<bool>, operators::logical_and<bool>, identities::logical_false,
	         identities::logical_true>
	    lor;
	Vector<bool> y6(m);
	rc = rc ? rc : mxv(y6, P, xb, lor);
	emit_vec("y_lor", y6);
	Vector<bool> one(n);
	rc = rc ? rc : setElement(one, true, 3);
	Vector<bool> y6b(m);
	rc = rc ? rc : mxv(y6b, P, one, lor);  // one nonzero in 29: the push path
	emit_vec("y_lor_push", y6b);
	Vector<bool> mask(m);
	for (size_t i = 0; i < m; i += 3) rc = rc ? rc : setElement(mask, true, i);
	Vector<bool> y6c(m);
	rc = rc ? rc : mxv<descriptors::structural_complement>(y6c, mask, P, xb, lor);
	emit_vec("mask", mask);
	emit_vec("y_lor_masked", y6c);
	// square pattern for add_identity: the leading n x n block of P
	Matrix<void> Q(n, n);
	{
		std::vector<size_t> qi, qj;
		for (size_t k = 0; k < I.size(); ++k)
			if (I[k] < n) {
				qi.push_back(I[k]);
				qj.push_back(J[k]);
			}
		rc = rc ? rc : buildMatrixUnique(Q, qi.data(), qj.data(), qi.size(), SEQUENTIAL);
	}
	Vector<bool> y6d(n);
	rc = rc ? rc : mxv<descriptors::add_identity>(y6d, Q, xb, lor);
	emit_vec("y_lor_add_identity", y6d);

	// 7. error codes: OVERLAP (output is the input), ILLEGAL (dense descriptor on
	//    a sparse input), MISMATCH (sizes)
	Vector<double> sq(n);
	rc = rc ? rc : set(sq, 1.0);
	Matrix<double> S(n, n);
	{
		std::vector<size_t> si, sj;
		std::vector<double> sv;
		for (size_t k = 0; k < n; ++k) {
			si.push_back(k);
			sj.push_back((k * 7) % n);
			sv.push_back(1.0);
		}
		rc = rc ? rc : buildMatrixUnique(S, si.data(), sj.data(), sv.data(), si.size(), SEQUENTIAL);
	}
	emit_num("rc_overlap", (double)mxv(sq, S, sq, plus_times));
	emit_num("rc_safe_overlap_ok", (double)(mxv<descriptors::safe_overlap>(y1, A, x, plus_times) == SUCCESS));
	emit_vec("y_plus_times_twice", y1);
	emit_num("rc_dense_sparse", (double)mxv<descriptors::dense>(y3, A, x, max_times));
	Vector<double> wrong(m + 1);
	emit_num("rc_mismatch", (double)mxv(wrong, A, x, plus_times));

	// 8. level 1 under sparsity: eWiseApply with an operator (intersection) and a
	//    monoid (union), eWiseMul and eWiseAdd in place, foldl, dot over max-times
	//    and a fold to a scalar
	Vector<double> u(n), w(n);
	for (size_t j = 0; j < n; ++j) {
		if (uniform() < 0.5) rc = rc ? rc : setElement(u, std::floor(uniform() * 7.0) - 3.0, j);
		if (uniform() < 0.5) rc = rc ? rc : setElement(w, std::floor(uniform() * 7.0) - 3.0, j);
	}
	emit_vec("u", u);
	emit_vec("w", w);
	Vector<double> z1(n), z2(n);
	rc = rc ? rc : eWiseApply(z1, u, w, operators::subtract<double>());
	rc = rc ? rc : eWiseApply(z2, u, w, Monoid<operators::max<double>, identities::negative_infinity>());
	emit_vec("ewise_op_sub", z1);
	emit_vec("ewise_monoid_max", z2);
	Vector<double> z3(u), z4(u);
	rc = rc ? rc : eWiseMul(z3, x, w, plus_times);
	rc = rc ? rc : eWiseAdd(z4, x, w, plus_times);
	emit_vec("ewise_mul_into_u", z3);
	emit_vec("ewise_add_into_u", z4);
	Vector<double> z5(u);
	rc = rc ? rc : foldl(z5, w, operators::mul<double>());
	emit_vec("foldl_op_mul", z5);
	double d1 = -INFINITY;
	rc = rc ? rc : dot(d1, u, w, max_times);
	emit_num("dot_max_times", d1);
	double d2 = 0.0;
	rc = rc ? rc : dot(d2, u, w, plus_times);
	emit_num("dot_plus_times", d2);
	int s1 = 0;
	rc = rc ? rc : foldl(s1, xi, Monoid<operators::add<int>, identities::zero>());
	emit_num("fold_sum_int", s1);
	double s2 = -INFINITY;
	rc = rc ? rc : foldl(s2, u, Monoid<operators::max<double>, identities::negative_infinity>());
	emit_num("fold_max", s2);

	// 9. eWiseLambda: a device lambda over u's nonzeroes, writing u and reading w
	{
		Vector<double> ul(u);
		auto uv = view(ul);
		auto wv = view(w);
		rc = rc ? rc : eWiseLambda([=] __device__(const size_t i) { uv[i] = 2.0 * uv[i] + (wv.assigned(i) ? wv[i] : 100.0); },
		                           ul, w);
		emit_vec("lambda_u", ul);
	}

	// 10. mxm over plus-min on int (generic SpGEMM) and over plus-times on double
	//     (the library's), A times A^T; masked set and clear
	Matrix<int> Ci(m, m);
	rc = rc ? rc : mxm<descriptors::transpose_right>(Ci, Ai, Ai, plus_min);
	emit_mat("mxm_plus_min_int", Ci);
	Matrix<double> Cd(m, m);
	rc = rc ? rc : mxm<descriptors::transpose_right>(Cd, A, A, plus_times);
	emit_mat("mxm_plus_times", Cd);
	Vector<double> ms(m);
	rc = rc ? rc : set(ms, mask, 4.0);
	emit_vec("masked_set", ms);

	// 11. a chain of sparse primitives with nothing read in between: each result
	//     count stays on the device until a later primitive needs it (the mxv)
	//     or the results are read at the end
	Vector<double> c1(n), c2(n), c3(n), c4(n), yc(m);
	rc = rc ? rc : eWiseApply(c1, u, w, operators::subtract<double>());
	rc = rc ? rc : eWiseAdd(c2, c1, u, plus_times);
	rc = rc ? rc : set(c3, c2);
	rc = rc ? rc : foldl(c3, c1, Monoid<operators::max<double>, identities::negative_infinity>());
	rc = rc ? rc : set(c4, c1, 7.0);  // a value mask: c1's nonzero entries
	rc = rc ? rc : eWiseMul(c4, c3, w, plus_times);
	rc = rc ? rc : mxv(yc, A, c3, plus_times);
	emit_vec("chain_c1", c1);
	emit_vec("chain_c2", c2);
	emit_vec("chain_c3", c3);
	emit_vec("chain_c4", c4);
	emit_vec("chain_y", yc);

	// 12. boolean vectors built from index and value ranges and from a dense
	//     range, and a count on the device over 2^21 entries (4096 blocks, so
	//     every lane of the count slot takes a share)
	{
		std::vector<size_t> bi;
		std::vector<bool> bv;
		for (size_t i = 0; i < n; i += 2) {
			bi.push_back(i);
			bv.push_back(i % 4 == 0);
		}
		Vector<bool> b1(n), b2(n);
		rc = rc ? rc : buildVector(b1, bi.begin(), bi.end(), bv.begin(), bv.end(), SEQUENTIAL);
		std::vector<bool> dv(n);
		for (size_t i = 0; i < n; ++i) dv[i] = i % 3 == 0;
		rc = rc ? rc : buildVector(b2, dv.begin(), dv.end(), SEQUENTIAL);
		emit_vec("bool_sparse", b1);
		emit_vec("bool_dense", b2);
		const size_t L = size_t(1) << 21;
		std::vector<bool> lmv(L);
		for (size_t i = 0; i < L; ++i) lmv[i] = i % 5 == 0;
		Vector<bool> lm(L);
		Vector<double> lv(L);
		rc = rc ? rc : buildVector(lm, lmv.begin(), lmv.end(), SEQUENTIAL);
		rc = rc ? rc : set(lv, lm, 1.0);  // a value mask: counted on the device
		emit_num("long_n", (double)L);
		emit_num("long_nnz", (double)nnz(lv));
		// repeated indices fold with Dup in input order, into an empty vector and
		// into one that holds entries; subtraction does not commute, so the order shows
		const std::vector<size_t> ri = {3, 5, 3, 3, 5, 0};
		const std::vector<double> rv = {10, 1, 2, 4, 8, 0.5};
		Vector<double> r1(n), r2(n);
		rc = rc ? rc : buildVector(r1, ri.begin(), ri.end(), rv.begin(), rv.end(), SEQUENTIAL,
		                           operators::subtract<double>());
		rc = rc ? rc : setElement(r2, 100.0, 3);
		rc = rc ? rc : setElement(r2, -1.0, 7);
		rc = rc ? rc : buildVector(r2, ri.begin(), ri.end(), rv.begin(), rv.end(), SEQUENTIAL,
		                           operators::subtract<double>());
		emit_vec("dup_empty", r1);
		emit_vec("dup_held", r2);
		// a long build with no repeats: every third index, value index / 3
		std::vector<size_t> li;
		std::vector<double> lvv;
		for (size_t i = L - 1; i < L; i -= 3) {  // descending: input order is not index order
			li.push_back(i);
			lvv.push_back((double)(i / 3));
		}
		Vector<double> lb(L);
		rc = rc ? rc : buildVector(lb, li.begin(), li.end(), lvv.begin(), lvv.end(), SEQUENTIAL);
		double lsum = 0;
		size_t lbad = 0;
		for (const auto &p : lb) {
			lsum += p.second;
			if (p.first % 3 != (L - 1) % 3 || p.second != (double)(p.first / 3)) ++lbad;
		}
		emit_num("long_build_nnz", (double)nnz(lb));
		emit_num("long_build_sum", lsum);
		emit_num("long_build_bad", (double)lbad);
	}
	emit_num("rc", (double)rc);
	out += "}";
	std::printf("%s\n", out.c_str());
	finalize();
	return rc == SUCCESS ? 0 : 2;
}
