// Lazy execution and grb::Pipeline on the hip backend: each loop below runs
// once with its primitives issued one by one and once as the body of a
// grb::Pipeline, from the same start; the results must agree bit for bit.
// Prints one JSON line.
//
//   grb_pipeline_program <nx> <ny> <nz> <sweeps>
//
// A. Jacobi sweeps on the 27-point stencil: t = A x; r = b - t; x += D^-1 r.
//    Dense vectors only: one segment per call, a graph updated in place.
// B. A body whose scalars change every call (x = beta_k x + alpha_k y).
// C. Power iteration with the norm inside the body: the dot is a forcing point,
//    so each call launches two segments.
// D. A value-masked set whose count only the device knows, consumed inside the
//    body by an mxv that asks for it (a forcing point), and read after the call.
// E. Jacobi-preconditioned CG written in grb:: primitives
//    (algorithms::preconditioned_conjugate_gradient): time per iteration.
// F. Eight Jacobi sweeps per call (32 launches in one segment): time per call
//    against the same sweeps launched directly.
#include <graphblas.hpp>
#include <graphblas/algorithms/conjugate_gradient.hpp>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <vector>

using namespace grb;

typedef Semiring<operators::add<double>, operators::mul<double>, identities::zero, identities::one> Ring;

static double now_us() {
	return std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

template <typename T>
static bool same_bits(const Vector<T> &a, const Vector<T> &b) {
	const std::vector<T> va = a.values(), vb = b.values();
	return va.size() == vb.size() && std::memcmp(va.data(), vb.data(), sizeof(T) * va.size()) == 0 &&
	       nnz(a) == nnz(b);
}

int main(int argc, char **argv) {
	const size_t nx = argc > 1 ? std::atoi(argv[1]) : 16, ny = argc > 2 ? std::atoi(argv[2]) : 16,
	             nz = argc > 3 ? std::atoi(argv[3]) : 16;
	const int sweeps = argc > 4 ? std::atoi(argv[4]) : 50;
	const size_t n = nx * ny * nz;
	if (init() != SUCCESS) return 1;
	// the 27-point stencil, rows in order, columns ascending
	std::vector<size_t> I, J;
	std::vector<double> V;
	for (size_t z = 0; z < nz; ++z)
		for (size_t y = 0; y < ny; ++y)
			for (size_t x = 0; x < nx; ++x)
				for (long dz = -1; dz <= 1; ++dz)
					for (long dy = -1; dy <= 1; ++dy)
						for (long dx = -1; dx <= 1; ++dx) {
							const long cx = (long)x + dx, cy = (long)y + dy, cz = (long)z + dz;
							if (cx < 0 || cy < 0 || cz < 0 || cx >= (long)nx || cy >= (long)ny || cz >= (long)nz)
								continue;
							I.push_back(x + nx * (y + ny * z));
							J.push_back((size_t)cx + nx * ((size_t)cy + ny * (size_t)cz));
							V.push_back(dx == 0 && dy == 0 && dz == 0 ? 26.0 : -1.0);
						}
	Matrix<double> A(n, n);
	RC rc = buildMatrixUnique(A, I.data(), J.data(), V.data(), I.size(), SEQUENTIAL);
	Ring ring;
	operators::subtract<double> minus;
	Vector<double> ones(n), b(n), dinv(n), x(n), t(n), r(n);
	rc = rc ? rc : set(ones, 1.0);
	rc = rc ? rc : set(b, 0.0);
	rc = rc ? rc : mxv(b, A, ones, ring);
	rc = rc ? rc : set(dinv, 1.0 / 26.0);

	// ---- A: Jacobi sweeps ------------------------------------------------------------
	auto sweep = [&]() -> RC {
		RC e = set(t, 0.0);
		e = e ? e : mxv(t, A, x, ring);
		e = e ? e : eWiseApply(r, b, t, minus);
		e = e ? e : eWiseMul(x, dinv, r, ring);
		return e;
	};
	rc = rc ? rc : set(x, 0.0);
	for (int k = 0; k < sweeps && !rc; ++k) rc = sweep();
	Vector<double> x_eager(x);
	Pipeline pa;
	rc = rc ? rc : set(x, 0.0);
	for (int k = 0; k < sweeps && !rc; ++k) rc = pa(sweep);
	const bool a_equal = !rc && same_bits(x, x_eager);
	const size_t a_segments = pa.segments(), a_updated = pa.updated(), a_inst = pa.instantiated();
	// time both forms from here on (the device work is the same)
	rc = rc ? rc : wait();
	double t0 = now_us();
	for (int k = 0; k < sweeps && !rc; ++k) rc = sweep();
	rc = rc ? rc : wait();
	const double a_eager_us = (now_us() - t0) / sweeps;
	t0 = now_us();
	for (int k = 0; k < sweeps && !rc; ++k) rc = pa(sweep);
	rc = rc ? rc : wait();
	const double a_pipe_us = (now_us() - t0) / sweeps;

	// ---- B: scalars that change every call -----------------------------------------
	Vector<double> y(n), u(n);
	rc = rc ? rc : set(y, 0.0);
	rc = rc ? rc : mxv(y, A, ones, ring);
	int kb = 0;
	auto scaled = [&]() -> RC {
		const double alpha = 1.0 / (kb + 2), beta = 0.5 + 0.01 * kb;
		++kb;
		RC e = foldl(u, beta, operators::mul<double>());
		return e ? e : eWiseMul(u, alpha, y, ring);
	};
	rc = rc ? rc : set(u, 1.0);
	for (kb = 0; kb < 20 && !rc;) rc = scaled();
	Vector<double> u_eager(u);
	Pipeline pb;
	rc = rc ? rc : set(u, 1.0);
	for (kb = 0; kb < 20 && !rc;) rc = pb(scaled);
	const bool b_equal = !rc && same_bits(u, u_eager);
	const size_t b_updated = pb.updated();

	// ---- C: power iteration, a forcing dot inside the body --------------------------
	double lambda = 0.0;
	auto power = [&]() -> RC {
		RC e = set(t, 0.0);
		e = e ? e : mxv(t, A, x, ring);
		double s = 0.0;
		e = e ? e : dot(s, t, t, ring);
		lambda = std::sqrt(s);
		e = e ? e : set(x, t);
		return e ? e : foldl(x, 1.0 / lambda, operators::mul<double>());
	};
	rc = rc ? rc : set(x, 1.0);
	for (int k = 0; k < 10 && !rc; ++k) rc = power();
	const double lambda_eager = lambda;
	Vector<double> p_eager(x);
	Pipeline pc;
	rc = rc ? rc : set(x, 1.0);
	for (int k = 0; k < 10 && !rc; ++k) rc = pc(power);
	const bool c_equal = !rc && same_bits(x, p_eager) && lambda == lambda_eager;
	const size_t c_segments = pc.segments();

	// ---- D: a count only the device knows, consumed inside the body -----------------
	Vector<bool> sel(n);
	Vector<double> s(n), w(n);
	auto masked = [&]() -> RC {
		RC e = eWiseApply(sel, x, 1.0 / std::sqrt((double)n), operators::greater_than<double, double, bool>());
		e = e ? e : clear(s);
		e = e ? e : set(s, sel, 2.0);  // the count depends on the values
		e = e ? e : set(w, 0.0);
		return e ? e : mxv(w, A, s, ring);  // asks for s's count: a forcing point
	};
	rc = rc ? rc : masked();
	const size_t d_nnz_eager = nnz(s);
	Vector<double> w_eager(w);
	Pipeline pd;
	rc = rc ? rc : pd(masked);
	rc = rc ? rc : pd(masked);
	const size_t d_nnz = nnz(s);
	size_t d_count = 0;
	for (bool v : sel.values()) d_count += v ? 1 : 0;
	const bool d_equal = !rc && same_bits(w, w_eager) && d_nnz == d_nnz_eager && d_nnz == d_count;
	const size_t d_segments = pd.segments();

	// ---- E: PCG in grb:: primitives -----------------------------------------------
	Vector<double> xs(n), rs(n), us(n), ts(n), zs(n);
	auto jacobi = [&](Vector<double> &z_, const Vector<double> &r_) -> RC {
		RC e = set(z_, 0.0);
		return e ? e : eWiseMul(z_, dinv, r_, ring);
	};
	size_t pcg_iters = 0;
	double pcg_res = 0.0;
	rc = rc ? rc : set(xs, 0.0);
	rc = rc ? rc : wait();
	t0 = now_us();
	const std::function<RC(Vector<double> &, const Vector<double> &)> minv = jacobi;
	const RC pcg = rc ? rc
	                  : algorithms::preconditioned_conjugate_gradient(xs, A, b, minv, 1000, 1e-10, pcg_iters, pcg_res,
	                                                                  rs, us, ts, zs);
	rc = rc ? rc : wait();
	const double pcg_us = (now_us() - t0) / (pcg_iters ? pcg_iters : 1);
	double err = 0.0;
	for (double v : xs.values()) err = std::fmax(err, std::fabs(v - 1.0));

	// ---- F: a long segment ------------------------------------------------------------
	auto sweeps8 = [&]() -> RC {
		RC e = SUCCESS;
		for (int j = 0; j < 8 && !e; ++j) e = sweep();
		return e;
	};
	const int calls = sweeps / 4 > 2 ? sweeps / 4 : 2;
	rc = rc ? rc : set(x, 0.0);
	for (int k = 0; k < calls && !rc; ++k) rc = sweeps8();
	Vector<double> f_eager(x);
	Pipeline pf;
	rc = rc ? rc : set(x, 0.0);
	for (int k = 0; k < calls && !rc; ++k) rc = pf(sweeps8);
	const bool f_equal = !rc && same_bits(x, f_eager);
	const size_t f_segments = pf.segments();
	rc = rc ? rc : wait();
	t0 = now_us();
	for (int k = 0; k < calls && !rc; ++k) rc = sweeps8();
	rc = rc ? rc : wait();
	const double f_eager_us = (now_us() - t0) / calls;
	t0 = now_us();
	for (int k = 0; k < calls && !rc; ++k) rc = pf(sweeps8);
	rc = rc ? rc : wait();
	const double f_pipe_us = (now_us() - t0) / calls;

	std::printf("{\"rc\": %d, \"n\": %zu, \"sweeps\": %d, \"a_equal\": %d, \"a_segments\": %zu, "
	            "\"a_updated\": %zu, \"a_instantiated\": %zu, \"a_eager_us\": %.2f, \"a_pipeline_us\": %.2f, "
	            "\"f_equal\": %d, \"f_segments\": %zu, \"f_eager_us\": %.2f, \"f_pipeline_us\": %.2f, "
	            "\"b_equal\": %d, \"b_updated\": %zu, \"c_equal\": %d, \"c_segments\": %zu, \"lambda\": %.17g, "
	            "\"d_equal\": %d, \"d_nnz\": %zu, \"d_segments\": %zu, \"pcg_rc\": %d, \"pcg_iterations\": %zu, "
	            "\"pcg_residual\": %.6e, \"pcg_max_err\": %.3e, \"pcg_us_per_iteration\": %.2f}\n",
	            (int)rc, n, sweeps, (int)a_equal, a_segments, a_updated, a_inst, a_eager_us, a_pipe_us,
	            (int)f_equal, f_segments, f_eager_us, f_pipe_us,
	            (int)b_equal, b_updated, (int)c_equal, c_segments, lambda, (int)d_equal, d_nnz, d_segments,
	            (int)pcg, pcg_iters, pcg_res, err, pcg_us);
	finalize();
	return rc == SUCCESS ? 0 : 2;
}
