// An ALP/GraphBLAS program in the grb:: style of the reference's own tests
// (tests/smoke/conjugate_gradient.cpp, tests/unit/dot.cpp, tests/smoke/knn.cpp),
// built against the hip backend's graphblas.hpp. Prints one JSON line.
//
//   grb_program <nx> <ny> <nz>
#include <graphblas.hpp>
#include <graphblas/algorithms/conjugate_gradient.hpp>
#include <graphblas/algorithms/knn.hpp>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

using namespace grb;

int main(int argc, char **argv) {
	const size_t nx = argc > 1 ? std::atoi(argv[1]) : 8, ny = argc > 2 ? std::atoi(argv[2]) : 8,
	             nz = argc > 3 ? std::atoi(argv[3]) : 8;
	const size_t n = nx * ny * nz;
	if (init() != SUCCESS) return 1;
	// the 27-point stencil as the reference's generator emits it: rows in order,
	// columns ascending, x fastest (ndim_matrix_builders.hpp:251-264)
	std::vector<size_t> I, J;
	std::vector<double> V;
	for (size_t z = 0; z < nz; ++z)
		for (size_t y = 0; y < ny; ++y)
			for (size_t x = 0; x < nx; ++x) {
				const size_t row = x + nx * (y + ny * z);
				for (long dz = -1; dz <= 1; ++dz)
					for (long dy = -1; dy <= 1; ++dy)
						for (long dx = -1; dx <= 1; ++dx) {
							const long cx = (long)x + dx, cy = (long)y + dy, cz = (long)z + dz;
							if (cx < 0 || cy < 0 || cz < 0 || cx >= (long)nx || cy >= (long)ny || cz >= (long)nz)
								continue;
							I.push_back(row);
							J.push_back((size_t)cx + nx * ((size_t)cy + ny * (size_t)cz));
							V.push_back(dx == 0 && dy == 0 && dz == 0 ? 26.0 : -1.0);
						}
			}
	Matrix<double> A(n, n);
	RC rc = buildMatrixUnique(A, I.data(), J.data(), V.data(), I.size(), SEQUENTIAL);
	Semiring<operators::add<double>, operators::mul<double>, identities::zero, identities::one> ring;
	Vector<double> x(n), b(n), ones(n), r(n), u(n), temp(n);
	rc = rc ? rc : set(ones, 1.0);
	rc = rc ? rc : set(b, 0.0);
	rc = rc ? rc : mxv(b, A, ones, ring);  // b = A 1
	double bb = 0.0;
	rc = rc ? rc : dot(bb, b, b, ring);
	const size_t nnz_b = nnz(b);
	// x = 0; CG to 1e-10 relative (the reference's smoke test uses these settings)
	rc = rc ? rc : set(x, 0.0);
	size_t iters = 0;
	double residual = 0.0;
	const RC cg = rc ? rc : algorithms::conjugate_gradient(x, A, b, 1000, 1e-10, iters, residual, r, u, temp);
	const std::vector<double> xv = x.values();
	double err = 0.0;
	for (double v : xv) err = std::fmax(err, std::fabs(v - 1.0));
	// BLAS1: temp = b - A x via eWiseApply and foldl
	rc = rc ? rc : set(temp, 0.0);
	rc = rc ? rc : mxv(temp, A, x, ring);
	rc = rc ? rc : foldl(temp, b, operators::subtract<double>());  // temp = A x - b
	double rr = 0.0;
	rc = rc ? rc : dot(rr, temp, temp, ring);
	// 2-hop neighbourhood of vertex 0 over the lor-land semiring (knn.hpp)
	Matrix<void> P(n, n);
	rc = rc ? rc : buildMatrixUnique(P, I.data(), J.data(), I.size(), SEQUENTIAL);
	Vector<bool> hop(n), buf(n);
	rc = rc ? rc : algorithms::knn(hop, P, 0, 2, buf);
	const std::vector<bool> hv = hop.values();
	size_t reached = 0;
	for (bool v : hv) reached += v ? 1 : 0;
	// masked lor-land product: frontier expansion into unvisited vertices
	Semiring<operators::logical_or<bool>, operators::logical_and<bool>, identities::logical_false,
	         identities::logical_true>
	    lor;
	Vector<bool> frontier(n), next(n);
	rc = rc ? rc : setElement(frontier, true, 0);
	rc = rc ? rc : set(next, false);
	rc = rc ? rc : mxv<descriptors::invert_mask>(next, frontier, P, frontier, lor);  // skip the frontier itself
	const std::vector<bool> nv = next.values();
	size_t next_count = 0;
	for (bool v : nv) next_count += v ? 1 : 0;
	std::printf("{\"rc\": %d, \"cg_rc\": %d, \"n\": %zu, \"nnz\": %zu, \"bb\": %.17g, \"nnz_b\": %zu, "
	            "\"iterations\": %zu, \"residual\": %.17g, \"max_err\": %.3e, \"true_rr\": %.6e, "
	            "\"two_hop\": %zu, \"next\": %zu}\n",
	            (int)rc, (int)cg, n, nnz(A), bb, nnz_b, iters, residual, err, rr, reached, next_count);
	finalize();
	return rc == SUCCESS && cg == SUCCESS ? 0 : 2;
}
