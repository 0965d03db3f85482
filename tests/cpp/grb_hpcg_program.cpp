// An ALP/GraphBLAS HPCG program in the form of the reference's smoke test
// (tests/smoke/hpcg.cpp): the system comes from build_hpcg_system with
// hpcg_system_params, b = A 1, x = 0, and grb::algorithms::hpcg runs under a
// grb::Launcher; the output struct carries the iterations, the residual, the
// squared distance ||x - 1||^2 and a PinnedVector of x. Prints one JSON line.
//
//   grb_hpcg_program <nx> <ny> <nz> <max_coarsening_levels> <colours> <max_iterations>
#include <graphblas.hpp>
#include <graphblas/algorithms/hpcg/hpcg.hpp>
#include <graphblas/algorithms/hpcg/system_building_utils.hpp>

#include <array>
#include <cstdio>
#include <cstdlib>
#include <memory>

using namespace grb;
using namespace grb::algorithms;

struct input {
	size_t nx, ny, nz, levels, colours, max_iterations;
};

struct output {
	RC error_code = SUCCESS;
	size_t iterations = 0;
	double residual = 0.0;
	double square_norm_diff = 0.0;
	size_t pinned_nonzeroes = 0;
	double time_ms = 0.0;
};

static void grbProgram(const input &in, output &out) {
	std::unique_ptr<hpcg_data<double, double, double>> state;
	const std::array<size_t, 3> sizes = {in.nx, in.ny, in.nz};
	hpcg_system_params<3, double> params{sizes, 1, in.colours, 26.0, -1.0, 4, in.levels, 2};
	RC rc = build_hpcg_system<3, double>(state, params);
	if (rc != SUCCESS) {
		out.error_code = rc;
		return;
	}
	Matrix<double> &A = state->A;
	Vector<double> &x = state->x;
	Vector<double> &b = state->b;
	Semiring<operators::add<double>, operators::mul<double>, identities::zero, identities::one> ring;
	rc = rc ? rc : set(x, 1.0);
	rc = rc ? rc : set(b, 0.0);
	rc = rc ? rc : mxv(b, A, x, ring);
	rc = rc ? rc : set(x, 0.0);
	utils::Timer timer;
	rc = rc ? rc : hpcg(*state, true, 1, 1, in.max_iterations, 0.0, out.iterations, out.residual);
	out.time_ms = timer.time();
	// the smoke test's verification: ||x - 1||^2 through eWiseMul and dot
	rc = rc ? rc : set(b, 1.0);
	rc = rc ? rc : eWiseMul(b, -1.0, x, ring);
	rc = rc ? rc : dot(out.square_norm_diff, b, b, ring);
	PinnedVector<double> pinned(x, SEQUENTIAL);
	out.pinned_nonzeroes = pinned.nonzeroes();
	out.error_code = rc;
}

int main(int argc, char **argv) {
	if (argc < 7) return 2;
	input in{(size_t)std::atoll(argv[1]), (size_t)std::atoll(argv[2]), (size_t)std::atoll(argv[3]),
	         (size_t)std::atoll(argv[4]), (size_t)std::atoll(argv[5]), (size_t)std::atoll(argv[6])};
	output out;
	Launcher<AUTOMATIC> launcher;
	const RC rc = launcher.exec(&grbProgram, in, out, true);
	std::printf("{\"launcher_rc\": %d, \"rc\": %d, \"iterations\": %zu, \"residual\": %.17g, "
	            "\"square_norm_diff\": %.17g, \"pinned_nonzeroes\": %zu, \"ms\": %.3f, \"nprocs\": %zu, "
	            "\"pid\": %zu}\n",
	            (int)rc, (int)out.error_code, out.iterations, out.residual, out.square_norm_diff,
	            out.pinned_nonzeroes, out.time_ms, spmd<>::nprocs(), spmd<>::pid());
	return rc == SUCCESS && out.error_code == SUCCESS ? 0 : 1;
}
