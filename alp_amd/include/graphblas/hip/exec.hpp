/*
 * exec.hpp -- the launch and measurement surface of an ALP program on the hip
 * backend: grb::Launcher and grb::Benchmarker (reference/exec.hpp:38,
 * reference/benchmark.hpp:41, base/exec.hpp), grb::spmd and grb::collectives
 * (reference/spmd.hpp, reference/collectives.hpp:82-94) and grb::utils::Timer
 * (utils/timer.hpp).
 *
 * One process drives one GPU, as the reference backends run one user process
 * (reference/exec.hpp:82-92 rejects P != 1): spmd reports process 0 of 1 and the
 * collectives are identities.
 */
#ifndef ALP_HIP_GRB_EXEC_HPP
#define ALP_HIP_GRB_EXEC_HPP

#include <chrono>
#include <cmath>
#include <cstdio>
#include <iostream>

#include "algebra.hpp"

namespace grb {

enum EXEC_MODE { AUTOMATIC = 0, MANUAL, FROM_MPI };

namespace utils {

// wall-clock milliseconds since the last reset (utils/timer.hpp)
class Timer {
	std::chrono::steady_clock::time_point t0_ = std::chrono::steady_clock::now();

public:
	void reset() { t0_ = std::chrono::steady_clock::now(); }
	double time() const {
		return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0_).count();
	}
};

struct TimerResults {
	double io = 0, preamble = 0, useful = 0, postamble = 0;
	void set(double v) { io = preamble = useful = postamble = v; }
	void accum(const TimerResults &o) {
		io += o.io;
		preamble += o.preamble;
		useful += o.useful;
		postamble += o.postamble;
	}
};

}  // namespace utils

template <enum Backend impl = config::default_backend>
class spmd {
public:
	static size_t nprocs() { return 1; }
	static size_t pid() { return 0; }
	static RC sync(const size_t = 0, const size_t = 0) { return wait(); }
	static RC barrier() { return wait(); }
};

template <enum Backend impl = config::default_backend>
class collectives {
public:
	template <Descriptor descr = descriptors::no_operation, class Operator, typename IOType>
	static RC allreduce(IOType &, const Operator & = Operator()) { return SUCCESS; }
	template <Descriptor descr = descriptors::no_operation, class Operator, typename IOType>
	static RC reduce(IOType &, const size_t root = 0, const Operator & = Operator()) {
		return root == 0 ? SUCCESS : ILLEGAL;
	}
	template <typename IOType>
	static RC broadcast(IOType &, const size_t root = 0) { return root == 0 ? SUCCESS : ILLEGAL; }
};

// grb::Launcher (reference/exec.hpp:38): runs a program once on this process
// after grb::init, then grb::finalize
template <enum EXEC_MODE mode = AUTOMATIC, enum Backend impl = config::default_backend>
class Launcher {
public:
	Launcher(const size_t process_id = 0, const size_t nprocs = 1, const std::string = "localhost",
	         const std::string = "0", const bool = false) {
		if (process_id != 0 || nprocs != 1) throw std::invalid_argument("grb::Launcher (hip): one process");
	}
	template <typename T, typename U>
	RC exec(void (*grb_program)(const T &, U &), const T &data_in, U &data_out, const bool = false) const {
		RC r = init();
		if (r) return r;
		(*grb_program)(data_in, data_out);
		r = wait();
		const RC f = finalize();
		return r ? r : f;
	}
	template <typename U>
	RC exec(void (*grb_program)(const void *, const size_t, U &), const void *data_in, const size_t in_size,
	        U &data_out, const bool = false) const {
		RC r = init();
		if (r) return r;
		(*grb_program)(data_in, in_size, data_out);
		r = wait();
		const RC f = finalize();
		return r ? r : f;
	}
	static RC finalize() { return SUCCESS; }
};

// grb::Benchmarker (reference/benchmark.hpp:41): runs the program inner x outer
// times and prints the mean and spread of the per-call time
template <enum EXEC_MODE mode = AUTOMATIC, enum Backend impl = config::default_backend>
class Benchmarker {
	static void report(const std::vector<double> &t, const size_t inner) {
		double mean = 0, mn = 1e300, mx = 0;
		for (double v : t) {
			mean += v;
			mn = v < mn ? v : mn;
			mx = v > mx ? v : mx;
		}
		mean /= (double)(t.size() ? t.size() : 1);
		double var = 0;
		for (double v : t) var += (v - mean) * (v - mean);
		const double sd = t.size() > 1 ? std::sqrt(var / (double)(t.size() - 1)) : 0.0;
		std::cout << "Overall timings (io, preamble, useful, postamble):\n"
		          << "Avg: 0, 0, " << mean << ", 0\nMin: 0, 0, " << mn << ", 0\nMax: 0, 0, " << mx
		          << ", 0\nStd: 0, 0, " << sd << ", 0\n(" << t.size() << " outer x " << inner << " inner repetitions)"
		          << std::endl;
	}

public:
	Benchmarker(const size_t process_id = 0, const size_t nprocs = 1, const std::string = "localhost",
	            const std::string = "0") {
		if (process_id != 0 || nprocs != 1) throw std::invalid_argument("grb::Benchmarker (hip): one process");
	}
	template <typename T, typename U>
	RC exec(void (*grb_program)(const T &, U &), const T &data_in, U &data_out, const size_t inner,
	        const size_t outer, const bool = false) const {
		RC r = init();
		if (r) return r;
		std::vector<double> t;
		for (size_t o = 0; o < outer && !r; ++o) {
			utils::Timer tm;
			for (size_t i = 0; i < inner; ++i) (*grb_program)(data_in, data_out);
			r = wait();
			t.push_back(tm.time() / (double)(inner ? inner : 1));
		}
		report(t, inner);
		const RC f = finalize();
		return r ? r : f;
	}
	static RC finalize() { return SUCCESS; }
};

}  // namespace grb

#endif
