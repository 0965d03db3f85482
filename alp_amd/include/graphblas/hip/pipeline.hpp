/*
 * pipeline.hpp -- grb::Pipeline: a sequence of primitives that a program runs
 * again and again (a loop body), launched as hipGraphs.
 *
 * ALP's nonblocking backend (docs/Nonblocking_backend.md, "Lazy evaluation";
 * include/graphblas/nonblocking/pipeline.hpp) defers each primitive as a stage of
 * a pipeline and runs the stages together when a primitive forces execution.
 * On this backend every primitive is already deferred: it enqueues device work
 * and returns, and the host waits only at forcing points (base.hpp). A Pipeline
 * further turns a recurring sequence into hipGraph launches:
 *
 *     grb::Pipeline step;
 *     for (...) {
 *         rc = step([&]() -> grb::RC {
 *             grb::RC r = grb::mxv(t, A, x, ring);
 *             r = r ? r : grb::eWiseApply(x, t, b, minus);
 *             return r;
 *         });
 *     }
 *
 * Each call runs the body's host code as written -- every argument, scalar and
 * sparsity decision is the one of this call -- while the work the primitives
 * enqueue is captured (hipStreamBeginCapture on the library's stream) instead of
 * launched. A forcing point inside the body (a scalar output, a count only the
 * device knows, a host copy, an allocation or release) ends the open segment:
 * the segment is launched as one graph, the forcing primitive runs as usual, and
 * the next enqueued work opens the next segment. At the end of the body the last
 * segment is launched. Segment k of a call updates segment k of the previous call
 * in place (hipGraphExecUpdate: same kernels, new arguments) and is instantiated
 * anew only when its shape changed. Results are those of running the body
 * without the Pipeline, bit for bit: the same kernels run in the same order.
 *
 * On MI355X with ROCm 7.2 this is not faster than the direct (lazy) launches: a
 * graph launch costs ~16 us whatever its size and the capture plus update ~2 us
 * per kernel on the host, against ~2.7 us per direct launch. A Jacobi sweep on
 * 16^3 (four primitives) takes 27 us per call this way against 8 us direct, and
 * eight sweeps in one call 86 against 64 us. The Pipeline is for callers
 * that want a loop's device work as graphs; the speed of user sequences comes
 * from the lazy launches and the fusion of element-wise runs.
 *
 * The call returns once the work is enqueued (as any primitive); grb::wait or any
 * forcing point waits for it. Calls may not nest, and may not run inside a
 * forcing scope. Device memory the body frees is freed at a forcing point, so a
 * body that creates and destroys containers runs correctly, in more segments.
 */
#ifndef ALP_HIP_GRB_PIPELINE_HPP
#define ALP_HIP_GRB_PIPELINE_HPP

#include "base.hpp"

namespace grb {

class Pipeline : private internal::hip::SegmentSink {
public:
	Pipeline() = default;
	Pipeline(const Pipeline &) = delete;
	Pipeline &operator=(const Pipeline &) = delete;
	~Pipeline() {
		for (hipGraphExec_t x : execs_)
			if (x) (void)hipGraphExecDestroy(x);
	}

	// runs body(), a callable returning grb::RC, with its device work captured
	// and launched segment by segment; returns the body's code, else the first
	// capture or launch error
	template <class F>
	RC operator()(F &&body) {
		internal::hip::LazyState &l = internal::hip::lazy();
		if (l.sink || l.forcing) return ILLEGAL;
		l.sink = this;
		seg_ = 0;
		err_ = SUCCESS;
		open_ = false;
		RC r = SUCCESS;
		try {
			r = body();
		} catch (...) {
			internal::hip::flush_stages();
			(void)close();
			l.sink = nullptr;
			throw;
		}
		internal::hip::flush_stages();  // the body's last fused pass joins its last segment
		const RC c = close();
		l.sink = nullptr;
		segments_ = seg_;
		++calls_;
		return r ? r : (c ? c : err_);
	}

	size_t calls() const { return calls_; }
	size_t segments() const { return segments_; }  // graphs the last call launched
	size_t updated() const { return updated_; }    // graph launches that reused an earlier graph
	size_t instantiated() const { return instantiated_; }

private:
	void open() override {
		if (open_ || err_) return;
		open_ = true;
		const hipError_t e = hipStreamBeginCapture(internal::hip::library_stream(), hipStreamCaptureModeThreadLocal);
		if (e != hipSuccess) {
			open_ = false;
			err_ = internal::hip::from_hip(e);
		}
	}

	RC close() override {
		if (!open_) return err_;
		open_ = false;
		const hipStream_t s = internal::hip::library_stream();
		hipGraph_t g = nullptr;
		hipError_t e = hipStreamEndCapture(s, &g);
		if (e != hipSuccess || !g) {
			// a call the capture does not allow invalidated it: nothing of the
			// segment ran
			(void)hipGetLastError();
			if (g) (void)hipGraphDestroy(g);
			err_ = PANIC;
			return err_;
		}
		size_t nodes = 0;
		(void)hipGraphGetNodes(g, nullptr, &nodes);
		if (nodes == 0) {
			(void)hipGraphDestroy(g);
			return err_;
		}
		if (seg_ >= execs_.size()) execs_.resize(seg_ + 1, nullptr);
		hipGraphExec_t &x = execs_[seg_];
		if (x) {
			hipGraphNode_t bad = nullptr;
			hipGraphExecUpdateResult res = hipGraphExecUpdateError;
			if (hipGraphExecUpdate(x, g, &bad, &res) == hipSuccess && res == hipGraphExecUpdateSuccess) {
				++updated_;
			} else {
				(void)hipGetLastError();
				(void)hipGraphExecDestroy(x);
				x = nullptr;
			}
		}
		if (!x) {
			e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
			if (e != hipSuccess) {
				x = nullptr;
				(void)hipGraphDestroy(g);
				err_ = internal::hip::from_hip(e);
				return err_;
			}
			++instantiated_;
		}
		(void)hipGraphDestroy(g);
		++seg_;
		e = hipGraphLaunch(x, s);
		if (e != hipSuccess) err_ = internal::hip::from_hip(e);
		return err_;
	}

	std::vector<hipGraphExec_t> execs_;
	size_t seg_ = 0, segments_ = 0, calls_ = 0, updated_ = 0, instantiated_ = 0;
	bool open_ = false;
	RC err_ = SUCCESS;
};

}  // namespace grb

#endif
