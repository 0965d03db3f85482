/*
 * grb::algorithms::hpcg on the hip backend: the signature and results of ALP's
 * algorithms/hpcg/hpcg.hpp:97-223. The preconditioned CG (multigrid V-cycle with
 * symmetric red-black Gauss-Seidel on index-mod-k colours, injection coarsening,
 * one smoother sweep on the coarsest level; stopping when ||r|| / ||r0|| <=
 * tolerance or after max_iterations) runs on the device in alp_hip_hpcg_solve
 * over the hierarchy of data.device, built here on first use or when the
 * smoother steps change. norm_residual is ||r||, as in the reference; like the
 * reference it returns SUCCESS also when the tolerance was not reached.
 */
#ifndef ALP_HIP_GRAPHBLAS_ALGORITHMS_HPCG_HPP
#define ALP_HIP_GRAPHBLAS_ALGORITHMS_HPCG_HPP

#include "hpcg_data.hpp"

namespace grb {
namespace algorithms {

template <typename IOType, typename ResidualType, typename NonzeroType, typename InputType,
          class Ring = Semiring<grb::operators::add<IOType>, grb::operators::mul<IOType>, grb::identities::zero,
                                grb::identities::one>,
          class Minus = operators::subtract<IOType>>
grb::RC hpcg(hpcg_data<IOType, NonzeroType, InputType> &data, bool with_preconditioning,
             const size_t presmoother_steps, const size_t postsmoother_steps, const size_t max_iterations,
             const ResidualType tolerance, size_t &iterations, ResidualType &norm_residual,
             const Ring & = Ring(), const Minus & = Minus()) {
	static_assert(std::is_same<IOType, double>::value && std::is_same<NonzeroType, double>::value &&
	                  std::is_same<InputType, double>::value,
	              "hpcg on the hip backend: double systems");
	grb::internal::hip::Forcing force;  // the library's solver is blocking
	if (force.rc) return force.rc;
	if (!data.device || data.presmooth != presmoother_steps || data.postsmooth != postsmoother_steps) {
		if (data.device) alp_hip_hpcg_destroy(data.device);
		data.device = nullptr;
		RC rc = grb::internal::rc(alp_hip_hpcg_create(&data.device, data.sizes[0], data.sizes[1], data.sizes[2],
		                                              data.max_levels, data.num_colors, presmoother_steps,
		                                              postsmoother_steps, 0));
		if (rc) return rc;
		data.presmooth = presmoother_steps;
		data.postsmooth = postsmoother_steps;
	}
	if (!data.b.dense()) return ILLEGAL;
	RC rc = SUCCESS;
	if (!data.x.dense()) rc = set<descriptors::invert_mask | descriptors::structural>(data.x, data.x, 0.0);
	if (rc) return rc;
	size_t it = 0;
	double res = 0.0;
	rc = grb::internal::rc(alp_hip_hpcg_solve(data.device, data.b.raw(), data.x.raw(), max_iterations,
	                                          (double)tolerance, 0, with_preconditioning ? 1 : 0, &it, &res,
	                                          nullptr));
	iterations = it;
	norm_residual = static_cast<ResidualType>(res);
	return rc;
}

}  // namespace algorithms
}  // namespace grb

#endif
