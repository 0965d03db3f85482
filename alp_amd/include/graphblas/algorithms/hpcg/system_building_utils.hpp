/*
 * build_hpcg_system and hpcg_system_params of ALP's
 * algorithms/hpcg/system_building_utils.hpp:84-150 on the hip backend: the fine
 * 27-point system (build_ndims_system_matrix, generated on the device) and the
 * parameters of the multigrid hierarchy, which is built on the device at the
 * first hpcg call (its smoother steps are arguments of that call). The device
 * hierarchy covers the reference's HPCG problems: three dimensions, halo 1,
 * coarsening step 2, minimum coarse side 4; other parameters give UNSUPPORTED.
 */
#ifndef ALP_HIP_GRAPHBLAS_ALGORITHMS_SYSTEM_BUILDING_UTILS_HPP
#define ALP_HIP_GRAPHBLAS_ALGORITHMS_SYSTEM_BUILDING_UTILS_HPP

#include <memory>

#include "hpcg_data.hpp"

namespace grb {
namespace algorithms {

template <std::size_t DIMS, typename T>
struct hpcg_system_params {
	const std::array<std::size_t, DIMS> &physical_sys_sizes;
	const std::size_t halo_size;
	const std::size_t num_colors;
	const T diag_value;
	const T non_diag_value;
	const std::size_t min_phys_size;
	const std::size_t max_levels;
	const std::size_t coarsening_step;
};

template <std::size_t DIMS, typename T = double>
grb::RC build_hpcg_system(std::unique_ptr<hpcg_data<T, T, T>> &holder, hpcg_system_params<DIMS, T> &params) {
	static_assert(DIMS == 3, "build_hpcg_system on the hip backend: three-dimensional systems");
	static_assert(std::is_same<T, double>::value, "build_hpcg_system on the hip backend: double values");
	if (params.halo_size != 1 || params.coarsening_step != 2 || params.min_phys_size != 4 ||
	    params.diag_value != 26.0 || params.non_diag_value != -1.0)
		return UNSUPPORTED;
	const auto &s = params.physical_sys_sizes;
	const std::size_t n = s[0] * s[1] * s[2];
	hpcg_data<T, T, T> *data = new hpcg_data<T, T, T>(n);
	holder = std::unique_ptr<hpcg_data<T, T, T>>(data);
	alp_hip_matrix_t h = nullptr;
	grb::internal::hip::Forcing force;  // a library build is blocking
	if (force.rc) return force.rc;
	RC rc = grb::internal::rc(alp_hip_matrix_stencil27(&h, s[0], s[1], s[2], params.diag_value,
	                                                   params.non_diag_value, 0));
	if (rc) return rc;
	data->A.reset(h);
	rc = set(data->A_diagonal, params.diag_value);
	data->sizes = {{s[0], s[1], s[2]}};
	data->max_levels = params.max_levels;
	data->num_colors = params.num_colors;
	return rc;
}

}  // namespace algorithms
}  // namespace grb

#endif
