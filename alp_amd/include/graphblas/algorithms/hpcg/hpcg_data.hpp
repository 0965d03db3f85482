/*
 * The HPCG problem containers of ALP's algorithms/hpcg/hpcg_data.hpp (system_data,
 * multi_grid_data, hpcg_data) on the hip backend.
 *
 * hpcg_data holds the fine system A, the vectors b, x, r, z, u, p and, beside
 * them, the device multigrid hierarchy (alp_hip_hpcg_t) that the solve runs
 * on: the coarse systems, injection coarseners and colourings are built on the
 * device as build_hpcg_system builds them (system_building_utils.hpp:84-150).
 * The per-level host containers of the reference (coarser_level, color_masks)
 * are therefore not materialised: coarser_level stays null and color_masks
 * empty.
 */
#ifndef ALP_HIP_GRAPHBLAS_ALGORITHMS_HPCG_DATA_HPP
#define ALP_HIP_GRAPHBLAS_ALGORITHMS_HPCG_DATA_HPP

#include <array>
#include <cstddef>
#include <vector>

#include "../../../graphblas.hpp"

namespace grb {
namespace algorithms {

template <typename IOType, typename NonzeroType>
struct system_data {
	const std::size_t system_size;
	grb::Matrix<NonzeroType> A;
	grb::Vector<IOType> A_diagonal;
	grb::Vector<IOType> z;
	grb::Vector<IOType> r;
	grb::Vector<IOType> smoother_temp;
	std::vector<grb::Vector<bool>> color_masks;
	system_data(std::size_t n) : system_size(n), A(n, n), A_diagonal(n), z(n), r(n), smoother_temp(n) {}
	system_data(const system_data &) = delete;
	system_data &operator=(const system_data &) = delete;
};

template <typename IOType, typename NonzeroType>
struct multi_grid_data : public system_data<IOType, NonzeroType> {
	const std::size_t finer_size;
	grb::Vector<IOType> Ax_finer;
	grb::Matrix<NonzeroType> coarsening_matrix;
	multi_grid_data<IOType, NonzeroType> *coarser_level = nullptr;
	multi_grid_data(std::size_t coarser_size, std::size_t fine)
	    : system_data<IOType, NonzeroType>(coarser_size), finer_size(fine), Ax_finer(fine),
	      coarsening_matrix(coarser_size, fine) {}
	virtual ~multi_grid_data() { delete coarser_level; }
};

template <typename IOType, typename NonzeroType, typename InputType>
struct hpcg_data : public system_data<IOType, NonzeroType> {
	grb::Vector<InputType> b;
	grb::Vector<IOType> u;
	grb::Vector<IOType> p;
	grb::Vector<IOType> x;
	multi_grid_data<IOType, NonzeroType> *coarser_level = nullptr;

	// the device hierarchy and what it was built with
	alp_hip_hpcg_t device = nullptr;
	std::array<std::size_t, 3> sizes{{0, 0, 0}};
	std::size_t max_levels = 0, num_colors = 0, presmooth = 0, postsmooth = 0;

	hpcg_data(std::size_t n) : system_data<IOType, NonzeroType>(n), b(n), u(n), p(n), x(n) {}
	virtual ~hpcg_data() {
		delete coarser_level;
		if (device) {
			grb::internal::hip::Forcing force;
			alp_hip_hpcg_destroy(device);
		}
	}
};

}  // namespace algorithms
}  // namespace grb

#endif
