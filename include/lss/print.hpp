// print.hpp -- the reference's text formats (common/lss_print.hpp), at the stream's precision (default 6
// significant digits):
//   2-D surface (:119-160): SPACE_POINTS_X / SPACE_POINTS_Y lines with the physical grid X(zeta_i),
//     Y(eta_j), then VALUES with one comma-separated row per S index;
//   1-D level (:82-110): SPACE_POINTS with X(zeta_i), then VALUES;
//   1-D surface in time (:170-220): SPACE_POINTS, TIME_POINTS (t_lo + m k), then VALUES, one row per
//     time index.
#pragma once

#include <ostream>

#include "configs.hpp"
#include "containers.hpp"
#include "lss_adi_sampler.hpp"
#include "lss_heat1d_sampler.hpp"

namespace lss_print {

template <typename fp_type, template <typename, typename> typename container, typename allocator>
void print(lss_pde_solvers::pde_discretization_config_2d_ptr<fp_type> const &pde_discretization_cfg,
           lss_grids::grid_config_hints_2d_ptr<fp_type> const &grid_hints_cfg,
           lss_containers::container_2d<lss_enumerations::by_enum::Row, fp_type, container, allocator> const &cont,
           std::ostream &out = std::cout)
{
    const auto np = pde_discretization_cfg->number_of_space_points();
    LSS_ASSERT((cont.columns() == np.second) && (cont.rows() == np.first),
               "The input cont container must have the correct size");
    lss_amd_detail::grid_spec<fp_type> g{};
    const auto sr = pde_discretization_cfg->space_range();
    g.x_lo = sr.first.lower();
    g.x_hi = sr.first.upper();
    g.y_lo = sr.second.lower();
    g.y_hi = sr.second.upper();
    g.accumulation_point = grid_hints_cfg->accumulation_point();
    g.alpha_scale = grid_hints_cfg->alpha_scale();
    g.beta_scale = grid_hints_cfg->beta_scale();
    g.nonuniform = grid_hints_cfg->grid() == lss_enumerations::grid_enum::Nonuniform;
    const lss_amd_detail::grid_map<fp_type> gm(g);
    const fp_type h1 = fp_type(1) / static_cast<fp_type>(np.first - 1);
    const fp_type h2 = fp_type(1) / static_cast<fp_type>(np.second - 1);
    out << "SPACE_POINTS_X\n";
    for (std::size_t t = 0; t < np.first; ++t)
        out << gm.x(h1 * static_cast<fp_type>(t)) << (t + 1 < np.first ? "," : "\nSPACE_POINTS_Y\n");
    for (std::size_t t = 0; t < np.second; ++t)
        out << gm.y(h2 * static_cast<fp_type>(t)) << (t + 1 < np.second ? "," : "\nVALUES\n");
    for (std::size_t r = 0; r < cont.rows(); ++r) {
        for (std::size_t c = 0; c < cont.columns(); ++c)
            out << cont(r, c) << (c + 1 < cont.columns() ? "," : "\n");
    }
}

namespace detail {
template <typename fp_type>
void space_points_1d(lss_pde_solvers::pde_discretization_config_1d_ptr<fp_type> const &pde_discretization_cfg,
                     lss_grids::grid_config_hints_1d_ptr<fp_type> const &grid_hints_cfg, std::ostream &out)
{
    lss_amd_detail::heat1d_spec<fp_type> g{};
    g.x_lo = pde_discretization_cfg->space_range().lower();
    g.x_hi = pde_discretization_cfg->space_range().upper();
    g.accumulation_point = grid_hints_cfg->accumulation_point();
    g.alpha_scale = grid_hints_cfg->alpha_scale();
    g.nonuniform = grid_hints_cfg->grid() == lss_enumerations::grid_enum::Nonuniform;
    const lss_amd_detail::grid_map1<fp_type> gm(g);
    const std::size_t n = pde_discretization_cfg->number_of_space_points();
    const fp_type h = fp_type(1) / static_cast<fp_type>(n - 1);
    out << "SPACE_POINTS\n";
    for (std::size_t t = 0; t < n; ++t)
        out << gm.x(h * static_cast<fp_type>(t)) << (t + 1 < n ? "," : "");
}
}  // namespace detail

template <typename fp_type, template <typename, typename> typename container, typename allocator>
void print(lss_pde_solvers::pde_discretization_config_1d_ptr<fp_type> const &pde_discretization_cfg,
           lss_grids::grid_config_hints_1d_ptr<fp_type> const &grid_hints_cfg, container<fp_type, allocator> const &cont,
           std::ostream &out = std::cout)
{
    const std::size_t space_size = pde_discretization_cfg->number_of_space_points();
    LSS_ASSERT(cont.size() == space_size, "Container size differs from passed discretization");
    detail::space_points_1d(pde_discretization_cfg, grid_hints_cfg, out);
    out << "\nVALUES\n";
    for (std::size_t t = 0; t < space_size; ++t)
        out << cont[t] << (t + 1 < space_size ? "," : "");
}

template <typename fp_type, template <typename, typename> typename container, typename allocator>
void print(lss_pde_solvers::pde_discretization_config_1d_ptr<fp_type> const &pde_discretization_cfg,
           lss_grids::grid_config_hints_1d_ptr<fp_type> const &grid_hints_cfg,
           lss_containers::container_2d<lss_enumerations::by_enum::Row, fp_type, container, allocator> const &cont_2d,
           std::ostream &out = std::cout)
{
    const std::size_t space_size = pde_discretization_cfg->number_of_space_points();
    const std::size_t time_size = pde_discretization_cfg->number_of_time_points();
    LSS_ASSERT((cont_2d.columns() == space_size) && (cont_2d.rows() == time_size),
               "Container size differs from passed discretization");
    detail::space_points_1d(pde_discretization_cfg, grid_hints_cfg, out);
    const fp_type k = pde_discretization_cfg->time_step();
    const fp_type time_start = pde_discretization_cfg->time_range().lower();
    out << "\nTIME_POINTS\n";
    for (std::size_t t = 0; t < time_size; ++t)
        out << (time_start + static_cast<fp_type>(t) * k) << (t + 1 < time_size ? "," : "");
    out << "\nVALUES\n";
    for (std::size_t t = 0; t < time_size; ++t)
        for (std::size_t i = 0; i < space_size; ++i)
            out << cont_2d(t, i) << (i + 1 < space_size ? "," : "\n");
}

}  // namespace lss_print
