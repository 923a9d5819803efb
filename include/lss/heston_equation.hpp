// heston_equation.hpp -- drop-in implicit_solvers::general_heston_equation backed by the MI355X library.
//
// Same class, constructor and solve(container_2d<Row>&) as the reference
// (pde_solvers/two_dimensional/heat_type/lss_2d_general_heston_equation.hpp:83-307).  solve() samples
// the user's closures on the reference grid (lss_adi_sampler.hpp), builds a one-problem plan through
// the C-ABI (include/lss_amd.h) and runs the whole time loop on the GPU:
//   * every memory_space / tridiagonal_method selects the GPU stepper; the line solver keeps the
//     operation order of the selected host solver (DoubleSweep -> double sweep, every other choice ->
//     Thomas LU), so a Host/ThomasLU or Host/DoubleSweep configuration returns the bits the
//     reference's host solver returns;
//   * boundaries are read through value() whatever their type, as heston_boundary_transform does
//     (lss_heston_boundary_transform.hpp:36-66): Dirichlet at S_min, Neumann at S_max, Dirichlet at v_max;
//   * a heat source is stepped: the first stage of every scheme adds theta rho F(t_next) +
//     (1 - theta) rho F(t) (rhs_intermed_1_source, lss_heat_douglas_rachford_method.hpp:70-104, and
//     the same function of CS/MCS/HV).  This is the one place where the drop-in does not return the
//     reference's result: the reference defines those functions but skips the time loop when a
//     source is set (implicit_kernel.hpp:1032-1044) and returns the zero surface (Host) or the
//     sampled initial surface (Device).
// Errors of the GPU library are thrown as std::runtime_error; size mismatches terminate through
// LSS_ASSERT as in the reference.
#pragma once

#include <map>
#include <stdexcept>
#include <string>

#include "boundaries.hpp"
#include "configs.hpp"
#include "containers.hpp"
#include "lss_adi_sampler.hpp"
#include "lss_amd.h"

namespace lss_pde_solvers {
namespace two_dimensional {
namespace implicit_solvers {

using lss_boundary::boundary_2d_pair;
using lss_boundary::boundary_2d_ptr;
using lss_containers::container_2d;
using lss_containers::container_3d;
using lss_enumerations::by_enum;
using lss_enumerations::grid_enum;
using lss_grids::grid_config_hints_2d_ptr;

namespace detail {
inline void check(int rc, const char *what)
{
    if (rc != LSS_OK)
        throw std::runtime_error(std::string(what) + ": " + lss_last_error());
}
}  // namespace detail

template <typename fp_type, template <typename, typename> typename container = std::vector,
          typename allocator = std::allocator<fp_type>>
class general_heston_equation {
  private:
    heat_data_config_2d_ptr<fp_type> heat_data_cfg_;
    pde_discretization_config_2d_ptr<fp_type> discretization_cfg_;
    boundary_2d_ptr<fp_type> vertical_upper_;
    boundary_2d_pair<fp_type> horizontal_pair_;
    splitting_method_config_ptr<fp_type> splitting_method_cfg_;
    grid_config_hints_2d_ptr<fp_type> grid_hints_;
    heat_implicit_solver_config_ptr solver_cfg_;
    std::map<std::string, fp_type> solver_config_details_;

    void initialize()
    {
        LSS_VERIFY(heat_data_cfg_, "heat_data_config must not be null");
        LSS_VERIFY(discretization_cfg_, "discretization_config must not be null");
        LSS_VERIFY(splitting_method_cfg_, "splitting_method_config must not be null");
        LSS_VERIFY(solver_cfg_, "solver_config must not be null");
        LSS_VERIFY(grid_hints_, "grid_config_hints must not be null");
        if (!solver_config_details_.empty()) {
            auto const &it = solver_config_details_.find("sor_omega");
            LSS_ASSERT(it != solver_config_details_.end(), "sor_omega is not defined");
        }
    }

  public:
    explicit general_heston_equation() = delete;
    explicit general_heston_equation(
        heat_data_config_2d_ptr<fp_type> const &heat_data_config,
        pde_discretization_config_2d_ptr<fp_type> const &discretization_config,
        boundary_2d_ptr<fp_type> const &vertical_upper_boundary_ptr,
        boundary_2d_pair<fp_type> const &horizontal_boundary_pair,
        splitting_method_config_ptr<fp_type> const &splitting_method_config,
        grid_config_hints_2d_ptr<fp_type> const &grid_config_hints,
        heat_implicit_solver_config_ptr const &solver_config =
            default_heat_solver_configs::host_fwd_dssolver_euler_solver_config_ptr,
        std::map<std::string, fp_type> const &solver_config_details = std::map<std::string, fp_type>())
        : heat_data_cfg_{heat_data_config}, discretization_cfg_{discretization_config},
          vertical_upper_{vertical_upper_boundary_ptr}, horizontal_pair_{horizontal_boundary_pair},
          splitting_method_cfg_{splitting_method_config}, grid_hints_{grid_config_hints},
          solver_cfg_{solver_config}, solver_config_details_{solver_config_details}
    {
        initialize();
    }
    general_heston_equation(general_heston_equation const &) = delete;
    general_heston_equation(general_heston_equation &&) = delete;
    general_heston_equation &operator=(general_heston_equation const &) = delete;
    general_heston_equation &operator=(general_heston_equation &&) = delete;

    // grid description of this equation in the sampler's terms
    lss_amd_detail::grid_spec<fp_type> grid() const
    {
        lss_amd_detail::grid_spec<fp_type> g;
        const auto sr = discretization_cfg_->space_range();
        const auto np = discretization_cfg_->number_of_space_points();
        g.x_lo = sr.first.lower();
        g.x_hi = sr.first.upper();
        g.y_lo = sr.second.lower();
        g.y_hi = sr.second.upper();
        g.t_lo = discretization_cfg_->time_range().lower();
        g.t_hi = discretization_cfg_->time_range().upper();
        g.nx = np.first;
        g.ny = np.second;
        g.nt = discretization_cfg_->number_of_time_points();
        g.accumulation_point = grid_hints_->accumulation_point();
        g.alpha_scale = grid_hints_->alpha_scale();
        g.beta_scale = grid_hints_->beta_scale();
        g.nonuniform = grid_hints_->grid() == grid_enum::Nonuniform;
        g.splitting = static_cast<int>(splitting_method_cfg_->splitting_method());
        g.weighting = splitting_method_cfg_->weighting_value();
        g.implicit_scheme = static_cast<int>(solver_cfg_->implicit_pde_scheme());
        g.tridiagonal = solver_cfg_->tridiagonal_method() == tridiagonal_method_enum::DoubleSweepSolver
                            ? LSS_DOUBLE_SWEEP
                            : LSS_THOMAS_LU;
        g.direction = solver_cfg_->traverse_direction() == traverse_direction_enum::Forward ? LSS_FORWARD
                                                                                          : LSS_BACKWARD;
        return g;
    }

    lss_amd_detail::user_data<fp_type> data() const
    {
        lss_amd_detail::user_data<fp_type> u;
        u.A = heat_data_cfg_->a_coefficient();
        u.B = heat_data_cfg_->b_coefficient();
        u.C = heat_data_cfg_->c_coefficient();
        u.D = heat_data_cfg_->d_coefficient();
        u.E = heat_data_cfg_->e_coefficient();
        u.F = heat_data_cfg_->f_coefficient();
        u.init = heat_data_cfg_->initial_condition();
        auto lo = horizontal_pair_.first, hi = horizontal_pair_.second, vu = vertical_upper_;
        u.h_lower = [lo](fp_type t, fp_type y) { return lo->value(t, y); };
        u.h_upper = [hi](fp_type t, fp_type y) { return hi->value(t, y); };
        u.v_upper = [vu](fp_type t, fp_type x) { return vu->value(t, x); };
        if (heat_data_cfg_->source_data_config())
            u.src = heat_data_cfg_->source_data_config()->heat_source();
        return u;
    }

    /**
     * Get the final solution of the PDE (heston_equation.hpp:184-307)
     *
     * \param solution - 2D container for the solution, nx rows (S) by ny columns (v)
     */
    void solve(container_2d<by_enum::Row, fp_type, container, allocator> &solution)
    {
        LSS_ASSERT((solution.rows()) > 0 && (solution.columns() > 0),
                   "The input solution container must be initialized");
        const auto np = discretization_cfg_->number_of_space_points();
        LSS_ASSERT((solution.columns() == np.second) && (solution.rows() == np.first),
                   "The input solution container must have the correct size");
        const auto g = grid();
        const auto u = data();
        lss_amd_detail::sampled<fp_type> s;
        lss_amd_detail::sample_fields(u, g, s, lss_amd_detail::TIME_AUTO, lss_amd_detail::TIME_AUTO);
        const lss_adi_desc d = s.desc(LSS_COEFF_FIELDS, LSS_LINE_EXACT);
        lss_adi_plan *plan = nullptr;
        detail::check(lss_adi_plan_create(&d, 1, &plan), "lss_adi_plan_create");
        const void *in[1] = {s.init.data()};
        void *out[1] = {solution.raw()};
        const int rc = lss_adi_run_host(plan, in, out);
        lss_adi_plan_destroy(plan);
        detail::check(rc, "lss_adi_run_host");
    }

    /**
     * Get every time level of the PDE (heston_equation.hpp:180, 309-313; the reference declares this
     * solve and leaves its body empty).  Layers are time indices as the reference's 1-D
     * run_with_stepping stores them (lss_1d_general_heat_equation_implicit_kernel.hpp:163-257): a
     * Forward loop puts the initial surface in layer 0 and the surface after s steps in layer s; a
     * Backward loop puts the terminal surface in the last layer and the surface after s steps in layer
     * number_of_time_points - 1 - s.  Layer time_idx thus holds the solution at time t_lo + time_idx k.
     *
     * \param solutions - 3D container, nx rows (S) by ny columns (v) by number_of_time_points layers
     */
    void solve(container_3d<by_enum::Row, fp_type, container, allocator> &solutions)
    {
        const auto np = discretization_cfg_->number_of_space_points();
        const std::size_t nt = discretization_cfg_->number_of_time_points();
        LSS_ASSERT((solutions.rows() == np.first) && (solutions.columns() == np.second) && (solutions.layers() == nt),
                   "The input solution container must have the correct size");
        const auto g = grid();
        const auto u = data();
        lss_amd_detail::sampled<fp_type> s;
        lss_amd_detail::sample_fields(u, g, s, lss_amd_detail::TIME_AUTO, lss_amd_detail::TIME_AUTO);
        const lss_adi_desc d = s.desc(LSS_COEFF_FIELDS, LSS_LINE_EXACT);
        lss_adi_plan *plan = nullptr;
        detail::check(lss_adi_plan_create(&d, 1, &plan), "lss_adi_plan_create");
        const std::size_t nxy = np.first * np.second;
        std::vector<fp_type> levels(nt * nxy);
        const void *in[1] = {s.init.data()};
        void *lv[1] = {levels.data()};
        const int rc = lss_adi_run_levels(plan, in, lv, nullptr, 1);
        lss_adi_plan_destroy(plan);
        detail::check(rc, "lss_adi_run_levels");
        const bool fwd = g.direction == LSS_FORWARD;
        for (std::size_t st = 0; st < nt; ++st) {
            auto &layer = solutions.layer(fwd ? st : nt - 1 - st);
            std::copy(levels.begin() + st * nxy, levels.begin() + (st + 1) * nxy, layer.raw());
        }
    }
};

template <typename fp_type, template <typename, typename> typename container = std::vector,
          typename allocator = std::allocator<fp_type>>
using general_heston_equation_ptr = lss_utility::sptr_t<general_heston_equation<fp_type, container, allocator>>;

}  // namespace implicit_solvers

// -------------------------------------------------------------------------------------------------
// Drop-in explicit_solvers::general_heston_equation: same class, constructor and
// solve(container_2d<Row>&) as the reference (lss_2d_general_heston_equation.hpp:317-510).  The Euler
// scheme runs on the GPU for both memory spaces (LSS_EXPLICIT_EULER; FIELDS coefficients, so the
// result has the bits of the reference's host loop); the ADE schemes throw "Not currently supported"
// as the reference's kernels do (explicit_kernel.hpp:88-99, 157-168).  A heat source leaves the
// reference's explicit loop unrun (lss_heston_euler_solver_scheme.hpp:150-155), so the result is
// then the zero surface, as there.
// -------------------------------------------------------------------------------------------------
namespace explicit_solvers {

using lss_boundary::boundary_2d_pair;
using lss_boundary::boundary_2d_ptr;
using lss_containers::container_2d;
using lss_enumerations::by_enum;
using lss_enumerations::grid_enum;
using lss_grids::grid_config_hints_2d_ptr;

template <typename fp_type, template <typename, typename> typename container = std::vector,
          typename allocator = std::allocator<fp_type>>
class general_heston_equation {
  private:
    heat_data_config_2d_ptr<fp_type> heat_data_cfg_;
    pde_discretization_config_2d_ptr<fp_type> discretization_cfg_;
    boundary_2d_ptr<fp_type> vertical_upper_;
    boundary_2d_pair<fp_type> horizontal_pair_;
    grid_config_hints_2d_ptr<fp_type> grid_hints_;
    heat_explicit_solver_config_ptr solver_cfg_;

  public:
    explicit general_heston_equation() = delete;
    explicit general_heston_equation(heat_data_config_2d_ptr<fp_type> const &heat_data_config,
                                     pde_discretization_config_2d_ptr<fp_type> const &discretization_config,
                                     boundary_2d_ptr<fp_type> const &vertical_upper_boundary_ptr,
                                     boundary_2d_pair<fp_type> const &horizontal_boundary_pair,
                                     grid_config_hints_2d_ptr<fp_type> const &grid_config_hints,
                                     heat_explicit_solver_config_ptr const &solver_config =
                                         default_heat_solver_configs::dev_expl_fwd_euler_solver_config_ptr)
        : heat_data_cfg_{heat_data_config}, discretization_cfg_{discretization_config},
          vertical_upper_{vertical_upper_boundary_ptr}, horizontal_pair_{horizontal_boundary_pair},
          grid_hints_{grid_config_hints}, solver_cfg_{solver_config}
    {
        LSS_VERIFY(heat_data_cfg_, "heat_data_config must not be null");
        LSS_VERIFY(discretization_cfg_, "discretization_config must not be null");
        LSS_VERIFY(solver_cfg_, "solver_config must not be null");
        LSS_VERIFY(grid_hints_, "grid_config_hints must not be null");
    }
    general_heston_equation(general_heston_equation const &) = delete;
    general_heston_equation(general_heston_equation &&) = delete;
    general_heston_equation &operator=(general_heston_equation const &) = delete;
    general_heston_equation &operator=(general_heston_equation &&) = delete;

    /**
     * Get the final solution of the PDE (heston_equation.hpp:430-503)
     *
     * \param solution - 2D container for the solution, nx rows (S) by ny columns (v)
     */
    void solve(container_2d<by_enum::Row, fp_type, container, allocator> &solution)
    {
        LSS_ASSERT((solution.rows()) > 0 && (solution.columns() > 0),
                   "The input solution container must be initialized");
        const auto np = discretization_cfg_->number_of_space_points();
        LSS_ASSERT((solution.columns() == np.second) && (solution.rows() == np.first),
                   "The input solution container must have the correct size");
        if (solver_cfg_->explicit_pde_scheme() != explicit_pde_schemes_enum::Euler)
            throw std::runtime_error("Not currently supported");
        if (heat_data_cfg_->source_data_config()) {
            for (std::size_t i = 0; i < np.first; ++i)
                for (std::size_t j = 0; j < np.second; ++j)
                    solution(i, j, fp_type{});
            return;
        }
        // the implicit class's sampling, with the explicit scheme selected
        lss_amd_detail::grid_spec<fp_type> g;
        const auto sr = discretization_cfg_->space_range();
        g.x_lo = sr.first.lower();
        g.x_hi = sr.first.upper();
        g.y_lo = sr.second.lower();
        g.y_hi = sr.second.upper();
        g.t_lo = discretization_cfg_->time_range().lower();
        g.t_hi = discretization_cfg_->time_range().upper();
        g.nx = np.first;
        g.ny = np.second;
        g.nt = discretization_cfg_->number_of_time_points();
        g.accumulation_point = grid_hints_->accumulation_point();
        g.alpha_scale = grid_hints_->alpha_scale();
        g.beta_scale = grid_hints_->beta_scale();
        g.nonuniform = grid_hints_->grid() == grid_enum::Nonuniform;
        g.splitting = LSS_EXPLICIT_EULER;
        g.weighting = fp_type(0.5);  // (unused by the explicit scheme)
        g.implicit_scheme = LSS_CRANK_NICOLSON;
        g.tridiagonal = LSS_THOMAS_LU;
        g.direction = solver_cfg_->traverse_direction() == traverse_direction_enum::Forward ? LSS_FORWARD
                                                                                          : LSS_BACKWARD;
        lss_amd_detail::user_data<fp_type> u;
        u.A = heat_data_cfg_->a_coefficient();
        u.B = heat_data_cfg_->b_coefficient();
        u.C = heat_data_cfg_->c_coefficient();
        u.D = heat_data_cfg_->d_coefficient();
        u.E = heat_data_cfg_->e_coefficient();
        u.F = heat_data_cfg_->f_coefficient();
        u.init = heat_data_cfg_->initial_condition();
        auto lo = horizontal_pair_.first, hi = horizontal_pair_.second, vu = vertical_upper_;
        u.h_lower = [lo](fp_type t, fp_type y) { return lo->value(t, y); };
        u.h_upper = [hi](fp_type t, fp_type y) { return hi->value(t, y); };
        u.v_upper = [vu](fp_type t, fp_type x) { return vu->value(t, x); };
        lss_amd_detail::sampled<fp_type> s;
        lss_amd_detail::sample_fields(u, g, s, lss_amd_detail::TIME_AUTO, lss_amd_detail::TIME_AUTO);
        const lss_adi_desc d = s.desc(LSS_COEFF_FIELDS, LSS_LINE_EXACT);
        lss_adi_plan *plan = nullptr;
        implicit_solvers::detail::check(lss_adi_plan_create(&d, 1, &plan), "lss_adi_plan_create");
        const void *in[1] = {s.init.data()};
        void *out[1] = {solution.raw()};
        const int rc = lss_adi_run_host(plan, in, out);
        lss_adi_plan_destroy(plan);
        implicit_solvers::detail::check(rc, "lss_adi_run_host");
    }
};

template <typename fp_type, template <typename, typename> typename container = std::vector,
          typename allocator = std::allocator<fp_type>>
using general_heston_equation_ptr = lss_utility::sptr_t<general_heston_equation<fp_type, container, allocator>>;

}  // namespace explicit_solvers
}  // namespace two_dimensional
}  // namespace lss_pde_solvers
