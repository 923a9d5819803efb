// heat_equation_1d.hpp -- drop-in one_dimensional::implicit_solvers::general_heat_equation backed by the
// MI355X library, with the reference's 1-D heat data configuration types.
//
// Same classes, constructors and solve(container&) / solve(container_2d<Row>&) as the reference
// (pde_solvers/lss_heat_data_config.hpp:27-340 for the data, pde_solvers/one_dimensional/heat_type/
// lss_1d_general_heat_equation.hpp:64-528 for the solver).  solve() samples the user's closures on the
// reference grid (lss_heat1d_sampler.hpp), builds a one-problem plan through the C-ABI
// (include/lss_heat1d.h) and runs the whole time loop on the GPU in the operation order of the selected
// host solver (DoubleSweep -> double sweep, every other choice -> Thomas LU), so the result has the
// bits of the reference's host solver.  A heat source is honoured, as the reference's 1-D kernels do
// (heat_time_loop::run with the source, lss_1d_general_heat_equation_implicit_kernel.hpp:83-122).
// The reference's 1-D implicit kernel takes the Euler and Crank-Nicolson schemes only and throws
// "Unreachable" for the others; so does this one.  Errors of the GPU library are thrown as
// std::runtime_error; size mismatches terminate through LSS_ASSERT as in the reference.
#pragma once

#include <map>
#include <stdexcept>
#include <string>

#include "boundaries.hpp"
#include "configs.hpp"
#include "containers.hpp"
#include "lss_heat1d.h"
#include "lss_heat1d_sampler.hpp"

namespace lss_pde_solvers {

// ---------------------------------------------------------------------------------------------
// heat data (1-D part, lss_heat_data_config.hpp:27-340)
// ---------------------------------------------------------------------------------------------
template <typename fp_type> struct heat_coefficient_data_config<dimension_enum::One, fp_type> {
  private:
    std::function<fp_type(fp_type, fp_type)> a_coeff_, b_coeff_, c_coeff_;

  public:
    explicit heat_coefficient_data_config(std::function<fp_type(fp_type, fp_type)> const &a_coefficient,
                                          std::function<fp_type(fp_type, fp_type)> const &b_coefficient,
                                          std::function<fp_type(fp_type, fp_type)> const &c_coefficient)
        : a_coeff_{a_coefficient}, b_coeff_{b_coefficient}, c_coeff_{c_coefficient}
    {
        LSS_VERIFY(a_coeff_, "a_coefficient must not be null");
        LSS_VERIFY(b_coeff_, "b_coefficient must not be null");
        LSS_VERIFY(c_coeff_, "c_coefficient must not be null");
    }
    std::function<fp_type(fp_type, fp_type)> const &a_coefficient() const { return a_coeff_; }
    std::function<fp_type(fp_type, fp_type)> const &b_coefficient() const { return b_coeff_; }
    std::function<fp_type(fp_type, fp_type)> const &c_coefficient() const { return c_coeff_; }
};

template <typename fp_type> struct heat_initial_data_config<dimension_enum::One, fp_type> {
  private:
    std::function<fp_type(fp_type)> initial_condition_;

  public:
    explicit heat_initial_data_config(std::function<fp_type(fp_type)> const &initial_condition)
        : initial_condition_{initial_condition}
    {
        LSS_VERIFY(initial_condition_, "initial_condition must not be null");
    }
    std::function<fp_type(fp_type)> const &initial_condition() const { return initial_condition_; }
};

template <typename fp_type> struct heat_source_data_config<dimension_enum::One, fp_type> {
  private:
    std::function<fp_type(fp_type, fp_type)> heat_source_;

  public:
    explicit heat_source_data_config(std::function<fp_type(fp_type, fp_type)> const &heat_source)
        : heat_source_{heat_source}
    {
        LSS_VERIFY(heat_source_, "heat_source must not be null");
    }
    std::function<fp_type(fp_type, fp_type)> const &heat_source() const { return heat_source_; }
};

template <typename fp_type>
using heat_coefficient_data_config_1d = heat_coefficient_data_config<dimension_enum::One, fp_type>;
template <typename fp_type> using heat_initial_data_config_1d = heat_initial_data_config<dimension_enum::One, fp_type>;
template <typename fp_type> using heat_source_data_config_1d = heat_source_data_config<dimension_enum::One, fp_type>;
template <typename fp_type>
using heat_coefficient_data_config_1d_ptr = sptr_t<heat_coefficient_data_config_1d<fp_type>>;
template <typename fp_type> using heat_initial_data_config_1d_ptr = sptr_t<heat_initial_data_config_1d<fp_type>>;
template <typename fp_type> using heat_source_data_config_1d_ptr = sptr_t<heat_source_data_config_1d<fp_type>>;

template <typename fp_type> struct heat_data_config<dimension_enum::One, fp_type> {
  private:
    heat_coefficient_data_config_1d_ptr<fp_type> coefficient_data_cfg_;
    heat_initial_data_config_1d_ptr<fp_type> initial_data_cfg_;
    heat_source_data_config_1d_ptr<fp_type> source_data_cfg_;

  public:
    explicit heat_data_config(heat_coefficient_data_config_1d_ptr<fp_type> const &coefficient_data_config,
                              heat_initial_data_config_1d_ptr<fp_type> const &initial_data_config,
                              heat_source_data_config_1d_ptr<fp_type> const &source_data_config = nullptr)
        : coefficient_data_cfg_{coefficient_data_config}, initial_data_cfg_{initial_data_config},
          source_data_cfg_{source_data_config}
    {
        LSS_VERIFY(coefficient_data_cfg_, "coefficient_data_config must not be null");
        LSS_VERIFY(initial_data_cfg_, "initial_data_config must not be null");
    }
    heat_source_data_config_1d_ptr<fp_type> const &source_data_config() const { return source_data_cfg_; }
    std::function<fp_type(fp_type)> const &initial_condition() const { return initial_data_cfg_->initial_condition(); }
    std::function<fp_type(fp_type, fp_type)> const &a_coefficient() const
    {
        return coefficient_data_cfg_->a_coefficient();
    }
    std::function<fp_type(fp_type, fp_type)> const &b_coefficient() const
    {
        return coefficient_data_cfg_->b_coefficient();
    }
    std::function<fp_type(fp_type, fp_type)> const &c_coefficient() const
    {
        return coefficient_data_cfg_->c_coefficient();
    }
};

template <typename fp_type> using heat_data_config_1d = heat_data_config<dimension_enum::One, fp_type>;
template <typename fp_type> using heat_data_config_1d_ptr = sptr_t<heat_data_config_1d<fp_type>>;

namespace one_dimensional {
namespace implicit_solvers {

using lss_boundary::boundary_1d_pair;
using lss_boundary::boundary_1d_ptr;
using lss_containers::container_2d;
using lss_enumerations::by_enum;
using lss_enumerations::grid_enum;
using lss_grids::grid_config_hints_1d_ptr;

namespace detail {
inline void check1(int rc, const char *what)
{
    if (rc != LSS_OK)
        throw std::runtime_error(std::string(what) + ": " + lss_last_error());
}

// boundary type and closures of one end (the transform's X' scaling is applied by the sampler)
template <typename fp_type>
void end_of(boundary_1d_ptr<fp_type> const &b, int &type, std::function<fp_type(fp_type)> &val,
            std::function<fp_type(fp_type)> &lin)
{
    using lss_boundary::dirichlet_boundary_1d;
    using lss_boundary::neumann_boundary_1d;
    using lss_boundary::robin_boundary_1d;
    if (auto const &d = std::dynamic_pointer_cast<dirichlet_boundary_1d<fp_type>>(b)) {
        type = LSS_BC_DIRICHLET;
        val = [d](fp_type t) { return d->value(t); };
    } else if (auto const &nb = std::dynamic_pointer_cast<neumann_boundary_1d<fp_type>>(b)) {
        type = LSS_BC_NEUMANN;
        val = [nb](fp_type t) { return nb->value(t); };
    } else if (auto const &r = std::dynamic_pointer_cast<robin_boundary_1d<fp_type>>(b)) {
        type = LSS_BC_ROBIN;
        val = [r](fp_type t) { return r->value(t); };
        lin = [r](fp_type t) { return r->linear_value(t); };
    } else {
        throw std::runtime_error("Unreachable");
    }
}
}  // namespace detail

template <typename fp_type, template <typename, typename> typename container = std::vector,
          typename allocator = std::allocator<fp_type>>
class general_heat_equation {
  private:
    heat_data_config_1d_ptr<fp_type> heat_data_cfg_;
    pde_discretization_config_1d_ptr<fp_type> discretization_cfg_;
    boundary_1d_pair<fp_type> boundary_pair_;
    grid_config_hints_1d_ptr<fp_type> grid_hints_;
    heat_implicit_solver_config_ptr solver_cfg_;
    std::map<std::string, fp_type> solver_config_details_;

    void initialize()
    {
        LSS_VERIFY(heat_data_cfg_, "heat_data_config must not be null");
        LSS_VERIFY(discretization_cfg_, "discretization_config must not be null");
        LSS_VERIFY(std::get<0>(boundary_pair_), "boundary_pair.first must not be null");
        LSS_VERIFY(std::get<1>(boundary_pair_), "boundary_pair.second must not be null");
        LSS_VERIFY(solver_cfg_, "solver_config must not be null");
        LSS_VERIFY(grid_hints_, "grid_config_hints must not be null");
        if (!solver_config_details_.empty()) {
            auto const &it = solver_config_details_.find("sor_omega");
            LSS_ASSERT(it != solver_config_details_.end(), "sor_omega is not defined");
        }
    }

    lss_amd_detail::heat1d_spec<fp_type> spec() const
    {
        lss_amd_detail::heat1d_spec<fp_type> g;
        g.x_lo = discretization_cfg_->space_range().lower();
        g.x_hi = discretization_cfg_->space_range().upper();
        g.t_lo = discretization_cfg_->time_range().lower();
        g.t_hi = discretization_cfg_->time_range().upper();
        g.n = discretization_cfg_->number_of_space_points();
        g.nt = discretization_cfg_->number_of_time_points();
        g.accumulation_point = grid_hints_->accumulation_point();
        g.alpha_scale = grid_hints_->alpha_scale();
        g.nonuniform = grid_hints_->grid() == grid_enum::Nonuniform;
        g.implicit_scheme = static_cast<int>(solver_cfg_->implicit_pde_scheme());
        g.tridiagonal = solver_cfg_->tridiagonal_method() == tridiagonal_method_enum::DoubleSweepSolver
                            ? LSS_DOUBLE_SWEEP
                            : LSS_THOMAS_LU;
        g.direction = solver_cfg_->traverse_direction() == traverse_direction_enum::Forward ? LSS_FORWARD
                                                                                          : LSS_BACKWARD;
        return g;
    }

    lss_amd_detail::heat1d_data<fp_type> data() const
    {
        lss_amd_detail::heat1d_data<fp_type> u;
        u.a = heat_data_cfg_->a_coefficient();
        u.b = heat_data_cfg_->b_coefficient();
        u.c = heat_data_cfg_->c_coefficient();
        u.init = heat_data_cfg_->initial_condition();
        if (heat_data_cfg_->source_data_config())
            u.src = heat_data_cfg_->source_data_config()->heat_source();
        detail::end_of(boundary_pair_.first, u.lo_type, u.lo_val, u.lo_lin);
        detail::end_of(boundary_pair_.second, u.hi_type, u.hi_val, u.hi_lin);
        return u;
    }

    // samples the problem and runs it on the GPU; surfaces: nullptr or (steps + 1) * n values
    void run(fp_type *final_level, fp_type *surfaces) const
    {
        lss_amd_detail::heat1d_sampled<fp_type> s;
        lss_amd_detail::sample_heat1d(data(), spec(), s, lss_amd_detail::TIME_AUTO, lss_amd_detail::TIME_AUTO);
        const lss_heat1d_desc d = s.desc(LSS_LINE_EXACT);
        lss_heat1d_plan *plan = nullptr;
        detail::check1(lss_heat1d_plan_create(&d, 1, &plan), "lss_heat1d_plan_create");
        const void *in[1] = {s.init.data()};
        void *out[1] = {final_level};
        void *sf[1] = {surfaces};
        const int rc = lss_heat1d_run_host(plan, in, out, surfaces ? sf : nullptr);
        lss_heat1d_plan_destroy(plan);
        detail::check1(rc, "lss_heat1d_run_host");
    }

  public:
    explicit general_heat_equation() = delete;
    explicit general_heat_equation(
        heat_data_config_1d_ptr<fp_type> const &heat_data_config,
        pde_discretization_config_1d_ptr<fp_type> const &discretization_config,
        boundary_1d_pair<fp_type> const &boundary_pair, grid_config_hints_1d_ptr<fp_type> const &grid_config_hints,
        heat_implicit_solver_config_ptr const &solver_config =
            default_heat_solver_configs::host_fwd_dssolver_euler_solver_config_ptr,
        std::map<std::string, fp_type> const &solver_config_details = std::map<std::string, fp_type>())
        : heat_data_cfg_{heat_data_config}, discretization_cfg_{discretization_config}, boundary_pair_{boundary_pair},
          grid_hints_{grid_config_hints}, solver_cfg_{solver_config}, solver_config_details_{solver_config_details}
    {
        initialize();
    }
    general_heat_equation(general_heat_equation const &) = delete;
    general_heat_equation(general_heat_equation &&) = delete;
    general_heat_equation &operator=(general_heat_equation const &) = delete;
    general_heat_equation &operator=(general_heat_equation &&) = delete;

    /**
     * Get the final solution of the PDE (lss_1d_general_heat_equation.hpp:134-245)
     *
     * \param solution - container for the solution, number_of_space_points values
     */
    void solve(container<fp_type, allocator> &solution)
    {
        LSS_ASSERT(solution.size() > 0, "The input solution container must be initialized");
        const std::size_t space_size = discretization_cfg_->number_of_space_points();
        LSS_ASSERT(solution.size() == space_size, "The input solution container must have the correct size");
        std::vector<fp_type> out(space_size);
        run(out.data(), nullptr);
        std::copy(out.begin(), out.end(), solution.begin());
    }

    /**
     * Get all solutions in time (surface) of the PDE (lss_1d_general_heat_equation.hpp:312-528)
     *
     * \param solutions - 2D container, number_of_time_points rows by number_of_space_points columns
     */
    void solve(container_2d<by_enum::Row, fp_type, container, allocator> &solutions)
    {
        LSS_ASSERT(((solutions.columns() > 0) && (solutions.rows() > 0)),
                   "The input solution 2D container must be initialized");
        const std::size_t space_size = discretization_cfg_->number_of_space_points();
        const std::size_t time_size = discretization_cfg_->number_of_time_points();
        LSS_ASSERT((solutions.rows() == time_size) && (solutions.columns() == space_size),
                   "The input solution 2D container must have the correct size");
        std::vector<fp_type> out(space_size);
        run(out.data(), solutions.raw());
    }
};

template <typename fp_type, template <typename, typename> typename container = std::vector,
          typename allocator = std::allocator<fp_type>>
using general_heat_equation_ptr = lss_utility::sptr_t<general_heat_equation<fp_type, container, allocator>>;

}  // namespace implicit_solvers
}  // namespace one_dimensional
}  // namespace lss_pde_solvers
