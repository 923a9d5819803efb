// heston_adjoint.hpp -- C++ front ends of the batched book and of the hand-written adjoint (no
// counterpart in the reference, whose solvers step one problem at a time and have no derivatives).
//
//   batched_heston_book    many general_heston_equation problems, as the reference's builders make them
//                          (builders/lss_2d_general_heston_equation_builder.hpp:57-128), stepped together
//                          in the same launches: one plan, one solve for the whole book.
//   heston_adjoint         the Heston call of the reference's tests (unit_tests/common/lss_print_t.hpp:
//                          2156-2186) with an autograd-Function-shaped pair: forward(u_in) -> u_out keeps
//                          checkpoints, backward(dJ/du_out) -> dJ/du_in and dJ/d(kappa, theta_v, sigma,
//                          rho, r).
// Both sit on the C-ABI of include/lss_amd.h; errors are thrown as std::runtime_error.
#pragma once

#include <array>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "heston_equation.hpp"

namespace lss_pde_solvers {
namespace two_dimensional {
namespace implicit_solvers {

// -------------------------------------------------------------------------------------------------
// batched_heston_book
// -------------------------------------------------------------------------------------------------
template <typename fp_type, template <typename, typename> typename container = std::vector,
          typename allocator = std::allocator<fp_type>>
class batched_heston_book {
  private:
    lss_adi_plan *plan_ = nullptr;
    std::vector<std::size_t> nx_, ny_;
    std::vector<std::vector<fp_type>> init_;

  public:
    typedef general_heston_equation<fp_type, container, allocator> equation_t;
    typedef container_2d<by_enum::Row, fp_type, container, allocator> surface_t;

    // Samples every equation's closures on its own grid (coefficient fields in the reference's
    // rounding).  exact_lines: each line in the host solver's sequential order (the reference's bits);
    // otherwise the partitioned line solver (lines of at most 1026 points).
    explicit batched_heston_book(std::vector<std::shared_ptr<equation_t>> const &equations, bool exact_lines = false)
    {
        LSS_VERIFY(!equations.empty(), "the book must hold at least one equation");
        std::vector<lss_amd_detail::sampled<fp_type>> s(equations.size());
        std::vector<lss_adi_desc> d(equations.size());
        for (std::size_t p = 0; p < equations.size(); ++p) {
            LSS_VERIFY(equations[p], "equation must not be null");
            const auto g = equations[p]->grid();
            lss_amd_detail::sample_fields(equations[p]->data(), g, s[p], lss_amd_detail::TIME_AUTO,
                                          lss_amd_detail::TIME_AUTO);
            d[p] = s[p].desc(LSS_COEFF_FIELDS, exact_lines ? LSS_LINE_EXACT : LSS_LINE_PARTITIONED);
            nx_.push_back(g.nx);
            ny_.push_back(g.ny);
            init_.push_back(std::move(s[p].init));
        }
        detail::check(lss_adi_plan_create(d.data(), (int64_t)d.size(), &plan_), "lss_adi_plan_create");
    }
    ~batched_heston_book()
    {
        if (plan_)
            lss_adi_plan_destroy(plan_);
    }
    batched_heston_book(batched_heston_book const &) = delete;
    batched_heston_book &operator=(batched_heston_book const &) = delete;

    std::size_t size() const { return nx_.size(); }

    // The final surface of every problem (solutions[p] pre-sized nx_p x ny_p, as general_heston_equation::solve
    // requires), all problems in one solve.
    void solve(std::vector<surface_t> &solutions)
    {
        LSS_ASSERT(solutions.size() == size(), "one solution container per problem");
        std::vector<const void *> in(size());
        std::vector<void *> out(size());
        for (std::size_t p = 0; p < size(); ++p) {
            LSS_ASSERT(solutions[p].rows() == nx_[p] && solutions[p].columns() == ny_[p],
                       "The input solution container must have the correct size");
            in[p] = init_[p].data();
            out[p] = solutions[p].raw();
        }
        detail::check(lss_adi_run_host(plan_, in.data(), out.data()), "lss_adi_run_host");
    }
};

// -------------------------------------------------------------------------------------------------
// heston_adjoint
// -------------------------------------------------------------------------------------------------
// The Heston call data: a = v S^2 / 2, b = sigma^2 v / 2, c = rho sigma v S, d = r S, e = kappa (theta - v),
// f = -r, payoff max(S - K, 0); Dirichlet lower_value at S_min, Neumann upper_value at S_max (-1: dU/dS =
// 1), Dirichlet U = S at v_max.
template <typename fp_type> struct heston_parameters {
    fp_type strike, rate, sigma, kappa, theta, rho;
    fp_type lower_value = fp_type(0.0);
    fp_type upper_value = fp_type(-1.0);
};

template <typename fp_type, template <typename, typename> typename container = std::vector,
          typename allocator = std::allocator<fp_type>>
class heston_adjoint {
  public:
    typedef container_2d<by_enum::Row, fp_type, container, allocator> surface_t;
    // dJ/du_in and dJ/dparam for param = kappa, theta_v, sigma, rho, r
    struct gradient {
        surface_t d_initial;
        std::array<double, 5> d_params;
        double d_kappa() const { return d_params[0]; }
        double d_theta() const { return d_params[1]; }
        double d_sigma() const { return d_params[2]; }
        double d_rho() const { return d_params[3]; }
        double d_rate() const { return d_params[4]; }
    };

  private:
    lss_adi_plan *plan_ = nullptr;
    std::size_t nx_ = 0, ny_ = 0;
    int ckpt_ = 32;
    bool forwarded_ = false;

  public:
    heston_adjoint(heston_parameters<fp_type> const &params,
                   pde_discretization_config_2d_ptr<fp_type> const &discretization_config,
                   splitting_method_config_ptr<fp_type> const &splitting_method_config,
                   grid_config_hints_2d_ptr<fp_type> const &grid_config_hints,
                   heat_implicit_solver_config_ptr const &solver_config =
                       default_heat_solver_configs::host_bwd_tlusolver_cn_solver_config_ptr,
                   int checkpoint_every = 32)
        : ckpt_{checkpoint_every}
    {
        LSS_VERIFY(discretization_config, "discretization_config must not be null");
        LSS_VERIFY(splitting_method_config, "splitting_method_config must not be null");
        LSS_VERIFY(grid_config_hints, "grid_config_hints must not be null");
        LSS_VERIFY(solver_config, "solver_config must not be null");
        lss_model m;
        std::memset(&m, 0, sizeof(m));
        m.kind = LSS_MODEL_HESTON;
        m.strike = params.strike;
        m.rate = params.rate;
        m.sig_sig = params.sigma;
        m.sig_kappa = params.kappa;
        m.sig_theta = params.theta;
        m.rho = params.rho;
        m.bnd_low = params.lower_value;
        m.bnd_high = params.upper_value;
        const auto sr = discretization_config->space_range();
        const auto np = discretization_config->number_of_space_points();
        m.x_lo = sr.first.lower();
        m.x_hi = sr.first.upper();
        m.y_lo = sr.second.lower();
        m.y_hi = sr.second.upper();
        m.t_lo = discretization_config->time_range().lower();
        m.t_hi = discretization_config->time_range().upper();
        m.maturity = m.t_hi;
        m.nx = (int64_t)np.first;
        m.ny = (int64_t)np.second;
        m.nt = (int64_t)discretization_config->number_of_time_points();
        m.accumulation_point = grid_config_hints->accumulation_point();
        m.alpha_scale = grid_config_hints->alpha_scale();
        m.beta_scale = grid_config_hints->beta_scale();
        m.grid = grid_config_hints->grid() == grid_enum::Nonuniform ? LSS_GRID_NONUNIFORM : LSS_GRID_UNIFORM;
        m.splitting = static_cast<int32_t>(splitting_method_config->splitting_method());
        m.weighting = splitting_method_config->weighting_value();
        m.implicit_scheme = static_cast<int32_t>(solver_config->implicit_pde_scheme());
        m.tridiagonal = LSS_THOMAS_LU;
        m.direction = solver_config->traverse_direction() == traverse_direction_enum::Forward ? LSS_FORWARD
                                                                                            : LSS_BACKWARD;
        m.coeff_mode = LSS_COEFF_SEPARABLE;  // the adjoint differentiates the separable coefficient factors
        m.line_solver = LSS_LINE_PARTITIONED;
        nx_ = np.first;
        ny_ = np.second;
        detail::check(lss_model_plan_create(&m, 1, sizeof(fp_type) == 8 ? LSS_FP64 : LSS_FP32, &plan_),
                      "lss_model_plan_create");
    }
    ~heston_adjoint()
    {
        if (plan_)
            lss_adi_plan_destroy(plan_);
    }
    heston_adjoint(heston_adjoint const &) = delete;
    heston_adjoint &operator=(heston_adjoint const &) = delete;

    // the sampled payoff on the reference grid
    surface_t initial() const
    {
        surface_t u(nx_, ny_);
        detail::check(lss_model_initial(plan_, 0, u.raw()), "lss_model_initial");
        return u;
    }
    // physical grid: S (or F) values X(zeta_i) and v values Y(eta_j)
    std::pair<std::vector<fp_type>, std::vector<fp_type>> grid() const
    {
        std::vector<fp_type> xs(nx_), ys(ny_);
        detail::check(lss_model_grid(plan_, 0, xs.data(), ys.data()), "lss_model_grid");
        return {xs, ys};
    }

    // u_out = the time loop applied to u_in; keeps every checkpoint_every-th level for backward()
    void forward(surface_t const &u_in, surface_t &u_out)
    {
        LSS_ASSERT(u_in.rows() == nx_ && u_in.columns() == ny_ && u_out.rows() == nx_ && u_out.columns() == ny_,
                   "surfaces must be nx x ny");
        const void *in[1] = {u_in.raw()};
        void *out[1] = {u_out.raw()};
        detail::check(lss_adi_forward_host(plan_, in, out, ckpt_), "lss_adi_forward_host");
        forwarded_ = true;
    }

    // gradient of J from dJ/du_out (the adjoint of the last forward())
    gradient backward(surface_t const &dJ_du_out)
    {
        if (!forwarded_)
            throw std::runtime_error("heston_adjoint::backward needs a preceding forward()");
        LSS_ASSERT(dJ_du_out.rows() == nx_ && dJ_du_out.columns() == ny_, "surfaces must be nx x ny");
        gradient g{surface_t(nx_, ny_), {}};
        const void *lo[1] = {dJ_du_out.raw()};
        void *li[1] = {g.d_initial.raw()};
        double dp[LSS_MAX_PARAMS] = {};
        detail::check(lss_adi_backward_host(plan_, lo, li, dp), "lss_adi_backward_host");
        for (int k = 0; k < 5; ++k)
            g.d_params[k] = dp[k];
        return g;
    }
};

}  // namespace implicit_solvers
}  // namespace two_dimensional
}  // namespace lss_pde_solvers
