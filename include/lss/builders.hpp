// builders.hpp -- fluent builders of the drop-in API, named and shaped as the reference's builders/
// (lss_range_builder.hpp, lss_{dirichlet,neumann,robin}_boundary_builder.hpp,
// lss_grid_config_hints_builder.hpp, lss_heat_data_config_builder.hpp,
// lss_pde_discretization_config_builder.hpp, lss_splitting_method_config_builder.hpp:26-49,
// lss_heat_solver_config_builder.hpp, lss_2d_general_heston_equation_builder.hpp:57-128).
// Difference: splitting_method_config_builder starts from DouglasRachford / 0.5 instead of leaving
// its fields uninitialised when a setter is not called.
#pragma once

#include "boundaries.hpp"
#include "configs.hpp"
#include "heston_equation.hpp"

namespace lss_utility {
template <typename fp_type> struct range_builder {
  private:
    fp_type l_{}, u_{};

  public:
    range_builder &lower(fp_type value)
    {
        l_ = value;
        return *this;
    }
    range_builder &upper(fp_type value)
    {
        u_ = value;
        return *this;
    }
    range_ptr<fp_type> build() { return std::make_shared<range<fp_type>>(l_, u_); }
};
}  // namespace lss_utility

namespace lss_boundary {
template <typename fp_type> struct dirichlet_boundary_2d_builder {
  private:
    std::function<fp_type(fp_type, fp_type)> value_;

  public:
    dirichlet_boundary_2d_builder &value(const std::function<fp_type(fp_type, fp_type)> &value)
    {
        value_ = value;
        return *this;
    }
    dirichlet_boundary_2d_ptr<fp_type> build() { return std::make_shared<dirichlet_boundary_2d<fp_type>>(value_); }
};
template <typename fp_type> struct neumann_boundary_2d_builder {
  private:
    std::function<fp_type(fp_type, fp_type)> value_;

  public:
    neumann_boundary_2d_builder &value(const std::function<fp_type(fp_type, fp_type)> &value)
    {
        value_ = value;
        return *this;
    }
    neumann_boundary_2d_ptr<fp_type> build() { return std::make_shared<neumann_boundary_2d<fp_type>>(value_); }
};
template <typename fp_type> struct robin_boundary_2d_builder {
  private:
    std::function<fp_type(fp_type, fp_type)> linear_value_, value_;

  public:
    robin_boundary_2d_builder &linear_value(const std::function<fp_type(fp_type, fp_type)> &linear_value)
    {
        linear_value_ = linear_value;
        return *this;
    }
    robin_boundary_2d_builder &value(const std::function<fp_type(fp_type, fp_type)> &value)
    {
        value_ = value;
        return *this;
    }
    robin_boundary_2d_ptr<fp_type> build()
    {
        return std::make_shared<robin_boundary_2d<fp_type>>(linear_value_, value_);
    }
};
}  // namespace lss_boundary

namespace lss_grids {
template <dimension_enum dimension, typename fp_type> struct grid_config_hints_builder;
template <typename fp_type> struct grid_config_hints_builder<dimension_enum::Two, fp_type> {
  private:
    fp_type accumulation_point_{fp_type(0.0)}, alpha_scale_{fp_type(3.0)}, beta_scale_{fp_type(50.0)};
    grid_enum grid_{grid_enum::Uniform};

  public:
    grid_config_hints_builder &accumulation_point(fp_type v)
    {
        accumulation_point_ = v;
        return *this;
    }
    grid_config_hints_builder &alpha_scale(fp_type v)
    {
        alpha_scale_ = v;
        return *this;
    }
    grid_config_hints_builder &beta_scale(fp_type v)
    {
        beta_scale_ = v;
        return *this;
    }
    grid_config_hints_builder &grid(grid_enum g)
    {
        grid_ = g;
        return *this;
    }
    grid_config_hints_2d_ptr<fp_type> build()
    {
        return std::make_shared<grid_config_hints_2d<fp_type>>(accumulation_point_, alpha_scale_, beta_scale_, grid_);
    }
};
template <typename fp_type> using grid_config_hints_2d_builder = grid_config_hints_builder<dimension_enum::Two, fp_type>;
}  // namespace lss_grids

namespace lss_pde_solvers {

template <dimension_enum dimension, typename fp_type> struct heat_coefficient_data_config_builder;
template <typename fp_type> struct heat_coefficient_data_config_builder<dimension_enum::Two, fp_type> {
  private:
    std::function<fp_type(fp_type, fp_type, fp_type)> a_, b_, c_, d_, e_, f_;

  public:
#define LSS_COEF_SETTER(name, member)                                                                   \
    heat_coefficient_data_config_builder &name(std::function<fp_type(fp_type, fp_type, fp_type)> const &f) \
    {                                                                                                   \
        member = f;                                                                                     \
        return *this;                                                                                   \
    }
    LSS_COEF_SETTER(a_coefficient, a_)
    LSS_COEF_SETTER(b_coefficient, b_)
    LSS_COEF_SETTER(c_coefficient, c_)
    LSS_COEF_SETTER(d_coefficient, d_)
    LSS_COEF_SETTER(e_coefficient, e_)
    LSS_COEF_SETTER(f_coefficient, f_)
#undef LSS_COEF_SETTER
    heat_coefficient_data_config_2d_ptr<fp_type> build()
    {
        return std::make_shared<heat_coefficient_data_config_2d<fp_type>>(a_, b_, c_, d_, e_, f_);
    }
};
template <typename fp_type>
using heat_coefficient_data_config_2d_builder = heat_coefficient_data_config_builder<dimension_enum::Two, fp_type>;

template <dimension_enum dimension, typename fp_type> struct heat_initial_data_config_builder;
template <typename fp_type> struct heat_initial_data_config_builder<dimension_enum::Two, fp_type> {
  private:
    std::function<fp_type(fp_type, fp_type)> initial_condition_;

  public:
    heat_initial_data_config_builder &initial_condition(std::function<fp_type(fp_type, fp_type)> const &f)
    {
        initial_condition_ = f;
        return *this;
    }
    heat_initial_data_config_2d_ptr<fp_type> build()
    {
        return std::make_shared<heat_initial_data_config_2d<fp_type>>(initial_condition_);
    }
};
template <typename fp_type>
using heat_initial_data_config_2d_builder = heat_initial_data_config_builder<dimension_enum::Two, fp_type>;

template <dimension_enum dimension, typename fp_type> struct heat_source_data_config_builder;
template <typename fp_type> struct heat_source_data_config_builder<dimension_enum::Two, fp_type> {
  private:
    std::function<fp_type(fp_type, fp_type, fp_type)> heat_source_;

  public:
    heat_source_data_config_builder &heat_source(std::function<fp_type(fp_type, fp_type, fp_type)> const &f)
    {
        heat_source_ = f;
        return *this;
    }
    heat_source_data_config_2d_ptr<fp_type> build()
    {
        return std::make_shared<heat_source_data_config_2d<fp_type>>(heat_source_);
    }
};
template <typename fp_type>
using heat_source_data_config_2d_builder = heat_source_data_config_builder<dimension_enum::Two, fp_type>;

template <dimension_enum dimension, typename fp_type> struct heat_data_config_builder;
template <typename fp_type> struct heat_data_config_builder<dimension_enum::Two, fp_type> {
  private:
    heat_coefficient_data_config_2d_ptr<fp_type> coefficient_data_config_;
    heat_initial_data_config_2d_ptr<fp_type> initial_data_config_;
    heat_source_data_config_2d_ptr<fp_type> source_data_config_;

  public:
    heat_data_config_builder &coefficient_data_config(heat_coefficient_data_config_2d_ptr<fp_type> const &c)
    {
        coefficient_data_config_ = c;
        return *this;
    }
    heat_data_config_builder &initial_data_config(heat_initial_data_config_2d_ptr<fp_type> const &c)
    {
        initial_data_config_ = c;
        return *this;
    }
    // the reference names the source setter initial_data_config too (lss_heat_data_config_builder.hpp:329)
    heat_data_config_builder &initial_data_config(heat_source_data_config_2d_ptr<fp_type> const &c)
    {
        source_data_config_ = c;
        return *this;
    }
    heat_data_config_builder &source_data_config(heat_source_data_config_2d_ptr<fp_type> const &c)
    {
        source_data_config_ = c;
        return *this;
    }
    heat_data_config_2d_ptr<fp_type> build()
    {
        return std::make_shared<heat_data_config_2d<fp_type>>(coefficient_data_config_, initial_data_config_,
                                                              source_data_config_);
    }
};
template <typename fp_type> using heat_data_config_2d_builder = heat_data_config_builder<dimension_enum::Two, fp_type>;

template <dimension_enum dimension, typename fp_type> struct pde_discretization_config_builder;
template <typename fp_type> struct pde_discretization_config_builder<dimension_enum::Two, fp_type> {
  private:
    range<fp_type> space_range_1_, space_range_2_, time_range_;
    std::size_t number_of_space_points_1_ = 0, number_of_space_points_2_ = 0, number_of_time_points_ = 0;

  public:
    pde_discretization_config_builder &space_range_1(range<fp_type> const &r)
    {
        space_range_1_ = r;
        return *this;
    }
    pde_discretization_config_builder &space_range_2(range<fp_type> const &r)
    {
        space_range_2_ = r;
        return *this;
    }
    pde_discretization_config_builder &number_of_space_points_1(std::size_t const &n)
    {
        number_of_space_points_1_ = n;
        return *this;
    }
    pde_discretization_config_builder &number_of_space_points_2(std::size_t const &n)
    {
        number_of_space_points_2_ = n;
        return *this;
    }
    pde_discretization_config_builder &time_range(range<fp_type> const &r)
    {
        time_range_ = r;
        return *this;
    }
    pde_discretization_config_builder &number_of_time_points(std::size_t const &n)
    {
        number_of_time_points_ = n;
        return *this;
    }
    pde_discretization_config_2d_ptr<fp_type> build()
    {
        return std::make_shared<pde_discretization_config_2d<fp_type>>(space_range_1_, space_range_2_,
                                                                       number_of_space_points_1_,
                                                                       number_of_space_points_2_, time_range_,
                                                                       number_of_time_points_);
    }
};
template <typename fp_type>
using pde_discretization_config_2d_builder = pde_discretization_config_builder<dimension_enum::Two, fp_type>;

template <typename fp_type> struct splitting_method_config_builder {
  private:
    splitting_method_enum splitting_method_{splitting_method_enum::DouglasRachford};
    fp_type weighting_value_{fp_type(0.5)};

  public:
    splitting_method_config_builder &splitting_method(splitting_method_enum m)
    {
        splitting_method_ = m;
        return *this;
    }
    splitting_method_config_builder &weighting_value(fp_type v)
    {
        weighting_value_ = v;
        return *this;
    }
    splitting_method_config_ptr<fp_type> build()
    {
        return std::make_shared<splitting_method_config<fp_type>>(splitting_method_, weighting_value_);
    }
};

struct heat_implicit_solver_config_builder {
  private:
    memory_space_enum memory_space_{memory_space_enum::Device};
    traverse_direction_enum traverse_direction_{traverse_direction_enum::Backward};
    tridiagonal_method_enum tridiagonal_method_{tridiagonal_method_enum::CUDASolver};
    factorization_enum tridiagonal_factorization_{factorization_enum::None};
    implicit_pde_schemes_enum implicit_pde_scheme_{implicit_pde_schemes_enum::CrankNicolson};

  public:
    heat_implicit_solver_config_builder &memory_space(memory_space_enum v)
    {
        memory_space_ = v;
        return *this;
    }
    heat_implicit_solver_config_builder &traverse_direction(traverse_direction_enum v)
    {
        traverse_direction_ = v;
        return *this;
    }
    heat_implicit_solver_config_builder &tridiagonal_method(tridiagonal_method_enum v)
    {
        tridiagonal_method_ = v;
        return *this;
    }
    heat_implicit_solver_config_builder &tridiagonal_factorization(factorization_enum v)
    {
        tridiagonal_factorization_ = v;
        return *this;
    }
    heat_implicit_solver_config_builder &implicit_pde_scheme(implicit_pde_schemes_enum v)
    {
        implicit_pde_scheme_ = v;
        return *this;
    }
    heat_implicit_solver_config_ptr build()
    {
        return std::make_shared<heat_implicit_solver_config>(memory_space_, traverse_direction_, tridiagonal_method_,
                                                             tridiagonal_factorization_, implicit_pde_scheme_);
    }
};

namespace two_dimensional {
namespace implicit_solvers {

template <typename fp_type, template <typename, typename> typename container = std::vector,
          typename allocator = std::allocator<fp_type>>
class general_heston_equation_builder {
  private:
    heat_data_config_2d_ptr<fp_type> heat_data_config_;
    pde_discretization_config_2d_ptr<fp_type> discretization_config_;
    boundary_2d_ptr<fp_type> vertical_upper_boundary_;
    boundary_2d_pair<fp_type> horizontal_boundary_pair_;
    splitting_method_config_ptr<fp_type> splitting_method_config_;
    grid_config_hints_2d_ptr<fp_type> grid_hints_cfg_;
    heat_implicit_solver_config_ptr solver_config_;
    std::map<std::string, fp_type> solver_config_details_;

  public:
    general_heston_equation_builder &heat_data_config(const heat_data_config_2d_ptr<fp_type> &c)
    {
        heat_data_config_ = c;
        return *this;
    }
    general_heston_equation_builder &discretization_config(const pde_discretization_config_2d_ptr<fp_type> &c)
    {
        discretization_config_ = c;
        return *this;
    }
    general_heston_equation_builder &vertical_upper_boundary(const boundary_2d_ptr<fp_type> &b)
    {
        vertical_upper_boundary_ = b;
        return *this;
    }
    general_heston_equation_builder &horizontal_boundary(const boundary_2d_pair<fp_type> &b)
    {
        horizontal_boundary_pair_ = b;
        return *this;
    }
    general_heston_equation_builder &splitting_method_config(const splitting_method_config_ptr<fp_type> &c)
    {
        splitting_method_config_ = c;
        return *this;
    }
    general_heston_equation_builder &grid_hints(const grid_config_hints_2d_ptr<fp_type> &h)
    {
        grid_hints_cfg_ = h;
        return *this;
    }
    general_heston_equation_builder &solver_config(const heat_implicit_solver_config_ptr &c)
    {
        solver_config_ = c;
        return *this;
    }
    general_heston_equation_builder &solver_config_details(const std::map<std::string, fp_type> &d)
    {
        solver_config_details_ = d;
        return *this;
    }
    lss_utility::sptr_t<general_heston_equation<fp_type, container, allocator>> build()
    {
        return std::make_shared<general_heston_equation<fp_type, container, allocator>>(
            heat_data_config_, discretization_config_, vertical_upper_boundary_, horizontal_boundary_pair_,
            splitting_method_config_, grid_hints_cfg_, solver_config_, solver_config_details_);
    }
};

}  // namespace implicit_solvers
}  // namespace two_dimensional
}  // namespace lss_pde_solvers
