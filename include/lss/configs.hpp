// configs.hpp -- configuration objects of the drop-in API (the reference's pde_solvers/*.hpp and
// discretization/lss_grid_config_hints.hpp), with the same names, constructors, accessors and
// default objects, so reference user code builds against them unchanged:
//   grid_config_hints<One|Two>          discretization/lss_grid_config_hints.hpp:15-88
//   pde_discretization_config<One|Two>  pde_solvers/lss_pde_discretization_config.hpp:18-144
//   heat_*_data_config<Two>, heat_data_config<Two>   pde_solvers/lss_heat_data_config.hpp:71-411
//   splitting_method_config             pde_solvers/lss_splitting_method_config.hpp:18-44
//   pde_implicit_solver_config, heat_implicit_solver_config, default_heat_solver_configs
//                                       pde_solvers/lss_pde_solver_config.hpp:25-115,
//                                       pde_solvers/lss_heat_solver_config.hpp:26-267
#pragma once

#include "common.hpp"

namespace lss_grids {

using lss_enumerations::dimension_enum;
using lss_enumerations::grid_enum;
using lss_utility::sptr_t;

template <dimension_enum dimension, typename fp_type> struct grid_config_hints;

template <typename fp_type> struct grid_config_hints<dimension_enum::One, fp_type> {
  private:
    fp_type accumulation_point_, alpha_scale_;
    grid_enum grid_;

  public:
    explicit grid_config_hints(fp_type accumulation_point = fp_type(0.0), fp_type alpha_scale = fp_type(3.0),
                               grid_enum grid_type = grid_enum::Uniform)
        : accumulation_point_{accumulation_point}, alpha_scale_{alpha_scale}, grid_{grid_type}
    {
    }
    fp_type accumulation_point() const { return accumulation_point_; }
    fp_type alpha_scale() const { return alpha_scale_; }
    grid_enum grid() const { return grid_; }
};

template <typename fp_type> struct grid_config_hints<dimension_enum::Two, fp_type> {
  private:
    fp_type accumulation_point_, alpha_scale_, beta_scale_;
    grid_enum grid_;

  public:
    explicit grid_config_hints(fp_type accumulation_point = fp_type(0.0), fp_type alpha_scale = fp_type(3.0),
                               fp_type beta_scale = fp_type(50.0), grid_enum grid_type = grid_enum::Uniform)
        : accumulation_point_{accumulation_point}, alpha_scale_{alpha_scale}, beta_scale_{beta_scale},
          grid_{grid_type}
    {
    }
    fp_type accumulation_point() const { return accumulation_point_; }
    fp_type alpha_scale() const { return alpha_scale_; }
    fp_type beta_scale() const { return beta_scale_; }
    grid_enum grid() const { return grid_; }
};

template <typename fp_type> using grid_config_hints_1d = grid_config_hints<dimension_enum::One, fp_type>;
template <typename fp_type> using grid_config_hints_2d = grid_config_hints<dimension_enum::Two, fp_type>;
template <typename fp_type> using grid_config_hints_1d_ptr = sptr_t<grid_config_hints_1d<fp_type>>;
template <typename fp_type> using grid_config_hints_2d_ptr = sptr_t<grid_config_hints_2d<fp_type>>;

}  // namespace lss_grids

namespace lss_pde_solvers {

using lss_enumerations::dimension_enum;
using lss_enumerations::explicit_pde_schemes_enum;
using lss_enumerations::factorization_enum;
using lss_enumerations::implicit_pde_schemes_enum;
using lss_enumerations::memory_space_enum;
using lss_enumerations::splitting_method_enum;
using lss_enumerations::traverse_direction_enum;
using lss_enumerations::tridiagonal_method_enum;
using lss_utility::range;
using lss_utility::sptr_t;

// ---------------------------------------------------------------------------------------------
// pde_discretization_config
// ---------------------------------------------------------------------------------------------
template <dimension_enum dimension, typename fp_type> struct pde_discretization_config;

template <typename fp_type> struct pde_discretization_config<dimension_enum::One, fp_type> {
  private:
    range<fp_type> space_range_, time_range_;
    std::size_t number_of_space_points_, number_of_time_points_;

  public:
    explicit pde_discretization_config(range<fp_type> const &space_range, std::size_t const &number_of_space_points,
                                       range<fp_type> const &time_range, std::size_t const &number_of_time_points)
        : space_range_{space_range}, time_range_{time_range}, number_of_space_points_{number_of_space_points},
          number_of_time_points_{number_of_time_points}
    {
    }
    range<fp_type> const &space_range() const { return space_range_; }
    range<fp_type> const &time_range() const { return time_range_; }
    std::size_t number_of_space_points() const { return number_of_space_points_; }
    std::size_t number_of_time_points() const { return number_of_time_points_; }
    fp_type space_step() const
    {
        return ((space_range_.spread()) / static_cast<fp_type>(number_of_space_points_ - 1));
    }
    fp_type time_step() const { return ((time_range_.spread()) / static_cast<fp_type>(number_of_time_points_ - 1)); }
};

template <typename fp_type> struct pde_discretization_config<dimension_enum::Two, fp_type> {
  private:
    range<fp_type> space_range_1_, space_range_2_, time_range_;
    std::size_t number_of_space_points_1_, number_of_space_points_2_, number_of_time_points_;

  public:
    explicit pde_discretization_config(range<fp_type> const &space_range_1, range<fp_type> const &space_range_2,
                                       std::size_t const &number_of_space_points_1,
                                       std::size_t const &number_of_space_points_2, range<fp_type> const &time_range,
                                       std::size_t const &number_of_time_points)
        : space_range_1_{space_range_1}, space_range_2_{space_range_2}, time_range_{time_range},
          number_of_space_points_1_{number_of_space_points_1}, number_of_space_points_2_{number_of_space_points_2},
          number_of_time_points_{number_of_time_points}
    {
    }
    sptr_t<pde_discretization_config<dimension_enum::One, fp_type>> const pde_discretization_1() const
    {
        return std::make_shared<pde_discretization_config<dimension_enum::One, fp_type>>(
            space_range_1_, number_of_space_points_1_, time_range_, number_of_time_points_);
    }
    sptr_t<pde_discretization_config<dimension_enum::One, fp_type>> const pde_discretization_2() const
    {
        return std::make_shared<pde_discretization_config<dimension_enum::One, fp_type>>(
            space_range_2_, number_of_space_points_2_, time_range_, number_of_time_points_);
    }
    std::pair<range<fp_type>, range<fp_type>> const space_range() const
    {
        return std::make_pair(space_range_1_, space_range_2_);
    }
    range<fp_type> const &time_range() const { return time_range_; }
    std::pair<std::size_t, std::size_t> const number_of_space_points() const
    {
        return std::make_pair(number_of_space_points_1_, number_of_space_points_2_);
    }
    std::size_t number_of_time_points() const { return number_of_time_points_; }
    std::pair<fp_type, fp_type> space_step() const
    {
        return std::make_pair((space_range_1_.spread()) / static_cast<fp_type>(number_of_space_points_1_ - 1),
                              (space_range_2_.spread()) / static_cast<fp_type>(number_of_space_points_2_ - 1));
    }
    fp_type time_step() const { return ((time_range_.spread()) / static_cast<fp_type>(number_of_time_points_ - 1)); }
};

template <typename fp_type> using pde_discretization_config_1d = pde_discretization_config<dimension_enum::One, fp_type>;
template <typename fp_type> using pde_discretization_config_2d = pde_discretization_config<dimension_enum::Two, fp_type>;
template <typename fp_type> using pde_discretization_config_1d_ptr = sptr_t<pde_discretization_config_1d<fp_type>>;
template <typename fp_type> using pde_discretization_config_2d_ptr = sptr_t<pde_discretization_config_2d<fp_type>>;

// ---------------------------------------------------------------------------------------------
// heat data (2-D part)
// ---------------------------------------------------------------------------------------------
template <dimension_enum dimension, typename fp_type> struct heat_coefficient_data_config;

template <typename fp_type> struct heat_coefficient_data_config<dimension_enum::Two, fp_type> {
  private:
    std::function<fp_type(fp_type, fp_type, fp_type)> a_, b_, c_, d_, e_, f_;

  public:
    explicit heat_coefficient_data_config(std::function<fp_type(fp_type, fp_type, fp_type)> const &a_coefficient,
                                          std::function<fp_type(fp_type, fp_type, fp_type)> const &b_coefficient,
                                          std::function<fp_type(fp_type, fp_type, fp_type)> const &c_coefficient,
                                          std::function<fp_type(fp_type, fp_type, fp_type)> const &d_coefficient,
                                          std::function<fp_type(fp_type, fp_type, fp_type)> const &e_coefficient,
                                          std::function<fp_type(fp_type, fp_type, fp_type)> const &f_coefficient)
        : a_{a_coefficient}, b_{b_coefficient}, c_{c_coefficient}, d_{d_coefficient}, e_{e_coefficient},
          f_{f_coefficient}
    {
        LSS_VERIFY(a_, "a_coefficient must not be null");
        LSS_VERIFY(b_, "b_coefficient must not be null");
        LSS_VERIFY(c_, "c_coefficient must not be null");
        LSS_VERIFY(d_, "d_coefficient must not be null");
        LSS_VERIFY(e_, "e_coefficient must not be null");
        LSS_VERIFY(f_, "f_coefficient must not be null");
    }
    std::function<fp_type(fp_type, fp_type, fp_type)> const &a_coefficient() const { return a_; }
    std::function<fp_type(fp_type, fp_type, fp_type)> const &b_coefficient() const { return b_; }
    std::function<fp_type(fp_type, fp_type, fp_type)> const &c_coefficient() const { return c_; }
    std::function<fp_type(fp_type, fp_type, fp_type)> const &d_coefficient() const { return d_; }
    std::function<fp_type(fp_type, fp_type, fp_type)> const &e_coefficient() const { return e_; }
    std::function<fp_type(fp_type, fp_type, fp_type)> const &f_coefficient() const { return f_; }
};

template <dimension_enum dimension, typename fp_type> struct heat_initial_data_config;
template <typename fp_type> struct heat_initial_data_config<dimension_enum::Two, fp_type> {
  private:
    std::function<fp_type(fp_type, fp_type)> initial_condition_;

  public:
    explicit heat_initial_data_config(std::function<fp_type(fp_type, fp_type)> const &initial_condition)
        : initial_condition_{initial_condition}
    {
        LSS_VERIFY(initial_condition_, "initial_condition must not be null");
    }
    std::function<fp_type(fp_type, fp_type)> const &initial_condition() const { return initial_condition_; }
};

template <dimension_enum dimension, typename fp_type> struct heat_source_data_config;
template <typename fp_type> struct heat_source_data_config<dimension_enum::Two, fp_type> {
  private:
    std::function<fp_type(fp_type, fp_type, fp_type)> heat_source_;

  public:
    explicit heat_source_data_config(std::function<fp_type(fp_type, fp_type, fp_type)> const &heat_source)
        : heat_source_{heat_source}
    {
        LSS_VERIFY(heat_source_, "heat_source must not be null");
    }
    std::function<fp_type(fp_type, fp_type, fp_type)> const &heat_source() const { return heat_source_; }
};

template <typename fp_type>
using heat_coefficient_data_config_2d = heat_coefficient_data_config<dimension_enum::Two, fp_type>;
template <typename fp_type> using heat_initial_data_config_2d = heat_initial_data_config<dimension_enum::Two, fp_type>;
template <typename fp_type> using heat_source_data_config_2d = heat_source_data_config<dimension_enum::Two, fp_type>;
template <typename fp_type>
using heat_coefficient_data_config_2d_ptr = sptr_t<heat_coefficient_data_config_2d<fp_type>>;
template <typename fp_type> using heat_initial_data_config_2d_ptr = sptr_t<heat_initial_data_config_2d<fp_type>>;
template <typename fp_type> using heat_source_data_config_2d_ptr = sptr_t<heat_source_data_config_2d<fp_type>>;

template <dimension_enum dimension, typename fp_type> struct heat_data_config;
template <typename fp_type> struct heat_data_config<dimension_enum::Two, fp_type> {
  private:
    heat_coefficient_data_config_2d_ptr<fp_type> coefficient_data_cfg_;
    heat_initial_data_config_2d_ptr<fp_type> initial_data_cfg_;
    heat_source_data_config_2d_ptr<fp_type> source_data_cfg_;

  public:
    explicit heat_data_config(heat_coefficient_data_config_2d_ptr<fp_type> const &coefficient_data_config,
                              heat_initial_data_config_2d_ptr<fp_type> const &initial_data_config,
                              heat_source_data_config_2d_ptr<fp_type> const &source_data_config = nullptr)
        : coefficient_data_cfg_{coefficient_data_config}, initial_data_cfg_{initial_data_config},
          source_data_cfg_{source_data_config}
    {
        LSS_VERIFY(coefficient_data_cfg_, "coefficient_data_config must not be null");
        LSS_VERIFY(initial_data_cfg_, "initial_data_config must not be null");
    }
    heat_source_data_config_2d_ptr<fp_type> const &source_data_config() const { return source_data_cfg_; }
    std::function<fp_type(fp_type, fp_type)> const &initial_condition() const
    {
        return initial_data_cfg_->initial_condition();
    }
    std::function<fp_type(fp_type, fp_type, fp_type)> const &a_coefficient() const
    {
        return coefficient_data_cfg_->a_coefficient();
    }
    std::function<fp_type(fp_type, fp_type, fp_type)> const &b_coefficient() const
    {
        return coefficient_data_cfg_->b_coefficient();
    }
    std::function<fp_type(fp_type, fp_type, fp_type)> const &c_coefficient() const
    {
        return coefficient_data_cfg_->c_coefficient();
    }
    std::function<fp_type(fp_type, fp_type, fp_type)> const &d_coefficient() const
    {
        return coefficient_data_cfg_->d_coefficient();
    }
    std::function<fp_type(fp_type, fp_type, fp_type)> const &e_coefficient() const
    {
        return coefficient_data_cfg_->e_coefficient();
    }
    std::function<fp_type(fp_type, fp_type, fp_type)> const &f_coefficient() const
    {
        return coefficient_data_cfg_->f_coefficient();
    }
};

template <typename fp_type> using heat_data_config_2d = heat_data_config<dimension_enum::Two, fp_type>;
template <typename fp_type> using heat_data_config_2d_ptr = sptr_t<heat_data_config_2d<fp_type>>;

// ---------------------------------------------------------------------------------------------
// splitting method config
// ---------------------------------------------------------------------------------------------
template <typename fp_type> struct splitting_method_config {
  private:
    splitting_method_enum splitting_method_;
    fp_type weighting_value_;

  public:
    explicit splitting_method_config() = delete;
    explicit splitting_method_config(splitting_method_enum splittitng_method, fp_type weighting_value = fp_type(0.5))
        : splitting_method_{splittitng_method}, weighting_value_{weighting_value}
    {
    }
    splitting_method_enum splitting_method() const { return splitting_method_; }
    fp_type weighting_value() const { return weighting_value_; }
};
template <typename fp_type> using splitting_method_config_ptr = sptr_t<splitting_method_config<fp_type>>;

// ---------------------------------------------------------------------------------------------
// solver configs (validation of lss_pde_solver_config.hpp:35-80 kept)
// ---------------------------------------------------------------------------------------------
struct pde_implicit_solver_config {
  private:
    memory_space_enum memory_space_;
    traverse_direction_enum traverse_direction_;
    tridiagonal_method_enum tridiagonal_method_;
    factorization_enum tridiagonal_factorization_;

  public:
    explicit pde_implicit_solver_config(memory_space_enum const &memory_space,
                                        traverse_direction_enum const &traverse_direction,
                                        tridiagonal_method_enum const &tridiagonal_method,
                                        factorization_enum const &tridiagonal_factorization)
        : memory_space_{memory_space}, traverse_direction_{traverse_direction},
          tridiagonal_method_{tridiagonal_method}, tridiagonal_factorization_{tridiagonal_factorization}
    {
        if (memory_space_ == memory_space_enum::Device) {
            LSS_VERIFY(!(tridiagonal_method_ == tridiagonal_method_enum::DoubleSweepSolver),
                       "No support for Double Sweep Solver on Device");
            LSS_VERIFY(!(tridiagonal_method_ == tridiagonal_method_enum::ThomasLUSolver),
                       "No support for Tomas LU Solver on Device");
        }
        if (tridiagonal_method_ == tridiagonal_method_enum::DoubleSweepSolver ||
            tridiagonal_method_ == tridiagonal_method_enum::ThomasLUSolver ||
            tridiagonal_method_ == tridiagonal_method_enum::SORSolver) {
            LSS_VERIFY(!(tridiagonal_factorization_ == factorization_enum::CholeskyMethod),
                       "No support for Cholesky Method factorization");
            LSS_VERIFY(!(tridiagonal_factorization_ == factorization_enum::LUMethod),
                       "No support for LU Method factorization");
            LSS_VERIFY(!(tridiagonal_factorization_ == factorization_enum::QRMethod),
                       "No support for QR Method factorization");
        }
    }
    virtual ~pde_implicit_solver_config() {}
    memory_space_enum memory_space() const { return memory_space_; }
    traverse_direction_enum traverse_direction() const { return traverse_direction_; }
    tridiagonal_method_enum tridiagonal_method() const { return tridiagonal_method_; }
    factorization_enum tridiagonal_factorization() const { return tridiagonal_factorization_; }
};

struct heat_implicit_solver_config : public pde_implicit_solver_config {
  private:
    implicit_pde_schemes_enum implicit_pde_scheme_;

  public:
    explicit heat_implicit_solver_config(memory_space_enum const &memory_space,
                                         traverse_direction_enum const &traverse_direction,
                                         tridiagonal_method_enum const &tridiagonal_method,
                                         factorization_enum const &tridiagonal_factorization,
                                         implicit_pde_schemes_enum const &implicit_pde_scheme)
        : pde_implicit_solver_config{memory_space, traverse_direction, tridiagonal_method, tridiagonal_factorization},
          implicit_pde_scheme_{implicit_pde_scheme}
    {
    }
    implicit_pde_schemes_enum implicit_pde_scheme() const { return implicit_pde_scheme_; }
};

using pde_implicit_solver_config_ptr = sptr_t<pde_implicit_solver_config>;
using heat_implicit_solver_config_ptr = sptr_t<heat_implicit_solver_config>;

// explicit solver configs (lss_pde_solver_config.hpp:120-147, lss_heat_solver_config.hpp:56-78)
struct pde_explicit_solver_config {
  private:
    memory_space_enum memory_space_;
    traverse_direction_enum traverse_direction_;

  public:
    explicit pde_explicit_solver_config(memory_space_enum const &memory_space,
                                        traverse_direction_enum const &traverse_direction)
        : memory_space_{memory_space}, traverse_direction_{traverse_direction}
    {
    }
    virtual ~pde_explicit_solver_config() {}
    memory_space_enum memory_space() const { return memory_space_; }
    traverse_direction_enum traverse_direction() const { return traverse_direction_; }
};

struct heat_explicit_solver_config : public pde_explicit_solver_config {
  private:
    explicit_pde_schemes_enum explicit_pde_scheme_;

  public:
    explicit heat_explicit_solver_config(memory_space_enum const &memory_space,
                                         traverse_direction_enum const &traverse_direction,
                                         explicit_pde_schemes_enum const &explicit_pde_scheme)
        : pde_explicit_solver_config{memory_space, traverse_direction}, explicit_pde_scheme_{explicit_pde_scheme}
    {
    }
    explicit_pde_schemes_enum explicit_pde_scheme() const { return explicit_pde_scheme_; }
};

using pde_explicit_solver_config_ptr = sptr_t<pde_explicit_solver_config>;
using heat_explicit_solver_config_ptr = sptr_t<heat_explicit_solver_config>;

// The reference's named default configurations (lss_heat_solver_config.hpp:84-239).
namespace default_heat_solver_configs {
#define LSS_DEFAULT_CFG(name, mem, dir, tri, fac, sch)                                                          \
    inline const heat_implicit_solver_config_ptr name = std::make_shared<heat_implicit_solver_config>(       \
        memory_space_enum::mem, traverse_direction_enum::dir, tridiagonal_method_enum::tri,                   \
        factorization_enum::fac, implicit_pde_schemes_enum::sch);
LSS_DEFAULT_CFG(dev_fwd_cusolver_qr_euler_solver_config_ptr, Device, Forward, CUDASolver, QRMethod, Euler)
LSS_DEFAULT_CFG(dev_fwd_cusolver_qr_cn_solver_config_ptr, Device, Forward, CUDASolver, QRMethod, CrankNicolson)
LSS_DEFAULT_CFG(dev_fwd_cusolver_lu_euler_solver_config_ptr, Device, Forward, CUDASolver, LUMethod, Euler)
LSS_DEFAULT_CFG(dev_fwd_cusolver_lu_cn_solver_config_ptr, Device, Forward, CUDASolver, LUMethod, CrankNicolson)
LSS_DEFAULT_CFG(dev_fwd_cusolver_euler_solver_config_ptr, Device, Forward, CUDASolver, None, Euler)
LSS_DEFAULT_CFG(host_fwd_cusolver_qr_euler_solver_config_ptr, Host, Forward, CUDASolver, QRMethod, Euler)
LSS_DEFAULT_CFG(host_fwd_cusolver_qr_cn_solver_config_ptr, Host, Forward, CUDASolver, QRMethod, CrankNicolson)
LSS_DEFAULT_CFG(host_fwd_cusolver_lu_euler_solver_config_ptr, Host, Forward, CUDASolver, LUMethod, Euler)
LSS_DEFAULT_CFG(host_fwd_cusolver_lu_cn_solver_config_ptr, Host, Forward, CUDASolver, LUMethod, CrankNicolson)
LSS_DEFAULT_CFG(host_fwd_cusolver_euler_solver_config_ptr, Host, Forward, CUDASolver, None, Euler)
LSS_DEFAULT_CFG(dev_fwd_sorsolver_euler_solver_config_ptr, Device, Forward, SORSolver, None, Euler)
LSS_DEFAULT_CFG(dev_fwd_sorsolver_cn_solver_config_ptr, Device, Forward, SORSolver, None, CrankNicolson)
LSS_DEFAULT_CFG(host_fwd_sorsolver_euler_solver_config_ptr, Host, Forward, SORSolver, None, Euler)
LSS_DEFAULT_CFG(host_fwd_sorsolver_cn_solver_config_ptr, Host, Forward, SORSolver, None, CrankNicolson)
LSS_DEFAULT_CFG(host_fwd_dssolver_euler_solver_config_ptr, Host, Forward, DoubleSweepSolver, None, Euler)
LSS_DEFAULT_CFG(host_fwd_dssolver_cn_solver_config_ptr, Host, Forward, DoubleSweepSolver, None, CrankNicolson)
LSS_DEFAULT_CFG(host_fwd_tlusolver_euler_solver_config_ptr, Host, Forward, ThomasLUSolver, None, Euler)
LSS_DEFAULT_CFG(host_fwd_tlusolver_cn_solver_config_ptr, Host, Forward, ThomasLUSolver, None, CrankNicolson)
LSS_DEFAULT_CFG(dev_bwd_cusolver_qr_euler_solver_config_ptr, Device, Backward, CUDASolver, QRMethod, Euler)
LSS_DEFAULT_CFG(dev_bwd_cusolver_qr_cn_solver_config_ptr, Device, Backward, CUDASolver, QRMethod, CrankNicolson)
LSS_DEFAULT_CFG(dev_bwd_cusolver_lu_euler_solver_config_ptr, Device, Backward, CUDASolver, LUMethod, Euler)
LSS_DEFAULT_CFG(dev_bwd_cusolver_lu_cn_solver_config_ptr, Device, Backward, CUDASolver, LUMethod, CrankNicolson)
LSS_DEFAULT_CFG(dev_bwd_cusolver_euler_solver_config_ptr, Device, Backward, CUDASolver, None, Euler)
LSS_DEFAULT_CFG(host_bwd_cusolver_qr_euler_solver_config_ptr, Host, Backward, CUDASolver, QRMethod, Euler)
LSS_DEFAULT_CFG(host_bwd_cusolver_qr_cn_solver_config_ptr, Host, Backward, CUDASolver, QRMethod, CrankNicolson)
LSS_DEFAULT_CFG(host_bwd_cusolver_lu_euler_solver_config_ptr, Host, Backward, CUDASolver, LUMethod, Euler)
LSS_DEFAULT_CFG(host_bwd_cusolver_lu_cn_solver_config_ptr, Host, Backward, CUDASolver, LUMethod, CrankNicolson)
LSS_DEFAULT_CFG(host_bwd_cusolver_euler_solver_config_ptr, Host, Backward, CUDASolver, None, Euler)
LSS_DEFAULT_CFG(dev_bwd_sorsolver_euler_solver_config_ptr, Device, Backward, SORSolver, None, Euler)
LSS_DEFAULT_CFG(dev_bwd_sorsolver_cn_solver_config_ptr, Device, Backward, SORSolver, None, CrankNicolson)
LSS_DEFAULT_CFG(host_bwd_sorsolver_euler_solver_config_ptr, Host, Backward, SORSolver, None, Euler)
LSS_DEFAULT_CFG(host_bwd_sorsolver_cn_solver_config_ptr, Host, Backward, SORSolver, None, CrankNicolson)
LSS_DEFAULT_CFG(host_bwd_dssolver_euler_solver_config_ptr, Host, Backward, DoubleSweepSolver, None, Euler)
LSS_DEFAULT_CFG(host_bwd_dssolver_cn_solver_config_ptr, Host, Backward, DoubleSweepSolver, None, CrankNicolson)
LSS_DEFAULT_CFG(host_bwd_tlusolver_euler_solver_config_ptr, Host, Backward, ThomasLUSolver, None, Euler)
LSS_DEFAULT_CFG(host_bwd_tlusolver_cn_solver_config_ptr, Host, Backward, ThomasLUSolver, None, CrankNicolson)
LSS_DEFAULT_CFG(host_bwd_tlusolver_o8_solver_config_ptr, Host, Backward, ThomasLUSolver, None, Theta_80)
#undef LSS_DEFAULT_CFG
// explicit ones (lss_heat_solver_config.hpp:244-266)
#define LSS_DEFAULT_XCFG(name, mem, dir, sch)                                                                   \
    inline const heat_explicit_solver_config_ptr name = std::make_shared<heat_explicit_solver_config>(       \
        memory_space_enum::mem, traverse_direction_enum::dir, explicit_pde_schemes_enum::sch);
LSS_DEFAULT_XCFG(dev_expl_fwd_euler_solver_config_ptr, Device, Forward, Euler)
LSS_DEFAULT_XCFG(host_expl_fwd_euler_solver_config_ptr, Host, Forward, Euler)
LSS_DEFAULT_XCFG(dev_expl_bwd_euler_solver_config_ptr, Device, Backward, Euler)
LSS_DEFAULT_XCFG(host_expl_bwd_euler_solver_config_ptr, Host, Backward, Euler)
LSS_DEFAULT_XCFG(host_expl_fwd_bc_solver_config_ptr, Host, Forward, ADEBarakatClark)
LSS_DEFAULT_XCFG(host_expl_bwd_bc_solver_config_ptr, Host, Backward, ADEBarakatClark)
LSS_DEFAULT_XCFG(host_expl_fwd_s_solver_config_ptr, Host, Forward, ADESaulyev)
LSS_DEFAULT_XCFG(host_expl_bwd_s_solver_config_ptr, Host, Backward, ADESaulyev)
#undef LSS_DEFAULT_XCFG
}  // namespace default_heat_solver_configs

}  // namespace lss_pde_solvers
