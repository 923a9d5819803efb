// dropin_heston.cpp -- reference user code run against the drop-in headers.
//
// Each case is written the way the reference's own tests write it (builders / constructor,
// closures, default solver configs) and prints the surface in the reference's text format, so the
// pytest driver can compare it with the committed reference surfaces and with the oracle.
//   case 0: unit_tests/builders/lss_2d_general_heston_equation_builder_t.hpp:19-163 (builders, DS)
//   case 1: unit_tests/common/lss_print_t.hpp:2104-2210 (Heston TLU DR, default Uniform hints)
//   case 2: unit_tests/common/lss_print_t.hpp:1976-2090 (SABR DS DR, time-dependent coefficients)
//   case 3: unit_tests/common/lss_print_t.hpp:1724-1836 (Device cuSOLVER-QR config)
//   case 4: HV with Host ThomasLU (unit_tests/common/lss_print_t.hpp:2470-2579)
//   case 5: explicit Euler, Host, backward (unit_tests/common/lss_print_t.hpp:2599-2703)
// usage: dropin_heston <case> <double|float>  (surface to stdout)
#include <cmath>
#include <cstdlib>
#include <iomanip>
#include <iostream>
#include <string>

#include "lss/lss.hpp"

using namespace lss_enumerations;

template <typename T> int run(int which)
{
    using lss_boundary::dirichlet_boundary_2d;
    using lss_boundary::dirichlet_boundary_2d_builder;
    using lss_boundary::neumann_boundary_2d;
    using lss_boundary::neumann_boundary_2d_builder;
    using lss_containers::container_2d;
    using lss_grids::grid_config_hints_2d;
    using lss_grids::grid_config_hints_2d_builder;
    using lss_pde_solvers::heat_coefficient_data_config_2d;
    using lss_pde_solvers::heat_coefficient_data_config_2d_builder;
    using lss_pde_solvers::heat_data_config_2d;
    using lss_pde_solvers::heat_data_config_2d_builder;
    using lss_pde_solvers::heat_initial_data_config_2d;
    using lss_pde_solvers::heat_initial_data_config_2d_builder;
    using lss_pde_solvers::pde_discretization_config_2d;
    using lss_pde_solvers::pde_discretization_config_2d_builder;
    using lss_pde_solvers::splitting_method_config;
    using lss_pde_solvers::splitting_method_config_builder;
    using lss_pde_solvers::two_dimensional::implicit_solvers::general_heston_equation;
    using lss_pde_solvers::two_dimensional::implicit_solvers::general_heston_equation_builder;
    using lss_utility::range;
    using lss_utility::range_builder;
    namespace dflt = lss_pde_solvers::default_heat_solver_configs;
    typedef container_2d<by_enum::Row, T, std::vector, std::allocator<T>> rcontainer_2d_t;

    if (which == 0) {
        auto const &strike = 10.0;
        auto const &maturity = 1.0;
        auto const &rate = 0.03;
        auto const &sig_sig = 0.3;
        auto const &sig_kappa = 2.0;
        auto const &sig_theta = 0.2;
        auto const &rho = 0.2;
        std::size_t const Sd = 100, Vd = 50, Td = 150;
        auto const &stock_range = range_builder<T>().lower(T(0.0)).upper(T(20.0)).build();
        auto const &vol_range = range_builder<T>().lower(T(0.0)).upper(T(1.5)).build();
        auto const &time_range = range_builder<T>().lower(T(0.0)).upper(maturity).build();
        auto const &discretization_ptr = pde_discretization_config_2d_builder<T>()
                                             .space_range_1(*stock_range)
                                             .space_range_2(*vol_range)
                                             .number_of_space_points_1(Sd)
                                             .number_of_space_points_2(Vd)
                                             .time_range(*time_range)
                                             .number_of_time_points(Td)
                                             .build();
        auto a = [=](T t, T s, T v) { return (0.5 * v * s * s); };
        auto b = [=](T t, T s, T v) { return (0.5 * sig_sig * sig_sig * v); };
        auto c = [=](T t, T s, T v) { return (rho * sig_sig * v * s); };
        auto d = [=](T t, T s, T v) { return (rate * s); };
        auto e = [=](T t, T s, T v) { return (sig_kappa * (sig_theta - v)); };
        auto f = [=](T t, T s, T v) { return (-rate); };
        auto const &coefficients_ptr = heat_coefficient_data_config_2d_builder<T>()
                                           .a_coefficient(a)
                                           .b_coefficient(b)
                                           .c_coefficient(c)
                                           .d_coefficient(d)
                                           .e_coefficient(e)
                                           .f_coefficient(f)
                                           .build();
        auto terminal_condition = [=](T s, T v) { return std::max<T>(0.0, s - strike); };
        auto const &init_data_ptr =
            heat_initial_data_config_2d_builder<T>().initial_condition(terminal_condition).build();
        auto const &data_ptr =
            heat_data_config_2d_builder<T>().coefficient_data_config(coefficients_ptr).initial_data_config(init_data_ptr).build();
        auto const &grid_config_hints_ptr = grid_config_hints_2d_builder<T>()
                                                .accumulation_point(strike)
                                                .alpha_scale(T(3.))
                                                .beta_scale(T(60.))
                                                .grid(grid_enum::Nonuniform)
                                                .build();
        auto const &boundary_low_ptr = dirichlet_boundary_2d_builder<T>().value([=](T t, T v) { return 0.0; }).build();
        auto const &boundary_high_ptr = neumann_boundary_2d_builder<T>().value([=](T t, T v) { return -1.0; }).build();
        auto const &horizontal_boundary_pair = std::make_pair(boundary_low_ptr, boundary_high_ptr);
        auto const &vertical_upper_boundary_ptr =
            dirichlet_boundary_2d_builder<T>().value([=](T t, T s) { return s; }).build();
        auto const &splitting_config_ptr = splitting_method_config_builder<T>()
                                               .splitting_method(splitting_method_enum::DouglasRachford)
                                               .weighting_value(T(0.5))
                                               .build();
        auto const &pde_solver = general_heston_equation_builder<T, std::vector, std::allocator<T>>()
                                     .heat_data_config(data_ptr)
                                     .discretization_config(discretization_ptr)
                                     .vertical_upper_boundary(vertical_upper_boundary_ptr)
                                     .horizontal_boundary(horizontal_boundary_pair)
                                     .splitting_method_config(splitting_config_ptr)
                                     .grid_hints(grid_config_hints_ptr)
                                     .solver_config(dflt::host_bwd_dssolver_cn_solver_config_ptr)
                                     .build();
        rcontainer_2d_t solution(Sd, Vd, T{});
        pde_solver->solve(solution);
        lss_print::print(discretization_ptr, grid_config_hints_ptr, solution, std::cout);
        return 0;
    }

    if (which == 5) {
        using lss_pde_solvers::two_dimensional::explicit_solvers::general_heston_equation;
        auto const &strike = 100.0;
        auto const &maturity = 1.0;
        auto const &rate = 0.03;
        auto const &sig_sig = 0.3;
        auto const &sig_kappa = 2.0;
        auto const &sig_theta = 0.2;
        auto const &rho = 0.8;
        std::size_t const Sd = 50, Vd = 30, Td = 8000;
        range<T> spacex_range(static_cast<T>(50.0), static_cast<T>(200.0));
        range<T> spacey_range(static_cast<T>(0.0), static_cast<T>(1.0));
        range<T> time_range(static_cast<T>(0.0), static_cast<T>(maturity));
        auto const discretization_ptr =
            std::make_shared<pde_discretization_config_2d<T>>(spacex_range, spacey_range, Sd, Vd, time_range, Td);
        auto a = [=](T t, T s, T v) { return (0.5 * v * s * s); };
        auto b = [=](T t, T s, T v) { return (0.5 * sig_sig * sig_sig * v); };
        auto c = [=](T t, T s, T v) { return (rho * sig_sig * v * s); };
        auto d = [=](T t, T s, T v) { return (rate * s); };
        auto e = [=](T t, T s, T v) { return (sig_kappa * (sig_theta - v)); };
        auto f = [=](T t, T s, T v) { return (-rate); };
        auto const heat_coeffs_data_ptr = std::make_shared<heat_coefficient_data_config_2d<T>>(a, b, c, d, e, f);
        auto terminal_condition = [=](T s, T v) { return std::max<T>(0.0, s - strike); };
        auto const heat_init_data_ptr = std::make_shared<heat_initial_data_config_2d<T>>(terminal_condition);
        auto const heat_data_ptr = std::make_shared<heat_data_config_2d<T>>(heat_coeffs_data_ptr, heat_init_data_ptr);
        auto const &dirichlet_low = [=](T t, T v) { return 0.0; };
        auto const &neumann_high = [=](T t, T s) { return -1.0; };
        auto const &boundary_low_ptr = std::make_shared<dirichlet_boundary_2d<T>>(dirichlet_low);
        auto const &boundary_high_ptr = std::make_shared<neumann_boundary_2d<T>>(neumann_high);
        auto const &horizontal_boundary_pair = std::make_pair(boundary_low_ptr, boundary_high_ptr);
        auto const &dirichlet_high = [=](T t, T s) { return s; };
        auto const &vertical_upper_boundary_ptr = std::make_shared<dirichlet_boundary_2d<T>>(dirichlet_high);
        auto const alpha_scale = static_cast<T>(3.);
        auto const beta_scale = static_cast<T>(50.);
        auto const &grid_config_hints_ptr =
            std::make_shared<grid_config_hints_2d<T>>(strike, alpha_scale, beta_scale, grid_enum::Nonuniform);
        general_heston_equation<T, std::vector, std::allocator<T>> pdesolver(
            heat_data_ptr, discretization_ptr, vertical_upper_boundary_ptr, horizontal_boundary_pair,
            grid_config_hints_ptr, dflt::host_expl_bwd_euler_solver_config_ptr);
        rcontainer_2d_t solution(Sd, Vd, T{});
        pdesolver.solve(solution);
        lss_print::print(discretization_ptr, grid_config_hints_ptr, solution, std::cout);
        return 0;
    }

    // constructor-style cases (lss_print_t.hpp)
    double strike = 100.0, rate = 0.03, sig_sig = 0.3, sig_kappa = 2.0, sig_theta = 0.2, rho = 0.2, beta = 0.7;
    const double maturity = 1.0;
    std::size_t Sd = 150, Vd = 50, Td = 200;
    T x_lo = 50, x_hi = 150, y_hi = 1.2;
    auto hints = std::make_shared<grid_config_hints_2d<T>>();
    auto split = std::make_shared<splitting_method_config<T>>(splitting_method_enum::DouglasRachford);
    auto solver_cfg = dflt::host_bwd_tlusolver_cn_solver_config_ptr;
    bool sabr = false;
    if (which == 2) {
        sabr = true;
        sig_sig = 0.081;
        rho = 0.6;
        Sd = 100;
        x_hi = 200;
        hints = std::make_shared<grid_config_hints_2d<T>>(strike, T(3.), T(50.), grid_enum::Nonuniform);
        solver_cfg = dflt::host_bwd_dssolver_cn_solver_config_ptr;
    } else if (which == 3) {
        strike = 10.0;
        Sd = 100;
        Td = 150;
        x_lo = 0;
        x_hi = 20;
        y_hi = 1.0;
        hints = std::make_shared<grid_config_hints_2d<T>>(strike, static_cast<T>(10. / 3.), static_cast<T>(1.0),
                                                          grid_enum::Nonuniform);
        solver_cfg = dflt::dev_bwd_cusolver_qr_cn_solver_config_ptr;
    } else if (which == 4) {
        x_hi = 200;
        hints = std::make_shared<grid_config_hints_2d<T>>(strike, T(3.), T(50.), grid_enum::Nonuniform);
        split = std::make_shared<splitting_method_config<T>>(splitting_method_enum::HundsdorferVerwer);
    }
    range<T> spacex_range(x_lo, x_hi), spacey_range(static_cast<T>(0.0), y_hi),
        time_range(static_cast<T>(0.0), static_cast<T>(maturity));
    auto const discretization_ptr =
        std::make_shared<pde_discretization_config_2d<T>>(spacex_range, spacey_range, Sd, Vd, time_range, Td);
    std::shared_ptr<heat_coefficient_data_config_2d<T>> coeffs;
    if (!sabr) {
        auto a = [=](T t, T s, T v) { return (0.5 * v * s * s); };
        auto b = [=](T t, T s, T v) { return (0.5 * sig_sig * sig_sig * v); };
        auto c = [=](T t, T s, T v) { return (rho * sig_sig * v * s); };
        auto d = [=](T t, T s, T v) { return (rate * s); };
        auto e = [=](T t, T s, T v) { return (sig_kappa * (sig_theta - v)); };
        auto f = [=](T t, T s, T v) { return (-rate); };
        coeffs = std::make_shared<heat_coefficient_data_config_2d<T>>(a, b, c, d, e, f);
    } else {
        auto D = [=](T t, T s, T alpha) { return std::exp(-rate * (maturity - t)); };
        auto a = [=](T t, T s, T alpha) {
            return (0.5 * alpha * alpha * std::pow(s, 2.0 * beta) * std::pow(D(t, s, alpha), 2.0 * (1.0 - beta)));
        };
        auto b = [=](T t, T s, T alpha) { return (0.5 * sig_sig * sig_sig * alpha * alpha); };
        auto c = [=](T t, T s, T alpha) {
            return (rho * sig_sig * alpha * alpha * std::pow(s, beta) * std::pow(D(t, s, alpha), (1.0 - beta)));
        };
        auto d = [=](T t, T s, T alpha) { return (rate * s); };
        auto e = [=](T t, T s, T alpha) { return 0.0; };
        auto f = [=](T t, T s, T alpha) { return (-rate); };
        coeffs = std::make_shared<heat_coefficient_data_config_2d<T>>(a, b, c, d, e, f);
    }
    auto terminal_condition = [=](T s, T v) { return std::max<T>(0.0, s - strike); };
    auto const heat_init_data_ptr = std::make_shared<heat_initial_data_config_2d<T>>(terminal_condition);
    auto const heat_data_ptr = std::make_shared<heat_data_config_2d<T>>(coeffs, heat_init_data_ptr);
    auto const &boundary_low_ptr = std::make_shared<dirichlet_boundary_2d<T>>([=](T t, T v) { return 0.0; });
    auto const &boundary_high_ptr = std::make_shared<neumann_boundary_2d<T>>([=](T t, T s) { return -1.0; });
    auto const &horizontal_boundary_pair = std::make_pair(boundary_low_ptr, boundary_high_ptr);
    auto const &vertical_upper_boundary_ptr = std::make_shared<dirichlet_boundary_2d<T>>([=](T t, T s) { return s; });
    general_heston_equation<T, std::vector, std::allocator<T>> pdesolver(
        heat_data_ptr, discretization_ptr, vertical_upper_boundary_ptr, horizontal_boundary_pair, split, hints,
        solver_cfg);
    rcontainer_2d_t solution(Sd, Vd, T{});
    pdesolver.solve(solution);
    lss_print::print(discretization_ptr, hints, solution, std::cout);
    return 0;
}

int main(int argc, char **argv)
{
    if (argc < 3) {
        std::cerr << "usage: dropin_heston <case 0..5> <double|float> [full]\n";
        return 2;
    }
    const int which = std::atoi(argv[1]);
    // "full": enough significant digits to round-trip every value (17 double, 9 float) instead of
    // the reference's 6
    if (argc > 3 && std::string(argv[3]) == "full")
        std::cout << std::setprecision(std::string(argv[2]) == "float" ? 9 : 17);
    try {
        return std::string(argv[2]) == "float" ? run<float>(which) : run<double>(which);
    } catch (const std::exception &e) {
        std::cerr << "error: " << e.what() << "\n";
        return 1;
    }
}
