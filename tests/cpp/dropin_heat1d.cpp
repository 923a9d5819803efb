// dropin_heat1d.cpp -- reference-style user code for the 1-D implicit heat solver, run against the
// drop-in headers (one_dimensional::implicit_solvers::general_heat_equation on the GPU).
//
// Cases 0-6 are written the way the reference's print tests write them and print in the reference's
// 1-D text formats, so the pytest driver compares them with the committed outputs and the oracle:
//   case 0: unit_tests/common/lss_print_t.hpp:105-207  Black-Scholes, Thomas LU, Euler, final level
//   case 1: unit_tests/common/lss_print_t.hpp:209-311  Black-Scholes, Thomas LU, Crank-Nicolson
//   case 2: unit_tests/common/lss_print_t.hpp:327-437  Black-Scholes, Euler, every level
//   case 3: unit_tests/common/lss_print_t.hpp:439-552  Black-Scholes, Crank-Nicolson, sinh grid, every level
//   case 4: unit_tests/common/lss_print_t.hpp:568-683  pure heat, Device cuSOLVER QR config, Euler
//   case 5: unit_tests/common/lss_print_t.hpp:685-802  pure heat, Device cuSOLVER QR config, CN
//   case 6: unit_tests/common/lss_print_t.hpp:949-1075 advection-diffusion, Thomas LU, CN
// Cases 7 and 8 exercise what those leave out: a heat source, Robin and Neumann ends, the double
// sweep and the Forward / Backward loops (the oracle holds the same data).
// usage: dropin_heat1d <case> <double|float> [full]
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <iomanip>
#include <iostream>
#include <string>

#include "lss/lss.hpp"

using namespace lss_enumerations;

template <typename T> int run(int which)
{
    using lss_boundary::dirichlet_boundary_1d;
    using lss_boundary::neumann_boundary_1d;
    using lss_boundary::robin_boundary_1d;
    using lss_containers::container_2d;
    using lss_grids::grid_config_hints_1d;
    using lss_pde_solvers::heat_coefficient_data_config_1d;
    using lss_pde_solvers::heat_data_config_1d;
    using lss_pde_solvers::heat_initial_data_config_1d;
    using lss_pde_solvers::heat_source_data_config_1d;
    using lss_pde_solvers::pde_discretization_config_1d;
    using lss_pde_solvers::one_dimensional::implicit_solvers::general_heat_equation;
    using lss_print::print;
    using lss_utility::range;
    namespace dflt = lss_pde_solvers::default_heat_solver_configs;
    typedef container_2d<by_enum::Row, T, std::vector, std::allocator<T>> container_2d_t;
    typedef general_heat_equation<T, std::vector, std::allocator<T>> pde_solver;

    if (which <= 3) {
        auto const &strike = 10;
        auto const &maturity = 1.0;
        auto const &rate = 0.2;
        auto const &sig = 0.25;
        std::size_t const Sd = 100;
        std::size_t const Td = 100;
        range<T> space_range(static_cast<T>(0.0), static_cast<T>(20.0));
        range<T> time_range(static_cast<T>(0.0), static_cast<T>(maturity));
        auto const discretization_ptr =
            std::make_shared<pde_discretization_config_1d<T>>(space_range, Sd, time_range, Td);
        auto a = [=](T t, T x) { return 0.5 * sig * sig * x * x; };
        auto b = [=](T t, T x) { return rate * x; };
        auto c = [=](T t, T x) { return -rate; };
        auto const heat_coeffs_data_ptr = std::make_shared<heat_coefficient_data_config_1d<T>>(a, b, c);
        auto terminal_condition = [=](T x) { return std::max<T>(0.0, x - strike); };
        auto const heat_init_data_ptr = std::make_shared<heat_initial_data_config_1d<T>>(terminal_condition);
        auto const heat_data_ptr = std::make_shared<heat_data_config_1d<T>>(heat_coeffs_data_ptr, heat_init_data_ptr);
        auto const &dirichlet_low = [=](T t) { return 0.0; };
        auto const &dirichlet_high = [=](T t) { return (20.0 - strike * std::exp(-rate * (maturity - t))); };
        auto const &boundary_low_ptr = std::make_shared<dirichlet_boundary_1d<T>>(dirichlet_low);
        auto const &boundary_high_ptr = std::make_shared<dirichlet_boundary_1d<T>>(dirichlet_high);
        auto const &boundary_pair = std::make_pair(boundary_low_ptr, boundary_high_ptr);
        if (which <= 1) {
            auto const &grid_config_hints_ptr = std::make_shared<grid_config_hints_1d<T>>();
            pde_solver pdesolver(heat_data_ptr, discretization_ptr, boundary_pair, grid_config_hints_ptr,
                                 which == 0 ? dflt::host_bwd_tlusolver_euler_solver_config_ptr
                                            : dflt::host_bwd_tlusolver_cn_solver_config_ptr);
            std::vector<T> solution(Sd, T{});
            pdesolver.solve(solution);
            print(discretization_ptr, grid_config_hints_ptr, solution, std::cout);
        } else {
            auto const alpha_scale = static_cast<T>(3.);
            auto const &grid_config_hints_ptr =
                which == 2 ? std::make_shared<grid_config_hints_1d<T>>(strike)
                           : std::make_shared<grid_config_hints_1d<T>>(strike, alpha_scale, grid_enum::Nonuniform);
            pde_solver pdesolver(heat_data_ptr, discretization_ptr, boundary_pair, grid_config_hints_ptr,
                                 which == 2 ? dflt::host_bwd_tlusolver_euler_solver_config_ptr
                                            : dflt::host_bwd_tlusolver_cn_solver_config_ptr);
            container_2d_t solutions(Td, Sd);
            pdesolver.solve(solutions);
            print(discretization_ptr, grid_config_hints_ptr, solutions, std::cout);
        }
        return 0;
    }
    if (which <= 6) {
        std::size_t const Sd = 100;
        std::size_t const Td = 100;
        range<T> space_range(static_cast<T>(0.0), static_cast<T>(1.0));
        range<T> time_range(static_cast<T>(0.0), static_cast<T>(0.1));
        auto const discretization_ptr =
            std::make_shared<pde_discretization_config_1d<T>>(space_range, Sd, time_range, Td);
        auto a = [](T t, T x) { return 1.0; };
        auto other = [](T t, T x) { return 0.0; };
        auto const heat_coeffs_data_ptr =
            which == 6 ? std::make_shared<heat_coefficient_data_config_1d<T>>(a, [](T t, T x) { return -1.0; }, other)
                       : std::make_shared<heat_coefficient_data_config_1d<T>>(a, other, other);
        auto initial_condition = [=](T x) { return which == 6 ? 1.0 : static_cast<double>(x); };
        auto const heat_init_data_ptr = std::make_shared<heat_initial_data_config_1d<T>>(initial_condition);
        auto const heat_data_ptr = std::make_shared<heat_data_config_1d<T>>(heat_coeffs_data_ptr, heat_init_data_ptr);
        auto const &dirichlet = [](T t) { return 0.0; };
        auto const &boundary_ptr = std::make_shared<dirichlet_boundary_1d<T>>(dirichlet);
        auto const &boundary_pair = std::make_pair(boundary_ptr, boundary_ptr);
        auto const alpha_scale = static_cast<T>(3.);
        auto const &grid_config_hints_ptr =
            which == 4 ? std::make_shared<grid_config_hints_1d<T>>()
                       : std::make_shared<grid_config_hints_1d<T>>(T(0.5), alpha_scale, grid_enum::Nonuniform);
        auto const &cfg = which == 4   ? dflt::dev_fwd_cusolver_qr_euler_solver_config_ptr
                          : which == 5 ? dflt::dev_fwd_cusolver_qr_cn_solver_config_ptr
                                       : dflt::host_fwd_tlusolver_cn_solver_config_ptr;
        pde_solver pdesolver(heat_data_ptr, discretization_ptr, boundary_pair, grid_config_hints_ptr, cfg);
        container_2d_t solutions(Td, Sd);
        pdesolver.solve(solutions);
        print(discretization_ptr, grid_config_hints_ptr, solutions, std::cout);
        return 0;
    }
    // 7: source, Robin lower, Neumann upper, double sweep, Crank-Nicolson, Forward, every level
    // 8: Neumann lower, Robin upper, Thomas LU, Euler, Backward, final level
    std::size_t const Sd = 60;
    std::size_t const Td = 80;
    range<T> space_range(static_cast<T>(0.0), static_cast<T>(1.0));
    range<T> time_range(static_cast<T>(0.0), static_cast<T>(0.5));
    auto const discretization_ptr = std::make_shared<pde_discretization_config_1d<T>>(space_range, Sd, time_range, Td);
    auto a = [](T t, T x) { return 1.0 + 0.5 * x; };
    auto b = [](T t, T x) { return 0.3; };
    auto c = [](T t, T x) { return -0.2; };
    auto const heat_coeffs_data_ptr = std::make_shared<heat_coefficient_data_config_1d<T>>(a, b, c);
    auto init = [](T x) { return x * x; };
    auto const heat_init_data_ptr = std::make_shared<heat_initial_data_config_1d<T>>(init);
    auto const alpha_scale = static_cast<T>(3.);
    auto const &grid_config_hints_ptr =
        std::make_shared<grid_config_hints_1d<T>>(T(0.5), alpha_scale, grid_enum::Nonuniform);
    if (which == 7) {
        auto src = [](T t, T x) { return t * x * (1.0 - x); };
        auto const heat_source_ptr = std::make_shared<heat_source_data_config_1d<T>>(src);
        auto const heat_data_ptr =
            std::make_shared<heat_data_config_1d<T>>(heat_coeffs_data_ptr, heat_init_data_ptr, heat_source_ptr);
        auto const &lo = std::make_shared<robin_boundary_1d<T>>([](T t) { return 0.5; }, [](T t) { return -0.1 * t; });
        auto const &hi = std::make_shared<neumann_boundary_1d<T>>([](T t) { return 0.3; });
        pde_solver pdesolver(heat_data_ptr, discretization_ptr, std::make_pair(lo, hi), grid_config_hints_ptr,
                             dflt::host_fwd_dssolver_cn_solver_config_ptr);
        container_2d_t solutions(Td, Sd);
        pdesolver.solve(solutions);
        print(discretization_ptr, grid_config_hints_ptr, solutions, std::cout);
    } else {
        auto const heat_data_ptr = std::make_shared<heat_data_config_1d<T>>(heat_coeffs_data_ptr, heat_init_data_ptr);
        auto const &lo = std::make_shared<neumann_boundary_1d<T>>([](T t) { return -0.4; });
        auto const &hi = std::make_shared<robin_boundary_1d<T>>([](T t) { return 0.8; }, [](T t) { return 0.25 * t; });
        pde_solver pdesolver(heat_data_ptr, discretization_ptr, std::make_pair(lo, hi), grid_config_hints_ptr,
                             dflt::host_bwd_tlusolver_euler_solver_config_ptr);
        std::vector<T> solution(Sd, T{});
        pdesolver.solve(solution);
        print(discretization_ptr, grid_config_hints_ptr, solution, std::cout);
    }
    return 0;
}

int main(int argc, char **argv)
{
    if (argc < 3) {
        std::cerr << "usage: dropin_heat1d <case 0..8> <double|float> [full]\n";
        return 2;
    }
    const int which = std::atoi(argv[1]);
    // "full": enough significant digits to round-trip every value (17 double, 9 float)
    if (argc > 3 && std::string(argv[3]) == "full")
        std::cout << std::setprecision(std::string(argv[2]) == "float" ? 9 : 17);
    try {
        return std::string(argv[2]) == "float" ? run<float>(which) : run<double>(which);
    } catch (const std::exception &e) {
        std::cerr << "error: " << e.what() << "\n";
        return 1;
    }
}
