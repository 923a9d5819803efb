// dropin_extras.cpp -- reference-style user code for what the drop-in adds beside general_heston_equation::
// solve(container_2d&): every time level (solve(container_3d&)), the batched book and the adjoint pair.
//
//   levels <double|float>  a Heston DR problem (the data of unit_tests/common/lss_print_t.hpp:2156-2186 on a
//                          48 x 24 grid, 30 steps, sinh grid) solved into a container_3d; prints every layer
//                          (full precision), then "FINAL" and the container_2d solve of the same problem
//   book                   three general_heston_equation problems of different sizes, schemes and data,
//                          solved one by one (solve(container_2d&)) and as one batched_heston_book with
//                          exact and with partitioned lines; prints the largest difference of each book
//                          against the single solves
//   adjoint                heston_adjoint on Heston MCS 40 x 24, 12 steps: forward, the bilinear price at
//                          (S, v) = (100, 0.04), backward; prints the price, the five parameter gradients
//                          and <dJ/du_in, u_in>
//   source <scheme> <fwd|bwd> <double|float> <tlu|ds>
//                          a Heston problem with the heat source F = 0.05 v S (1 + t) (40 x 20 grid, 16
//                          steps, sinh grid, Crank-Nicolson), scheme 0-3 = DR, CS, MCS, HV; prints the
//                          final surface (full precision)
//   source_book            the four schemes with that source, solved one by one and as one partitioned
//                          batched_heston_book; prints the largest relative difference of each
//   source_bench           a partitioned batched_heston_book of 64 problems (100 x 50 points, 100 steps, the
//                          four schemes in turn, strikes 80..120) without and with that source; prints
//                          "source_bench <without|with> <best solve ms of 5> <grid-point-steps per solve>"
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <iomanip>
#include <iostream>
#include <string>

#include "lss/heston_adjoint.hpp"
#include "lss/lss.hpp"

using namespace lss_enumerations;
using lss_boundary::dirichlet_boundary_2d;
using lss_boundary::neumann_boundary_2d;
using lss_containers::container_2d;
using lss_containers::container_3d;
using lss_grids::grid_config_hints_2d;
using lss_pde_solvers::heat_coefficient_data_config_2d;
using lss_pde_solvers::heat_data_config_2d;
using lss_pde_solvers::heat_initial_data_config_2d;
using lss_pde_solvers::heat_source_data_config_2d;
using lss_pde_solvers::pde_discretization_config_2d;
using lss_pde_solvers::splitting_method_config;
using lss_pde_solvers::two_dimensional::implicit_solvers::batched_heston_book;
using lss_pde_solvers::two_dimensional::implicit_solvers::general_heston_equation;
using lss_pde_solvers::two_dimensional::implicit_solvers::heston_adjoint;
using lss_pde_solvers::two_dimensional::implicit_solvers::heston_parameters;
using lss_utility::range;
namespace dflt = lss_pde_solvers::default_heat_solver_configs;

template <typename T>
std::shared_ptr<general_heston_equation<T>> heston(std::size_t Sd, std::size_t Vd, std::size_t Td, double strike,
                                                   double sig_kappa, double rho, splitting_method_enum scheme,
                                                   lss_pde_solvers::heat_implicit_solver_config_ptr cfg,
                                                   bool with_source = false)
{
    const double rate = 0.03, sig_sig = 0.3, sig_theta = 0.2, maturity = 0.25;
    range<T> spacex_range(static_cast<T>(0.0), static_cast<T>(2.0 * strike)), spacey_range(T(0.0), T(1.2)),
        time_range(T(0.0), T(maturity));
    auto const discretization_ptr =
        std::make_shared<pde_discretization_config_2d<T>>(spacex_range, spacey_range, Sd, Vd, time_range, Td);
    auto a = [=](T t, T s, T v) { return (0.5 * v * s * s); };
    auto b = [=](T t, T s, T v) { return (0.5 * sig_sig * sig_sig * v); };
    auto c = [=](T t, T s, T v) { return (rho * sig_sig * v * s); };
    auto d = [=](T t, T s, T v) { return (rate * s); };
    auto e = [=](T t, T s, T v) { return (sig_kappa * (sig_theta - v)); };
    auto f = [=](T t, T s, T v) { return (-rate); };
    auto coeffs = std::make_shared<heat_coefficient_data_config_2d<T>>(a, b, c, d, e, f);
    auto terminal_condition = [=](T s, T v) { return std::max<T>(0.0, s - strike); };
    auto src = [=](T t, T s, T v) { return (0.05 * v * s * (1.0 + t)); };
    auto const heat_data_ptr = std::make_shared<heat_data_config_2d<T>>(
        coeffs, std::make_shared<heat_initial_data_config_2d<T>>(terminal_condition),
        with_source ? std::make_shared<heat_source_data_config_2d<T>>(src) : nullptr);
    auto const &lo = std::make_shared<dirichlet_boundary_2d<T>>([=](T t, T v) { return 0.0; });
    auto const &hi = std::make_shared<neumann_boundary_2d<T>>([=](T t, T s) { return -1.0; });
    auto const &vu = std::make_shared<dirichlet_boundary_2d<T>>([=](T t, T s) { return s; });
    auto hints = std::make_shared<grid_config_hints_2d<T>>(static_cast<T>(strike), T(3.), T(15.), grid_enum::Nonuniform);
    auto split = std::make_shared<splitting_method_config<T>>(scheme, T(0.5));
    return std::make_shared<general_heston_equation<T>>(heat_data_ptr, discretization_ptr, vu, std::make_pair(lo, hi),
                                                        split, hints, cfg);
}

template <typename T> int levels()
{
    typedef container_2d<by_enum::Row, T, std::vector, std::allocator<T>> c2_t;
    typedef container_3d<by_enum::Row, T, std::vector, std::allocator<T>> c3_t;
    const std::size_t Sd = 48, Vd = 24, Td = 31;
    auto eq = heston<T>(Sd, Vd, Td, 100.0, 2.0, 0.2, splitting_method_enum::DouglasRachford,
                        dflt::host_bwd_tlusolver_cn_solver_config_ptr);
    c3_t all(Sd, Vd, Td);
    eq->solve(all);
    c2_t last(Sd, Vd);
    eq->solve(last);
    for (std::size_t l = 0; l < Td; ++l) {
        std::cout << "LAYER " << l << "\n";
        for (std::size_t r = 0; r < Sd; ++r)
            for (std::size_t c = 0; c < Vd; ++c)
                std::cout << all(r, c, l) << (c + 1 < Vd ? "," : "\n");
    }
    std::cout << "FINAL\n";
    for (std::size_t r = 0; r < Sd; ++r)
        for (std::size_t c = 0; c < Vd; ++c)
            std::cout << last(r, c) << (c + 1 < Vd ? "," : "\n");
    return 0;
}

int book()
{
    typedef container_2d<by_enum::Row, double, std::vector, std::allocator<double>> c2_t;
    std::vector<std::shared_ptr<general_heston_equation<double>>> eqs = {
        heston<double>(60, 30, 21, 100.0, 2.0, 0.2, splitting_method_enum::DouglasRachford,
                       dflt::host_bwd_tlusolver_cn_solver_config_ptr),
        heston<double>(90, 40, 35, 80.0, 3.0, -0.5, splitting_method_enum::ModifiedCraigSneyd,
                       dflt::host_bwd_dssolver_cn_solver_config_ptr),
        heston<double>(130, 26, 12, 120.0, 1.0, -0.3, splitting_method_enum::HundsdorferVerwer,
                       dflt::host_fwd_tlusolver_cn_solver_config_ptr)};
    std::vector<c2_t> single, exact, part;
    for (auto &e : eqs) {
        const auto g = e->grid();
        single.emplace_back(g.nx, g.ny);
        exact.emplace_back(g.nx, g.ny);
        part.emplace_back(g.nx, g.ny);
        e->solve(single.back());
    }
    batched_heston_book<double> b_exact(eqs, true), b_part(eqs, false);
    b_exact.solve(exact);
    b_part.solve(part);
    for (std::size_t p = 0; p < eqs.size(); ++p) {
        double de = 0, dp = 0, mx = 0;
        for (std::size_t r = 0; r < single[p].rows(); ++r)
            for (std::size_t c = 0; c < single[p].columns(); ++c) {
                de = std::max(de, std::abs(exact[p](r, c) - single[p](r, c)));
                dp = std::max(dp, std::abs(part[p](r, c) - single[p](r, c)));
                mx = std::max(mx, std::abs(single[p](r, c)));
            }
        std::cout << "problem " << p << " exact_book_maxdiff " << de << " partitioned_book_rel " << dp / mx << "\n";
    }
    return 0;
}

int adjoint()
{
    typedef container_2d<by_enum::Row, double, std::vector, std::allocator<double>> c2_t;
    heston_parameters<double> hp{100.0, 0.02, 0.8, 3.0, 0.2, 0.2};
    const std::size_t Sd = 40, Vd = 24, Td = 13;
    const double k = std::min(0.01, 1.5 / (Vd - 1));
    range<double> sx(0.0, 200.0), sy(0.0, 1.2), tr(0.0, k * (Td - 1));
    auto disc = std::make_shared<pde_discretization_config_2d<double>>(sx, sy, Sd, Vd, tr, Td);
    auto split = std::make_shared<splitting_method_config<double>>(splitting_method_enum::ModifiedCraigSneyd, 0.5);
    auto hints = std::make_shared<grid_config_hints_2d<double>>(100.0, 3.0, 50.0, grid_enum::Nonuniform);
    heston_adjoint<double> adj(hp, disc, split, hints, dflt::host_bwd_tlusolver_cn_solver_config_ptr, 4);
    const c2_t u0 = adj.initial();
    c2_t uT(Sd, Vd);
    adj.forward(u0, uT);
    // bilinear price at (100, 0.04) and its seed dJ/du_out
    const auto xy = adj.grid();
    const auto &xs = xy.first, &ys = xy.second;
    std::size_t i = 0, j = 0;
    while (i + 2 < Sd && xs[i + 1] <= 100.0)
        ++i;
    while (j + 2 < Vd && ys[j + 1] <= 0.04)
        ++j;
    const double wx = (100.0 - xs[i]) / (xs[i + 1] - xs[i]), wy = (0.04 - ys[j]) / (ys[j + 1] - ys[j]);
    c2_t seed(Sd, Vd, 0.0);
    seed(i, j, (1 - wx) * (1 - wy));
    seed(i + 1, j, wx * (1 - wy));
    seed(i, j + 1, (1 - wx) * wy);
    seed(i + 1, j + 1, wx * wy);
    double price = 0;
    for (std::size_t r = 0; r < Sd; ++r)
        for (std::size_t c = 0; c < Vd; ++c)
            price += seed(r, c) * uT(r, c);
    const auto g = adj.backward(seed);
    double dot = 0;
    for (std::size_t r = 0; r < Sd; ++r)
        for (std::size_t c = 0; c < Vd; ++c)
            dot += g.d_initial(r, c) * u0(r, c);
    std::cout << "price " << price << "\n";
    std::cout << "grad " << g.d_kappa() << " " << g.d_theta() << " " << g.d_sigma() << " " << g.d_rho() << " "
              << g.d_rate() << "\n";
    std::cout << "dot " << dot << "\n";
    return 0;
}

const splitting_method_enum SCHEMES[4] = {
    splitting_method_enum::DouglasRachford, splitting_method_enum::CraigSneyd,
    splitting_method_enum::ModifiedCraigSneyd, splitting_method_enum::HundsdorferVerwer};

template <typename T> int source(int scheme, bool fwd, bool ds)
{
    typedef container_2d<by_enum::Row, T, std::vector, std::allocator<T>> c2_t;
    const std::size_t Sd = 40, Vd = 20, Td = 17;
    auto cfg = fwd ? (ds ? dflt::host_fwd_dssolver_cn_solver_config_ptr : dflt::host_fwd_tlusolver_cn_solver_config_ptr)
                   : (ds ? dflt::host_bwd_dssolver_cn_solver_config_ptr : dflt::host_bwd_tlusolver_cn_solver_config_ptr);
    auto eq = heston<T>(Sd, Vd, Td, 100.0, 2.0, -0.4, SCHEMES[scheme], cfg, true);
    c2_t sol(Sd, Vd);
    eq->solve(sol);
    for (std::size_t r = 0; r < Sd; ++r)
        for (std::size_t c = 0; c < Vd; ++c)
            std::cout << sol(r, c) << (c + 1 < Vd ? "," : "\n");
    return 0;
}

int source_book()
{
    typedef container_2d<by_enum::Row, double, std::vector, std::allocator<double>> c2_t;
    std::vector<std::shared_ptr<general_heston_equation<double>>> eqs;
    for (int s = 0; s < 4; ++s)
        eqs.push_back(heston<double>(40 + 8 * s, 20 + 2 * s, 17, 100.0, 2.0, -0.4, SCHEMES[s],
                                     s % 2 ? dflt::host_fwd_tlusolver_cn_solver_config_ptr
                                           : dflt::host_bwd_tlusolver_cn_solver_config_ptr,
                                     true));
    std::vector<c2_t> single, part;
    for (auto &e : eqs) {
        const auto g = e->grid();
        single.emplace_back(g.nx, g.ny);
        part.emplace_back(g.nx, g.ny);
        e->solve(single.back());
    }
    batched_heston_book<double> b_part(eqs, false);
    b_part.solve(part);
    for (std::size_t p = 0; p < eqs.size(); ++p) {
        double dp = 0, mx = 0;
        for (std::size_t r = 0; r < single[p].rows(); ++r)
            for (std::size_t c = 0; c < single[p].columns(); ++c) {
                dp = std::max(dp, std::abs(part[p](r, c) - single[p](r, c)));
                mx = std::max(mx, std::abs(single[p](r, c)));
            }
        std::cout << "problem " << p << " partitioned_book_rel " << dp / mx << "\n";
    }
    return 0;
}

int source_bench()
{
    typedef container_2d<by_enum::Row, double, std::vector, std::allocator<double>> c2_t;
    for (const bool with : {false, true}) {
        std::vector<std::shared_ptr<general_heston_equation<double>>> eqs;
        for (int q = 0; q < 64; ++q)
            eqs.push_back(heston<double>(100, 50, 101, 80.0 + 40.0 * q / 63.0, 2.0, -0.4, SCHEMES[q % 4],
                                         dflt::host_bwd_tlusolver_cn_solver_config_ptr, with));
        batched_heston_book<double> b(eqs, false);
        std::vector<c2_t> out;
        for (auto &e : eqs)
            out.emplace_back(e->grid().nx, e->grid().ny);
        b.solve(out);  // (first solve: autotuning and graph capture)
        double best = 1e30;
        for (int r = 0; r < 5; ++r) {
            const auto t0 = std::chrono::steady_clock::now();
            b.solve(out);
            best = std::min(best, std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
        }
        std::cout << "source_bench " << (with ? "with" : "without") << " " << best << " " << 64 * 100 * 50 * 100
                  << "\n";
    }
    return 0;
}

int main(int argc, char **argv)
{
    if (argc < 2) {
        std::cerr << "usage: dropin_extras <levels double|float | book | adjoint | source_book | source_bench |\n"
                     "                     source <0-3> <fwd|bwd> <double|float> <tlu|ds>>\n";
        return 2;
    }
    const std::string what = argv[1];
    const bool flt = argc > 2 && std::string(argv[2]) == "float";
    std::cout << std::setprecision(flt ? 9 : 17);
    try {
        if (what == "levels")
            return flt ? levels<float>() : levels<double>();
        if (what == "book")
            return book();
        if (what == "adjoint")
            return adjoint();
        if (what == "source_book")
            return source_book();
        if (what == "source_bench")
            return source_bench();
        if (what == "source" && argc > 5) {
            std::cout << std::setprecision(std::string(argv[4]) == "float" ? 9 : 17);
            const int sch = std::atoi(argv[2]);
            const bool fwd = std::string(argv[3]) == "fwd", ds = argc > 5 && std::string(argv[5]) == "ds";
            if (sch < 0 || sch > 3)
                return 2;
            return std::string(argv[4]) == "float" ? source<float>(sch, fwd, ds) : source<double>(sch, fwd, ds);
        }
    } catch (const std::exception &e) {
        std::cerr << "error: " << e.what() << "\n";
        return 1;
    }
    std::cerr << "unknown mode\n";
    return 2;
}
