// lss_heat1d_sampler.hpp -- turns reference-API 1-D heat problem data (closures) into an
// lss_heat1d_desc.
//
// general_heat_equation::solve evaluates its transformed coefficients, boundaries, source and initial
// condition through std::function closures at the grid values zeta_i = h*i and the loop times.  The
// GPU cannot call closures, so the host samples them here, in precision T and in the reference's
// operand order, so that the sampled values carry the bits the reference computes:
//   grid map and Jacobians          discretization/lss_grid_transform_config.hpp:45-97
//   transformed a, b, c, init, src  pde_solvers/transformation/lss_heat_data_transform.hpp:47-90
//   A, B, D                         one_dimensional/heat_type/implicit_coefficients/
//                                   lss_1d_general_heat_equation_coefficients.hpp:36-67
//   boundary transform              transformation/lss_boundary_transform.hpp:54-96
//   loop times                      lss_1d_general_heat_equation_implicit_kernel.hpp:83-122
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <stdexcept>
#include <vector>

#include "lss_adi_sampler.hpp"
#include "lss_heat1d.h"

namespace lss_amd_detail {

// u_t = a u_xx + b u_x + c u (+ src), a boundary pair of types LSS_BC_* (value(t); Robin also lin(t))
template <typename T> struct heat1d_data {
    std::function<T(T, T)> a, b, c;
    std::function<T(T, T)> src;  // empty: no heat source
    std::function<T(T)> init;
    int lo_type = LSS_BC_DIRICHLET, hi_type = LSS_BC_DIRICHLET;
    std::function<T(T)> lo_val, lo_lin, hi_val, hi_lin;
};

template <typename T> struct heat1d_spec {
    T x_lo, x_hi, t_lo, t_hi;
    std::size_t n, nt;  // number of space points, number of time points
    T accumulation_point, alpha_scale;
    bool nonuniform;
    int implicit_scheme;  // LSS_EULER | LSS_CRANK_NICOLSON
    int tridiagonal;
    int direction;
};

// grid_transform_config<One> (lss_grid_transform_config.hpp:45-97)
template <typename T> struct grid_map1 {
    T init, alpha, c0, c1, c_diff;
    explicit grid_map1(const heat1d_spec<T> &g)
    {
        const T low = g.x_lo, high = g.x_hi, point = g.accumulation_point;
        alpha = high - low;
        if (g.nonuniform)
            alpha = (point - low) / g.alpha_scale;
        init = point;
        c0 = std::asinh((high - point) / alpha);
        c1 = std::asinh((low - point) / alpha);
        c_diff = (c0 - c1);
    }
    T arg(T zeta) const { return c0 * zeta + c1 * (T(1) - zeta); }
    T x(T zeta) const { return (init + alpha * std::sinh(arg(zeta))); }
    T dx(T zeta) const { return (alpha * c_diff * std::cosh(arg(zeta))); }
    T ddx(T zeta) const { return (alpha * c_diff * c_diff * std::sinh(arg(zeta))); }
};

template <typename T> struct heat1d_sampled {
    std::size_t n = 0, steps = 0;
    int tridiagonal = LSS_THOMAS_LU, direction = LSS_BACKWARD, lo_type = LSS_BC_DIRICHLET, hi_type = LSS_BC_DIRICHLET;
    T theta{}, k{};
    std::vector<T> times, next_times;  // loop time and next time of each step
    std::vector<T> xs;                 // X(zeta_i)
    std::vector<T> init;
    std::vector<T> coeff, bnd, source;
    std::size_t coeff_sets = 0, bnd_sets = 0;

    lss_heat1d_desc desc(int line_solver) const
    {
        lss_heat1d_desc d;
        std::memset(&d, 0, sizeof(d));
        d.fp = sizeof(T) == 8 ? LSS_FP64 : LSS_FP32;
        d.tridiagonal = tridiagonal;
        d.line_solver = line_solver;
        d.lo_type = lo_type;
        d.hi_type = hi_type;
        d.direction = direction;
        d.n = (int64_t)n;
        d.steps = (int64_t)steps;
        d.theta = theta;
        d.k = k;
        d.coeff_sets = (int64_t)coeff_sets;
        d.coeff = coeff.data();
        d.bnd_sets = (int64_t)bnd_sets;
        d.bnd = bnd.data();
        d.source = source.empty() ? nullptr : source.data();
        return d;
    }
};

// Theta, time step, grid, loop times and initial level (general_heat_equation::solve,
// lss_1d_general_heat_equation.hpp:148-160; theta: implicit_kernel.hpp:931-943, which accepts Euler
// and Crank-Nicolson only).
template <typename T> void heat1d_grid(const heat1d_data<T> &u, const heat1d_spec<T> &g, heat1d_sampled<T> &s)
{
    const T one = T(1);
    if (g.n < 4 || g.nt < 2)
        throw std::invalid_argument("1-D heat problem needs at least 4 space points and 2 time points");
    if (g.implicit_scheme == LSS_EULER)
        s.theta = one;
    else if (g.implicit_scheme == LSS_CRANK_NICOLSON)
        s.theta = T(0.5);
    else
        throw std::invalid_argument("Unreachable");
    s.n = g.n;
    s.steps = g.nt - 1;
    s.tridiagonal = g.tridiagonal;
    s.direction = g.direction;
    s.lo_type = u.lo_type;
    s.hi_type = u.hi_type;
    s.k = ((g.t_hi - g.t_lo) / static_cast<T>(g.nt - 1));
    const grid_map1<T> gm(g);
    const T step = one / static_cast<T>(g.n - 1);
    s.xs.resize(g.n);
    s.init.resize(g.n);
    for (std::size_t i = 0; i < g.n; ++i) {
        s.xs[i] = gm.x(step * static_cast<T>(i));
        s.init[i] = u.init(s.xs[i]);
    }
    // heat_time_loop::run: time and next_time advance by k each step (:83-122)
    s.times.resize(s.steps);
    s.next_times.resize(s.steps);
    const T k = s.k;
    T t = g.direction == LSS_FORWARD ? g.t_lo + k : g.t_hi - k;
    T tn = g.direction == LSS_FORWARD ? t + k : t - k;
    for (std::size_t q = 0; q < s.steps; ++q) {
        s.times[q] = t;
        s.next_times[q] = tn;
        if (g.direction == LSS_FORWARD) {
            t += k;
            tn += k;
        } else {
            t -= k;
            tn -= k;
        }
    }
}

// One coefficient set at loop time t: A, B, D at every grid point (n each).
template <typename T>
void heat1d_coeff_set(const heat1d_data<T> &u, const heat1d_spec<T> &g, T t, T *out)
{
    const T one = T(1), two = T(2), half = T(0.5);
    const grid_map1<T> gm(g);
    const std::size_t n = g.n;
    const T k = ((g.t_hi - g.t_lo) / static_cast<T>(g.nt - 1));
    const T h = one / static_cast<T>(n - 1);
    const T lambda = k / (h * h), gamma = k / (two * h), delta = half * k;
    for (std::size_t i = 0; i < n; ++i) {
        const T zeta = h * static_cast<T>(i);
        const T x = gm.x(zeta), a_der = gm.dx(zeta), b_der = gm.ddx(zeta);
        const T A = u.a(t, x), B = u.b(t, x), C = u.c(t, x);
        // heat_data_transform<One> (lss_heat_data_transform.hpp:66-84)
        const T ta = (A / (a_der * a_der));
        const T first = B / a_der;
        const T second = (A * b_der) / (a_der * a_der * a_der);
        const T tb = (first - second);
        const T tc = C;
        out[i] = (lambda * ta - gamma * tb);
        out[n + i] = (lambda * ta - delta * tc);
        out[2 * n + i] = (lambda * ta + gamma * tb);
    }
}

// One boundary set at time t: lower lin, lower value, upper lin, upper value (Neumann and Robin parts
// times X'(0) / X'(1), boundary_transform<One>)
template <typename T>
void heat1d_bnd_set(const heat1d_data<T> &u, const heat1d_spec<T> &g, T t, T *out)
{
    const grid_map1<T> gm(g);
    const T d0 = gm.dx(T(0)), d1 = gm.dx(T(1));
    out[0] = out[2] = T(0);
    out[1] = u.lo_val(t);
    out[3] = u.hi_val(t);
    if (u.lo_type == LSS_BC_NEUMANN) {
        out[1] = out[1] * d0;
    } else if (u.lo_type == LSS_BC_ROBIN) {
        out[0] = u.lo_lin(t) * d0;
        out[1] = out[1] * d0;
    }
    if (u.hi_type == LSS_BC_NEUMANN) {
        out[3] = out[3] * d1;
    } else if (u.hi_type == LSS_BC_ROBIN) {
        out[2] = u.hi_lin(t) * d1;
        out[3] = out[3] * d1;
    }
}

// Everything of one problem; coefficient and boundary sets follow the time mode (TIME_AUTO: one set
// when every loop time gives the first set's bits).  The source, when set, is sampled at every step's
// time and next time.
template <typename T>
void sample_heat1d(const heat1d_data<T> &u, const heat1d_spec<T> &g, heat1d_sampled<T> &s, time_mode coeff_time,
                   time_mode bnd_time)
{
    heat1d_grid(u, g, s);
    const std::size_t n = g.n, steps = s.steps;
    auto sample = [&](time_mode mode, std::size_t stride, auto fill, std::vector<T> &out) -> std::size_t {
        if (mode == TIME_INDEPENDENT || steps == 1) {
            out.assign(stride, T(0));
            fill(s.times[0], out.data());
            return 1;
        }
        std::size_t q = 1;
        if (mode == TIME_AUTO) {
            std::vector<T> first(stride), cur(stride);
            fill(s.times[0], first.data());
            for (; q < steps; ++q) {
                fill(s.times[q], cur.data());
                if (std::memcmp(first.data(), cur.data(), stride * sizeof(T)) != 0)
                    break;
            }
            if (q == steps) {
                out = std::move(first);
                return 1;
            }
            out.assign(steps * stride, T(0));
            for (std::size_t r = 0; r < q; ++r)
                std::memcpy(out.data() + r * stride, first.data(), stride * sizeof(T));
            std::memcpy(out.data() + q * stride, cur.data(), stride * sizeof(T));
            ++q;
        } else {
            out.assign(steps * stride, T(0));
            fill(s.times[0], out.data());
        }
        for (; q < steps; ++q)
            fill(s.times[q], out.data() + q * stride);
        return steps;
    };
    s.coeff_sets = sample(coeff_time, 3 * n, [&](T t, T *o) { heat1d_coeff_set(u, g, t, o); }, s.coeff);
    s.bnd_sets = sample(bnd_time, 4, [&](T t, T *o) { heat1d_bnd_set(u, g, t, o); }, s.bnd);
    s.source.clear();
    if (u.src) {
        const grid_map1<T> gm(g);
        const T h = T(1) / static_cast<T>(n - 1);
        s.source.resize(steps * 2 * n);
        for (std::size_t q = 0; q < steps; ++q)
            for (std::size_t i = 0; i < n; ++i) {
                const T x = gm.x(h * static_cast<T>(i));
                s.source[q * 2 * n + i] = u.src(s.times[q], x);
                s.source[q * 2 * n + n + i] = u.src(s.next_times[q], x);
            }
    }
}

}  // namespace lss_amd_detail
