// lss_adi_sampler.hpp -- turns reference-API problem data (closures) into an lss_adi_desc.
//
// general_heston_equation::solve evaluates its transformed coefficients, boundaries and initial
// condition through std::function closures at grid points (ζ_i = h1*i, η_j = h2*j) and loop times.
// The GPU cannot call closures, so the host samples them once here, in precision T and with the
// reference's operand order, so that the sampled values carry the same bits the reference computes:
//   grid map and Jacobians         discretization/lss_grid_transform_config.hpp:118-187
//   transformed a..f, initial data pde_solvers/transformation/lss_heat_data_transform.hpp:174-220
//   M, M~, P, P~, Z, W, C          implicit_coefficients/lss_2d_general_heston_equation_coefficients.hpp:62-98
//   boundary transform             two_dimensional/heat_type/transformation/lss_heston_boundary_transform.hpp:48-62
//   index maps                     discretization/lss_grid_config.hpp:56-59, 110-118
//   loop times                     lss_2d_general_heston_equation_implicit_kernel.hpp:171-206
// Time dependence: coefficient and boundary sets are sampled at every loop time unless the caller
// states they are time independent, or (auto mode) the samples at every loop time agree bit for bit
// with the first, in which case one set is kept.  A heat source is sampled in auto mode.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <vector>

#include "lss_amd.h"

namespace lss_amd_detail {

template <typename T> struct user_data {
    std::function<T(T, T, T)> A, B, C, D, E, F;  // u_t = A u_xx + B u_yy + C u_xy + D u_x + E u_y + F u
    std::function<T(T, T)> init;                 // G(x, y)
    std::function<T(T, T)> h_lower;              // Dirichlet at x_lo: value(t, y)
    std::function<T(T, T)> h_upper;              // Neumann at x_hi: value(t, y)
    std::function<T(T, T)> v_upper;              // Dirichlet at y_hi: value(t, x)
    std::function<T(T, T, T)> src;               // heat source F(t, x, y), or empty
};

template <typename T> struct grid_spec {
    T x_lo, x_hi, y_lo, y_hi, t_lo, t_hi;
    std::size_t nx, ny, nt;
    T accumulation_point, alpha_scale, beta_scale;
    bool nonuniform;
    int splitting;
    T weighting;
    int implicit_scheme;  // implicit_pde_schemes_enum
    int tridiagonal;      // LSS_THOMAS_LU / LSS_DOUBLE_SWEEP / LSS_CUDA_SOLVER (Thomas order)
    int direction;        // LSS_FORWARD / LSS_BACKWARD
};

// How closures are sampled over the loop times: TIME_AUTO samples them at every loop time and keeps
// one set when every set equals the first bit for bit (else all of them); TIME_INDEPENDENT samples
// once at the first loop time; TIME_DEPENDENT keeps one set per step.
enum time_mode { TIME_AUTO = 0, TIME_INDEPENDENT = 1, TIME_DEPENDENT = 2 };

// grid_transform_config_2d (lss_grid_transform_config.hpp:118-149)
template <typename T> struct grid_map {
    T init1, init2, alpha, beta, c0, c1, d, c_diff;
    explicit grid_map(const grid_spec<T> &g)
    {
        const T low1 = g.x_lo, high1 = g.x_hi, low2 = g.y_lo, high2 = g.y_hi, point = g.accumulation_point;
        alpha = high1 - low1;
        beta = high2 - low2;
        if (g.nonuniform) {
            alpha = (point - low1) / g.alpha_scale;
            beta = (high2 - low2) / g.beta_scale;
        }
        init1 = point;
        init2 = low2;
        c0 = std::asinh((high1 - point) / alpha);
        c1 = std::asinh((low1 - point) / alpha);
        d = std::asinh((high2 - low2) / beta);
        c_diff = (c0 - c1);
    }
    T arg1(T zeta) const { return c0 * zeta + c1 * (T(1) - zeta); }
    T x(T zeta) const { return (init1 + alpha * std::sinh(arg1(zeta))); }
    T y(T eta) const { return (init2 + beta * std::sinh(d * eta)); }
    T dx(T zeta) const { return (alpha * c_diff * std::cosh(arg1(zeta))); }
    T dy(T eta) const { return (beta * d * std::cosh(d * eta)); }
    T ddx(T zeta) const { return (alpha * c_diff * c_diff * std::sinh(arg1(zeta))); }
    T ddy(T eta) const { return (beta * d * d * std::sinh(d * eta)); }
};

template <typename T> struct sampled {
    std::size_t nx = 0, ny = 0, steps = 0;
    int splitting = 0, tridiagonal = LSS_THOMAS_LU;
    T theta{}, zeta{}, alpha{}, beta{}, gamma{}, delta{}, ni{}, rho{}, h1{}, k{};
    std::vector<uint32_t> idx1, idx2;
    std::vector<T> times;   // loop time of each step
    std::vector<T> xs, ys;  // X(ζ_i), Y(η_j)
    std::vector<T> coeff;
    std::size_t coeff_sets = 0;
    std::vector<T> bnd;
    std::size_t bnd_sets = 0;
    std::vector<T> src;     // heat source sets [F(t) | F(t_next)] (empty without a source)
    std::size_t src_sets = 0;
    std::vector<T> init;    // nx*ny initial surface

    lss_adi_desc desc(int coeff_mode, int line_solver = LSS_LINE_EXACT) const
    {
        lss_adi_desc d;
        std::memset(&d, 0, sizeof(d));
        d.fp = sizeof(T) == 8 ? LSS_FP64 : LSS_FP32;
        d.splitting = splitting;
        d.tridiagonal = tridiagonal;
        d.coeff_mode = coeff_mode;
        d.line_solver = line_solver;
        d.nx = (int64_t)nx;
        d.ny = (int64_t)ny;
        d.steps = (int64_t)steps;
        d.theta = theta;
        d.weighting = zeta;
        d.alpha = alpha;
        d.beta = beta;
        d.gamma = gamma;
        d.delta = delta;
        d.ni = ni;
        d.rho = rho;
        d.h1 = h1;
        d.idx1 = idx1.data();
        d.idx2 = idx2.data();
        d.coeff_sets = (int64_t)coeff_sets;
        d.coeff = coeff.data();
        d.bnd_sets = (int64_t)bnd_sets;
        d.bnd = bnd.data();
        d.src_sets = (int64_t)src_sets;
        d.src = src_sets ? src.data() : nullptr;
        return d;
    }
};

// Scheme scalars, grid data, index maps and loop times (no closure calls).
template <typename T> void sample_grid(const grid_spec<T> &g, sampled<T> &s)
{
    const T one = T(1), half = T(0.5), quarter = T(0.25);
    s.nx = g.nx;
    s.ny = g.ny;
    s.steps = g.nt - 1;
    s.splitting = g.splitting;
    s.tridiagonal = g.tridiagonal;
    // implicit_kernel.hpp:978-995
    switch (g.implicit_scheme) {
    case LSS_EULER: s.theta = one; break;
    case LSS_THETA_30: s.theta = T(0.3); break;
    case LSS_THETA_80: s.theta = T(0.8); break;
    default: s.theta = half; break;
    }
    // pde_discretization_config::time_step (lss_pde_discretization_config.hpp:140-143)
    const T k = ((g.t_hi - g.t_lo) / static_cast<T>(g.nt - 1));
    const T h_1 = one / static_cast<T>(g.nx - 1), h_2 = one / static_cast<T>(g.ny - 1);
    s.k = k;
    s.alpha = k / (h_1 * h_1);
    s.beta = k / (h_2 * h_2);
    s.gamma = quarter * k / (h_1 * h_2);
    s.delta = half * k / h_1;
    s.ni = half * k / h_2;
    s.rho = k;
    s.zeta = g.weighting;
    s.h1 = h_1;
    const grid_map<T> gm(g);
    s.idx1.resize(g.nx);
    s.xs.resize(g.nx);
    for (std::size_t i = 0; i < g.nx; ++i) {
        const T z = h_1 * static_cast<T>(i);
        s.idx1[i] = static_cast<uint32_t>(static_cast<std::size_t>(z / h_1));
        s.xs[i] = gm.x(z);
    }
    s.idx2.resize(g.ny);
    s.ys.resize(g.ny);
    for (std::size_t j = 0; j < g.ny; ++j) {
        const T e = h_2 * static_cast<T>(j);
        s.idx2[j] = static_cast<uint32_t>(static_cast<std::size_t>(e / h_2));
        s.ys[j] = gm.y(e);
    }
    s.times.resize(s.steps);
    if (g.direction == LSS_FORWARD) {
        T t = g.t_lo + k;
        for (std::size_t q = 0; q < s.steps; ++q) {
            s.times[q] = t;
            t += k;
        }
    } else {
        T t = g.t_hi - k;
        for (std::size_t q = 0; q < s.steps; ++q) {
            s.times[q] = t;
            t -= k;
        }
    }
}

// One FIELDS set at loop time t: 7 nx*ny fields then D, E, F on the v = 0 row.
template <typename T>
void sample_field_set(const user_data<T> &u, const grid_spec<T> &g, const sampled<T> &s, T t, T *out)
{
    const grid_map<T> gm(g);
    const std::size_t nx = g.nx, ny = g.ny, nxy = nx * ny;
    const T two = T(2), half = T(0.5);
    const T h_1 = T(1) / static_cast<T>(nx - 1), h_2 = T(1) / static_cast<T>(ny - 1);
    std::vector<T> z(nx), ax(nx), cx(nx), e(ny), by(ny), dy(ny);
    for (std::size_t i = 0; i < nx; ++i) {
        z[i] = h_1 * static_cast<T>(i);
        ax[i] = gm.dx(z[i]);
        cx[i] = gm.ddx(z[i]);
    }
    for (std::size_t j = 0; j < ny; ++j) {
        e[j] = h_2 * static_cast<T>(j);
        by[j] = gm.dy(e[j]);
        dy[j] = gm.ddy(e[j]);
    }
    for (std::size_t i = 0; i < nx; ++i) {
        const T x = s.xs[i], a_val = ax[i];
        for (std::size_t j = 0; j < ny; ++j) {
            const T y = s.ys[j], b_val = by[j];
            const T A = u.A(t, x, y), B = u.B(t, x, y), C = u.C(t, x, y);
            const T D = u.D(t, x, y), E = u.E(t, x, y), F = u.F(t, x, y);
            // lss_heat_data_transform.hpp:174-214
            const T ta = (A / (a_val * a_val));
            const T tb = (B / (b_val * b_val));
            const T tc = (C / (a_val * b_val));
            const T td = (D / a_val) - ((A * cx[i]) / (a_val * a_val * a_val));
            const T te = (E / b_val) - ((B * dy[j]) / (b_val * b_val * b_val));
            const T tf = F;
            const std::size_t q = i * ny + j;
            // coefficients.hpp:89-95
            out[q] = (s.alpha * ta - s.delta * td);
            out[nxy + q] = (s.beta * tb - s.ni * te);
            out[2 * nxy + q] = (s.alpha * ta + s.delta * td);
            out[3 * nxy + q] = (s.beta * tb + s.ni * te);
            out[4 * nxy + q] = (two * s.alpha * ta - half * s.rho * tf);
            out[5 * nxy + q] = (two * s.beta * tb - half * s.rho * tf);
            out[6 * nxy + q] = tc;
            if (j == 0) {
                out[7 * nxy + i] = td;
                out[7 * nxy + nx + i] = te;
                out[7 * nxy + 2 * nx + i] = tf;
            }
        }
    }
}

// One boundary set at loop time t: v_upper(t, X(ζ_i)), h_lower(t, Y(η_j)), h_upper(t, Y(η_j)) * X'(1).
template <typename T>
void sample_bnd_set(const user_data<T> &u, const grid_spec<T> &g, const sampled<T> &s, T t, T *out)
{
    const grid_map<T> gm(g);
    const std::size_t nx = g.nx, ny = g.ny;
    const T xd1 = gm.dx(T(1));
    for (std::size_t i = 0; i < nx; ++i)
        out[i] = u.v_upper(t, s.xs[i]);
    for (std::size_t j = 0; j < ny; ++j) {
        out[nx + j] = u.h_lower(t, s.ys[j]);
        out[nx + ny + j] = (u.h_upper(t, s.ys[j]) * xd1);
    }
}

template <typename T>
void sample_fields(const user_data<T> &u, const grid_spec<T> &g, sampled<T> &s, time_mode coeff_time,
                   time_mode bnd_time)
{
    sample_grid(g, s);
    const std::size_t nx = g.nx, ny = g.ny;
    // initial condition (discretization.hpp:177-193 of the transformed init, heat_data_transform.hpp:216-220)
    s.init.resize(nx * ny);
    for (std::size_t i = 0; i < nx; ++i)
        for (std::size_t j = 0; j < ny; ++j)
            s.init[i * ny + j] = u.init(s.xs[i], s.ys[j]);

    const std::size_t cstride = 7 * nx * ny + 3 * nx, bstride = nx + 2 * ny, n = s.steps;
    // TIME_AUTO samples every loop time and compares each set with the first: one set is kept only
    // when all of them agree bit for bit (a closure that changes at any loop time is sampled per step)
    auto sample = [&](time_mode mode, std::size_t stride, auto fill, std::vector<T> &out) -> std::size_t {
        if (mode == TIME_INDEPENDENT || n == 1) {
            out.assign(stride, T(0));
            fill(s.times[0], out.data());
            return 1;
        }
        std::size_t q = 1;
        if (mode == TIME_AUTO) {
            std::vector<T> first(stride), cur(stride);
            fill(s.times[0], first.data());
            for (; q < n; ++q) {
                fill(s.times[q], cur.data());
                if (std::memcmp(first.data(), cur.data(), stride * sizeof(T)) != 0)
                    break;
            }
            if (q == n) {
                out = std::move(first);
                return 1;
            }
            out.assign(n * stride, T(0));
            std::memcpy(out.data(), first.data(), stride * sizeof(T));
            std::memcpy(out.data() + q * stride, cur.data(), stride * sizeof(T));
            for (std::size_t r = 1; r < q; ++r)  // equal to the first set
                std::memcpy(out.data() + r * stride, first.data(), stride * sizeof(T));
            ++q;
        } else {
            out.assign(n * stride, T(0));
            fill(s.times[0], out.data());
        }
        for (; q < n; ++q)
            fill(s.times[q], out.data() + q * stride);
        return n;
    };
    auto fill_c = [&](T t, T *o) { sample_field_set(u, g, s, t, o); };
    auto fill_b = [&](T t, T *o) { sample_bnd_set(u, g, s, t, o); };
    s.coeff_sets = sample(coeff_time, cstride, fill_c, s.coeff);
    s.bnd_sets = sample(bnd_time, bstride, fill_b, s.bnd);
    if (u.src) {
        // the source at the loop time t and at the loop's next time (t + k Forward, t - k Backward),
        // at the physical points (X(zeta_i), Y(eta_j)) (lss_heat_data_transform.hpp:164-173)
        auto fill_s = [&](T t, T *o) {
            const T tn = g.direction == LSS_FORWARD ? t + s.k : t - s.k;
            for (std::size_t i = 0; i < nx; ++i)
                for (std::size_t j = 0; j < ny; ++j) {
                    o[i * ny + j] = u.src(t, s.xs[i], s.ys[j]);
                    o[nx * ny + i * ny + j] = u.src(tn, s.xs[i], s.ys[j]);
                }
        };
        s.src_sets = sample(TIME_AUTO, 2 * nx * ny, fill_s, s.src);
    }
}

}  // namespace lss_amd_detail
