// lss_heat1d.hip -- the 1-D implicit heat / Black-Scholes solver: the exact kernel (one lane per problem,
// the reference's sequential Thomas LU / double-sweep order), the host plan and the C-ABI of
// include/lss_heat1d.h.
//
// The exact kernel runs a problem's whole time loop in one lane: per step it forms each row of the
// theta scheme (implicit_heat_scheme::rhs / rhs_source and heat_implicit_solver_method::split,
// pde_solvers/one_dimensional/heat_type/solver_method/lss_heat_implicit_solver_method.hpp:36-210) and
// sweeps it as thomas_lu_solver (lss_thomas_lu_solver.hpp:125-191 with lss_thomas_lu_boundary.hpp:
// 114-205) or double_sweep_solver (lss_double_sweep_solver.hpp:112-155 with
// lss_double_sweep_boundary.hpp:88-190) does, every operation in the reference's order (this
// translation unit is compiled without contraction), so the result equals the reference's host
// result bit for bit.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "lss/lss_heat1d_sampler.hpp"
#include "lss_capi_common.hpp"
#include "lss_heat1d.h"
#include "lss_heat1d_device.hpp"

namespace lssamd {

template <typename T, int M>
__global__ void k_h1_wave(const H1Prob<T> *, int, const T *const *, T *const *, T *const *);

// boundary value pair of a step
template <typename T> struct H1Bc {
    int type;
    T lin, val;
};

template <typename T>
__global__ void __launch_bounds__(H1_EXACT_TILE) k_h1_exact(const H1Prob<T> *__restrict__ probs, int count,
                                                             const T *const *__restrict__ uin,
                                                             T *const *__restrict__ uout, T *const *__restrict__ surf)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= count)
        return;
    const H1Prob<T> &P = probs[p];
    const int n = P.n, N = n - 1, steps = P.steps;
    const T one = (T)1.0, two = (T)2.0, mone = (T)-1.0, theta = P.theta, k = P.k, h = P.step;
    T *prev = P.w0, *next = P.w1, *aux = P.w2;
    const T *in = uin[p];
    for (int g = 0; g < n; ++g)
        prev[g] = in[g];
    T *sf = surf ? surf[p] : nullptr;
    const bool fwd = P.direction == H1_FORWARD;
    if (sf)
        for (int g = 0; g < n; ++g)
            sf[(size_t)(fwd ? 0 : steps) * n + g] = prev[g];
    for (int st = 0; st < steps; ++st) {
        const T *cs = P.coeff + (size_t)(P.coeff_sets > 1 ? st : 0) * 3 * (size_t)n;
        const T *A = cs, *B = cs + n, *D = cs + 2 * n;
        const T *bs = P.bnd + (P.bnd_sets > 1 ? st : 0) * 4;
        const H1Bc<T> lo{P.lo_type, bs[0], bs[1]}, hi{P.hi_type, bs[2], bs[3]};
        const T *src = P.src ? P.src + (size_t)st * 2 * (size_t)n : nullptr;
        const T *srcn = src ? src + n : nullptr;
        // rows of split (:200-210): low = -theta A, diag = 1 + 2 theta B, high = -theta D
        auto ra = [&](int t) { return (-theta * A[t]); };
        auto rb = [&](int t) { return (one + two * theta * B[t]); };
        auto rc = [&](int t) { return (-theta * D[t]); };
        // right-hand side of row t (rhs / rhs_source, :36-171); rows 0 and N only for Neumann / Robin ends
        auto rf = [&](int t) -> T {
            T r;
            if (t == 0) {
                const T beta = two * h * lo.val;
                if (lo.type == BC_NEUMANN) {
                    r = (one - theta) * beta * A[0] + (one - two * (one - theta) * B[0]) * prev[0] +
                        (one - theta) * (A[0] + D[0]) * prev[1];
                } else {
                    const T alpha = two * h * lo.lin;
                    r = (one - theta) * beta * A[0] + (one - (one - theta) * (two * B[0] - alpha * A[0])) * prev[0] +
                        (one - theta) * (A[0] + D[0]) * prev[1];
                }
                if (src)
                    r = r + theta * k * srcn[0] + (one - theta) * k * src[0];
            } else if (t == N) {
                const T delta = two * h * hi.val;
                if (hi.type == BC_NEUMANN) {
                    r = (one - theta) * (A[N] + D[N]) * prev[N - 1] + (one - two * (one - theta) * B[N]) * prev[N] -
                        (one - theta) * delta * D[N];
                } else {
                    const T gamma = two * h * hi.lin;
                    r = (one - theta) * (A[N] + D[N]) * prev[N - 1] +
                        (one - (one - theta) * (two * B[N] + gamma * D[N])) * prev[N] - (one - theta) * delta * D[N];
                }
                if (src)
                    r = r + theta * k * srcn[N] + (one - theta) * k * src[N];
            } else {
                r = (D[t] * (one - theta) * prev[t + 1]) + ((one - two * B[t] * (one - theta)) * prev[t]) +
                    (A[t] * (one - theta) * prev[t - 1]);
                if (src)
                    r = r + k * (theta * srcn[t] + (one - theta) * src[t]);
            }
            return r;
        };
        if (P.tri == TRI_DSWEEP) {
            // double_sweep_solver: L in aux, K in next, then the solution in place in next
            T *L = aux, *K = next;
            int start;
            T k0, l0;
            if (lo.type == BC_DIRICHLET) {
                start = 1;
                k0 = lo.val;
                l0 = (T)0;
            } else if (lo.type == BC_NEUMANN) {
                start = 0;
                k0 = (rf(0) - ra(0) * h * two * lo.val) / rb(0);
                l0 = mone * (ra(0) + rc(0)) / rb(0);
            } else {
                const T tmp = rb(0) + ra(0) * h * two * lo.lin;
                start = 0;
                k0 = (rf(0) - ra(0) * h * two * lo.val) / tmp;
                l0 = mone * (ra(0) + rc(0)) / tmp;
            }
            const int end = hi.type == BC_DIRICHLET ? n - 2 : n - 1;
            L[0] = l0;
            K[0] = k0;
            for (int t = 1; t <= end; ++t) {
                const T at = ra(t);
                const T tmp = rb(t) + (at * L[t - 1]);
                L[t] = mone * rc(t) / tmp;
                K[t] = (rf(t) - (at * K[t - 1])) / tmp;
            }
            if (hi.type == BC_DIRICHLET) {
                next[N] = hi.val;
            } else if (hi.type == BC_NEUMANN) {
                const T cst = two * h * hi.val;
                next[N] = (L[N] * (K[N - 1] - cst) + K[N]) / (one - L[N] * L[N - 1]);
            } else {
                const T lin = two * h * hi.lin, cst = two * h * hi.val;
                next[N] = (L[N] * (K[N - 1] - cst) + K[N]) / (one - L[N] * (L[N - 1] - lin));
            }
            for (int t = N - 1; t >= start; --t)
                next[t] = (L[t] * next[t + 1]) + K[t];
            if (start == 1)
                next[0] = lo.val;
        } else {
            // thomas_lu_solver: gamma in aux, x in next, back-substitution in place in next
            T *gam = aux, *x = next;
            int start, end;
            T beta0, gamma0, r0, z0, an, bn, rn;
            if (lo.type == BC_DIRICHLET) {
                start = 1;
                beta0 = rb(1);
                gamma0 = rc(1) / beta0;
                r0 = rf(1) - ra(1) * lo.val;
                z0 = r0 / beta0;
            } else if (lo.type == BC_NEUMANN) {
                const T cst = two * h * lo.val;
                start = 0;
                beta0 = rb(0);
                gamma0 = (ra(0) + rc(0)) / beta0;
                r0 = rf(0) - ra(0) * cst;
                z0 = r0 / beta0;
            } else {
                const T lin = two * h * lo.lin, cst = two * h * lo.val;
                start = 0;
                beta0 = rb(0) + ra(0) * lin;
                gamma0 = (ra(0) + rc(0)) / beta0;
                r0 = rf(0) - ra(0) * cst;
                z0 = r0 / beta0;
            }
            if (hi.type == BC_DIRICHLET) {
                end = n - 2;
                an = ra(N - 1);
                bn = rb(N - 1);
                rn = rf(N - 1) - rc(N - 1) * hi.val;
            } else if (hi.type == BC_NEUMANN) {
                const T cst = two * h * hi.val;
                end = n - 1;
                an = ra(N) + rc(N);
                bn = rb(N);
                rn = rf(N) + rc(N) * cst;
            } else {
                const T lin = two * h * hi.lin, cst = two * h * hi.val;
                end = n - 1;
                an = ra(N) + rc(N);
                bn = rb(N) - rc(N) * lin;
                rn = rf(N) + rc(N) * cst;
            }
            gam[start] = gamma0;
            x[start] = z0;
            for (int t = start + 1; t < end; ++t) {
                const T at = ra(t);
                const T beta = rb(t) - (at * gam[t - 1]);
                gam[t] = rc(t) / beta;
                x[t] = (rf(t) - (at * x[t - 1])) / beta;
            }
            const T beta_end = bn - (an * gam[end - 1]);
            x[end] = (rn - (an * x[end - 1])) / beta_end;
            for (int t = end - 1; t >= start; --t)
                x[t] = x[t] - (gam[t] * x[t + 1]);
            if (start == 1)
                x[0] = lo.val;
            if (end == N - 1)
                x[N] = hi.val;
        }
        T *tmp = prev;
        prev = next;
        next = tmp;
        if (sf)
            for (int g = 0; g < n; ++g)
                sf[(size_t)(fwd ? st + 1 : steps - 1 - st) * n + g] = prev[g];
    }
    T *out = uout[p];
    for (int g = 0; g < n; ++g)
        out[g] = prev[g];
}

template __global__ void k_h1_exact<double>(const H1Prob<double> *, int, const double *const *, double *const *,
                                            double *const *);
template __global__ void k_h1_exact<float>(const H1Prob<float> *, int, const float *const *, float *const *,
                                           float *const *);

}  // namespace lssamd

using namespace lssamd;
using namespace lssamd_host;

struct lss_heat1d_plan {
    virtual ~lss_heat1d_plan() = default;
    int fp = LSS_FP64;
    int64_t count = 0;
    std::vector<int64_t> n, steps;
    std::vector<std::vector<char>> init_host, grid_host;  // model plans only
    virtual void run(const void *const *u_in, void *const *u_out, void *const *surf, hipStream_t st, bool host) = 0;
};

namespace {

template <typename T> struct H1Plan final : lss_heat1d_plan {
    std::vector<H1Prob<T>> probs;
    dev_buffer mem, probs_dev, table_dev;
    int line_solver = LSS_LINE_EXACT;
    int pm = 1;  // chunk bound of the partitioned kernel
    std::vector<const void *> table;  // u_in, u_out, surfaces pointers of the last device run
    bool table_valid = false;

    static void validate(const lss_heat1d_desc &d)
    {
        require(d.fp == (sizeof(T) == 8 ? LSS_FP64 : LSS_FP32), "all problems of a plan share fp");
        require(d.n >= 4 && d.n <= (1 << 24), "n must be in [4, 2^24]");
        require(d.steps >= 1 && d.steps < (1 << 30), "steps must be >= 1");
        require(d.coeff_sets == 1 || d.coeff_sets == d.steps, "coeff_sets must be 1 or steps");
        require(d.bnd_sets == 1 || d.bnd_sets == d.steps, "bnd_sets must be 1 or steps");
        require(d.coeff && d.bnd, "null coefficient or boundary array");
        for (int t : {d.lo_type, d.hi_type})
            require(t == LSS_BC_DIRICHLET || t == LSS_BC_NEUMANN || t == LSS_BC_ROBIN,
                    "boundary type must be LSS_BC_DIRICHLET, LSS_BC_NEUMANN or LSS_BC_ROBIN");
        require(d.direction == LSS_FORWARD || d.direction == LSS_BACKWARD, "direction must be Forward or Backward");
        require(d.line_solver == LSS_LINE_EXACT || d.line_solver == LSS_LINE_PARTITIONED, "unknown line_solver");
        if (d.line_solver == LSS_LINE_PARTITIONED) {
            const int64_t nu = d.n - (d.lo_type == LSS_BC_DIRICHLET) - (d.hi_type == LSS_BC_DIRICHLET);
            require(nu <= 64 * 16, "the partitioned 1-D solver takes at most 1024 unknowns per line");
        }
    }

    void build(const lss_heat1d_desc *d, int64_t cnt)
    {
        require(cnt >= 1 && cnt < (1 << 30), "count must be in [1, 2^30)");
        count = cnt;
        line_solver = d[0].line_solver;
        size_t total = 0;
        std::vector<size_t> off(cnt);
        auto round = [](size_t x) { return (x + 63) / 64 * 64; };
        for (int64_t q = 0; q < cnt; ++q) {
            validate(d[q]);
            require(d[q].line_solver == line_solver, "all problems of a plan share line_solver");
            const size_t nq = (size_t)d[q].n;
            off[q] = total;
            total += round(3 * nq * (size_t)d[q].coeff_sets) + round(4 * (size_t)d[q].bnd_sets) +
                     (d[q].source ? round(2 * nq * (size_t)d[q].steps) : 0) +
                     (line_solver == LSS_LINE_EXACT ? 3 * round(nq) : 0);
            const int64_t nu = d[q].n - (d[q].lo_type == LSS_BC_DIRICHLET) - (d[q].hi_type == LSS_BC_DIRICHLET);
            while (pm * 64 < nu && pm < 16)
                pm *= 2;
        }
        mem.alloc(total * sizeof(T));
        probs.resize(cnt);
        n.resize(cnt);
        steps.resize(cnt);
        T *base = (T *)mem.p;
        for (int64_t q = 0; q < cnt; ++q) {
            const lss_heat1d_desc &e = d[q];
            const size_t nq = (size_t)e.n;
            H1Prob<T> &P = probs[q];
            std::memset(&P, 0, sizeof(P));
            P.n = (int32_t)e.n;
            P.steps = (int32_t)e.steps;
            P.tri = e.tridiagonal == LSS_DOUBLE_SWEEP ? TRI_DSWEEP : TRI_THOMAS;
            P.lo_type = e.lo_type;
            P.hi_type = e.hi_type;
            P.direction = e.direction == LSS_FORWARD ? H1_FORWARD : H1_BACKWARD;
            P.coeff_sets = (int32_t)e.coeff_sets;
            P.bnd_sets = (int32_t)e.bnd_sets;
            P.theta = (T)e.theta;
            P.k = (T)e.k;
            P.step = (T)1 / (T)(e.n - 1);  // grid_config<One>: one / (n - 1)
            T *o = base + off[q];
            T *c = o;
            o += round(3 * nq * (size_t)e.coeff_sets);
            T *b = o;
            o += round(4 * (size_t)e.bnd_sets);
            LSS_HIPCHK(hipMemcpy(c, e.coeff, 3 * nq * (size_t)e.coeff_sets * sizeof(T), hipMemcpyHostToDevice));
            LSS_HIPCHK(hipMemcpy(b, e.bnd, 4 * (size_t)e.bnd_sets * sizeof(T), hipMemcpyHostToDevice));
            P.coeff = c;
            P.bnd = b;
            if (e.source) {
                LSS_HIPCHK(hipMemcpy(o, e.source, 2 * nq * (size_t)e.steps * sizeof(T), hipMemcpyHostToDevice));
                P.src = o;
                o += round(2 * nq * (size_t)e.steps);
            }
            if (line_solver == LSS_LINE_EXACT) {
                P.w0 = o;
                P.w1 = o + round(nq);
                P.w2 = o + 2 * round(nq);
            }
            n[q] = e.n;
            steps[q] = e.steps;
        }
        probs_dev.alloc(cnt * sizeof(H1Prob<T>));
        LSS_HIPCHK(hipMemcpy(probs_dev.p, probs.data(), cnt * sizeof(H1Prob<T>), hipMemcpyHostToDevice));
        table_dev.alloc(3 * cnt * sizeof(void *));
    }

    template <int M> void launch_wave(hipStream_t st, const T *const *ui, T *const *uo, T *const *sf)
    {
        const unsigned blocks = (unsigned)((count + H1_WAVES - 1) / H1_WAVES);
        hipLaunchKernelGGL((k_h1_wave<T, M>), dim3(blocks), dim3(64 * H1_WAVES), 0, st,
                           (const H1Prob<T> *)probs_dev.p, (int)count, ui, uo, sf);
    }

    void launch(hipStream_t st, bool with_surf)
    {
        void **tb = (void **)table_dev.p;
        const T *const *ui = (const T *const *)tb;
        T *const *uo = (T *const *)(tb + count);
        T *const *sf = with_surf ? (T *const *)(tb + 2 * count) : nullptr;
        if (line_solver == LSS_LINE_EXACT) {
            const unsigned blocks = (unsigned)((count + H1_EXACT_TILE - 1) / H1_EXACT_TILE);
            hipLaunchKernelGGL(k_h1_exact<T>, dim3(blocks), dim3(H1_EXACT_TILE), 0, st, (const H1Prob<T> *)probs_dev.p,
                               (int)count, ui, uo, sf);
        } else {
            switch (pm) {
            case 1: launch_wave<1>(st, ui, uo, sf); break;
            case 2: launch_wave<2>(st, ui, uo, sf); break;
            case 4: launch_wave<4>(st, ui, uo, sf); break;
            case 8: launch_wave<8>(st, ui, uo, sf); break;
            default: launch_wave<16>(st, ui, uo, sf); break;
            }
        }
        LSS_HIPCHK(hipGetLastError());
    }

    // the pointer table (u_in, u_out, surfaces) on the device; rewritten only when it changes
    void set_table(const void *const *u_in, void *const *u_out, void *const *surf, hipStream_t st)
    {
        std::vector<const void *> t(3 * count, nullptr);
        for (int64_t q = 0; q < count; ++q) {
            t[q] = u_in[q];
            t[count + q] = u_out[q];
            t[2 * count + q] = surf ? surf[q] : nullptr;
        }
        if (table_valid && t == table)
            return;
        LSS_HIPCHK(hipStreamSynchronize(st));  // an earlier launch may still read the old table
        LSS_HIPCHK(hipMemcpy(table_dev.p, t.data(), t.size() * sizeof(void *), hipMemcpyHostToDevice));
        table = std::move(t);
        table_valid = true;
    }

    void run(const void *const *u_in, void *const *u_out, void *const *surf, hipStream_t st, bool host) override
    {
        require(u_in && u_out, "null u_in / u_out array");
        for (int64_t q = 0; q < count; ++q)
            require(u_in[q] && u_out[q], "null surface pointer");
        bool with_surf = false;
        if (surf)
            for (int64_t q = 0; q < count; ++q)
                with_surf = with_surf || surf[q] != nullptr;
        if (!host) {
            set_table(u_in, u_out, with_surf ? surf : nullptr, st);
            launch(st, with_surf);
            return;
        }
        // host pointers: stage everything through one device allocation
        size_t tot = 0;
        std::vector<size_t> oi(count), oo(count), os(count, 0);
        for (int64_t q = 0; q < count; ++q) {
            oi[q] = tot;
            tot += (size_t)n[q];
            oo[q] = tot;
            tot += (size_t)n[q];
            if (with_surf && surf[q]) {
                os[q] = tot;
                tot += (size_t)n[q] * (size_t)(steps[q] + 1);
            }
        }
        dev_buffer stage;
        stage.alloc(tot * sizeof(T));
        T *sb = (T *)stage.p;
        std::vector<const void *> di(count);
        std::vector<void *> dout(count), ds(count, nullptr);
        for (int64_t q = 0; q < count; ++q) {
            LSS_HIPCHK(hipMemcpy(sb + oi[q], u_in[q], (size_t)n[q] * sizeof(T), hipMemcpyHostToDevice));
            di[q] = sb + oi[q];
            dout[q] = sb + oo[q];
            if (with_surf && surf[q])
                ds[q] = sb + os[q];
        }
        set_table(di.data(), dout.data(), with_surf ? ds.data() : nullptr, nullptr);
        launch(nullptr, with_surf);
        LSS_HIPCHK(hipDeviceSynchronize());
        table_valid = false;  // the staging buffer goes away
        for (int64_t q = 0; q < count; ++q) {
            LSS_HIPCHK(hipMemcpy(u_out[q], sb + oo[q], (size_t)n[q] * sizeof(T), hipMemcpyDeviceToHost));
            if (with_surf && surf[q])
                LSS_HIPCHK(hipMemcpy(surf[q], sb + os[q], (size_t)n[q] * (size_t)(steps[q] + 1) * sizeof(T),
                                     hipMemcpyDeviceToHost));
        }
    }
};

lss_heat1d_plan *make_plan(const lss_heat1d_desc *d, int64_t cnt)
{
    if (d[0].fp == LSS_FP64) {
        auto *p = new H1Plan<double>();
        p->fp = LSS_FP64;
        try {
            p->build(d, cnt);
        } catch (...) {
            delete p;
            throw;
        }
        return p;
    }
    require(d[0].fp == LSS_FP32, "fp must be LSS_FP64 or LSS_FP32");
    auto *p = new H1Plan<float>();
    p->fp = LSS_FP32;
    try {
        p->build(d, cnt);
    } catch (...) {
        delete p;
        throw;
    }
    return p;
}

// Black-Scholes data of the reference's 1-D tests (lss_black_scholes_equation_t.hpp:66-80): the
// captured constants are double, each closure computes in double and rounds on return.
template <typename T> lss_amd_detail::heat1d_data<T> bs_data(const lss_bs_model &m)
{
    lss_amd_detail::heat1d_data<T> u;
    const double K = m.strike, r = m.rate, sig = m.sig, mat = m.maturity, xh = m.x_hi;
    u.a = [=](T, T x) { return (T)(0.5 * sig * sig * x * x); };
    u.b = [=](T, T x) { return (T)(r * x); };
    u.c = [=](T, T) { return (T)(-r); };
    u.init = [=](T x) { return std::max<T>((T)0.0, (T)(x - K)); };
    u.lo_type = LSS_BC_DIRICHLET;
    u.hi_type = LSS_BC_DIRICHLET;
    u.lo_val = [](T) { return (T)0.0; };
    u.hi_val = [=](T t) { return (T)(xh - K * std::exp(-r * (mat - t))); };
    return u;
}

template <typename T> lss_heat1d_plan *bs_plan(const lss_bs_model *models, int64_t cnt)
{
    std::vector<lss_amd_detail::heat1d_sampled<T>> s(cnt);
    std::vector<lss_heat1d_desc> d(cnt);
    for (int64_t q = 0; q < cnt; ++q) {
        const lss_bs_model &m = models[q];
        require(m.n >= 4 && m.nt >= 2, "model grid too small");
        require(m.implicit_scheme == LSS_EULER || m.implicit_scheme == LSS_CRANK_NICOLSON,
                "the 1-D solver takes the Euler or Crank-Nicolson scheme");
        lss_amd_detail::heat1d_spec<T> g;
        g.x_lo = (T)m.x_lo;
        g.x_hi = (T)m.x_hi;
        g.t_lo = (T)m.t_lo;
        g.t_hi = (T)m.t_hi;
        g.n = (size_t)m.n;
        g.nt = (size_t)m.nt;
        g.accumulation_point = (T)m.accumulation_point;
        g.alpha_scale = (T)m.alpha_scale;
        g.nonuniform = m.grid == LSS_GRID_NONUNIFORM;
        g.implicit_scheme = m.implicit_scheme;
        g.tridiagonal = m.tridiagonal;
        g.direction = m.direction;
        lss_amd_detail::sample_heat1d(bs_data<T>(m), g, s[q], lss_amd_detail::TIME_INDEPENDENT,
                                      lss_amd_detail::TIME_DEPENDENT);
        d[q] = s[q].desc(m.line_solver);
    }
    lss_heat1d_plan *plan = make_plan(d.data(), cnt);
    plan->init_host.resize(cnt);
    plan->grid_host.resize(cnt);
    for (int64_t q = 0; q < cnt; ++q) {
        plan->init_host[q].resize(s[q].init.size() * sizeof(T));
        std::memcpy(plan->init_host[q].data(), s[q].init.data(), plan->init_host[q].size());
        plan->grid_host[q].resize(s[q].xs.size() * sizeof(T));
        std::memcpy(plan->grid_host[q].data(), s[q].xs.data(), plan->grid_host[q].size());
    }
    return plan;
}

}  // namespace

extern "C" {

int lss_heat1d_plan_create(const lss_heat1d_desc *descs, int64_t count, lss_heat1d_plan **plan)
{
    return guarded([&] {
        require(descs && plan && count >= 1, "null descriptor/plan pointer or count < 1");
        *plan = nullptr;
        *plan = make_plan(descs, count);
    });
}

int lss_heat1d_plan_destroy(lss_heat1d_plan *plan)
{
    return guarded([&] { delete plan; });
}

int lss_heat1d_run_dev(lss_heat1d_plan *plan, const void *const *u_in, void *const *u_out, void *const *surfaces,
                       void *stream)
{
    return guarded([&] {
        require(plan != nullptr, "null plan");
        plan->run(u_in, u_out, surfaces, (hipStream_t)stream, false);
    });
}

int lss_heat1d_run_host(lss_heat1d_plan *plan, const void *const *u_in, void *const *u_out, void *const *surfaces)
{
    return guarded([&] {
        require(plan != nullptr, "null plan");
        plan->run(u_in, u_out, surfaces, nullptr, true);
    });
}

int lss_heat1d_problem_shape(const lss_heat1d_plan *plan, int64_t p, int64_t *n, int64_t *steps)
{
    return guarded([&] {
        require(plan != nullptr, "null plan");
        require(p >= 0 && p < plan->count, "problem index out of range");
        if (n)
            *n = plan->n[p];
        if (steps)
            *steps = plan->steps[p];
    });
}

int lss_bs_plan_create(const lss_bs_model *models, int64_t count, int32_t fp, lss_heat1d_plan **plan)
{
    return guarded([&] {
        require(models && plan && count >= 1, "null model/plan pointer or count < 1");
        *plan = nullptr;
        if (fp == LSS_FP64)
            *plan = bs_plan<double>(models, count);
        else if (fp == LSS_FP32)
            *plan = bs_plan<float>(models, count);
        else
            throw lss_error(LSS_ERR_INVALID, "fp must be LSS_FP64 or LSS_FP32");
    });
}

int lss_heat1d_initial(const lss_heat1d_plan *plan, int64_t p, void *u0_host)
{
    return guarded([&] {
        require(plan && u0_host, "null pointer");
        require(p >= 0 && p < plan->count, "problem index out of range");
        require((int64_t)plan->init_host.size() == plan->count, "plan was not created from models");
        std::memcpy(u0_host, plan->init_host[p].data(), plan->init_host[p].size());
    });
}

int lss_heat1d_grid(const lss_heat1d_plan *plan, int64_t p, void *xs_host)
{
    return guarded([&] {
        require(plan && xs_host, "null pointer");
        require(p >= 0 && p < plan->count, "problem index out of range");
        require((int64_t)plan->grid_host.size() == plan->count, "plan was not created from models");
        std::memcpy(xs_host, plan->grid_host[p].data(), plan->grid_host[p].size());
    });
}

}  // extern "C"
