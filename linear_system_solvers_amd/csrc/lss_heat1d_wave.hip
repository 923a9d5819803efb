// lss_heat1d_wave.hip -- the 1-D implicit heat / Black-Scholes time loop with one wavefront per
// problem and the whole loop in registers.
//
// Lane l holds the C consecutive unknowns of its chunk of the line (the partition of
// lss_adi_chunk.hpp), the level's values and, for coefficients that do not change in time, A, B, D of
// its points.  Each step forms the theta scheme's right-hand side from the lane's values and its
// neighbours' (one shuffle each way), folds the boundary pair into the end rows, and solves the
// tridiagonal system by the register-resident partitioned elimination with cyclic reduction of the
// chunk interfaces.  HBM traffic per problem is the first level, the last level (and every level when
// the caller asks for the surface) plus the coefficient and boundary sets that vary in time.
//
// The arithmetic of each row follows implicit_heat_scheme::rhs / rhs_source and
// heat_implicit_solver_method::split (pde_solvers/one_dimensional/heat_type/solver_method/
// lss_heat_implicit_solver_method.hpp:36-171, 200-210) and thomas_lu_solver_boundary's folding of the
// end rows (sparse_solvers/tridiagonal/thomas_lu_solver/lss_thomas_lu_boundary.hpp:114-205); multiply-
// adds may be fused here and the elimination order differs from the sequential Thomas sweep.
#pragma clang fp contract(fast)
#include "lss_adi_chunk.hpp"
#include "lss_heat1d_device.hpp"

namespace lssamd {

template <typename T, int M>
__global__ void __launch_bounds__(64 * H1_WAVES) k_h1_wave(const H1Prob<T> *__restrict__ probs, int count,
                                                           const T *const *__restrict__ uin,
                                                           T *const *__restrict__ uout, T *const *__restrict__ surf)
{
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int p = blockIdx.x * H1_WAVES + wave;
    if (p >= count)
        return;  // uniform over the wavefront
    const H1Prob<T> &P = probs[p];
    const int n = P.n, N = n - 1, steps = P.steps;
    const int gs = P.lo_type == BC_DIRICHLET ? 1 : 0, ge = P.hi_type == BC_DIRICHLET ? n - 2 : n - 1;
    const int Nu = ge - gs + 1;  // unknowns of the folded system
    const T one = (T)1, two = (T)2, theta = P.theta, omt = one - theta, k = P.k, h2 = two * P.step;
    const bool has_src = P.src != nullptr;
    const T *in = uin[p];
    T *out = uout[p];
    T *sf = surf ? surf[p] : nullptr;
    with_chunk_len<M>(chunk_len(Nu), [&](auto cc) {
        constexpr int C = decltype(cc)::value;
        const int s0 = 1 + lane * C;  // first unknown (1-based) of the lane's chunk
        // grid index of the lane's t-th point; points past the line read its last point (discarded)
        auto gidx = [&](int t) { return min(gs + s0 + t - 1, N); };
        T u[C], cA[C], cB[C], cD[C];
#pragma unroll
        for (int t = 0; t < C; ++t)
            u[t] = in[gidx(t)];
        T ulo = in[0], uhi = in[N];  // the level's end values (used where the ends are Dirichlet)
        auto load_coeff = [&](int st) {
            const T *cs = P.coeff + (size_t)(P.coeff_sets > 1 ? st : 0) * 3 * (size_t)n;
#pragma unroll
            for (int t = 0; t < C; ++t) {
                const int g = gidx(t);
                cA[t] = cs[g];
                cB[t] = cs[n + g];
                cD[t] = cs[2 * n + g];
            }
        };
        auto store_level = [&](T *row) {
#pragma unroll
            for (int t = 0; t < C; ++t)
                if (s0 + t <= Nu)
                    row[gs + s0 + t - 1] = u[t];
            if (lane == 0) {
                if (gs == 1)
                    row[0] = ulo;
                if (ge == n - 2)
                    row[N] = uhi;
            }
        };
        load_coeff(0);
        const bool fwd = P.direction == H1_FORWARD;
        if (sf)
            store_level(sf + (size_t)(fwd ? 0 : steps) * n);
        for (int st = 0; st < steps; ++st) {
            if (st > 0 && P.coeff_sets > 1)
                load_coeff(st);
            const T *bs = P.bnd + (P.bnd_sets > 1 ? st : 0) * 4;
            LineBC<T> bc;
            bc.lo = P.lo_type;
            bc.hi = P.hi_type;
            bc.lo_lin = bs[0];
            bc.lo_val = bs[1];
            bc.hi_lin = bs[2];
            bc.hi_val = bs[3];
            bc.h2 = h2;
            T sr[C], srn[C];
            if (has_src) {
                const T *q = P.src + (size_t)st * 2 * (size_t)n;
#pragma unroll
                for (int t = 0; t < C; ++t) {
                    sr[t] = q[gidx(t)];
                    srn[t] = q[n + gidx(t)];
                }
            }
            const T left0 = __shfl_up(u[C - 1], 1), rightC = __shfl_down(u[0], 1);
            auto eq = [&](int t, int kk) -> Eq<T> {
                const int g = gs + kk - 1;
                const T mid = u[t], A = cA[t], B = cB[t], D = cD[t];
                T left = t > 0 ? u[t - 1] : left0;
                T right = t < C - 1 ? u[t + 1] : rightC;
                if (kk == 1)
                    left = ulo;  // read only when the lower end is Dirichlet (g = 1)
                if (kk == Nu)
                    right = uhi;  // read only when the upper end is Dirichlet (g = N - 1)
                T r;
                if (g == 0) {  // Neumann / Robin lower row (lss_heat_implicit_solver_method.hpp:52-67)
                    const T beta = h2 * bc.lo_val;
                    if (P.lo_type == BC_NEUMANN)
                        r = omt * beta * A + (one - two * omt * B) * mid + omt * (A + D) * right;
                    else
                        r = omt * beta * A + (one - omt * (two * B - h2 * bc.lo_lin * A)) * mid + omt * (A + D) * right;
                    if (has_src)
                        r = r + theta * k * srn[t] + omt * k * sr[t];
                } else if (g == N) {  // Neumann / Robin upper row (:70-86)
                    const T delta = h2 * bc.hi_val;
                    if (P.hi_type == BC_NEUMANN)
                        r = omt * (A + D) * left + (one - two * omt * B) * mid - omt * delta * D;
                    else
                        r = omt * (A + D) * left + (one - omt * (two * B + h2 * bc.hi_lin * D)) * mid -
                            omt * delta * D;
                    if (has_src)
                        r = r + theta * k * srn[t] + omt * k * sr[t];
                } else {  // interior (:88-94)
                    r = (D * omt * right) + ((one - two * B * omt) * mid) + (A * omt * left);
                    if (has_src)
                        r = r + k * (theta * srn[t] + omt * sr[t]);
                }
                Eq<T> e;
                e.a = (-theta * A);
                e.b = (one + two * theta * B);
                e.c = (-theta * D);
                e.r = r;
                bc.fold(kk == 1, kk >= Nu, e.a, e.b, e.c, e.r);  // (c = 0 on the padding past the end too)
                return e;
            };
            T nu[C];
            // (the padding past the end restarts from zero every step, so that it stays finite)
            wave_solve_c<T, C, true>(Nu, lane, eq, [&](int t, int kk, T x) { nu[t] = kk <= Nu ? x : (T)0; });
#pragma unroll
            for (int t = 0; t < C; ++t)
                u[t] = nu[t];
            if (P.lo_type == BC_DIRICHLET)
                ulo = bc.lo_val;
            if (P.hi_type == BC_DIRICHLET)
                uhi = bc.hi_val;
            if (sf)
                store_level(sf + (size_t)(fwd ? st + 1 : steps - 1 - st) * n);
        }
        store_level(out);
    });
}

#define LSS_INST_H1W(T, M)                                                                                    \
    template __global__ void k_h1_wave<T, M>(const H1Prob<T> *, int, const T *const *, T *const *, T *const *);
LSS_INST_H1W(double, 1)
LSS_INST_H1W(double, 2)
LSS_INST_H1W(double, 4)
LSS_INST_H1W(double, 8)
LSS_INST_H1W(double, 16)
LSS_INST_H1W(float, 1)
LSS_INST_H1W(float, 2)
LSS_INST_H1W(float, 4)
LSS_INST_H1W(float, 8)
LSS_INST_H1W(float, 16)
#undef LSS_INST_H1W

}  // namespace lssamd
