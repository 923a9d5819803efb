// lss_adi_adjoint.hpp -- reverse mode of the ADI time step for gfx950 (hand-written adjoint).
//
// One forward step maps the level U_s to U_{s+1} through the stages of its scheme (boundary stage A,
// S-sweep X1, v-sweep V1, [HV explicit H], S-sweep X2, v-sweep V2, boundary stage B; DR stops after
// V1).  Every stage is linear in the state and in the transformed coefficients a..f of the grid
// points it touches, so the adjoint step runs the stages backwards:
//   k_adj_v (last)  B^T and the transposed solve of the last v-stage:   lambda(U_{s+1}) -> lv2, elo2
//   k_adj_x (2)     transposed solve of X2 (CS, MCS, HV):               lv2 -> lx2
//   k_adj_g (1)     lambda(Y2) gathered from lx2 (and lv2 for HV) (CS, MCS, HV): -> ly2
//   k_adj_v (1)     transposed solve of V1:                              ly2 -> lv1, elo1
//   k_adj_x (1)     lambda(Y1) gathered, transposed solve of X1:        -> lx1
//   k_adj_g (0)     lambda(U_s) gathered from every equation that reads U_s (the stencils of the
//                   right-hand sides, the Dirichlet ends of the S-lines through idx2, boundary stage A)
// A transposed solve of row equations (a_k, b_k, c_k) solves A^T l = r, whose row k is
// (c_{k-1}, b_k, a_{k+1}); it reuses the register-resident partitioned elimination of the forward
// sweeps.  Every right-hand side is written as a gather (each output point sums over the equation
// points whose stencil contains it), so no kernel scatters or uses atomics.
//
// Parameter gradients: an equation F(y, in; q) = 0 with multiplier l contributes -l dF/dq for each
// stencil coefficient (M, M~, P, P~, Z, W, C of coefficients.hpp:89-95); those map to the transformed
// coefficients a..f and, through the problem's derivative terms, to the model parameters.  Each
// workgroup reduces its contributions in double and adds them to its own slot, so the sums are
// deterministic.

// The fast path may fuse multiply-adds (the library is built with -ffp-contract=off so that the exact
// line solver keeps the reference's rounding; this translation unit opts back in).
#pragma once
#pragma clang fp contract(fast)
#include "lss_adi_chunk.hpp"
#include "lss_adi_stencil.hpp"

#include <type_traits>

namespace lssamd {

// gradient with respect to the stencil coefficients at one point
template <typename T> struct SG {
    T M, Mt, P, Pt, Z, W, C;
};

template <typename T> __device__ __forceinline__ SG<T> sg_zero()
{
    SG<T> g;
    g.M = g.Mt = g.P = g.Pt = g.Z = g.W = g.C = (T)0;
    return g;
}

// compile-time loop: f(std::integral_constant<int, t>) for t = B .. E - 1
template <int B, int E, typename F> __device__ __forceinline__ void static_for(F &&f)
{
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

// first term of TPat<PAT> with the same coefficient as term t and the same factor along the
// direction that varies within a thread (FIX 1: X, FIX 2: Y); such terms share one running sum
template <int PAT, int FIX> __device__ constexpr int term_rep(int t)
{
    using TP = TPat<PAT>;
    for (int u = 0; u < t; ++u)
        if (TP::tq[u] == TP::tq[t] && (FIX == 1 ? TP::fx[u] == TP::fx[t] : TP::fy[u] == TP::fy[t]))
            return u;
    return t;
}

template <typename T> __device__ __forceinline__ T pick6(int q, const T (&g)[6])
{
    return q == 0 ? g[0] : (q == 1 ? g[1] : (q == 2 ? g[2] : (q == 3 ? g[3] : (q == 4 ? g[4] : g[5]))));
}

// Parameter-term sums of one thread.  Term t of a problem contributes g_{tq[t]}(i, j) X_t[i] Y_t[j]
// to parameter tp[t], g_q being d J / d q at the point (q = a..f).  FIX 1 (S-line kernels: the
// thread's column j is fixed) sums g X_t[i] over the thread's points and multiplies by Y_t[j] once
// in finish(); FIX 2 (v-line kernels: row i fixed) the other way round; FIX 0 applies both factors
// at each point.  PAT 1, 2: the term table is TPat<PAT>, fixed at compile time; PAT 0: the
// problem's own (Adj::tp, tq, fx, fy), applied term by term at each point in a loop (compact code
// for the rarely used run-time tables), s[p] summing parameter p.  Sums are kept in double.
template <typename T, int PAT, int FIX> struct TermSums {
    double s[NTERM];
    gptr<const T> fX, fY;
    int nx, ny, n;

    __device__ __forceinline__ TermSums(const Adj<T> &A, int step, int nx_, int ny_)
    {
        nx = nx_;
        ny = ny_;
        fX = G(A.dfac + (A.dfac_sets > 1 ? (int64_t)step : 0) * A.dfac_stride);
        fY = fX + (int)A.nfx * nx;
        if constexpr (PAT == PAT_GENERIC)
            n = A.nterm;
        else
            n = TPat<PAT>::n;
#pragma unroll
        for (int t = 0; t < NTERM; ++t)
            s[t] = 0.0;
    }

    __device__ __forceinline__ double xf(int code, int i) const
    {
        return code >= 0 ? (double)fX[code * nx + i] : 1.0;
    }
    __device__ __forceinline__ double yf(int code, int j) const
    {
        return code >= 0 ? (double)fY[code * ny + j] : 1.0;
    }

    // one point (i, j) with d J / d (a, b, c, d, e, f) = g; QM: bit q set when g[q] may be nonzero
    // (the caller's equations do not depend on the other coefficients, whose terms are skipped)
    template <int QM = 63> __device__ __forceinline__ void add(const Adj<T> &A, int i, int j, const T (&g)[6])
    {
        if constexpr (PAT == PAT_GENERIC) {  // compact code: s[p] sums parameter p directly
#pragma unroll 1
            for (int t = 0; t < n; ++t) {
                if (!((QM >> A.tq[t]) & 1))
                    continue;
                const double v = (double)pick6(A.tq[t], g) * yf(A.fy[t], j) * xf(A.fx[t], i);
                const int k = A.tp[t];
#pragma unroll
                for (int z = 0; z < NPAR; ++z)
                    s[z] += (k == z) ? v : 0.0;
            }
            return;
        }
        static_for<0, NTERM>([&](auto tt) {
            constexpr int t = decltype(tt)::value;
            if constexpr (PAT != PAT_GENERIC) {
                using TP = TPat<PAT>;
                if constexpr (t < TP::n && ((QM >> TP::tq[t]) & 1)) {
                    if constexpr (FIX == 0 || term_rep<PAT, FIX>(t) == t) {
                        double v = (double)g[TP::tq[t]];
                        if constexpr (FIX != 1 && TP::fy[t] >= 0)
                            v *= (double)fY[TP::fy[t] * ny + j];
                        if constexpr (FIX != 2 && TP::fx[t] >= 0)
                            v *= (double)fX[TP::fx[t] * nx + i];
                        s[t] += v;
                    }
                }
            }
        });
    }

    // acc[p] += the sums of parameter p, with the fixed factor of each term applied at (i, j)
    __device__ __forceinline__ void finish(const Adj<T> &A, int i, int j, double *acc) const
    {
        if constexpr (PAT == PAT_GENERIC) {
#pragma unroll
            for (int z = 0; z < NPAR; ++z)
                acc[z] += s[z];
            return;
        }
        static_for<0, NTERM>([&](auto tt) {
            constexpr int t = decltype(tt)::value;
            if constexpr (PAT != PAT_GENERIC) {
                using TP = TPat<PAT>;
                if constexpr (t < TP::n) {
                    double v = s[FIX == 0 ? t : term_rep<PAT, FIX>(t)];
                    if constexpr (FIX == 1 && TP::fy[t] >= 0)
                        v *= (double)fY[TP::fy[t] * ny + j];
                    if constexpr (FIX == 2 && TP::fx[t] >= 0)
                        v *= (double)fX[TP::fx[t] * nx + i];
                    acc[TP::tp[t]] += v;
                }
            }
        });
    }
};

// d J / d (a..f) from d J / d (M, M~, P, P~, Z, W, C): M = alpha a - delta d, M~ = beta b - ni e,
// P = alpha a + delta d, P~ = beta b + ni e, Z = 2 alpha a - rho f / 2, W = 2 beta b - rho f / 2,
// C = c  (coefficients.hpp:89-95)
template <typename T> __device__ __forceinline__ void sg_to_q(const Prob<T> &p, const SG<T> &g, T (&q)[6])
{
    const T two = (T)2, half = (T)0.5;
    q[0] = p.alpha * (g.M + g.P) + two * p.alpha * g.Z;
    q[1] = p.beta * (g.Mt + g.Pt) + two * p.beta * g.W;
    q[2] = g.C;
    q[3] = p.delta * (g.P - g.M);
    q[4] = p.ni * (g.Pt - g.Mt);
    q[5] = -half * p.rho * (g.Z + g.W);
}

// sum acc over the workgroup and add it to slot[0..NPAR); every thread of the block must call it
template <int NWAVES>
__device__ __forceinline__ void block_acc_store(double *acc, double (*red)[NPAR], double *slot)
{
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int z = 0; z < NPAR; ++z) {
        double v = acc[z];
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1)
            v += __shfl_xor(v, d);
        if (lane == 0)
            red[wave][z] = v;
    }
    __syncthreads();
    if (threadIdx.x < NPAR) {
        double v = 0.0;
#pragma unroll
        for (int w = 0; w < NWAVES; ++w)
            v += red[w][threadIdx.x];
        slot[threadIdx.x] += v;
    }
}

// -------------------------------------------------------------------------------------------------
// Weights of the right-hand sides: the coefficient with which an equation at point e reads the
// value at e + (oi, oj).  Each restates the corresponding forward right-hand side term by term.
// -------------------------------------------------------------------------------------------------
// rhs_intermed_1 (lss_heat_douglas_rachford_method.hpp:37-68), on U
template <typename T> __device__ __forceinline__ T w_r1(const Prob<T> &p, const Stencil<T> &s, int oi, int oj)
{
    const T one = (T)1, th = p.theta, gC = p.gamma * s.C;
    if (oi == -1)
        return oj == -1 ? gC : (oj == 0 ? (one - th) * s.M : -gC);
    if (oi == 1)
        return oj == -1 ? -gC : (oj == 0 ? (one - th) * s.P : gC);
    return oj == -1 ? s.Mt : (oj == 0 ? (one - s.W - (one - th) * s.Z) : s.Pt);
}

// v-direction rhs (e.g. lss_heat_douglas_rachford_method.hpp:106-126), on `in` (same row)
template <typename T> __device__ __forceinline__ T w_rv(const Prob<T> &p, const Stencil<T> &s, int oj)
{
    const T th = p.theta;
    return oj == -1 ? -th * s.Mt : (oj == 0 ? th * s.W : -th * s.Pt);
}

// CS / MCS rhs_intermed_3 (lss_heat_craig_sneyd_method.hpp:128-156,
// lss_heat_modified_craig_sneyd_method.hpp:128-165), on U (onU) or on Y2 (!onU)
template <typename T>
__device__ __forceinline__ T w_r3(const Prob<T> &p, const Stencil<T> &s, int oi, int oj, bool mcs, bool onU)
{
    const T th = p.theta, zg = p.zeta * p.gamma * s.C, zt = mcs ? (p.zeta - th) : (T)0;
    T base = (T)0, d = (T)0;  // base: U-only terms; d: coefficient of Delta = Y2 - U
    if (oi != 0 && oj != 0) {
        d = (oi == oj) ? zg : -zg;
    } else if (oi == -1) {
        base = -th * s.M;
        d = zt * s.M;
    } else if (oi == 1) {
        base = -th * s.P;
        d = -zt * s.P;
    } else if (oj == 0) {
        base = th * s.Z;
        d = -zt * (s.Z + s.W);
    } else {
        d = zt * (oj == -1 ? s.Mt : s.Pt);
    }
    return onU ? base - d : d;
}

// HV explicit stage Y3 (lss_heat_hundsdorfer_verwer_method.hpp:128-172): which = 0 on Y1, 1 on U, 2 on Y2
template <typename T>
__device__ __forceinline__ T w_h(const Prob<T> &p, const Stencil<T> &s, int oi, int oj, int which)
{
    const T one = (T)1, th = p.theta, z = p.zeta, zg = p.zeta * p.gamma * s.C;
    T y1 = (T)0, u = (T)0, d = (T)0;
    if (oi != 0 && oj != 0) {
        d = (oi == oj) ? zg : -zg;
    } else if (oi == -1) {
        y1 = -th * s.M;
        u = th * s.M;
        d = z * s.M;
    } else if (oi == 1) {
        y1 = -th * s.P;
        u = th * s.P;
        d = z * s.P;
    } else if (oj == 0) {
        y1 = one + th * s.Z;
        u = -th * s.Z;
        d = -z * (s.W + s.Z);
    } else {
        d = z * (oj == -1 ? s.Mt : s.Pt);
    }
    return which == 0 ? y1 : (which == 1 ? u - d : d);
}

// HV rhs_intermed_4 (lss_heat_hundsdorfer_verwer_method.hpp:174-194), on Y2 (same column)
template <typename T> __device__ __forceinline__ T w_r4(const Prob<T> &p, const Stencil<T> &s, int oi)
{
    const T th = p.theta;
    return oi == -1 ? -th * s.M : (oi == 0 ? th * s.Z : -th * s.P);
}

template <typename T> __device__ __forceinline__ bool inside(const Prob<T> &p, int i, int j)
{
    return i >= 1 && i <= p.nx - 2 && j >= 1 && j <= p.ny - 2;
}

// -------------------------------------------------------------------------------------------------
// Adjoint of a v-sweep (one wavefront per row i).  last: the step's last v-stage, whose output went
// through boundary stage B into U_{s+1}; otherwise V1 of CS/MCS/HV.
// -------------------------------------------------------------------------------------------------
template <typename T, int M, int PAT>
__global__ void __launch_bounds__(64 * VROWS) k_adj_v(const Prob<T> *__restrict__ probs, const Adj<T> *__restrict__ adjs,
                                                      const Tile *__restrict__ tiles, int step, int last,
                                                      double *__restrict__ slots)
{
    __shared__ double red[VROWS][NPAR];
    // Y-block factor vectors of a compile-time pattern in LDS (read at every point of a v-line)
    constexpr int NVY = fpat_nvec<PAT>(12);
    __shared__ T ysh[NVY > 0 ? NVY * (NCH * M + 2) : 1];
    const Tile tl = tiles[blockIdx.x];
    const Prob<T> &p = probs[tl.prob];
    const Adj<T> &A = adjs[tl.prob];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double acc[NPAR];
#pragma unroll
    for (int z = 0; z < NPAR; ++z)
        acc[z] = 0.0;
    const bool wg_run = step < p.steps && (last || p.scheme != SCHEME_DR);  // uniform over the workgroup
    if constexpr (NVY > 0) {
        if (wg_run) {
            const GArr<T> yb = GArr<T>{coeff_set(p, step)} + 12 * (uint32_t)p.nx;
            for (int q = threadIdx.x; q < NVY * p.ny; q += blockDim.x)
                ysh[q] = yb[(uint32_t)q];
        }
        __syncthreads();
    }
    auto stn = [&](const GArr<T> &cs, int i, int j) -> Stencil<T> {
        if constexpr (NVY > 0)
            return stencil_xy<T, PAT>(p, cs, LArr<T>{(lptr<const T>)ysh}, i, j);
        else
            return stencil_at<T, PAT>(p, cs, i, j);
    };
    const bool run = wg_run && wave < tl.count;
    if (run) {
        const int i = __builtin_amdgcn_readfirstlane(tl.first + wave);  // the same for the wavefront
        const int nx = p.nx, ny = p.ny, N = ny - 2, P = nx - 1;
        const T *u = (step & 1) ? p.b : p.a;
        const T *un = (step & 1) ? p.a : p.b;
        const gptr<const T> lamn = G((const T *)(((step + 1) & 1) ? A.lb : A.la));
        const GArr<T> cs{coeff_set(p, step)};
        const T one = (T)1, theta = p.theta;
        const bool hv = p.scheme == SCHEME_HV;
        TermSums<T, PAT, 2> ts(A, step, nx, ny);
        const gptr<const T> yrow = G((last ? un : (const T *)p.y2) + (int)i * ny);
        const gptr<const T> irow = G(((last && hv) ? (const T *)p.y2 : u) + (int)i * ny);
        const gptr<T> L = G((last ? A.lv2 : A.lv1) + (int)i * ny);
        const gptr<const uint32_t> inv2o = G(A.inv2o), inv2i = G(A.inv2i);
        // lambda of the stage output at (i, j): B^T (explicit_boundary.hpp:202-238) for the last stage
        const gptr<const T> ly2 = G((const T *)A.ly2) + (int)i * ny;
        auto lam_out = [&](int j) -> T {
            if (!last)
                return ly2[j];
            T v = lamn[(int)i * ny + j];
            if (i == P - 1 || i == P - 2) {
                T sB = (T)0;
                for (uint32_t k = inv2o[j]; k < inv2o[j + 1]; ++k)
                    sB += lamn[(int)P * ny + inv2i[k]];
                v += (i == P - 1) ? ((T)4 / (T)3) * sB : -(sB / (T)3);
            }
            return v;
        };
        // column k + d, clamped to N + 1 (points past the line's end are discarded)
        const uint32_t jmax = (uint32_t)(N + 1);
        auto rc = [&](int k, int d) { return (int)min((uint32_t)(k + d), jmax); };
        auto eq = [&](int k) -> Eq<T> {
            Eq<T> e;
            e.b = one + theta * stn(cs, i, rc(k, 0)).W;
            e.a = (k >= 2) ? -theta * stn(cs, i, rc(k, -1)).Pt : (T)0;
            e.c = (k <= N - 1) ? -theta * stn(cs, i, rc(k, 1)).Mt : (T)0;
            e.r = lam_out(rc(k, 0));
            return e;
        };
        // the solution stays in registers until every load of the parameter terms is issued
        T xs[M];
        auto put = [&](int t, int k, T x) {
            xs[t] = x;
            if (ts.n == 0)
                return;
            const T xv = k <= N ? x : (T)0;  // a discarded point contributes nothing (x may be any value)
            SG<T> g = sg_zero<T>();
            g.Mt = theta * xv * (yrow[rc(k, -1)] - irow[rc(k, -1)]);
            g.W = -theta * xv * (yrow[rc(k, 0)] - irow[rc(k, 0)]);
            g.Pt = theta * xv * (yrow[rc(k, 1)] - irow[rc(k, 1)]);
            T gq[6];
            sg_to_q(p, g, gq);
            // the v-rows depend on M~, W, P~ only: on b, e and f
            ts.template add<(1 << 1) | (1 << 4) | (1 << 5)>(A, i, rc(k, 0), gq);
        };
        wave_solve<T, M, true>(N, lane, eq, put);
        int s0, cnt;
        chunk_span(N, lane, s0, cnt);
#pragma unroll
        for (int t = 0; t < M; ++t)
            if (t < cnt)
                L[s0 + t] = xs[t];
        if (lane == 0) {
            const T a1 = -theta * stn(cs, i, 1).Mt;
            G(last ? A.elo2 : A.elo1)[i] = lam_out(0) - a1 * xs[0];
        }
        ts.finish(A, i, 0, acc);
    }
    block_acc_store<VROWS>(acc, red, slots + (int64_t)blockIdx.x * NPAR);
}

// -------------------------------------------------------------------------------------------------
// Adjoint of an S-sweep (XTJ columns x NCH chunks per workgroup).  stage 2: X2 of CS/MCS/HV
// (with HV's explicit stage H, whose multiplier equals X2's); stage 1: X1.
// -------------------------------------------------------------------------------------------------
// KIND (compile time, so that the loads of neighbouring points are shared and issued together):
// stage 1 of DR 0, of CS/MCS 1, of HV 2; stage 2 of CS 3, of MCS 4, of HV 5.
template <typename T, int M, int W, int PAT, int KIND, int NC = NCH>
__device__ __forceinline__ void adj_x_lines(T (*sh)[4][NC][W], const Prob<T> &p, const Adj<T> &A, const Tile &tl,
                                            int step, double *acc, lptr<const T> xsh)
{
    constexpr bool S2 = KIND >= 3, HV = KIND == 2 || KIND == 5, MCS = KIND == 4;
    const int col = threadIdx.x % W, chn = threadIdx.x / W;
    const bool colok = col < tl.count;
    const int j = tl.first + (colok ? col : 0);
    const int nx = p.nx, ny = p.ny, N = nx - 2;
    const uint32_t kmax = (uint32_t)(N + 1), uj = (uint32_t)j, uny = (uint32_t)ny;
    const gptr<const T> u = G((const T *)((step & 1) ? p.b : p.a));
    const GArr<T> cs{coeff_set(p, step)};
    const T one = (T)1, theta = p.theta, zeta = p.zeta, gamma = p.gamma;
    TermSums<T, PAT, 1> ts(A, step, nx, ny);
    // stage output: Y1, or Y3 (CS/MCS) / Y4 (HV)
    const gptr<const T> y = G((const T *)(!S2 ? p.y1 : (HV ? p.y4 : p.y3)));
    const gptr<T> L = G(S2 ? A.lx2 : A.lx1);
    const gptr<const T> lv2 = G((const T *)A.lv2), lv1 = G((const T *)A.lv1), lx2 = G((const T *)A.lx2);
    const gptr<const T> y1 = G((const T *)p.y1), y2 = G((const T *)p.y2);
    // row k + d of column j + dc, rows clamped to N + 1 (points past the line's end are discarded)
    auto at = [&](gptr<const T> b, int k, int d, int dc) {
        return ldg(b, min((uint32_t)(k + d), kmax) * uny + (uj + (uint32_t)dc));
    };
    // lambda of the stage output at (k, j).  Stage 1 of HV: Y1 is also read by the explicit stage H,
    // whose weights on Y1 in column j (-theta P_{k-1}, 1 + theta Z_k, -theta M_{k+1}; see w_h) are
    // the coefficients of this row of the transposed system.
    // stencils: the X-block factors of a compile-time pattern come from LDS
    auto stn = [&](int r) -> Stencil<T> {
        if constexpr (PAT != PAT_GENERIC)
            return stencil_xy<T, PAT>(p, LArr<T>{xsh}, cs + 12 * (uint32_t)nx, r, j);
        else
            return stencil_at<T, PAT>(p, cs, r, j);
    };
    auto eq = [&](int k) -> Eq<T> {
        const int r0 = (int)min((uint32_t)k, kmax), rm = (int)min((uint32_t)(k - 1), kmax),
                  rp = (int)min((uint32_t)(k + 1), kmax);
        Eq<T> e;
        e.b = one + theta * stn(r0).Z;
        e.a = (k >= 2) ? -theta * stn(rm).P : (T)0;
        e.c = (k <= N - 1) ? -theta * stn(rp).M : (T)0;
        if constexpr (S2 || KIND == 0) {
            e.r = at(lv2, k, 0, 0);
        } else if constexpr (KIND == 1) {
            e.r = at(lx2, k, 0, 0) + at(lv1, k, 0, 0);
        } else {  // e.a = 0 at k = 1 and e.c = 0 at k >= N, so the clamped reads there are harmless
            e.r = at(lv1, k, 0, 0) + e.b * at(lx2, k, 0, 0) + e.a * at(lx2, k, -1, 0) + e.c * at(lx2, k, 1, 0);
        }
        return e;
    };
    // the solution stays in registers until every load of the parameter terms is issued
    T xs[M];
    auto put = [&](int t, int k, T x) {
        xs[t] = x;
        if (ts.n == 0)
            return;
        const T xv = k <= N ? x : (T)0;  // a discarded point contributes nothing (x may be any value)
        auto U = [&](int d, int dc) { return at(u, k, d, dc); };
        auto Y2 = [&](int d, int dc) { return at(y2, k, d, dc); };
        SG<T> g = sg_zero<T>();
        // implicit rows [-theta M, 1 + theta Z, -theta P] applied to the stage output
        g.M = theta * xv * at(y, k, -1, 0);
        g.Z = -theta * xv * at(y, k, 0, 0);
        g.P = theta * xv * at(y, k, 1, 0);
        if constexpr (!S2) {  // rhs_intermed_1 on U
            g.M += (one - theta) * xv * U(-1, 0);
            g.Z += -(one - theta) * xv * U(0, 0);
            g.W += -xv * U(0, 0);
            g.P += (one - theta) * xv * U(1, 0);
            g.Mt += xv * U(0, -1);
            g.Pt += xv * U(0, 1);
            g.C += xv * gamma * (U(-1, -1) - U(-1, 1) - U(1, -1) + U(1, 1));
        } else {
            auto D = [&](int d, int dc) { return Y2(d, dc) - U(d, dc); };
            const T cornerD = D(1, 1) - D(1, -1) - D(-1, 1) + D(-1, -1);
            if constexpr (!HV) {  // rhs_intermed_3
                g.M += -theta * xv * U(-1, 0);
                g.Z += theta * xv * U(0, 0);
                g.P += -theta * xv * U(1, 0);
                g.C += xv * zeta * gamma * cornerD;
                if constexpr (MCS) {
                    const T zt = zeta - theta;
                    g.M += xv * zt * D(-1, 0);
                    g.Z += -xv * zt * D(0, 0);
                    g.W += -xv * zt * D(0, 0);
                    g.P += -xv * zt * D(1, 0);
                    g.Mt += xv * zt * D(0, -1);
                    g.Pt += xv * zt * D(0, 1);
                }
            } else {  // rhs_intermed_4 on Y2, then the explicit stage H with the same multiplier
                g.M += -theta * xv * Y2(-1, 0);
                g.Z += theta * xv * Y2(0, 0);
                g.P += -theta * xv * Y2(1, 0);
                g.M += xv * (-theta * at(y1, k, -1, 0) + theta * U(-1, 0) + zeta * D(-1, 0));
                g.Z += xv * (theta * at(y1, k, 0, 0) - theta * U(0, 0) - zeta * D(0, 0));
                g.P += xv * (-theta * at(y1, k, 1, 0) + theta * U(1, 0) + zeta * D(1, 0));
                g.W += -xv * zeta * D(0, 0);
                g.Mt += xv * zeta * D(0, -1);
                g.Pt += xv * zeta * D(0, 1);
                g.C += xv * zeta * gamma * cornerD;
            }
        }
        T gq[6];
        sg_to_q(p, g, gq);
        // CS stage 2 reads M, Z, P and C only: a, c, d and f
        constexpr int QM = KIND == 3 ? (1 << 0) | (1 << 2) | (1 << 3) | (1 << 5) : 63;
        ts.template add<QM>(A, (int)min((uint32_t)k, kmax), j, gq);
    };
    block_solve<T, M, W, true, NC>(sh, N, col, chn, colok, eq, put);
    if (colok) {
        int s0, cnt;
        chunk_span(N, chn, s0, cnt, NC);
#pragma unroll
        for (int t = 0; t < M; ++t)
            if (t < cnt)
                L[(int)(s0 + t) * ny + j] = xs[t];
    }
    ts.finish(A, 0, j, acc);
}

template <typename T, int M, int W, int PAT, int STAGE, int NC>
__global__ void __launch_bounds__(W *NC) k_adj_x(const Prob<T> *__restrict__ probs, const Adj<T> *__restrict__ adjs,
                                                 const Tile *__restrict__ tiles, int step, double *__restrict__ slots)
{
    __shared__ T sh[2][4][NC][W];
    __shared__ double red[W * NC / 64][NPAR];
    // X-block factor vectors of a compile-time pattern in LDS (read at every point of an S-line)
    constexpr int NVX = fpat_nvec<PAT>(0);
    __shared__ T xsh[NVX > 0 ? NVX * (NC * M + 2) : 1];
    const Tile tl = tiles[blockIdx.x];
    const Prob<T> &p = probs[tl.prob];
    const Adj<T> &A = adjs[tl.prob];
    double acc[NPAR];
#pragma unroll
    for (int z = 0; z < NPAR; ++z)
        acc[z] = 0.0;
    const int sch = p.scheme;
    if (step >= p.steps || (STAGE == 2 && sch == SCHEME_DR))
        return;  // uniform over the workgroup
    if constexpr (NVX > 0) {
        const GArr<T> xb{coeff_set(p, step)};
        for (int q = threadIdx.x; q < NVX * p.nx; q += blockDim.x)
            xsh[q] = xb[(uint32_t)q];
        __syncthreads();
    }
    const lptr<const T> xs = (lptr<const T>)xsh;
    if constexpr (STAGE == 1) {
        if (sch == SCHEME_DR)
            adj_x_lines<T, M, W, PAT, 0, NC>(sh, p, A, tl, step, acc, xs);
        else if (sch == SCHEME_HV)
            adj_x_lines<T, M, W, PAT, 2, NC>(sh, p, A, tl, step, acc, xs);
        else
            adj_x_lines<T, M, W, PAT, 1, NC>(sh, p, A, tl, step, acc, xs);
    } else {
        if (sch == SCHEME_CS)
            adj_x_lines<T, M, W, PAT, 3, NC>(sh, p, A, tl, step, acc, xs);
        else if (sch == SCHEME_MCS)
            adj_x_lines<T, M, W, PAT, 4, NC>(sh, p, A, tl, step, acc, xs);
        else
            adj_x_lines<T, M, W, PAT, 5, NC>(sh, p, A, tl, step, acc, xs);
    }
    block_acc_store<W * NC / 64>(acc, red, slots + (int64_t)blockIdx.x * NPAR);
}

// -------------------------------------------------------------------------------------------------
// Gathers over two-dimensional tiles of AGR x AGC points (one thread per point of the tile and its
// halo ring).  Every equation point e of the (AGR + 2) x (AGC + 2) points around the tile first
// forms, once, the nine contributions it makes to the lambdas of the values it reads, c_e(oi, oj) for
// the value at e + (oi, oj) (the weight of that value in each of e's equations times the equation's
// lambda), and leaves them in LDS; each point of the tile then sums the nine contributions addressed
// to it.
//   WHAT 0: lambda(U_s) from every equation that reads U_s (the stencils of the right-hand sides,
//           the Dirichlet ends of the S-lines through idx2, boundary stage A), and the parameter
//           terms of boundary stage A (its D, E, F on the v = 0 row)
//   WHAT 1: lambda(Y2) (CS, MCS, HV), read by the adjoint of V1
// -------------------------------------------------------------------------------------------------
template <typename T, int WHAT, int PAT>
__global__ void __launch_bounds__(AGT) k_adj_g(const Prob<T> *__restrict__ probs, const Adj<T> *__restrict__ adjs,
                                                    const Tile *__restrict__ tiles, int step, double *__restrict__ slots)
{
    __shared__ T cst[9][AGR + 2][AGC + 2];  // c_e(oi, oj) at [(oi + 1) * 3 + oj + 1]
    __shared__ double red[AGT / 64][NPAR];
    const Tile tl = tiles[blockIdx.x];
    const Prob<T> &p = probs[tl.prob];
    const Adj<T> &A = adjs[tl.prob];
    double acc[NPAR];
#pragma unroll
    for (int z = 0; z < NPAR; ++z)
        acc[z] = 0.0;
    const bool run = step < p.steps && (WHAT == 0 || p.scheme != SCHEME_DR);  // uniform over the workgroup
    const int r0 = tl.first, c0 = tl.col;
    const int nx = p.nx, ny = p.ny, P = nx - 1;
    const GArr<T> cs{coeff_set(p, step)};
    // the scheme as a compile-time constant (uniform over the workgroup)
    auto with_scheme = [&](auto &&f) {
        switch (p.scheme) {
        case SCHEME_DR: f(std::integral_constant<int, SCHEME_DR>{}); break;
        case SCHEME_CS: f(std::integral_constant<int, SCHEME_CS>{}); break;
        case SCHEME_MCS: f(std::integral_constant<int, SCHEME_MCS>{}); break;
        default: f(std::integral_constant<int, SCHEME_HV>{}); break;
        }
    };
    if (run) {
        const gptr<const T> lv2 = G((const T *)A.lv2), lv1 = G((const T *)A.lv1), lx2 = G((const T *)A.lx2),
                            lx1 = G((const T *)A.lx1);
        do {  // one thread per point of the tile and its halo ring (continue: no equation point)
            const int lr = threadIdx.x / (AGC + 2), lc = threadIdx.x % (AGC + 2);
            const int ei = r0 - 1 + lr, ej = c0 - 1 + lc;
            if (!inside(p, ei, ej)) {  // not an equation point: no contributions
#pragma unroll
                for (int k = 0; k < 9; ++k)
                    cst[k][lr][lc] = (T)0;
                continue;
            }
            const Stencil<T> st = stencil_at<T, PAT>(p, cs, ei, ej);
            const uint32_t q = (uint32_t)ei * (uint32_t)ny + (uint32_t)ej;
            with_scheme([&](auto sc) {
                constexpr int SCH = decltype(sc)::value;
                constexpr bool DR = SCH == SCHEME_DR, HV = SCH == SCHEME_HV, MCS = SCH == SCHEME_MCS;
                if constexpr (WHAT == 0) {
                    const T l1 = ldg(lx1, q), l2 = DR ? (T)0 : ldg(lx2, q);
                    const T lv = (HV ? (T)0 : ldg(lv2, q)) + (DR ? (T)0 : ldg(lv1, q));
#pragma unroll
                    for (int oi = -1; oi <= 1; ++oi) {
#pragma unroll
                        for (int oj = -1; oj <= 1; ++oj) {
                            T t = w_r1(p, st, oi, oj) * l1;
                            if (oi == 0)
                                t += w_rv(p, st, oj) * lv;
                            if constexpr (!DR && !HV)
                                t += w_r3(p, st, oi, oj, MCS, true) * l2;
                            if constexpr (HV)
                                t += w_h(p, st, oi, oj, 1) * l2;
                            cst[(oi + 1) * 3 + oj + 1][lr][lc] = t;
                        }
                    }
                } else {  // lambda(Y2): the X2 equations (and, for HV, V2, whose input is Y2)
                    const T l2 = ldg(lx2, q), lv = HV ? ldg(lv2, q) : (T)0;
#pragma unroll
                    for (int oi = -1; oi <= 1; ++oi) {
#pragma unroll
                        for (int oj = -1; oj <= 1; ++oj) {
                            T t;
                            if constexpr (HV) {
                                T w = w_h(p, st, oi, oj, 2);
                                if (oj == 0)
                                    w += w_r4(p, st, oi);
                                t = w * l2;
                                if (oi == 0)
                                    t += w_rv(p, st, oj) * lv;
                            } else {
                                t = w_r3(p, st, oi, oj, MCS, false) * l2;
                            }
                            cst[(oi + 1) * 3 + oj + 1][lr][lc] = t;
                        }
                    }
                }
            });
        } while (false);
    }
    __syncthreads();
    const int tr = threadIdx.x / AGC, tc = threadIdx.x % AGC;
    const int r = r0 + tr, c = c0 + tc;
    if (run && (int)threadIdx.x < AGR * AGC && tr < tl.count && c < ny) {
        // the contributions addressed to (r, c): from e = (r + di, c + dj) at offset (-di, -dj)
        T v = (T)0;
#pragma unroll
        for (int di = -1; di <= 1; ++di) {
#pragma unroll
            for (int dj = -1; dj <= 1; ++dj)
                v += cst[(1 - di) * 3 + 1 - dj][tr + 1 + di][tc + 1 + dj];
        }
        if (WHAT == 1) {
            if (r >= 1 && r <= P - 1)
                G(A.ly2)[(int)r * ny + c] = v;
        } else {
            const gptr<const T> u = G((const T *)((step & 1) ? p.b : p.a));
            const gptr<T> lam = G((step & 1) ? A.lb : A.la);
            const gptr<const T> bs = bnd_set(p, step);
            const gptr<const T> lx2 = G((const T *)A.lx2), lx1 = G((const T *)A.lx1), elo2 = G((const T *)A.elo2),
                                elo1 = G((const T *)A.elo1);
            const gptr<const uint32_t> inv1o = G(A.inv1o), inv1i = G(A.inv1i), inv2o = G(A.inv2o),
                                       inv2i = G(A.inv2i);
            const T theta = p.theta, delta = p.delta, ni = p.ni, rho = p.rho;
            const bool dr = p.scheme == SCHEME_DR, hv = p.scheme == SCHEME_HV;
            // lambda of colA0[i]: every v-line whose lower end read it (through idx1)
            auto lamA = [&](int i) -> T {
                T v = (T)0;
                for (uint32_t k = inv1o[i]; k < inv1o[i + 1]; ++k) {
                    const uint32_t rr = inv1i[k];
                    if (rr >= 1 && (int)rr <= P - 1)
                        v += elo2[rr] + (dr ? (T)0 : elo1[rr]);
                }
                return v;
            };
            // Dirichlet ends of the S-lines j with idx2[j] == c (implicit_kernel.hpp:86-101)
            if (r == 0 || r == P) {
                const int ei = (r == 0) ? 1 : P - 1;
                for (uint32_t k = inv2o[c]; k < inv2o[c + 1]; ++k) {
                    const int j = (int)inv2i[k];
                    if (j < 1 || j > ny - 2)
                        continue;
                    const Stencil<T> st = stencil_at<T, PAT>(p, cs, ei, j);
                    const T w = (r == 0) ? theta * st.M : theta * st.P;  // -(folded coefficient)
                    const int q = (int)ei * ny + j;
                    T l = lx1[q];
                    if (!dr)
                        l += lx2[q];
                    v += w * l;
                    if (hv)  // H reads Y1 at rows 0 and P; Y1's ends are these U values
                        v += -w * lx2[q];
                }
            }
            // boundary stage A (explicit_boundary.hpp:39-84): the v = 0 column of U_{s+1} read U_s
            if (c <= 2) {
                if (c == 0) {
#pragma unroll
                    for (int di = -1; di <= 1; ++di) {
                        const int i = r + di;
                        if (i < 1 || i > P)
                            continue;
                        T D, E, F;
                        def_row0<T, PAT>(p, cs, i, D, E, F);
                        const T la = lamA(i);
                        if (di == 0)
                            v += ((T)1 - (T)3 * ni * E + rho * F) * la;
                        else if (i < P)
                            v += (di == 1 ? -delta * D : delta * D) * la;
                    }
                } else if (r >= 1) {
                    T D, E, F;
                    def_row0<T, PAT>(p, cs, r, D, E, F);
                    v += (c == 1 ? (T)4 * ni * E : -ni * E) * lamA(r);
                }
                if (c == 0 && r >= 1 && A.nterm > 0) {
                    TermSums<T, PAT, 0> ts(A, step, nx, ny);
                    const T la = lamA(r);
                    const gptr<const T> ur = u + (int)r * ny;
                    const T gd = (r < P) ? la * delta * (u[(int)(r + 1) * ny] - u[(int)(r - 1) * ny])
                                         : -la * delta * ((T)2 * p.h1 * bs[nx + ny]);
                    const T ge = la * ni * (-(T)3 * ur[0] + (T)4 * ur[1] - ur[2]);
                    const T gf = la * rho * ur[0];
                    const T gq[6] = {(T)0, (T)0, (T)0, gd, ge, gf};
                    ts.template add<(1 << 3) | (1 << 4) | (1 << 5)>(A, r, 0, gq);
                    ts.finish(A, r, 0, acc);
                }
            }
            lam[(int)r * ny + c] = v;
        }
    }
    // (only the tiles of column 0 hold parameter terms: boundary stage A's D, E, F on the v = 0 row)
    if (WHAT == 0 && c0 == 0)
        block_acc_store<AGT / 64>(acc, red, slots + (int64_t)blockIdx.x * NPAR);
}

// -------------------------------------------------------------------------------------------------
// Batched device copies: job blockIdx.y, grid-stride over its elements.
// -------------------------------------------------------------------------------------------------
template <typename T> __global__ void __launch_bounds__(256) k_copy_jobs(const CopyJob *__restrict__ jobs)
{
    const CopyJob jb = jobs[blockIdx.y];
    const T *src = (const T *)jb.src;
    T *dst = (T *)jb.dst;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < jb.n; k += (int64_t)gridDim.x * blockDim.x)
        dst[k] = src[k];
}

}  // namespace lssamd

// Instantiations (one translation unit per precision and pattern): LSS_INST_ADJ(T, PAT).
#define LSS_INST_AXN(T, M, W, PAT, NC)                                                                         \
    template __global__ void k_adj_x<T, M, W, PAT, 1, NC>(const Prob<T> *, const Adj<T> *, const Tile *, int,       \
                                                          double *);                                               \
    template __global__ void k_adj_x<T, M, W, PAT, 2, NC>(const Prob<T> *, const Adj<T> *, const Tile *, int,       \
                                                          double *);
#define LSS_INST_AX(T, M, W, PAT) LSS_INST_AXN(T, M, W, PAT, 64)
#define LSS_INST_AV(T, M, PAT) \
    template __global__ void k_adj_v<T, M, PAT>(const Prob<T> *, const Adj<T> *, const Tile *, int, int, double *);
// S-lines per adjoint S-sweep workgroup: 16 (64 chunks) for chunk bounds up to 8; lines of more than 512
// points take 128 chunks of at most 8 points, 8 lines per workgroup (the plan's rule)
#define LSS_INST_ADJ(T, PAT)                                                                                   \
    namespace lssamd {                                                                                         \
    LSS_INST_AV(T, 1, PAT) LSS_INST_AV(T, 2, PAT) LSS_INST_AV(T, 4, PAT) LSS_INST_AV(T, 8, PAT)                 \
    LSS_INST_AV(T, 16, PAT) LSS_INST_AX(T, 1, 16, PAT) LSS_INST_AX(T, 2, 16, PAT) LSS_INST_AX(T, 4, 16, PAT)    \
    LSS_INST_AX(T, 8, 16, PAT) LSS_INST_AXN(T, 8, 8, PAT, 128)                                                  \
    template __global__ void k_adj_g<T, 0, PAT>(const Prob<T> *, const Adj<T> *, const Tile *, int, double *);       \
    template __global__ void k_adj_g<T, 1, PAT>(const Prob<T> *, const Adj<T> *, const Tile *, int, double *);       \
    }
