// lss_adi_kernels.hip -- HIP kernels of the ADI time step for gfx950 (MI355X).
//
// One time step of every scheme is the same program of launches (the reference's loop body,
// pde_solvers/two_dimensional/heat_type/lss_2d_general_heston_equation_implicit_kernel.hpp:177-184):
//   bnd_a   boundary stage A: v = 0 column (explicit degenerate row) and v = v_max column
//           (lss_2d_general_heston_equation_explicit_boundary.hpp:181-200, :39-84)
//   xsweep  stage 1: S-lines j = 1..ny-2, rhs_intermed_1 of u, implicit solve along S  -> Y1
//   vsweep  stage 1: v-lines i = 1..nx-2, v-direction rhs, implicit solve along v     -> Y2 (DR: next)
//   hv_y3   HV only: full-grid explicit stage                                         -> Y3
//   xsweep  stage 2: CS/MCS rhs_intermed_3 -> Y3 ; HV rhs_intermed_4 -> Y4 (stored in Y1's buffer)
//   vsweep  stage 2: v-direction rhs of (u, Y3) or (Y2, Y4)                          -> next
//   bnd_b   boundary stage B: S_min row and S_max Neumann extrapolation (explicit_boundary.hpp:202-238)
// The explicit Euler solver's step is one launch of k_euler (interior and both boundary stages).
// Kernels are compiled with -ffp-contract=off and keep the reference's operand order, so with
// FIELDS coefficients every value is rounded exactly as the reference's host code rounds it.
//
// Line solves (v1): one lane per line, the forward sweep keeps (gamma_t, sol_t) in HBM
// (sol in the output array, gamma in a scratch array), the backward sweep overwrites the output.
// Lanes of a wavefront hold adjacent S-lines, so S-line sweeps read and write whole rows.

#include "lss_adi_stencil.hpp"

namespace lssamd {

constexpr int LB = 8;  // points per batch of the exact line sweeps

// -------------------------------------------------------------------------------------------------
// Boundary stage A: colA0 = new v = 0 column, colA1 = v_max column (explicit_boundary.hpp:39-84, :181-200)
// as a launch of its own (plans with exact lines; the partitioned stage-1 S-sweep evaluates it itself)
// -------------------------------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(TILE) k_bnd_a(const Prob<T> *__restrict__ probs, const Tile *__restrict__ tiles,
                                                int step)
{
    const Tile tl = tiles[blockIdx.x];
    const Prob<T> &p = probs[tl.prob];
    if (step >= p.steps || (int)threadIdx.x >= tl.count)
        return;
    bnd_a_row(p, step, tl.first + (int)threadIdx.x);
}

// -------------------------------------------------------------------------------------------------
// Boundary stage B: row 0 Dirichlet, row nx-1 Neumann extrapolation through idx2 (explicit_boundary.hpp:202-238)
// -------------------------------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(TILE) k_bnd_b(const Prob<T> *__restrict__ probs, const Tile *__restrict__ tiles,
                                                int step)
{
    const Tile tl = tiles[blockIdx.x];
    const Prob<T> &p = probs[tl.prob];
    if (step >= p.steps || (int)threadIdx.x >= tl.count)
        return;
    const int j = tl.first + threadIdx.x;
    const int nx = p.nx, ny = p.ny;
    const gptr<T> nxt = G((step & 1) ? p.a : p.b);
    const gptr<const T> bs = bnd_set(p, step);
    const T two = (T)2.0, three = (T)3.0, four = (T)4.0;
    const uint32_t jj = p.idx2[j];
    const T e = (((four * nxt[(int64_t)(nx - 2) * ny + jj]) - nxt[(int64_t)(nx - 3) * ny + jj] -
                  (two * p.h1 * bs[nx + ny + j])) /
                 three);
    nxt[j] = bs[nx + j];
    nxt[(int64_t)(nx - 1) * ny + j] = e;
}

// -------------------------------------------------------------------------------------------------
// Line solve step of the forward sweep.  Thomas LU keeps (gamma_t, sol_t); double sweep keeps
// (L_t, K_t).  first/last fold the Dirichlet ends exactly as thomas_lu_boundary.hpp:127-135,
// 176-183 and double_sweep_boundary.hpp:100-106 do.
// -------------------------------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void fwd_step(int tri, bool first, bool last, T low, T diag, T high, T r, T lo, T hi,
                                         T gprev, T sprev, T &g, T &s)
{
    if (tri == TRI_THOMAS) {
        if (first) {
            const T beta = diag;
            g = high / beta;
            const T rr = r - low * lo;
            s = rr / beta;
        } else if (!last) {
            const T beta = diag - (low * gprev);
            g = high / beta;
            s = (r - (low * sprev)) / beta;
        } else {
            const T beta = diag - (low * gprev);
            const T rn = r - high * hi;
            s = (rn - (low * sprev)) / beta;
            g = (T)0;
        }
    } else {
        const T mone = (T)-1.0;
        const T Lp = first ? (T)0 : gprev;
        const T Kp = first ? lo : sprev;
        const T tmp = diag + (low * Lp);
        g = mone * high / tmp;
        s = (r - (low * Kp)) / tmp;
    }
}

template <typename T> __device__ __forceinline__ T bwd_step(int tri, T g, T s, T xnext)
{
    return tri == TRI_THOMAS ? (s - (g * xnext)) : ((g * xnext) + s);
}

// -------------------------------------------------------------------------------------------------
// S-line sweep (one lane per line j).  stage 1: rhs_intermed_1(u) -> Y1.
// stage 2: CS/MCS rhs_intermed_3(u, Y1, Y2) -> Y3; HV rhs_intermed_4(Y2, Y3) -> Y4 (in Y1's buffer).
// Dirichlet ends: u(0, idx2[j]), u(nx-1, idx2[j]) (implicit_kernel.hpp:86-101, old time level).
// -------------------------------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(TILE) k_xsweep(const Prob<T> *__restrict__ probs, const Tile *__restrict__ tiles,
                                                 int step, int stage)
{
    const Tile tl = tiles[blockIdx.x];
    const Prob<T> &p = probs[tl.prob];
    if (step >= p.steps || (stage == 2 && p.scheme == SCHEME_DR) || (int)threadIdx.x >= tl.count)
        return;
    const int j = tl.first + threadIdx.x;
    const int nx = p.nx, ny = p.ny, end = nx - 2, tri = p.tri;
    const T *u0 = (step & 1) ? p.b : p.a;
    const gptr<const T> u = G(u0);
    const gptr<const T> cs = coeff_set(p, step);
    const T one = (T)1.0, theta = p.theta;
    const int kind = (stage == 1) ? 0 : (p.scheme == SCHEME_CS ? 1 : (p.scheme == SCHEME_MCS ? 2 : 3));
    const gptr<const T> in = G((kind == 3) ? (const T *)p.y2 : u0);
    const gptr<const T> inh = G((const T *)((kind == 3) ? p.y3 : p.y1));
    const gptr<const T> inhn = G((const T *)p.y2);
    const gptr<T> out = G((kind == 0) ? p.y1 : (kind == 3 ? p.y4 : p.y3));
    const gptr<T> scr = G(p.scr);
    const uint32_t jj = p.idx2[j];
    const T lo = u[jj], hi = u[(int64_t)(nx - 1) * ny + jj];
    auto at = [&](gptr<const T> b, int64_t q) { return PAt<gptr<const T>>{b + q, ny}; };
    const T *sp = kind == 0 ? src_set(p, step) : nullptr;
    const int64_t nxy = (int64_t)nx * ny;
    // The sweeps run in batches of LB points: the loads of a batch are issued together before its
    // stores (the arrays read and written are distinct), so a lane waits once per batch, not per point.
    T gprev = (T)0, sprev = (T)0;
    for (int t0 = 1; t0 <= end; t0 += LB) {
        T lw[LB], dg[LB], hg[LB], rr[LB];
#pragma unroll
        for (int k = 0; k < LB; ++k) {
            const int t = t0 + k;
            if (t > end)
                break;
            const int64_t q = (int64_t)t * ny + j;
            const Stencil<T> s = stencil_at(p, cs, t, j);
            lw[k] = (-theta * s.M);
            dg[k] = (one + theta * s.Z);
            hg[k] = (-theta * s.P);
            T r;
            if (kind == 0) {
                r = rhs_intermed_1(p, s, at(in, q));
                // rhs_intermed_1_source (lss_heat_douglas_rachford_method.hpp:70-104)
                if (sp)
                    r = r + (theta * p.rho * sp[nxy + q]) + ((one - theta) * p.rho * sp[q]);
            } else if (kind == 1)
                r = rhs_intermed_3<T, false>(p, s, at(in, q), at(inh, q), at(inhn, q));
            else if (kind == 2)
                r = rhs_intermed_3<T, true>(p, s, at(in, q), at(inh, q), at(inhn, q));
            else
                r = rhs_intermed_4(p, s, at(in, q), at(inh, q));
            rr[k] = r;
        }
#pragma unroll
        for (int k = 0; k < LB; ++k) {
            const int t = t0 + k;
            if (t > end)
                break;
            const int64_t q = (int64_t)t * ny + j;
            T g, sv;
            fwd_step(tri, t == 1, t == end, lw[k], dg[k], hg[k], rr[k], lo, hi, gprev, sprev, g, sv);
            scr[q] = g;
            out[q] = sv;
            gprev = g;
            sprev = sv;
        }
    }
    T x = (tri == TRI_THOMAS) ? sprev : bwd_step(tri, gprev, sprev, hi);
    out[(int64_t)end * ny + j] = x;
    for (int t0 = end - 1; t0 >= 1; t0 -= LB) {
        T gg[LB], ss[LB];
#pragma unroll
        for (int k = 0; k < LB; ++k) {
            const int t = t0 - k;
            if (t < 1)
                break;
            const int64_t q = (int64_t)t * ny + j;
            gg[k] = scr[q];
            ss[k] = out[q];
        }
#pragma unroll
        for (int k = 0; k < LB; ++k) {
            const int t = t0 - k;
            if (t < 1)
                break;
            x = bwd_step(tri, gg[k], ss[k], x);
            out[(int64_t)t * ny + j] = x;
        }
    }
    out[j] = lo;
    out[(int64_t)(nx - 1) * ny + j] = hi;
}

// -------------------------------------------------------------------------------------------------
// v-line sweep (one lane per line i).  stage 1: rhs_v(u, Y1) -> Y2 (DR: next).
// stage 2: CS/MCS rhs_v(u, Y3) -> next; HV rhs_v(Y2, Y4) -> next.
// Dirichlet ends colA0/colA1[idx1[i]]: the copy of `next` the vertical lambdas captured after
// boundary stage A (implicit_kernel.hpp:69-84).
// -------------------------------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(TILE) k_vsweep(const Prob<T> *__restrict__ probs, const Tile *__restrict__ tiles,
                                                 int step, int stage)
{
    const Tile tl = tiles[blockIdx.x];
    const Prob<T> &p = probs[tl.prob];
    if (step >= p.steps || (stage == 2 && p.scheme == SCHEME_DR) || (int)threadIdx.x >= tl.count)
        return;
    const int i = tl.first + threadIdx.x;
    const int ny = p.ny, end = ny - 2, tri = p.tri;
    const T *u = (step & 1) ? p.b : p.a;
    T *nxt = (step & 1) ? p.a : p.b;
    const gptr<const T> cs = coeff_set(p, step);
    const T one = (T)1.0, theta = p.theta;
    const T *in0, *inh0;
    T *out0;
    if (stage == 1) {
        in0 = u;
        inh0 = p.y1;
        out0 = (p.scheme == SCHEME_DR) ? nxt : p.y2;
    } else if (p.scheme == SCHEME_HV) {
        in0 = p.y2;
        inh0 = p.y4;
        out0 = nxt;
    } else {
        in0 = u;
        inh0 = p.y3;
        out0 = nxt;
    }
    const int64_t row = (int64_t)i * ny;
    const uint32_t ii = p.idx1[i];
    const T lo = G(p.colA0)[ii], hi = G(p.colA1)[ii];
    const gptr<T> scr = G(p.scr + row);
    const gptr<const T> in = G(in0 + row);
    const gptr<const T> inh = G(inh0 + row);
    const gptr<T> out = G(out0 + row);
    // (batches of LB points, as in k_xsweep)
    T gprev = (T)0, sprev = (T)0;
    for (int t0 = 1; t0 <= end; t0 += LB) {
        T lw[LB], dg[LB], hg[LB], rr[LB];
#pragma unroll
        for (int k = 0; k < LB; ++k) {
            const int t = t0 + k;
            if (t > end)
                break;
            const Stencil<T> s = stencil_at(p, cs, i, t);
            lw[k] = (-theta * s.Mt);
            dg[k] = (one + theta * s.W);
            hg[k] = (-theta * s.Pt);
            rr[k] = rhs_v(p, s, in + t, inh + t);
        }
#pragma unroll
        for (int k = 0; k < LB; ++k) {
            const int t = t0 + k;
            if (t > end)
                break;
            T g, sv;
            fwd_step(tri, t == 1, t == end, lw[k], dg[k], hg[k], rr[k], lo, hi, gprev, sprev, g, sv);
            scr[t] = g;
            out[t] = sv;
            gprev = g;
            sprev = sv;
        }
    }
    T x = (tri == TRI_THOMAS) ? sprev : bwd_step(tri, gprev, sprev, hi);
    out[end] = x;
    for (int t0 = end - 1; t0 >= 1; t0 -= LB) {
        T gg[LB], ss[LB];
#pragma unroll
        for (int k = 0; k < LB; ++k) {
            const int t = t0 - k;
            if (t < 1)
                break;
            gg[k] = scr[t];
            ss[k] = out[t];
        }
#pragma unroll
        for (int k = 0; k < LB; ++k) {
            const int t = t0 - k;
            if (t < 1)
                break;
            x = bwd_step(tri, gg[k], ss[k], x);
            out[t] = x;
        }
    }
    out[0] = lo;
    out[ny - 1] = hi;
}

// -------------------------------------------------------------------------------------------------
// HV full-grid explicit stage Y3 (lss_heat_hundsdorfer_verwer_method.hpp:128-172), interior only.
// One workgroup per row r; lanes over columns.
// -------------------------------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(TILE) k_hv_y3(const Prob<T> *__restrict__ probs, const Tile *__restrict__ tiles,
                                                int step)
{
    const Tile tl = tiles[blockIdx.x];
    const Prob<T> &p = probs[tl.prob];
    if (step >= p.steps || p.scheme != SCHEME_HV)
        return;
    const int r = tl.first;
    const int ny = p.ny;
    const gptr<const T> u = G((const T *)((step & 1) ? p.b : p.a));
    const gptr<const T> cs = coeff_set(p, step);
    const gptr<const T> y1 = G((const T *)p.y1), y2 = G((const T *)p.y2);
    const gptr<T> y3 = G(p.y3);
    for (int c = 1 + threadIdx.x; c < ny - 1; c += blockDim.x) {
        const int64_t q = (int64_t)r * ny + c;
        const Stencil<T> s = stencil_at(p, cs, r, c);
        const PAt<gptr<const T>> in{u + q, ny}, inh{y1 + q, ny}, inhn{y2 + q, ny};
        const T val = hv_y3_at(p, s, in, inh, inhn);
        y3[q] = val;
    }
}

// -------------------------------------------------------------------------------------------------
// Explicit Euler step of the explicit solver in one launch (lss_heston_euler_solver_scheme.hpp:63-67,
// 78-83: the interior, boundary stage A, boundary stage B, each from the level before):
//   interior  explicit_heston_scheme::rhs (lss_heston_euler_solver_method.hpp:26-52), whose C is
//             gamma * c (lss_heston_euler_coefficients.hpp:50-51);
//   column 0 and ny - 1  boundary stage A (explicit_boundary.hpp:39-84, 181-200), which writes them
//             for every row;
//   row 0 and nx - 1     boundary stage B (explicit_boundary.hpp:202-238); its S_max extrapolation
//             reads rows nx - 2 and nx - 3 of the new level through idx2, and the thread recomputes
//             those two values from u with the same arithmetic instead of waiting for them.
// Every value is rounded as the reference rounds it (this file is compiled without contraction), so
// with FIELDS coefficients the step is bit-identical to the three stages run one after another.
// PAT selects compile-time factor patterns of SEPARABLE plans.  Two-dimensional tiles of GTR * ERPT
// rows x GTC columns; a thread takes ERPT consecutive rows of one column (lanes along the contiguous v
// direction) and reads the rows of its three columns once.
// -------------------------------------------------------------------------------------------------
template <typename T, int PAT>
__global__ void __launch_bounds__(GTR *GTC) k_euler(const Prob<T> *__restrict__ probs,
                                                    const Tile *__restrict__ tiles, int step)
{
    const Tile tl = tiles[blockIdx.x];
    const Prob<T> &p = probs[tl.prob];
    if (step >= p.steps)
        return;
    const int tg = threadIdx.x / GTC, tc = threadIdx.x % GTC;
    const int j = tl.col + tc;
    const int nx = p.nx, ny = p.ny;
    const int r0 = tl.first + tg * ERPT;               // this thread's rows r0 .. r0 + rn - 1
    const int rn = min(ERPT, tl.first + tl.count - r0);
    if (j >= ny || rn <= 0)
        return;
    const gptr<const T> u = G((const T *)((step & 1) ? p.b : p.a));
    const gptr<T> nxt = G((step & 1) ? p.a : p.b);
    const gptr<const T> cs = coeff_set(p, step);
    const gptr<const T> bs = bnd_set(p, step);
    const T one = (T)1.0, two = (T)2.0, three = (T)3.0, four = (T)4.0;
    // the value stages "interior" and A give point (r, c), 1 <= r <= nx - 2, from direct loads
    auto level = [&](int r, int c) -> T {
        const gptr<const T> w = u + ((int64_t)r * ny + c);
        if (c == ny - 1)
            return bs[r];  // v_upper(t, X(zeta_r))
        if (c == 0) {      // the degenerate v = 0 row
            T D, E, F;
            def_row0<T, PAT>(p, cs, r, D, E, F);
            const T delta = p.delta, ni = p.ni, rho = p.rho;
            return (-delta * D * w[-ny]) + ((one - three * ni * E + rho * F) * w[0]) + (delta * D * w[ny]) -
                   (ni * E * w[2]) + (four * ni * E * w[1]);
        }
        const Stencil<T> s = stencil_at<T, PAT>(p, cs, r, c);
        const T C = p.gamma * s.C;
        return C * w[-ny - 1] + s.M * w[-ny] - C * w[-ny + 1] + s.Mt * w[-1] + (one - s.Z - s.W) * w[0] +
               s.Pt * w[1] - C * w[ny - 1] + s.P * w[ny] + C * w[ny + 1];
    };
    // rows r0 - 1 .. r0 + ERPT of columns j - 1, j, j + 1 (clamped into the grid), all loaded before
    // the first store, so that each value is read once and the interior points share them
    const int jm = max(j - 1, 0), jp = min(j + 1, ny - 1);
    T w[ERPT + 2][3];
#pragma unroll
    for (int t = 0; t < ERPT + 2; ++t) {
        const gptr<const T> row = u + (int64_t)min(max(r0 - 1 + t, 0), nx - 1) * ny;
        w[t][0] = row[jm];
        w[t][1] = row[j];
        w[t][2] = row[jp];
    }
    T out[ERPT];
#pragma unroll
    for (int t = 0; t < ERPT; ++t) {
        const int i = r0 + t;
        T v = (T)0;
        if (t < rn) {
            if (i == 0) {
                v = bs[nx + j];  // h_lower(t, Y(eta_j))
            } else if (i == nx - 1) {
                const int jj = (int)p.idx2[j];
                v = (((four * level(nx - 2, jj)) - level(nx - 3, jj) - (two * p.h1 * bs[nx + ny + j])) / three);
            } else if (j == 0 || j == ny - 1) {
                v = level(i, j);
            } else {
                const Stencil<T> s = stencil_at<T, PAT>(p, cs, i, j);
                const T C = p.gamma * s.C;
                v = C * w[t][0] + s.M * w[t][1] - C * w[t][2] + s.Mt * w[t + 1][0] + (one - s.Z - s.W) * w[t + 1][1] +
                    s.Pt * w[t + 1][2] - C * w[t + 2][0] + s.P * w[t + 2][1] + C * w[t + 2][2];
            }
        }
        out[t] = v;
    }
#pragma unroll
    for (int t = 0; t < ERPT; ++t)
        if (t < rn)
            nxt[(int64_t)(r0 + t) * ny + j] = out[t];
}

// -------------------------------------------------------------------------------------------------
// The whole explicit Euler time loop of one small problem in one workgroup of ERES_T threads, its two
// time levels resident in LDS (lss_heston_euler_solver_scheme.hpp:63-67, 78-83).  Per step: the
// interior and boundary stage A from the old level into the new one; a barrier; boundary stage B (row 0,
// and row nx - 1 extrapolated from rows nx - 2 and nx - 3 of the new level, as the reference reads
// them); a barrier.  Every value takes the arithmetic of k_euler in the same order.  The thread's
// points are numbered interior first, then the v = 0 column, then the v_max column, so that the
// wavefronts mostly take one branch; their stencil weights and boundary values are formed once (the
// problems have one coefficient set and one boundary set).  For problems of at most ERES_T * ERES_K
// points with nx >= 4 and ny <= ERES_T; HBM is read once and written once.
// -------------------------------------------------------------------------------------------------
template <typename T, int PAT>
__global__ void __launch_bounds__(ERES_T) __attribute__((amdgpu_waves_per_eu(4)))
k_euler_res(const Prob<T> *__restrict__ probs, const Tile *__restrict__ tiles)
{
    __shared__ T lv[2][ERES_T * ERES_K];
    const Tile tl = tiles[blockIdx.x];
    const Prob<T> &p = probs[tl.prob];
    const int nx = p.nx, ny = p.ny, n = nx * ny, steps = p.steps, tid = (int)threadIdx.x;
    const int ni_ = (nx - 2) * (ny - 2), na = ni_ + (nx - 2);  // interior points, then the v = 0 column
    const T one = (T)1.0, two = (T)2.0, three = (T)3.0, four = (T)4.0;
    {
        const gptr<const T> a = G((const T *)p.a);
        for (int q = tid; q < n; q += ERES_T)
            lv[0][q] = a[q];
    }
    // the thread's points: pt = kind << 16 | (r * ny + c); kind 1 interior, 2 v = 0, 3 v_max, 0 none
    int pt[ERES_K];
#pragma unroll
    for (int k = 0; k < ERES_K; ++k) {
        const int e = tid + k * ERES_T;
        int r = 0, c = 0, kd = 0;
        if (e < ni_) {
            r = 1 + e / (ny - 2);
            c = 1 + e % (ny - 2);
            kd = 1;
        } else if (e < na) {
            r = 1 + (e - ni_);
            kd = 2;
        } else if (e < na + (nx - 2)) {
            r = 1 + (e - na);
            c = ny - 1;
            kd = 3;
        }
        pt[k] = kd << 16 | (r * ny + c);
    }
    // interior: C, M, M~, 1 - Z - W, P~, P (C = gamma c); v = 0: -delta D, 1 - 3 ni E + rho F, delta D,
    // ni E, 4 ni E; v_max: the boundary value
    T wt[ERES_K][6];
    T lo_j = (T)0, hup_j = (T)0;  // boundary stage B of column tid: h_lower and 2 h_1 h_upper X'(1)
    int jj = 0;
    {
        const gptr<const T> cs = coeff_set(p, 0);
        const gptr<const T> bs = bnd_set(p, 0);
#pragma unroll
        for (int k = 0; k < ERES_K; ++k) {
            const int kd = pt[k] >> 16, o = pt[k] & 0xffff, r = o / ny, c = o - r * ny;
            if (kd == 1) {
                const Stencil<T> st = stencil_at<T, PAT>(p, cs, r, c);
                wt[k][0] = p.gamma * st.C;
                wt[k][1] = st.M;
                wt[k][2] = st.Mt;
                wt[k][3] = one - st.Z - st.W;
                wt[k][4] = st.Pt;
                wt[k][5] = st.P;
            } else if (kd == 2) {
                T D, E, F;
                def_row0<T, PAT>(p, cs, r, D, E, F);
                const T delta = p.delta, ni = p.ni, rho = p.rho;
                wt[k][0] = -delta * D;
                wt[k][1] = one - three * ni * E + rho * F;
                wt[k][2] = delta * D;
                wt[k][3] = ni * E;
                wt[k][4] = four * ni * E;
            } else if (kd == 3) {
                wt[k][0] = bs[r];  // v_upper(t, X(zeta_r))
            }
        }
        if (tid < ny) {
            lo_j = bs[nx + tid];  // h_lower(t, Y(eta_j))
            hup_j = two * p.h1 * bs[nx + ny + tid];
            jj = (int)p.idx2[tid];
        }
    }
    __syncthreads();
    for (int s = 0; s < steps; ++s) {
        const T *u = lv[s & 1];
        T *w = lv[(s & 1) ^ 1];
#pragma unroll
        for (int k = 0; k < ERES_K; ++k) {
            const int kd = pt[k] >> 16, o = pt[k] & 0xffff;
            if (kd == 1) {
                const T C = wt[k][0];
                w[o] = C * u[o - ny - 1] + wt[k][1] * u[o - ny] - C * u[o - ny + 1] + wt[k][2] * u[o - 1] +
                       wt[k][3] * u[o] + wt[k][4] * u[o + 1] - C * u[o + ny - 1] + wt[k][5] * u[o + ny] +
                       C * u[o + ny + 1];
            } else if (kd == 2) {
                w[o] = (wt[k][0] * u[o - ny]) + (wt[k][1] * u[o]) + (wt[k][2] * u[o + ny]) - (wt[k][3] * u[o + 2]) +
                       (wt[k][4] * u[o + 1]);
            } else if (kd == 3) {
                w[o] = wt[k][0];
            }
        }
        __syncthreads();
        if (tid < ny) {
            const T e = (((four * w[(nx - 2) * ny + jj]) - w[(nx - 3) * ny + jj] - hup_j) / three);
            w[tid] = lo_j;
            w[(nx - 1) * ny + tid] = e;
        }
        __syncthreads();
    }
    const gptr<T> out = G((steps & 1) ? p.b : p.a);
    for (int q = tid; q < n; q += ERES_T)
        out[q] = lv[steps & 1][q];
}

// -------------------------------------------------------------------------------------------------
// Batched tridiagonal systems (one lane per system), the reference's Thomas LU in its operation
// order with the end rows folded by `bc` (LineBC; PLAIN ignores a[0] and c[n-1]).
// Layout: contiguous x[k*n + i] (strided == 0) or interleaved x[i*B + k] (strided == 1).
// -------------------------------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) k_tridiag_lane(int64_t B, int n, int strided, const T *__restrict__ a,
                                                      const T *__restrict__ b, const T *__restrict__ c,
                                                      const T *__restrict__ d, T *__restrict__ x, T *__restrict__ scr,
                                                      LineBC<T> bc)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= B)
        return;
    const int64_t base = strided ? k : k * n, st = strided ? B : 1;
    const int s0 = bc.start(), e0 = bc.end(n);
    // row t with its end folded (thomas_lu_solver.hpp:125-191, thomas_lu_boundary.hpp:114-205)
    auto row = [&](int t, T &at, T &bt, T &ct, T &rt) {
        const int64_t q = base + t * st;
        at = a[q];
        bt = b[q];
        ct = c[q];
        rt = d[q];
        bc.fold(t == s0, t == e0, at, bt, ct, rt);
    };
    T at, bt, ct, rt;
    row(s0, at, bt, ct, rt);
    T beta = bt;
    T g = ct / beta;
    T s = rt / beta;
    scr[base + s0 * st] = g;
    x[base + s0 * st] = s;
    for (int t = s0 + 1; t <= e0; ++t) {
        row(t, at, bt, ct, rt);
        beta = bt - (at * g);
        g = ct / beta;
        s = (rt - (at * s)) / beta;
        scr[base + t * st] = g;
        x[base + t * st] = s;
    }
    T xv = s;
    for (int t = e0 - 1; t >= s0; --t) {
        const int64_t q = base + t * st;
        xv = x[q] - (scr[q] * xv);
        x[q] = xv;
    }
    if (bc.lo == BC_DIRICHLET)
        x[base] = bc.lo_val;
    if (bc.hi == BC_DIRICHLET)
        x[base + (int64_t)(n - 1) * st] = bc.hi_val;
}

// -------------------------------------------------------------------------------------------------
// explicit instantiations
// -------------------------------------------------------------------------------------------------
#define LSS_INST(T)                                                                                           \
    template __global__ void k_bnd_a<T>(const Prob<T> *, const Tile *, int);                                  \
    template __global__ void k_bnd_b<T>(const Prob<T> *, const Tile *, int);                                  \
    template __global__ void k_xsweep<T>(const Prob<T> *, const Tile *, int, int);                            \
    template __global__ void k_vsweep<T>(const Prob<T> *, const Tile *, int, int);                            \
    template __global__ void k_hv_y3<T>(const Prob<T> *, const Tile *, int);                                  \
    template __global__ void k_euler_res<T, PAT_GENERIC>(const Prob<T> *, const Tile *);                      \
    template __global__ void k_euler_res<T, PAT_HESTON>(const Prob<T> *, const Tile *);                       \
    template __global__ void k_euler_res<T, PAT_SABR0>(const Prob<T> *, const Tile *);                        \
    template __global__ void k_euler<T, PAT_GENERIC>(const Prob<T> *, const Tile *, int);                     \
    template __global__ void k_euler<T, PAT_HESTON>(const Prob<T> *, const Tile *, int);                      \
    template __global__ void k_euler<T, PAT_SABR0>(const Prob<T> *, const Tile *, int);                       \
    template __global__ void k_tridiag_lane<T>(int64_t, int, int, const T *, const T *, const T *, const T *, \
                                               T *, T *, LineBC<T>);
LSS_INST(double)
LSS_INST(float)
#undef LSS_INST

}  // namespace lssamd
