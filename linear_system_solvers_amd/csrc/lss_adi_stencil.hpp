// lss_adi_stencil.hpp -- device functions shared by the ADI kernels: stencil coefficients at a grid
// point (FIELDS or SEPARABLE) and the stage right-hand sides, each in the reference's operand order.
#pragma once

#include "lss_adi_device.hpp"

namespace lssamd {

template <typename T> struct Stencil {
    T M, Mt, P, Pt, Z, W, C;
};

// Factor patterns of SEPARABLE coefficient sets (codes of Prob::fcode: slots 2q + r of X, then of
// Y; -1 zero, -2 one, k >= 0 packed vector k).  PAT 0 reads a problem's codes at run time; PAT 1
// and 2 are the patterns of the reference models' data (Heston with r != 0, SABR with r = 0), fixed
// at compile time so that zero and unit factors cost nothing.  The plan selects a pattern only
// when every problem has exactly its codes.
enum : int32_t { PAT_GENERIC = 0, PAT_HESTON = 1, PAT_SABR0 = 2, PAT_COUNT = 3 };
template <int PAT> struct FPat;
template <> struct FPat<PAT_HESTON> {
    static constexpr int c[24] = {0, -1, -2, -1, 1, -1, 2, 3, -2, -1, -2, -1,
                                  0, -1, 1, -1, 2, -1, -2, 3, 4, -1, 5, -1};
};
template <> struct FPat<PAT_SABR0> {
    static constexpr int c[24] = {0, -1, -2, -1, 1, -1, -1, 2, -2, -1, -2, -1,
                                  0, -1, 1, -1, 2, -1, -2, 3, 4, -1, -1, -1};
};

// Derivative-term tables of the reference models' parameters in the packed form the plan stores
// (Adj::tp, tq, fx, fy; fx/fy: -2 one, k >= 0 packed factor vector k): Heston (kappa, theta_v,
// sigma, rho, r) and SABR (nu, rho, beta, scale of the alpha axis), in the order the plan generates them.  The adjoint
// kernels use them at compile time when every problem of the plan has exactly these tables and the
// same factor pattern.
template <int PAT> struct TPat;
template <> struct TPat<PAT_HESTON> {
    static constexpr int n = 8;
    static constexpr int tp[8] = {0, 1, 2, 2, 2, 3, 4, 4};
    static constexpr int tq[8] = {4, 4, 1, 2, 4, 2, 3, 5};
    static constexpr int fx[8] = {-2, -2, -2, 0, -2, 0, 0, -2};
    static constexpr int fy[8] = {0, 1, 2, 3, 4, 5, -2, 6};
};
template <> struct TPat<PAT_SABR0> {  // (with y_lo = 0 the alpha-axis terms of b and e vanish)
    static constexpr int n = 10;
    static constexpr int tp[10] = {0, 0, 0, 1, 2, 2, 2, 3, 3, 3};
    static constexpr int tq[10] = {1, 2, 4, 2, 0, 2, 3, 0, 2, 3};
    static constexpr int fx[10] = {-2, 0, -2, 0, 1, 2, 3, 4, 0, 5};
    static constexpr int fy[10] = {0, 1, 2, 3, 4, 5, 4, 6, 5, 6};
};

// factor slot SLOT of a packed SEPARABLE block B (length-n vectors) at index i
template <typename T, int PAT, int SLOT, typename P>
__device__ __forceinline__ T factor(const Prob<T> &p, P B, int n, int i)
{
    if constexpr (PAT == PAT_GENERIC) {
        const int code = p.fcode[SLOT];
        return code >= 0 ? B[code * n + i] : (code == -2 ? (T)1 : (T)0);
    } else {
        constexpr int code = FPat<PAT>::c[SLOT];
        if constexpr (code >= 0)
            return B[code * n + i];
        else
            return code == -2 ? (T)1 : (T)0;
    }
}

// number of packed vectors a pattern reads in its X block (first = 0) or Y block (first = 12); 0 for
// the run-time pattern
template <int PAT> __host__ __device__ constexpr int fpat_nvec(int first)
{
    if constexpr (PAT == PAT_GENERIC) {
        return 0;
    } else {
        int m = -1;
        for (int s = first; s < first + 12; ++s)
            m = FPat<PAT>::c[s] > m ? FPat<PAT>::c[s] : m;
        return m + 1;
    }
}

// q(i, j) = X[2K][i] Y[2K][j] + X[2K+1][i] Y[2K+1][j]
template <typename T, int PAT, int K, typename PX, typename PY>
__device__ __forceinline__ T sep(const Prob<T> &p, PX X, PY Y, int i, int j)
{
    return factor<T, PAT, 2 * K>(p, X, p.nx, i) * factor<T, PAT, 12 + 2 * K>(p, Y, p.ny, j) +
           factor<T, PAT, 2 * K + 1>(p, X, p.nx, i) * factor<T, PAT, 12 + 2 * K + 1>(p, Y, p.ny, j);
}

// Stencil coefficients at (zeta_i, eta_j) of SEPARABLE data whose X and Y blocks are read through
// the accessors X and Y (global memory, or a copy in LDS): general_heston_equation_coefficients
// (coefficients.hpp:89-95).
template <typename T, int PAT, typename PX, typename PY>
__device__ __forceinline__ Stencil<T> stencil_xy(const Prob<T> &p, PX X, PY Y, int i, int j)
{
    Stencil<T> s;
    const T a = sep<T, PAT, 0>(p, X, Y, i, j), b = sep<T, PAT, 1>(p, X, Y, i, j);
    const T c = sep<T, PAT, 2>(p, X, Y, i, j), d = sep<T, PAT, 3>(p, X, Y, i, j);
    const T e = sep<T, PAT, 4>(p, X, Y, i, j), f = sep<T, PAT, 5>(p, X, Y, i, j);
    const T two = (T)2.0, half = (T)0.5;
    s.M = (p.alpha * a - p.delta * d);
    s.Mt = (p.beta * b - p.ni * e);
    s.P = (p.alpha * a + p.delta * d);
    s.Pt = (p.beta * b + p.ni * e);
    s.Z = (two * p.alpha * a - half * p.rho * f);
    s.W = (two * p.beta * b - half * p.rho * f);
    s.C = c;
    return s;
}

// Stencil coefficients at (zeta_i, eta_j): general_heston_equation_coefficients (coefficients.hpp:89-95).
template <typename T, int PAT = PAT_GENERIC, typename P>
__device__ __forceinline__ Stencil<T> stencil_at(const Prob<T> &p, P cs, int i, int j)
{
    Stencil<T> s;
    if (PAT == PAT_GENERIC && p.coeff_mode == COEFF_FIELDS) {
        const int nxy = p.nx * p.ny, q = i * p.ny + j;
        s.M = cs[q];
        s.Mt = cs[nxy + q];
        s.P = cs[2 * nxy + q];
        s.Pt = cs[3 * nxy + q];
        s.Z = cs[4 * nxy + q];
        s.W = cs[5 * nxy + q];
        s.C = cs[6 * nxy + q];
    } else {
        s = stencil_xy<T, PAT>(p, cs, cs + 12 * p.nx, i, j);
    }
    return s;
}

// D, E, F on the v = 0 row at zeta_i (used by boundary stage A).
template <typename T, int PAT = PAT_GENERIC, typename P>
__device__ __forceinline__ void def_row0(const Prob<T> &p, P cs, int i, T &D, T &E, T &F)
{
    if (PAT == PAT_GENERIC && p.coeff_mode == COEFF_FIELDS) {
        const P r = cs + 7 * (int)p.nx * p.ny;
        D = r[i];
        E = r[p.nx + i];
        F = r[2 * p.nx + i];
    } else {
        const P X = cs, Y = cs + 12 * (int)p.nx;
        D = sep<T, PAT, 3>(p, X, Y, i, 0);
        E = sep<T, PAT, 4>(p, X, Y, i, 0);
        F = sep<T, PAT, 5>(p, X, Y, i, 0);
    }
}

template <typename T> __device__ __forceinline__ gptr<const T> coeff_set(const Prob<T> &p, int step)
{
    return G(p.coeff + (p.coeff_sets > 1 ? (int64_t)step : 0) * p.coeff_stride);
}
template <typename T> __device__ __forceinline__ gptr<const T> bnd_set(const Prob<T> &p, int step)
{
    return G(p.bnd + (p.bnd_sets > 1 ? (int64_t)step : 0) * p.bnd_stride);
}

// Boundary stage A of step `step` at row i (explicit_boundary.hpp:39-84, :181-200): colA0[i], the new
// v = 0 value from the explicit degenerate row, and colA1[i], the v_max value.  It reads the old
// level only; the v-sweeps of the step read what it writes.
template <typename T, int PAT = PAT_GENERIC> __device__ __forceinline__ void bnd_a_row(const Prob<T> &p, int step, int i)
{
    const int nx = p.nx, ny = p.ny, N = nx - 1;
    const gptr<const T> u = G((const T *)((step & 1) ? p.b : p.a));
    const gptr<const T> cs = coeff_set(p, step);
    const gptr<const T> bs = bnd_set(p, step);
    const T four = (T)4.0, three = (T)3.0, two = (T)2.0, one = (T)1.0;
    const T delta = p.delta, ni = p.ni, rho = p.rho;
    T v;
    if (i == 0) {
        v = bs[nx];  // h_lower(t, Y(0))
    } else {
        T D, E, F;
        def_row0<T, PAT>(p, cs, i, D, E, F);
        const gptr<const T> ui = u + (int64_t)i * ny;
        if (i == N) {
            const T delta_ = two * p.h1 * bs[nx + ny];  // 2 h_1 * h_upper(t, Y(0)) X'(1)
            v = ((one - three * ni * E + rho * F) * ui[0]) + (four * ni * E * ui[1]) - (ni * E * ui[2]) -
                (delta * delta_ * D);
        } else {
            v = (-delta * D * ui[-ny]) + ((one - three * ni * E + rho * F) * ui[0]) + (delta * D * ui[ny]) -
                (ni * E * ui[2]) + (four * ni * E * ui[1]);
        }
    }
    G(p.colA0)[i] = v;
    G(p.colA1)[i] = bs[i];  // v_upper(t, X(zeta_i))
}

// Two-dimensional point accessors: in(di, dj) is the value at (i + di, j + dj) of the point (i, j).
// PAt: a pointer at the point and the row stride.  GRC: a global base, the point's row k and column
// j; every access forms the offset of its row, min(k + di, kmax) * ny, so that neighbouring points
// of one line read their common rows through the same addresses (one load per value).
template <typename P> struct PAt {
    P b;
    int64_t ny;
    __device__ __forceinline__ auto operator()(int di, int dj) const { return b[di * ny + dj]; }
};
template <typename T> struct GRC {
    gptr<const T> b;
    int k;
    uint32_t j, ny, kmax;  // rows are clamped to kmax (for points past a line's end)
    __device__ __forceinline__ T operator()(int di, int dj) const
    {
        return ldg(b, min((uint32_t)(k + di), kmax) * ny + (j + (uint32_t)dj));
    }
};

// -------------------------------------------------------------------------------------------------
// Right-hand sides (operand order of the reference's rhs functions)
// -------------------------------------------------------------------------------------------------
// rhs_intermed_1 (lss_heat_douglas_rachford_method.hpp:37-68); in(di, dj) is the value at (t + di, j + dj)
template <typename T, typename P>
__device__ __forceinline__ T rhs_intermed_1(const Prob<T> &p, const Stencil<T> &s, P in)
{
    const T one = (T)1.0, gamma = p.gamma, theta = p.theta;
    return (gamma * s.C * in(-1, -1)) + ((one - theta) * s.M * in(-1, 0)) - (gamma * s.C * in(-1, 1)) +
           (s.Mt * in(0, -1)) + ((one - s.W - (one - theta) * s.Z) * in(0, 0)) + (s.Pt * in(0, 1)) -
           (gamma * s.C * in(1, -1)) + ((one - theta) * s.P * in(1, 0)) + (gamma * s.C * in(1, 1));
}

// CS rhs_intermed_3 (lss_heat_craig_sneyd_method.hpp:128-156), MCS adds the (zeta - theta) term
// of lss_heat_modified_craig_sneyd_method.hpp:158-163 (with its "- P")
template <typename T, bool MCS, typename P>
__device__ __forceinline__ T rhs_intermed_3(const Prob<T> &p, const Stencil<T> &s, P in, P inh, P inhn)
{
    const T gamma = p.gamma, theta = p.theta, zeta = p.zeta;
    T v = (-theta * s.M * in(-1, 0)) + (theta * s.Z * in(0, 0)) - (theta * s.P * in(1, 0)) + (inh(0, 0)) +
          (zeta * gamma * s.C *
           ((inhn(1, 1) - in(1, 1)) - (inhn(1, -1) - in(1, -1)) - (inhn(-1, 1) - in(-1, 1)) +
            (inhn(-1, -1) - in(-1, -1))));
    if (MCS) {
        v = v + ((zeta - theta) * (s.M * (inhn(-1, 0) - in(-1, 0)) - (s.Z + s.W) * (inhn(0, 0) - in(0, 0)) -
                                   s.P * (inhn(1, 0) - in(1, 0)) + s.Mt * (inhn(0, -1) - in(0, -1)) +
                                   s.Pt * (inhn(0, 1) - in(0, 1))));
    }
    return v;
}

// HV rhs_intermed_4 (lss_heat_hundsdorfer_verwer_method.hpp:174-194)
template <typename T, typename P>
__device__ __forceinline__ T rhs_intermed_4(const Prob<T> &p, const Stencil<T> &s, P in, P inh)
{
    const T theta = p.theta;
    return (-theta * s.M * in(-1, 0)) + (theta * s.Z * in(0, 0)) - (theta * s.P * in(1, 0)) + inh(0, 0);
}

// v-direction rhs shared by every scheme (e.g. lss_heat_douglas_rachford_method.hpp:106-126)
template <typename T, typename P>
__device__ __forceinline__ T rhs_v(const Prob<T> &p, const Stencil<T> &s, P in, P inh)
{
    const T theta = p.theta;
    return (-theta * s.Mt * in[-1]) + (theta * s.W * in[0]) - (theta * s.Pt * in[1]) + inh[0];
}

// HV explicit stage value Y3 at a point (lss_heat_hundsdorfer_verwer_method.hpp:146-168);
// in = u, inh = Y1, inhn = Y2, each an accessor at the point.
template <typename T, typename P>
__device__ __forceinline__ T hv_y3_at(const Prob<T> &p, const Stencil<T> &s, P in, P inh, P inhn)
{
    const T one = (T)1.0, gamma = p.gamma, theta = p.theta, zeta = p.zeta;
    return (-theta * s.M * inh(-1, 0)) + ((one + theta * s.Z) * inh(0, 0)) - (theta * s.P * inh(1, 0)) +
           (theta * s.M * in(-1, 0)) - (theta * s.Z * in(0, 0)) + (theta * s.P * in(1, 0)) +
           (zeta * gamma * s.C *
            ((inhn(1, 1) - in(1, 1)) - (inhn(1, -1) - in(1, -1)) - (inhn(-1, 1) - in(-1, 1)) +
             (inhn(-1, -1) - in(-1, -1)))) +
           (zeta * (s.M * (inhn(-1, 0) - in(-1, 0)) - (s.W + s.Z) * (inhn(0, 0) - in(0, 0)) +
                    s.P * (inhn(1, 0) - in(1, 0)) + s.Mt * (inhn(0, -1) - in(0, -1)) +
                    s.Pt * (inhn(0, 1) - in(0, 1))));
}

}  // namespace lssamd
