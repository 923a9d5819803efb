// lss_adi_chunk.hpp -- register-resident partitioned tridiagonal elimination shared by the forward
// line sweeps and their adjoints.
//
// A line with interior unknowns 1..N is cut into NCH = 64 chunks of at most M consecutive points.
// Chunk l owns points [s, e]; its last point e is the chunk's interface unknown xi_l.  With xi_{l-1}
// (left) and xi_l (right) as Dirichlet values, the chunk's interior points s..e-1 are eliminated by
// Thomas with three right-hand sides (the data and the two spikes):
//     x_k = p_k + q_k xi_{l-1} + r_k xi_l            (k = s..e-1, held in registers)
// Substituting the last interior point of chunk l and the first interior point of chunk l+1 into the
// equation of point e gives one tridiagonal equation per chunk in xi_{l-1}, xi_l, xi_{l+1}; that
// 64-unknown interface system is solved by parallel cyclic reduction (wave shuffles or LDS; each step
// exchanges the neighbours' rows divided by their diagonals), after which every chunk finishes its
// points.
//
// Every chunk is full: a line whose N is not a multiple of the chunk length is padded with points
// past its end, whose equations the caller supplies (finite, with c = 0 and b != 0; see wave_solve_c).
// The last real point has c = 0 (its end is folded), so nothing the padding holds reaches a real
// point: its spike coefficient g = c / b' is exactly zero, which zeroes every product with padded
// values in the back substitution and every coupling of a real interface row to a padded one.  With
// full chunks the point count of a chunk is a compile-time constant, and the elimination needs no
// per-point selects.
#pragma once

#include "lss_adi_device.hpp"

#include <type_traits>

namespace lssamd {

constexpr int NCH = 64;   // chunks per line
constexpr int XTJ = 16;   // S-lines per x-sweep workgroup
constexpr int VROWS = 4;  // v-lines (wavefronts) per v-sweep workgroup

template <typename T> struct Eq {
    T a, b, c, r;
};

// 1 / x from the hardware reciprocal refined by Newton steps (two in double, one in float): full
// working precision for the diagonals met here (normal numbers of modest range) in half the
// instructions of the IEEE division sequence.
__device__ __forceinline__ double fast_rcp(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-x, r, 1.0);
    return __builtin_fma(r, e, r);
}
__device__ __forceinline__ float fast_rcp(float x)
{
    const float r = __builtin_amdgcn_rcpf(x);
    const float e = __builtin_fmaf(-x, r, 1.0f);
    return __builtin_fmaf(r, e, r);
}

// Register-resident elimination of one chunk of M points (M - 1 interior points and the interface
// point M - 1).  After put() of every point and bwd(): g, zp, zq hold the relation
// x_t = zp[t] + zq[t] xi_{l-1} + g[t] xi_l of each interior point t, and (ae, be, ce, re) is the
// equation of the interface point.
template <typename T, int M> struct Chunk {
    T g[M], zp[M], zq[M];
    T ae, be, ce, re;
    T pl, ql, rl;  // relation of the last interior point (xi_{l-1} itself when M = 1)

    // Forward elimination of point t (a compile-time index after unrolling; every array entry is
    // addressed with a compile-time index, since a run-time one would move the arrays to scratch
    // memory).  x_t = zp[t] + zq[t] xi_{l-1} - g[t] x_{t+1} afterwards.
    __device__ __forceinline__ void put(int t, const Eq<T> &q)
    {
        if (t == M - 1) {
            ae = q.a;
            be = q.b;
            ce = q.c;
            re = q.r;
            return;
        }
#pragma unroll
        for (int u = 0; u < M - 1; ++u) {
            if (u == t) {
                const T rb = fast_rcp(u == 0 ? q.b : q.b - q.a * g[u > 0 ? u - 1 : 0]);
                g[u] = q.c * rb;
                zp[u] = (u == 0 ? q.r : q.r - q.a * zp[u > 0 ? u - 1 : 0]) * rb;
                zq[u] = (u == 0 ? -q.a : -q.a * zq[u > 0 ? u - 1 : 0]) * rb;
            }
        }
    }

    // Backward substitution of the spikes (x_{M-1} = xi_l).
    __device__ __forceinline__ void bwd()
    {
        if constexpr (M >= 2) {
            g[M - 2] = -g[M - 2];
#pragma unroll
            for (int u = M - 3; u >= 0; --u) {
                const T gu = g[u];
                g[u] = -gu * g[u + 1];
                zp[u] = zp[u] - gu * zp[u + 1];
                zq[u] = zq[u] - gu * zq[u + 1];
            }
            pl = zp[M - 2];
            ql = zq[M - 2];
            rl = g[M - 2];
        } else {
            pl = (T)0;
            ql = (T)1;
            rl = (T)0;
        }
    }

    // a chunk that holds no points of a line: its interface row is the identity and its relations
    // are zero
    __device__ __forceinline__ void set_identity()
    {
#pragma unroll
        for (int u = 0; u < M; ++u)
            g[u] = zp[u] = zq[u] = (T)0;
        ae = ce = re = (T)0;
        be = (T)1;
        pl = ql = rl = (T)0;
    }

    // (p, q, r) of the first interior point; a chunk of one point has none: seen from the left
    // neighbour, its first point is xi_l itself (0, 0, 1).
    __device__ __forceinline__ void first_rel(T &P, T &Q, T &R) const
    {
        if constexpr (M >= 2) {
            P = zp[0];
            Q = zq[0];
            R = g[0];
        } else {
            P = (T)0;
            Q = (T)0;
            R = (T)1;
        }
    }

    // interface equation of chunk l, given the next chunk's first relation (PF, QF, RF)
    __device__ __forceinline__ Eq<T> interface_eq(T PF, T QF, T RF) const
    {
        Eq<T> q;
        q.a = ae * ql;
        q.b = ae * rl + be + ce * QF;
        q.c = ce * RF;
        q.r = re - ae * pl - ce * PF;
        return q;
    }

    __device__ __forceinline__ T value(int t, T xl, T xr) const
    {
        T v = xr;
#pragma unroll
        for (int u = 0; u < M - 1; ++u)
            if (u == t)
                v = zp[u] + zq[u] * xl + g[u] * xr;
        return v;
    }
};

// A row divided by its diagonal, (a, c, r) / b: what a cyclic-reduction step needs of a neighbour
// (three values to exchange instead of four).
template <typename T> struct NEq {
    T a, c, r;
};
template <typename T> __device__ __forceinline__ NEq<T> normalized(const Eq<T> &e)
{
    const T rb = fast_rcp(e.b);
    NEq<T> n;
    n.a = e.a * rb;
    n.c = e.c * rb;
    n.r = e.r * rb;
    return n;
}
template <typename T> __device__ __forceinline__ NEq<T> zero_neq()
{
    NEq<T> n;
    n.a = n.c = n.r = (T)0;
    return n;
}

// one cyclic-reduction step: eliminate the neighbours' unknowns (lo, hi normalised; zero when absent)
template <typename T> __device__ __forceinline__ Eq<T> pcr_step(const Eq<T> &me, const NEq<T> &lo, const NEq<T> &hi)
{
    Eq<T> n;
    n.a = -me.a * lo.a;
    n.c = -me.c * hi.c;
    n.b = me.b - me.a * lo.c - me.c * hi.a;
    n.r = me.r - me.a * lo.r - me.c * hi.r;
    return n;
}

template <typename T> __device__ __forceinline__ Eq<T> identity_eq()
{
    Eq<T> q;
    q.a = (T)0;
    q.b = (T)1;
    q.c = (T)0;
    q.r = (T)0;
    return q;
}


// -------------------------------------------------------------------------------------------------
// Generic line solves.  A line of N interior unknowns is cut into chunks of C = chunk_len(N) points
// (the smallest of 1, 2, 3, 4, 6, 8, 16 with NCH * C >= N, so the partition depends on N alone); the chunks are padded past the line's end to C
// points each.  Each thread evaluates eq for all C points of its
// chunk without a per-point branch, so the loads of all its points are issued together.
//   eq(k) returns the equation of interior unknown k (1..N) with the ends already folded (a = 0 at
//   k = 1, c = 0 at k = N).  It is also called with N < k <= NCH * C (the padding): it must then read
//   only in-bounds data and return a finite equation with c = 0 and b != 0 (folding the upper end
//   at every k >= N gives this, as does repeating the folded row N).
//   put(t, k, x) receives the solution of unknown k (k <= N), the t-th point of the calling
//   thread's chunk (t is a compile-time index after unrolling, so a caller may keep the values in
//   registers and store them later).  With PUT_ALL, put is called for all C points without
//   per-point branches (so that the caller's loads can be issued together); then k > N marks a
//   padded point whose x is finite but meaningless.
// -------------------------------------------------------------------------------------------------
__device__ __forceinline__ int chunk_len(int N, int nch = NCH)
{
    int c = 1;
    while (c * nch < N)
        c <<= 1;
    // 3 or 6 when they suffice: fewer points past the end (12 is not used: its code path slowed the
    // 16-point kernels by more than the padding it saves)
    if (c >= 4 && c <= 8 && (3 * c / 4) * nch >= N)
        c = 3 * c / 4;
    return c;
}

// the equation of the t-th point (unknown k) of the calling thread's chunk: eq(t, k) when eq takes the
// position too (t is a compile-time index after unrolling, so eq may read per-point registers), else eq(k)
template <typename EqF> __device__ __forceinline__ auto eq_at(EqF &eq, int t, int k)
{
    if constexpr (std::is_invocable_v<EqF &, int, int>)
        return eq(t, k);
    else
        return eq(k);
}

// f(std::integral_constant<int, C>) for the chunk length c (a value of chunk_len <= M, M a power of
// two; uniform)
template <int M, typename F> __device__ __forceinline__ void with_chunk_len(int c, F &&f)
{
    if constexpr (M > 1) {
        if (c <= M / 2) {
            with_chunk_len<M / 2>(c, f);
            return;
        }
    }
    if constexpr (M >= 4 && M <= 8) {
        if (c <= 3 * M / 4) {
            f(std::integral_constant<int, 3 * M / 4>{});
            return;
        }
    }
    f(std::integral_constant<int, M>{});
}

// One line per NL lanes (NL = 64: one line per wavefront; 32: two), lane = chunk (its index within
// the line's lanes); the interface system is reduced with shuffles inside each group of NL lanes.
// Every lane of the wavefront must call it, and the lines of a wavefront must have the same N;
// wave_solve_c takes the chunk length C = chunk_len(N, NL) as a template argument.
template <typename T, int C, bool PUT_ALL = false, int NL = NCH, typename EqF, typename PutF>
__device__ __forceinline__ void wave_solve_c(int N, int lane, EqF &&eq, PutF &&put)
{
    const int s = 1 + lane * C;
    Chunk<T, C> ch;
    if (s <= N) {
#pragma unroll
        for (int t = 0; t < C; ++t)
            ch.put(t, eq_at(eq, t, s + t));
        ch.bwd();
    } else {  // a chunk wholly past the end: identity rows, and no loads
        ch.set_identity();
    }
    T PF, QF, RF;
    ch.first_rel(PF, QF, RF);
    // (the last lane takes its own relation: its interface row has c = 0, as every row at k >= N)
    const T nPF = __shfl_down(PF, 1, NL), nQF = __shfl_down(QF, 1, NL), nRF = __shfl_down(RF, 1, NL);
    Eq<T> e = ch.interface_eq(nPF, nQF, nRF);
    // A lane without a neighbour at distance d takes its own row from the shuffle instead of zeros:
    // its coupling to that side is already exactly zero (a = 0 on the first d lanes, c = 0 on the
    // last d: the line's ends are folded, the padding past the end holds rows with c = 0, and a step
    // keeps the zeros, n.a = -a * lo.a, n.c = -c * hi.c), so the finite row it meets is multiplied
    // by zero and every result is the one a zero row gives.
#pragma unroll
    for (int d = 1; d < NL; d <<= 1) {
        const NEq<T> m = normalized(e);
        NEq<T> l, h;
        l.a = __shfl_up(m.a, d, NL);
        l.c = __shfl_up(m.c, d, NL);
        l.r = __shfl_up(m.r, d, NL);
        h.a = __shfl_down(m.a, d, NL);
        h.c = __shfl_down(m.c, d, NL);
        h.r = __shfl_down(m.r, d, NL);
        e = pcr_step(e, l, h);
    }
    const T xi = e.r * fast_rcp(e.b);
    // (lane 0 takes its own xi: its relations' xi_{l-1} coefficients are zero, a = 0 at k = 1)
    const T xl = __shfl_up(xi, 1, NL);
    if constexpr (PUT_ALL) {
#pragma unroll
        for (int t = 0; t < C; ++t)
            put(t, s + t, ch.value(t, xl, xi));
    } else {
#pragma unroll
        for (int t = 0; t < C; ++t)
            if (s + t <= N)
                put(t, s + t, ch.value(t, xl, xi));
    }
}

// the same for the chunk length of N (chunk_len(N, NL) <= M)
template <typename T, int M, bool PUT_ALL = false, int NL = NCH, typename EqF, typename PutF>
__device__ __forceinline__ void wave_solve(int N, int lane, EqF &&eq, PutF &&put)
{
    with_chunk_len<M>(chunk_len(N, NL),
                      [&](auto cc) { wave_solve_c<T, decltype(cc)::value, PUT_ALL, NL>(N, lane, eq, put); });
}

// W lines per workgroup of W * NC threads (thread = (chunk chn, line col)), NC chunks per line; the
// interface system is reduced through LDS `sh`.  Every thread of the workgroup must call it (it
// holds barriers) with the same N (chunk_len(N, NC) <= M); threads of lines that do not exist pass
// active = false (their eq is not called).
template <typename T, int M, int W, bool PUT_ALL = false, int NC = NCH, typename EqF, typename PutF>
__device__ __forceinline__ void block_solve(T (*sh)[4][NC][W], int N, int col, int chn, bool active, EqF &&eq,
                                            PutF &&put)
{
    with_chunk_len<M>(chunk_len(N, NC), [&](auto cc) {
        constexpr int C = decltype(cc)::value;
        const int s = 1 + chn * C;
        Chunk<T, C> ch;
        if (active && s <= N) {
#pragma unroll
            for (int t = 0; t < C; ++t)
                ch.put(t, eq_at(eq, t, s + t));
            ch.bwd();
        } else {  // no line, or a chunk wholly past its end: identity rows, and no loads
            ch.set_identity();
        }
        {
            T PF, QF, RF;
            ch.first_rel(PF, QF, RF);
            sh[1][0][chn][col] = PF;
            sh[1][1][chn][col] = QF;
            sh[1][2][chn][col] = RF;
        }
        __syncthreads();
        // (the last chunk takes its own relation: its interface row has c = 0)
        const int cn = chn + 1 < NC ? chn + 1 : chn;
        const T nPF = sh[1][0][cn][col], nQF = sh[1][1][cn][col], nRF = sh[1][2][cn][col];
        Eq<T> e = ch.interface_eq(nPF, nQF, nRF);
        int buf = 0;
        {
            const NEq<T> m = normalized(e);
            sh[0][0][chn][col] = m.a;
            sh[0][1][chn][col] = m.c;
            sh[0][2][chn][col] = m.r;
        }
        __syncthreads();
        // (a chunk without a neighbour at distance d reads its own row: its coupling to that side
        // is exactly zero, as in wave_solve_c)
#pragma unroll
        for (int d = 1; d < NC; d <<= 1) {
            NEq<T> l, h;
            const int cl = chn >= d ? chn - d : chn, ch_ = chn + d < NC ? chn + d : chn;
            l.a = sh[buf][0][cl][col];
            l.c = sh[buf][1][cl][col];
            l.r = sh[buf][2][cl][col];
            h.a = sh[buf][0][ch_][col];
            h.c = sh[buf][1][ch_][col];
            h.r = sh[buf][2][ch_][col];
            e = pcr_step(e, l, h);
            buf ^= 1;
            const NEq<T> m = normalized(e);
            sh[buf][0][chn][col] = m.a;
            sh[buf][1][chn][col] = m.c;
            sh[buf][2][chn][col] = m.r;
            __syncthreads();
        }
        const T xi = e.r * fast_rcp(e.b);
        sh[buf ^ 1][0][chn][col] = xi;
        __syncthreads();
        const T xl = sh[buf ^ 1][0][chn > 0 ? chn - 1 : 0][col];
        if (!active)
            return;
        if constexpr (PUT_ALL) {
#pragma unroll
            for (int t = 0; t < C; ++t)
                put(t, s + t, ch.value(t, xl, xi));
        } else {
#pragma unroll
            for (int t = 0; t < C; ++t)
                if (s + t <= N)
                    put(t, s + t, ch.value(t, xl, xi));
        }
    });
}

// First point and point count of the chunk of thread position `pos` (lane or chunk index) of a line
// of N interior unknowns, as the solvers above partition it.
__device__ __forceinline__ void chunk_span(int N, int pos, int &s, int &cnt, int nch = NCH)
{
    const int c = chunk_len(N, nch);
    s = 1 + pos * c;
    cnt = max(0, min(c, N - s + 1));
}

}  // namespace lssamd
