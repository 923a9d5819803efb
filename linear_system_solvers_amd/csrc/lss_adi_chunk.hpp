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

// Register-resident chunk elimination.  After fwd()/bwd(): g holds r_k, zp holds p_k, zq holds q_k
// for the cnt - 1 interior points, and (ae, be, ce, re) is the equation of the chunk's last point.
template <typename T, int M> struct Chunk {
    T g[M], zp[M], zq[M];
    T ae, be, ce, re;
    int cnt;

    // Forward elimination of point t (a compile-time index after unrolling).  The array entries are
    // written unconditionally, so no two stores share an address computed at run time (which would
    // move the arrays to scratch memory); the entries of the chunk's last point are never read.
    __device__ __forceinline__ void put(int t, const Eq<T> &q)
    {
#pragma unroll
        for (int u = 0; u < M; ++u) {
            if (u == t) {
                const T rb = fast_rcp(u == 0 ? q.b : q.b - q.a * g[u > 0 ? u - 1 : 0]);
                g[u] = q.c * rb;
                zp[u] = (u == 0 ? q.r : q.r - q.a * zp[u > 0 ? u - 1 : 0]) * rb;
                zq[u] = (u == 0 ? -q.a : -q.a * zq[u > 0 ? u - 1 : 0]) * rb;
            }
        }
        if (t == cnt - 1) {
            ae = q.a;
            be = q.b;
            ce = q.c;
            re = q.r;
        }
    }

    // Backward substitution of the spikes.  Every array element is addressed with a compile-time
    // index (a runtime index would move the arrays to scratch memory); (pl, ql, rl) keep the
    // relation of the last interior point, whose position depends on cnt.
    T pl, ql, rl;

    __device__ __forceinline__ void bwd()
    {
        const int ni = cnt - 1;
        pl = (T)0;
        ql = (T)1;
        rl = (T)0;
#pragma unroll
        for (int u = M - 1; u >= 0; --u) {
            const bool valid = u < ni, lastp = u == ni - 1;
            const T gu = g[u];
            T gn = -gu, pn = zp[u], qn = zq[u];
            if (u + 1 < M) {
                const T g1 = g[u + 1], p1 = zp[u + 1], q1 = zq[u + 1];
                gn = lastp ? -gu : -gu * g1;
                pn = lastp ? zp[u] : zp[u] - gu * p1;
                qn = lastp ? zq[u] : zq[u] - gu * q1;
            }
            g[u] = valid ? gn : gu;
            zp[u] = valid ? pn : zp[u];
            zq[u] = valid ? qn : zq[u];
            pl = lastp ? pn : pl;
            ql = lastp ? qn : ql;
            rl = lastp ? gn : rl;
        }
    }

    // (p, q, r) of the first and of the last interior point; a chunk of one point has none:
    // then x_{e-1} = xi_{l-1} (0, 1, 0) and, seen from the left neighbour, x_s = xi_l (0, 0, 1).
    __device__ __forceinline__ void first_rel(T &P, T &Q, T &R) const
    {
        if (cnt >= 2) {
            P = zp[0];
            Q = zq[0];
            R = g[0];
        } else {
            P = (T)0;
            Q = (T)0;
            R = (T)1;
        }
    }
    __device__ __forceinline__ void last_rel(T &P, T &Q, T &R) const
    {
        P = pl;
        Q = ql;
        R = rl;
    }

    // interface equation of chunk l, given the next chunk's first relation (PF, QF, RF)
    __device__ __forceinline__ Eq<T> interface_eq(T PF, T QF, T RF) const
    {
        T PL, QL, RL;
        last_rel(PL, QL, RL);
        Eq<T> q;
        q.a = ae * QL;
        q.b = ae * RL + be + ce * QF;
        q.c = ce * RF;
        q.r = re - ae * PL - ce * PF;
        return q;
    }

    __device__ __forceinline__ T value(int t, T xl, T xr) const
    {
        T v = xr;
#pragma unroll
        for (int u = 0; u < M; ++u)
            if (u == t && u < cnt - 1)
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
// (the smallest power of two with NCH * C >= N, so the partition depends on N alone); the last
// non-empty chunk may be partial and the chunks after it are empty.  Each thread with a non-empty
// chunk evaluates eq for all C of its points without a per-point branch, so the loads of all its
// points are issued together.
//   eq(k) returns the equation of interior unknown k (1..N) with the ends already folded (a = 0 at
//   k = 1, c = 0 at k = N).  It is also called with N < k <= NCH * C (points past the end of the
//   last chunk, whose equations are discarded): it must then read only in-bounds data.
//   put(t, k, x) receives the solution of unknown k (k <= N), the t-th point of the calling
//   thread's chunk (t is a compile-time index after unrolling, so a caller may keep the values in
//   registers and store them later).  With PUT_ALL, put is called for all C points of a non-empty
//   chunk without per-point branches (so that the caller's loads can be issued together); then
//   k > N marks a discarded point whose x is arbitrary (possibly not finite).
// -------------------------------------------------------------------------------------------------
__device__ __forceinline__ int chunk_len(int N, int nch = NCH)
{
    int c = 1;
    while (c * nch < N)
        c <<= 1;
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

// f(std::integral_constant<int, C>) for the chunk length c (a power of two <= M; uniform)
template <int M, typename F> __device__ __forceinline__ void with_chunk_len(int c, F &&f)
{
    if constexpr (M > 1) {
        if (c < M) {
            with_chunk_len<M / 2>(c, f);
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
    ch.cnt = max(0, min(C, N - s + 1));
    const bool last_real = ch.cnt > 0 && s + ch.cnt - 1 == N;
    if (ch.cnt > 0) {
#pragma unroll
        for (int t = 0; t < C; ++t)
            ch.put(t, eq_at(eq, t, s + t));
        ch.bwd();
    }
    T PF, QF, RF;
    ch.first_rel(PF, QF, RF);
    T nPF = __shfl_down(PF, 1, NL), nQF = __shfl_down(QF, 1, NL), nRF = __shfl_down(RF, 1, NL);
    if (last_real || lane == NL - 1) {
        nPF = (T)0;
        nQF = (T)0;
        nRF = (T)0;
    }
    Eq<T> e = ch.cnt > 0 ? ch.interface_eq(nPF, nQF, nRF) : identity_eq<T>();
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
        if (lane < d)
            l = zero_neq<T>();
        if (lane + d >= NL)
            h = zero_neq<T>();
        e = pcr_step(e, l, h);
    }
    const T xi = e.r * fast_rcp(e.b);
    T xl = __shfl_up(xi, 1, NL);
    if (lane == 0)
        xl = (T)0;
    if constexpr (PUT_ALL) {
        if (ch.cnt > 0) {
#pragma unroll
            for (int t = 0; t < C; ++t)
                put(t, s + t, ch.value(t, xl, xi));
        }
    } else {
#pragma unroll
        for (int t = 0; t < C; ++t)
            if (t < ch.cnt)
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
        ch.cnt = active ? max(0, min(C, N - s + 1)) : 0;
        const bool last_real = ch.cnt > 0 && s + ch.cnt - 1 == N;
        if (ch.cnt > 0) {
#pragma unroll
            for (int t = 0; t < C; ++t)
                ch.put(t, eq_at(eq, t, s + t));
            ch.bwd();
        }
        {
            T PF, QF, RF;
            ch.first_rel(PF, QF, RF);
            sh[1][0][chn][col] = PF;
            sh[1][1][chn][col] = QF;
            sh[1][2][chn][col] = RF;
        }
        __syncthreads();
        T nPF = (T)0, nQF = (T)0, nRF = (T)0;
        if (!last_real && chn + 1 < NC) {
            nPF = sh[1][0][chn + 1][col];
            nQF = sh[1][1][chn + 1][col];
            nRF = sh[1][2][chn + 1][col];
        }
        Eq<T> e = ch.cnt > 0 ? ch.interface_eq(nPF, nQF, nRF) : identity_eq<T>();
        int buf = 0;
        {
            const NEq<T> m = normalized(e);
            sh[0][0][chn][col] = m.a;
            sh[0][1][chn][col] = m.c;
            sh[0][2][chn][col] = m.r;
        }
        __syncthreads();
#pragma unroll
        for (int d = 1; d < NC; d <<= 1) {
            NEq<T> l = zero_neq<T>(), h = zero_neq<T>();
            if (chn >= d) {
                l.a = sh[buf][0][chn - d][col];
                l.c = sh[buf][1][chn - d][col];
                l.r = sh[buf][2][chn - d][col];
            }
            if (chn + d < NC) {
                h.a = sh[buf][0][chn + d][col];
                h.c = sh[buf][1][chn + d][col];
                h.r = sh[buf][2][chn + d][col];
            }
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
        const T xl = chn > 0 ? sh[buf ^ 1][0][chn - 1][col] : (T)0;
        if (!active)
            return;
        if constexpr (PUT_ALL) {
            if (ch.cnt > 0) {
#pragma unroll
                for (int t = 0; t < C; ++t)
                    put(t, s + t, ch.value(t, xl, xi));
            }
        } else {
#pragma unroll
            for (int t = 0; t < C; ++t)
                if (t < ch.cnt)
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
