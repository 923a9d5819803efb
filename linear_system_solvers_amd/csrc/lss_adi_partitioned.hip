// lss_adi_partitioned.hip -- bandwidth-oriented line sweeps for gfx950: partitioned tridiagonal
// solves whose intermediates stay in registers and LDS.
//
// A line with interior unknowns 1..N (Dirichlet ends lo, hi) is cut into NCH = 64 chunks of at most
// M consecutive points.  Chunk l owns points [s, e]; its last point e is the chunk's interface
// unknown xi_l.  With xi_{l-1} (left) and xi_l (right) as Dirichlet values, the chunk's interior
// points s..e-1 are eliminated by Thomas with three right-hand sides (the data and the two spikes):
//     x_k = p_k + q_k xi_{l-1} + r_k xi_l            (k = s..e-1, held in registers)
// Substituting the last interior point of chunk l and the first interior point of chunk l+1 into
// the equation of point e gives one tridiagonal equation per chunk in xi_{l-1}, xi_l, xi_{l+1};
// that 64-unknown interface system is solved by parallel cyclic reduction, after which every chunk
// finishes its points.  Each input is read once and each output written once.
//
//   k_vsweep_p: one wavefront per v-line (contiguous in HBM); lane = chunk; the interface system
//               is reduced across lanes with shuffles.
//   k_xsweep_p: a workgroup covers XTJ = 16 adjacent S-lines (one 128-byte row segment per access)
//               and 64 row chunks per line (16 wavefronts); the interface system is reduced through LDS.
//               The HV stage-2 sweep evaluates the explicit Y3 stage pointwise inside its right-hand
//               side, so Y3 is never stored.
//
// The arithmetic of each point (coefficients, right-hand sides) is the reference's; only the
// elimination order of the line solve differs from the reference's sequential Thomas sweep.

// The fast path may fuse multiply-adds (the library is built with -ffp-contract=off so that the exact
// line solver keeps the reference's rounding; this translation unit opts back in).
#pragma clang fp contract(fast)
#include "lss_adi_chunk.hpp"
#include "lss_adi_stencil.hpp"

namespace lssamd {

// -------------------------------------------------------------------------------------------------
// v-line sweep.  stage 1: rhs_v(u, Y1) -> Y2 (DR: next); stage 2: rhs_v(u, Y3) (CS/MCS) or
// rhs_v(Y2, Y4) (HV) -> next.  Ends colA0/colA1[idx1[i]] (implicit_kernel.hpp:69-84).
// -------------------------------------------------------------------------------------------------
template <typename T, int M>
__global__ void __launch_bounds__(64 * VROWS) k_vsweep_p(const Prob<T> *__restrict__ probs,
                                                         const Tile *__restrict__ tiles, int step, int stage)
{
    const Tile tl = tiles[blockIdx.x];
    const Prob<T> &p = probs[tl.prob];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (step >= p.steps || (stage == 2 && p.scheme == SCHEME_DR) || wave >= tl.count)
        return;
    const int i = tl.first + wave;
    const int ny = p.ny, N = ny - 2;
    const int m = (N + NCH - 1) / NCH;
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
    const gptr<const T> in = G(in0 + row);
    const gptr<const T> inh = G(inh0 + row);
    const gptr<T> out = G(out0 + row);
    const uint32_t ii = p.idx1[i];
    const T lo = G(p.colA0)[ii], hi = G(p.colA1)[ii];
    const int s = 1 + lane * m;
    Chunk<T, M> ch;
    ch.cnt = max(0, min(m, N - s + 1));
    const bool last_real = ch.cnt > 0 && s + ch.cnt - 1 == N;
#pragma unroll
    for (int t = 0; t < M; ++t) {
        if (t < ch.cnt) {
            const int jj = s + t;
            const Stencil<T> st = stencil_at(p, cs, i, jj);
            Eq<T> q;
            q.a = (-theta * st.Mt);
            q.b = (one + theta * st.W);
            q.c = (-theta * st.Pt);
            q.r = rhs_v(p, st, in + jj, inh + jj);
            if (jj == 1) {  // fold the Dirichlet ends into the first / last equation
                q.r = q.r - q.a * lo;
                q.a = (T)0;
            }
            if (jj == N) {
                q.r = q.r - q.c * hi;
                q.c = (T)0;
            }
            ch.put(t, q);
        }
    }
    ch.bwd();
    T PF, QF, RF;
    ch.first_rel(PF, QF, RF);
    T nPF = __shfl_down(PF, 1), nQF = __shfl_down(QF, 1), nRF = __shfl_down(RF, 1);
    if (last_real || lane == 63) {
        nPF = (T)0;
        nQF = (T)0;
        nRF = (T)0;
    }
    Eq<T> e = ch.cnt > 0 ? ch.interface_eq(nPF, nQF, nRF) : identity_eq<T>();
#pragma unroll
    for (int d = 1; d < NCH; d <<= 1) {
        const T rb = (T)1 / e.b;
        Eq<T> l, h;
        l.a = __shfl_up(e.a, d);
        l.b = __shfl_up(rb, d);
        l.c = __shfl_up(e.c, d);
        l.r = __shfl_up(e.r, d);
        h.a = __shfl_down(e.a, d);
        h.b = __shfl_down(rb, d);
        h.c = __shfl_down(e.c, d);
        h.r = __shfl_down(e.r, d);
        if (lane < d)
            l = identity_eq<T>();
        if (lane + d >= NCH)
            h = identity_eq<T>();
        e = pcr_combine(e, l, h);
    }
    const T xi = e.r / e.b;
    T xl = __shfl_up(xi, 1);
    if (lane == 0)
        xl = (T)0;  // the first chunk's left neighbour is folded into its equations
#pragma unroll
    for (int t = 0; t < M; ++t)
        if (t < ch.cnt)
            out[s + t] = ch.value(t, xl, xi);
    if (lane == 0)
        out[0] = lo;
    if (last_real)
        out[ny - 1] = hi;
}

// -------------------------------------------------------------------------------------------------
// S-line sweep.  kind 0: rhs_intermed_1(u) -> Y1; kind 1/2: CS/MCS rhs_intermed_3(u, Y1, Y2) -> Y3;
// kind 3: HV rhs_intermed_4(Y2, Y3) with Y3 evaluated in place from (u, Y1, Y2) -> Y4 (= y4).
// Ends u(0, idx2[j]), u(nx-1, idx2[j]) (implicit_kernel.hpp:86-101).
// -------------------------------------------------------------------------------------------------
template <typename T, int M, int STAGE>
__global__ void __launch_bounds__(XTJ *NCH) k_xsweep_p(const Prob<T> *__restrict__ probs,
                                                       const Tile *__restrict__ tiles, int step)
{
    constexpr int stage = STAGE;
    __shared__ T sh[2][4][NCH][XTJ];
    const Tile tl = tiles[blockIdx.x];
    const Prob<T> &p = probs[tl.prob];
    if (step >= p.steps || (stage == 2 && p.scheme == SCHEME_DR))
        return;  // uniform over the workgroup
    const int col = threadIdx.x % XTJ, chn = threadIdx.x / XTJ;
    const bool colok = col < tl.count;
    const int j = tl.first + (colok ? col : 0);
    const int nx = p.nx, ny = p.ny, N = nx - 2;
    const int m = (N + NCH - 1) / NCH;
    const gptr<const T> u = G((const T *)((step & 1) ? p.b : p.a));
    const gptr<const T> cs = coeff_set(p, step);
    const gptr<const T> y1 = G((const T *)p.y1), y2 = G((const T *)p.y2);
    const T one = (T)1.0, theta = p.theta;
    const int kind = (stage == 1) ? 0 : (p.scheme == SCHEME_CS ? 1 : (p.scheme == SCHEME_MCS ? 2 : 3));
    const gptr<T> out = G((kind == 0) ? p.y1 : (kind == 3 ? p.y4 : p.y3));
    const uint32_t jj = p.idx2[j];
    const T lo = u[jj], hi = u[(int64_t)(nx - 1) * ny + jj];
    const int s = 1 + chn * m;
    Chunk<T, M> ch;
    ch.cnt = colok ? max(0, min(m, N - s + 1)) : 0;
    const bool last_real = ch.cnt > 0 && s + ch.cnt - 1 == N;
#pragma unroll
    for (int t = 0; t < M; ++t) {
        if (t < ch.cnt) {
            const int k = s + t;
            const int64_t q = (int64_t)k * ny + j;
            const Stencil<T> st = stencil_at(p, cs, k, j);
            Eq<T> e;
            e.a = (-theta * st.M);
            e.b = (one + theta * st.Z);
            e.c = (-theta * st.P);
            if (kind == 0) {
                e.r = rhs_intermed_1(p, st, u + q, ny);
            } else if (kind == 1) {
                e.r = rhs_intermed_3<T, false>(p, st, u + q, y1 + q, y2 + q, ny);
            } else if (kind == 2) {
                e.r = rhs_intermed_3<T, true>(p, st, u + q, y1 + q, y2 + q, ny);
            } else {
                const gptr<const T> y2q = y2 + q;
                const T y3 = hv_y3_at(p, st, u + q, y1 + q, y2q, ny);
                e.r = (-theta * st.M * y2q[-ny]) + (theta * st.Z * y2q[0]) - (theta * st.P * y2q[ny]) + y3;
            }
            if (k == 1) {
                e.r = e.r - e.a * lo;
                e.a = (T)0;
            }
            if (k == N) {
                e.r = e.r - e.c * hi;
                e.c = (T)0;
            }
            ch.put(t, e);
        }
    }
    ch.bwd();
    // next chunk's first relation through LDS
    {
        T PF, QF, RF;
        ch.first_rel(PF, QF, RF);
        sh[1][0][chn][col] = PF;
        sh[1][1][chn][col] = QF;
        sh[1][2][chn][col] = RF;
    }
    __syncthreads();
    T nPF = (T)0, nQF = (T)0, nRF = (T)0;
    if (!last_real && chn + 1 < NCH) {
        nPF = sh[1][0][chn + 1][col];
        nQF = sh[1][1][chn + 1][col];
        nRF = sh[1][2][chn + 1][col];
    }
    Eq<T> e = ch.cnt > 0 ? ch.interface_eq(nPF, nQF, nRF) : identity_eq<T>();
    int buf = 0;
    sh[0][0][chn][col] = e.a;
    sh[0][1][chn][col] = (T)1 / e.b;
    sh[0][2][chn][col] = e.c;
    sh[0][3][chn][col] = e.r;
    __syncthreads();
#pragma unroll
    for (int d = 1; d < NCH; d <<= 1) {
        Eq<T> l = identity_eq<T>(), h = identity_eq<T>();
        if (chn >= d) {
            l.a = sh[buf][0][chn - d][col];
            l.b = sh[buf][1][chn - d][col];
            l.c = sh[buf][2][chn - d][col];
            l.r = sh[buf][3][chn - d][col];
        }
        if (chn + d < NCH) {
            h.a = sh[buf][0][chn + d][col];
            h.b = sh[buf][1][chn + d][col];
            h.c = sh[buf][2][chn + d][col];
            h.r = sh[buf][3][chn + d][col];
        }
        e = pcr_combine(e, l, h);
        buf ^= 1;
        sh[buf][0][chn][col] = e.a;
        sh[buf][1][chn][col] = (T)1 / e.b;
        sh[buf][2][chn][col] = e.c;
        sh[buf][3][chn][col] = e.r;
        __syncthreads();
    }
    const T xi = e.r / e.b;
    sh[buf ^ 1][0][chn][col] = xi;
    __syncthreads();
    const T xl = chn > 0 ? sh[buf ^ 1][0][chn - 1][col] : (T)0;
    if (!colok)
        return;
#pragma unroll
    for (int t = 0; t < M; ++t)
        if (t < ch.cnt)
            out[(int64_t)(s + t) * ny + j] = ch.value(t, xl, xi);
    if (chn == 0)
        out[j] = lo;
    if (last_real)
        out[(int64_t)(nx - 1) * ny + j] = hi;
}

#define LSS_INST_P(T, M)                                                                  \
    template __global__ void k_vsweep_p<T, M>(const Prob<T> *, const Tile *, int, int);   \
    template __global__ void k_xsweep_p<T, M, 1>(const Prob<T> *, const Tile *, int);                     \
    template __global__ void k_xsweep_p<T, M, 2>(const Prob<T> *, const Tile *, int);
LSS_INST_P(double, 1)
LSS_INST_P(double, 2)
LSS_INST_P(double, 4)
LSS_INST_P(double, 8)
LSS_INST_P(double, 16)
LSS_INST_P(float, 1)
LSS_INST_P(float, 2)
LSS_INST_P(float, 4)
LSS_INST_P(float, 8)
LSS_INST_P(float, 16)
#undef LSS_INST_P

}  // namespace lssamd
