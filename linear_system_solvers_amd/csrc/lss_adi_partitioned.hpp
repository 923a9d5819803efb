// lss_adi_partitioned.hpp -- bandwidth-oriented line sweeps for gfx950: partitioned tridiagonal
// solves (lss_adi_chunk.hpp) whose intermediates stay in registers and LDS.  Each input is read
// once and each output written once.
//
//   k_vsweep_p<VR>: VR wavefronts per workgroup, one v-line (contiguous in HBM) per wavefront, or two
//                   for short lines; lane = chunk; the interface system is reduced across lanes with
//                   shuffles.
//   k_xsweep_p<W>:  a workgroup covers W adjacent S-lines (W = 16: one 128-byte row segment per
//                   access) and 64 row chunks per line; the interface system is reduced through LDS.
//                   The stage-1 sweep also evaluates boundary stage A (one launch less per step).
//                   The HV stage-2 sweep evaluates the explicit Y3 stage pointwise inside its
//                   right-hand side, so Y3 is never stored.
// W and VR are the tile/lane mappings the plan's autotuner chooses between; they do not change the
// arithmetic of a line.  (One wavefront per S-line, as in the v-sweep, spreads each access over 64
// rows; measured on the Heston book it ran the S-sweeps at about half the speed of W = 8.)
//
// The coefficients and right-hand sides of each point follow the reference's formulas (multiply-adds
// may be fused here); the elimination order of the line solve differs from the reference's
// sequential Thomas sweep.

// The fast path may fuse multiply-adds (the library is built with -ffp-contract=off so that the exact
// line solver keeps the reference's rounding; this translation unit opts back in).
#pragma once
#pragma clang fp contract(fast)
#include "lss_adi_chunk.hpp"
#include "lss_adi_stencil.hpp"

namespace lssamd {

// -------------------------------------------------------------------------------------------------
// v-line sweep.  stage 1: rhs_v(u, Y1) -> Y2 (DR: next); stage 2: rhs_v(u, Y3) (CS/MCS) or
// rhs_v(Y2, Y4) (HV) -> next.  Ends colA0/colA1[idx1[i]] (implicit_kernel.hpp:69-84).
// -------------------------------------------------------------------------------------------------
// LPW v-lines per wavefront: 1 (the row is wavefront-uniform, in scalar registers), or 2 for lines of
// at most 32 M interior points (lanes 0-31 and 32-63 each hold one line: half the reduction steps per
// line).  A half without a line of its own repeats its neighbour's line and stores nothing, so both
// halves take the same chunk length.
template <typename T, int M, int VR, int PAT, int LPW>
__device__ __forceinline__ void vsweep_lines(const Prob<T> &p, const Tile &tl, int step, int stage)
{
    const GArr<T> cs{coeff_set(p, step)};
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (wave * LPW >= tl.count)
        return;
    constexpr int NL = 64 / LPW;
    const int half = LPW == 2 ? lane >> 5 : 0, ll = lane & (NL - 1);
    const bool own = wave * LPW + half < tl.count;  // this half has a line of its own
    int i;
    if constexpr (LPW == 1)
        i = __builtin_amdgcn_readfirstlane(tl.first + wave);
    else
        i = tl.first + min(wave * LPW + half, tl.count - 1);
    const int ny = p.ny, N = ny - 2;
    const T *u = (step & 1) ? p.b : p.a;
    T *nxt = (step & 1) ? p.a : p.b;
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
    const uint32_t row = (uint32_t)i * (uint32_t)ny;
    const gptr<const T> in = G(in0), inh = G(inh0);
    const gptr<T> out = G(out0);
    const uint32_t ii = p.idx1[i];
    const T lo = G(p.colA0)[ii], hi = G(p.colA1)[ii];
    const uint32_t jmax = (uint32_t)(N + 1);  // points past the line's end read column N + 1
    auto eq = [&](int jj) -> Eq<T> {
        // (the Y factors are read from the caches: a copy in LDS per workgroup cost what it saved)
        const Stencil<T> st = stencil_at<T, PAT>(p, cs, i, (int)min((uint32_t)jj, jmax));
        Eq<T> q;
        q.a = (-theta * st.Mt);
        q.b = (one + theta * st.W);
        q.c = (-theta * st.Pt);
        q.r = rhs_v(p, st, GAtC<T>{in, row, jj, jmax}, GAtC<T>{inh, row, jj, jmax});
        if (jj == 1) {  // fold the Dirichlet ends into the first / last equation
            q.r = q.r - q.a * lo;
            q.a = (T)0;
        }
        if (jj >= N) {  // (and the padding past the end: c = 0 there too)
            q.r = q.r - q.c * hi;
            q.c = (T)0;
        }
        return q;
    };
    auto put = [&](int, int jj, T x) {
        if (own)
            stg(out, row + jj, x);
    };
    wave_solve<T, M, false, NL>(N, ll, eq, put);
    if (ll == 0 && own) {
        stg(out, row, lo);
        stg(out, row + ny - 1, hi);
    }
}

// (eight wavefronts per SIMD: the two-lines-per-wavefront body must not raise the register count of
// the kernel above 64, which costs a wavefront per SIMD on every line)
template <typename T, int M, int VR, int PAT>
__global__ void __launch_bounds__(64 * VR) __attribute__((amdgpu_waves_per_eu(M <= 4 ? 8 : 1)))
k_vsweep_p(const Prob<T> *__restrict__ probs, const Tile *__restrict__ tiles, int step, int stage, int swz)
{
    const Tile tl = tiles[swz ? xcd_tile(blockIdx.x, gridDim.x) : blockIdx.x];
    const Prob<T> &p = probs[tl.prob];
    if (step >= p.steps || (stage == 2 && p.scheme == SCHEME_DR))
        return;  // uniform over the workgroup
    if constexpr (M <= 4) {
        if (vlines_per_wave(p.ny - 2, M) == 2) {  // (the plan's tiles hold 2 VR lines for these problems)
            vsweep_lines<T, M, VR, PAT, 2>(p, tl, step, stage);
            return;
        }
    }
    vsweep_lines<T, M, VR, PAT, 1>(p, tl, step, stage);
}

// -------------------------------------------------------------------------------------------------
// S-line sweep.  STAGE 1: rhs_intermed_1(u) -> Y1.  STAGE 2: CS/MCS rhs_intermed_3(u, Y1, Y2) -> Y3;
// HV rhs_intermed_4(Y2, Y3) with Y3 evaluated in place from (u, Y1, Y2) -> Y4 (= y4).
// Ends u(0, idx2[j]), u(nx-1, idx2[j]) (implicit_kernel.hpp:86-101).
// -------------------------------------------------------------------------------------------------
// KIND 0: stage 1 (every scheme); 1: CS, 2: MCS, 3: HV stage 2 (compile time, so that the loads of
// neighbouring points are shared and issued together)
template <typename T, int KIND, int PAT> struct XLine {
    const Prob<T> &p;
    gptr<const T> u, y1, y2;
    GArr<T> cs;
    LArr<T> xl;  // the X-block factor vectors in LDS (compile-time patterns)
    gptr<T> out;
    int N, ny, step_;
    uint32_t j;
    T lo, hi;

    __device__ __forceinline__ XLine(const Prob<T> &p_, int step, int j_, lptr<const T> xsh) : p(p_)
    {
        step_ = step;
        xl = LArr<T>{xsh};
        ny = p.ny;
        N = p.nx - 2;
        j = (uint32_t)j_;
        u = G((const T *)((step & 1) ? p.b : p.a));
        cs = GArr<T>{coeff_set(p, step)};
        y1 = G((const T *)p.y1);
        y2 = G((const T *)p.y2);
        out = G(KIND == 0 ? p.y1 : (KIND == 3 ? p.y4 : p.y3));
        const uint32_t jj = p.idx2[j];
        lo = ldg(u, jj);
        hi = ldg(u, (uint32_t)(p.nx - 1) * (uint32_t)ny + jj);
    }

    __device__ __forceinline__ Eq<T> operator()(int k) const
    {
        const T one = (T)1.0, theta = p.theta;
        const uint32_t kmax = (uint32_t)(N + 1);  // points past the line's end read row N + 1
        Stencil<T> st;
        if constexpr (PAT != PAT_GENERIC)
            st = stencil_xy<T, PAT>(p, xl, cs + 12 * (uint32_t)(N + 2), (int)min((uint32_t)k, kmax), (int)j);
        else
            st = stencil_at<T, PAT>(p, cs, (int)min((uint32_t)k, kmax), (int)j);
        Eq<T> e;
        e.a = (-theta * st.M);
        e.b = (one + theta * st.Z);
        e.c = (-theta * st.P);
        const GRC<T> uq{u, k, j, (uint32_t)ny, kmax}, y1q{y1, k, j, (uint32_t)ny, kmax},
            y2q{y2, k, j, (uint32_t)ny, kmax};
        if constexpr (KIND == 0) {
            e.r = rhs_intermed_1(p, st, uq);
            // rhs_intermed_1_source (lss_heat_douglas_rachford_method.hpp:70-104), a uniform branch
            if (p.src_sets) {
                const gptr<const T> sc = G(src_set(p, step_));
                const uint32_t q = min((uint32_t)k, kmax) * (uint32_t)ny + j,
                               nxy = (uint32_t)(N + 2) * (uint32_t)ny;
                e.r = e.r + (theta * p.rho * ldg(sc, nxy + q)) + ((one - theta) * p.rho * ldg(sc, q));
            }
        } else if constexpr (KIND == 1) {
            e.r = rhs_intermed_3<T, false>(p, st, uq, y1q, y2q);
        } else if constexpr (KIND == 2) {
            e.r = rhs_intermed_3<T, true>(p, st, uq, y1q, y2q);
        } else {
            const T y3 = hv_y3_at(p, st, uq, y1q, y2q);
            e.r = (-theta * st.M * y2q(-1, 0)) + (theta * st.Z * y2q(0, 0)) - (theta * st.P * y2q(1, 0)) + y3;
        }
        if (k == 1) {
            e.r = e.r - e.a * lo;
            e.a = (T)0;
        }
        if (k >= N) {  // (and the padding past the end: c = 0 there too)
            e.r = e.r - e.c * hi;
            e.c = (T)0;
        }
        return e;
    }

    __device__ __forceinline__ void store(int k, T x) const { stg(out, (uint32_t)k * (uint32_t)ny + j, x); }
    __device__ __forceinline__ void store_ends() const
    {
        stg(out, j, lo);
        stg(out, (uint32_t)(N + 1) * (uint32_t)ny + j, hi);
    }
};

template <typename T, int M, int W, int KIND, int PAT, int NC>
__device__ __forceinline__ void xsweep_lines(T (*sh)[4][NC][W], const Prob<T> &p, const Tile &tl, int step,
                                             lptr<const T> xsh)
{
    const int col = threadIdx.x % W, chn = threadIdx.x / W;
    const bool colok = col < tl.count;
    const XLine<T, KIND, PAT> ln(p, step, tl.first + (colok ? col : 0), xsh);
    auto put = [&](int, int k, T x) { ln.store(k, x); };
    block_solve<T, M, W, false, NC>(sh, ln.N, col, chn, colok, ln, put);
    if (colok && chn == 0)
        ln.store_ends();
}

// W S-lines per workgroup of NC chunks each, the chunks spread over the workgroup (block_solve)
template <typename T, int M, int STAGE, int W, int PAT, int NC>
__global__ void __launch_bounds__(W *NC) k_xsweep_p(const Prob<T> *__restrict__ probs,
                                                    const Tile *__restrict__ tiles, int step, int swz)
{
    __shared__ T sh[2][4][NC][W];
    // the X-block factor vectors of a compile-time pattern, copied to LDS once per workgroup (every
    // point of an S-line reads them at its own row; the Y factors are the same for the whole line)
    constexpr int NVX = fpat_nvec<PAT>(0);
    __shared__ T xsh[NVX > 0 ? NVX * (NC * M + 2) : 1];
    const Tile tl = tiles[swz ? xcd_tile(blockIdx.x, gridDim.x) : blockIdx.x];
    const Prob<T> &p = probs[tl.prob];
    if (step >= p.steps || (STAGE == 2 && p.scheme == SCHEME_DR))
        return;  // uniform over the workgroup
    if constexpr (NVX > 0) {
        const GArr<T> xb{coeff_set(p, step)};
        for (int q = threadIdx.x; q < NVX * p.nx; q += blockDim.x)
            xsh[q] = xb[(uint32_t)q];
        __syncthreads();
    }
    const lptr<const T> xs = (lptr<const T>)xsh;
    if constexpr (STAGE == 1) {
        xsweep_lines<T, M, W, 0, PAT, NC>(sh, p, tl, step, xs);
        // boundary stage A of the step, its rows spread over the problem's S-line tiles (the v-sweeps
        // of the step read colA0 and colA1; this launch does not)
        const int ntl = (p.ny - 2 + W - 1) / W, k = (tl.first - 1) / W;
        for (int i = k * (W * NC) + (int)threadIdx.x; i < p.nx; i += ntl * (W * NC))
            bnd_a_row<T, PAT>(p, step, i);
    } else if (p.scheme == SCHEME_CS)
        xsweep_lines<T, M, W, 1, PAT, NC>(sh, p, tl, step, xs);
    else if (p.scheme == SCHEME_MCS)
        xsweep_lines<T, M, W, 2, PAT, NC>(sh, p, tl, step, xs);
    else
        xsweep_lines<T, M, W, 3, PAT, NC>(sh, p, tl, step, xs);
}

}  // namespace lssamd

// Instantiations (one translation unit per precision and coefficient pattern, so that the library
// builds in parallel): LSS_INST_PART(T, PAT) instantiates every sweep of precision T and pattern
// PAT for the chunk bounds the plans use.
#define LSS_INST_X(T, M, S, W, PAT, NC) \
    template __global__ void k_xsweep_p<T, M, S, W, PAT, NC>(const Prob<T> *, const Tile *, int, int);
#define LSS_INST_P(T, M, PAT)                                                                                  \
    template __global__ void k_vsweep_p<T, M, 4, PAT>(const Prob<T> *, const Tile *, int, int, int);            \
    template __global__ void k_vsweep_p<T, M, 16, PAT>(const Prob<T> *, const Tile *, int, int, int);           \
    LSS_INST_X(T, M, 1, 16, PAT, 64) LSS_INST_X(T, M, 2, 16, PAT, 64) LSS_INST_X(T, M, 1, 8, PAT, 64)          \
    LSS_INST_X(T, M, 2, 8, PAT, 64) LSS_INST_X(T, M, 1, 4, PAT, 64) LSS_INST_X(T, M, 2, 4, PAT, 64)
// stage 2 with 32 chunks per line (plans whose S-lines all have at most 512 points; chunk bound 2M)
#define LSS_INST_P32(T, M, PAT) \
    LSS_INST_X(T, M, 2, 16, PAT, 32) LSS_INST_X(T, M, 2, 8, PAT, 32) LSS_INST_X(T, M, 2, 4, PAT, 32)
// both stages with 128 chunks per line (plans with S-lines of more than 512 points; chunk bound 8)
#define LSS_INST_P128(T, PAT) \
    LSS_INST_X(T, 8, 1, 8, PAT, 128) LSS_INST_X(T, 8, 2, 8, PAT, 128) LSS_INST_X(T, 8, 1, 4, PAT, 128) \
    LSS_INST_X(T, 8, 2, 4, PAT, 128)
#define LSS_INST_PART(T, PAT)                                                                                  \
    namespace lssamd {                                                                                         \
    LSS_INST_P(T, 1, PAT) LSS_INST_P(T, 2, PAT) LSS_INST_P(T, 4, PAT) LSS_INST_P(T, 8, PAT) LSS_INST_P(T, 16, PAT) \
    LSS_INST_P32(T, 2, PAT) LSS_INST_P32(T, 4, PAT) LSS_INST_P32(T, 8, PAT) LSS_INST_P32(T, 16, PAT)          \
    LSS_INST_P128(T, PAT)                                                                                      \
    }
