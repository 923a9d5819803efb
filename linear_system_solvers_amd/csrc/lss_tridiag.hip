// lss_tridiag.hip -- batched plain tridiagonal systems for gfx950 with the register-resident
// partitioned elimination (no scratch arrays in HBM: every input is read once, the solution
// written once).
//
// System k has unknowns x_0..x_{n-1} and rows a_i x_{i-1} + b_i x_i + c_i x_{i+1} = d_i, whose end rows
// are folded with the end conditions `bc` (LineBC: Dirichlet / Neumann / Robin as the reference's
// thomas_lu_boundary.hpp:114-205 folds them, or PLAIN: a_0 and c_{n-1} ignored).  The folded system
// has unknowns start..end; Dirichlet end values are written to x_0 / x_{n-1}.
//   k_tridiag_rows  contiguous layout x[k*n + i]: one wavefront per system; the system is staged
//                   through LDS so that every global access is a coalesced row of the system.
//   k_tridiag_cols  strided layout x[i*B + k]: XTJ systems x NCH chunks per workgroup, so that each
//                   access reads XTJ consecutive systems of one row.

// The fast path may fuse multiply-adds (the library is built with -ffp-contract=off so that the exact
// line solver keeps the reference's rounding; this translation unit opts back in).
#pragma clang fp contract(fast)
#include "lss_adi_chunk.hpp"

namespace lssamd {

constexpr int TRI_WAVES = 4;  // systems per workgroup of k_tridiag_rows

template <typename T, int M>
__global__ void __launch_bounds__(64 * TRI_WAVES) k_tridiag_rows(int64_t B, int n, const T *__restrict__ a,
                                                                  const T *__restrict__ b, const T *__restrict__ c,
                                                                  const T *__restrict__ d, T *__restrict__ x,
                                                                  LineBC<T> bc)
{
    extern __shared__ unsigned char lds_raw[];
    T *lds = reinterpret_cast<T *>(lds_raw);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t k = (int64_t)blockIdx.x * TRI_WAVES + wave;
    const bool live = k < B;
    T *sa = lds + (size_t)wave * 4 * n, *sb = sa + n, *sc = sb + n, *sd = sc + n;
    if (live) {
        const int64_t base = k * n;
        for (int i = lane; i < n; i += 64) {
            sa[i] = a[base + i];
            sb[i] = b[base + i];
            sc[i] = c[base + i];
            sd[i] = d[base + i];
        }
    }
    __syncthreads();
    const int s0 = bc.start(), e0 = bc.end(n), nf = e0 - s0 + 1;
    if (live) {
        auto eq = [&](int kk) -> Eq<T> {
            const int i = s0 + min(kk, nf) - 1;  // points past the end repeat the last equation
            Eq<T> e;
            e.a = sa[i];
            e.b = sb[i];
            e.c = sc[i];
            e.r = sd[i];
            bc.fold(i == s0, i == e0, e.a, e.b, e.c, e.r);
            return e;
        };
        auto put = [&](int, int kk, T v) { sd[s0 + kk - 1] = v; };  // every eq() has run before the first put()
        wave_solve<T, M>(nf, lane, eq, put);
        if (lane == 0) {
            if (bc.lo == BC_DIRICHLET)
                sd[0] = bc.lo_val;
            if (bc.hi == BC_DIRICHLET)
                sd[n - 1] = bc.hi_val;
        }
    }
    __syncthreads();
    if (live) {
        const int64_t base = k * n;
        for (int i = lane; i < n; i += 64)
            x[base + i] = sd[i];
    }
}

template <typename T, int M>
__global__ void __launch_bounds__(XTJ *NCH) k_tridiag_cols(int64_t B, int n, const T *__restrict__ a,
                                                           const T *__restrict__ b, const T *__restrict__ c,
                                                           const T *__restrict__ d, T *__restrict__ x, LineBC<T> bc)
{
    __shared__ T sh[2][4][NCH][XTJ];
    const int col = threadIdx.x % XTJ, chn = threadIdx.x / XTJ;
    const int64_t k = (int64_t)blockIdx.x * XTJ + col;
    const bool live = k < B;
    const int s0 = bc.start(), e0 = bc.end(n), nf = e0 - s0 + 1;
    auto eq = [&](int kk) -> Eq<T> {
        const int i = s0 + min(kk, nf) - 1;  // points past the end repeat the last equation
        const int64_t q = (int64_t)i * B + k;
        Eq<T> e;
        e.a = a[q];
        e.b = b[q];
        e.c = c[q];
        e.r = d[q];
        bc.fold(i == s0, i == e0, e.a, e.b, e.c, e.r);
        return e;
    };
    auto put = [&](int, int kk, T v) { x[(int64_t)(s0 + kk - 1) * B + k] = v; };
    block_solve<T, M, XTJ>(sh, nf, col, chn, live, eq, put);
    if (live && chn == 0) {
        if (bc.lo == BC_DIRICHLET)
            x[k] = bc.lo_val;
        if (bc.hi == BC_DIRICHLET)
            x[(int64_t)(n - 1) * B + k] = bc.hi_val;
    }
}

#define LSS_INST_T(T, M)                                                                                        \
    template __global__ void k_tridiag_rows<T, M>(int64_t, int, const T *, const T *, const T *, const T *, T *,  \
                                                  LineBC<T>);                                                    \
    template __global__ void k_tridiag_cols<T, M>(int64_t, int, const T *, const T *, const T *, const T *, T *,  \
                                                  LineBC<T>);
LSS_INST_T(double, 1)
LSS_INST_T(double, 2)
LSS_INST_T(double, 4)
LSS_INST_T(double, 8)
LSS_INST_T(double, 16)
LSS_INST_T(float, 1)
LSS_INST_T(float, 2)
LSS_INST_T(float, 4)
LSS_INST_T(float, 8)
LSS_INST_T(float, 16)
#undef LSS_INST_T

}  // namespace lssamd
