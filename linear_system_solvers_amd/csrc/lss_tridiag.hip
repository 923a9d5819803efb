// lss_tridiag.hip -- batched plain tridiagonal systems for gfx950 with the register-resident
// partitioned elimination (no scratch arrays in HBM: every input is read once, the solution
// written once).
//
// System k has unknowns x_0..x_{n-1}: b_0 x_0 + c_0 x_1 = d_0, a_i x_{i-1} + b_i x_i + c_i x_{i+1} = d_i,
// a_{n-1} x_{n-2} + b_{n-1} x_{n-1} = d_{n-1}; a_0 and c_{n-1} are ignored.  This is the reference's
// Thomas LU with Neumann(0)/Neumann(0) ends (thomas_lu_boundary.hpp:136-144, 184-191).
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
                                                                  const T *__restrict__ d, T *__restrict__ x)
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
    if (live) {
        auto eq = [&](int kk) -> Eq<T> {
            const int i = min(kk, n) - 1;  // points past the end repeat the last equation
            Eq<T> e;
            e.a = i == 0 ? (T)0 : sa[i];
            e.b = sb[i];
            e.c = i == n - 1 ? (T)0 : sc[i];
            e.r = sd[i];
            return e;
        };
        auto put = [&](int, int kk, T v) { sd[kk - 1] = v; };  // every eq() has run before the first put()
        wave_solve<T, M>(n, lane, eq, put);
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
                                                           const T *__restrict__ d, T *__restrict__ x)
{
    __shared__ T sh[2][4][NCH][XTJ];
    const int col = threadIdx.x % XTJ, chn = threadIdx.x / XTJ;
    const int64_t k = (int64_t)blockIdx.x * XTJ + col;
    const bool live = k < B;
    auto eq = [&](int kk) -> Eq<T> {
        const int i = min(kk, n) - 1;  // points past the end repeat the last equation
        const int64_t q = (int64_t)i * B + k;
        Eq<T> e;
        e.a = i == 0 ? (T)0 : a[q];
        e.b = b[q];
        e.c = i == n - 1 ? (T)0 : c[q];
        e.r = d[q];
        return e;
    };
    auto put = [&](int, int kk, T v) { x[(int64_t)(kk - 1) * B + k] = v; };
    block_solve<T, M, XTJ>(sh, n, col, chn, live, eq, put);
}

#define LSS_INST_T(T, M)                                                                                        \
    template __global__ void k_tridiag_rows<T, M>(int64_t, int, const T *, const T *, const T *, const T *, T *); \
    template __global__ void k_tridiag_cols<T, M>(int64_t, int, const T *, const T *, const T *, const T *, T *);
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
