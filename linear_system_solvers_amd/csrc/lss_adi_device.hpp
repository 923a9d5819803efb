// lss_adi_device.hpp -- device-side problem layout shared by the ADI kernels and the host plan code.
//
// One "problem" is one Heston/SABR-type PDE on an nx x ny grid (row i = S/F index, column j = v/alpha
// index, row-major in HBM: u[i*ny + j]).  A plan holds an array of problems; the single-problem
// stepper is a plan with one problem, the batched book a plan with many.  Every kernel receives the
// problem table plus a tile table that maps each workgroup to (problem, first line), so one launch
// covers one stage of every active problem.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lssamd {

// splitting_method_enum (common/lss_enumerations.hpp:35-41)
enum : int32_t { SCHEME_DR = 0, SCHEME_CS = 1, SCHEME_MCS = 2, SCHEME_HV = 3, SCHEME_EULER = 4 };
// operation order of the line solve: Thomas LU (lss_thomas_lu_solver.hpp:125-191) or
// double sweep (lss_double_sweep_solver.hpp:112-155)
enum : int32_t { TRI_THOMAS = 0, TRI_DSWEEP = 1 };
// how the stencil coefficients reach the kernels
//   FIELDS:    M, M~, P, P~, Z, W, C sampled per grid point (7 nx*ny arrays) plus D, E, F on the
//              v = 0 row (3 nx arrays), exactly as the reference evaluates them.
//   SEPARABLE: each transformed coefficient q in {a,b,c,d,e,f} is q(i,j) = X0q[i]*Y0q[j] + X1q[i]*Y1q[j]
//              (12 nx vectors, 12 ny vectors); M..W are formed in the kernel.
enum : int32_t { COEFF_FIELDS = 0, COEFF_SEPARABLE = 1 };
// line solver of the sweeps: the reference's sequential order (one lane per line) or partitioned
enum : int32_t { LINE_EXACT = 0, LINE_PARTITIONED = 1 };

constexpr int TILE = 64;  // lines (or points) per workgroup tile: one wavefront

// Global-memory pointers.  The arrays are reached through pointers stored in the problem tables, so
// the compiler cannot see that they point to global memory and would emit flat instructions (extra
// latency, and they count against the LDS wait counter too); kernels convert each pointer once.
template <typename T> using gptr = __attribute__((address_space(1))) T *;
template <typename T> __device__ __forceinline__ gptr<T> G(T *p) { return (gptr<T>)p; }

// Loads and stores at an unsigned 32-bit element index from a base pointer: with the base in scalar
// registers this is the global_load/store "saddr + 32-bit offset" form, free of 64-bit address
// arithmetic (a problem's arrays hold at most 2^28 elements).
template <typename T> __device__ __forceinline__ T ldg(gptr<const T> base, uint32_t idx)
{
    return *(gptr<const T>)((__attribute__((address_space(1))) const char *)base + idx * (uint32_t)sizeof(T));
}
template <typename T> __device__ __forceinline__ void stg(gptr<T> base, uint32_t idx, T v)
{
    *(gptr<T>)((__attribute__((address_space(1))) char *)base + idx * (uint32_t)sizeof(T)) = v;
}
// read-only array and read-only view at one grid point (offsets in elements, may be negative)
template <typename T> struct GArr {
    gptr<const T> b;
    __device__ __forceinline__ T operator[](uint32_t i) const { return ldg(b, i); }
    __device__ __forceinline__ GArr operator+(uint32_t o) const { return GArr{b + o}; }
};
// read-only array in LDS (address space 3, so that accesses are ds_read, not flat)
template <typename T> using lptr = __attribute__((address_space(3))) T *;
template <typename T> struct LArr {
    lptr<const T> b;
    __device__ __forceinline__ T operator[](uint32_t i) const { return b[i]; }
};
template <typename T> struct GAt {
    gptr<const T> b;
    uint32_t q;
    __device__ __forceinline__ T operator[](int d) const { return ldg(b, q + (uint32_t)d); }
};
// view at element k of a row starting at `row`, every access clamped to element kmax of the row
// (for points past a line's end, whose values are discarded); within the line it equals GAt
template <typename T> struct GAtC {
    gptr<const T> b;
    uint32_t row;
    int k;
    uint32_t kmax;
    __device__ __forceinline__ T operator[](int d) const { return ldg(b, row + min((uint32_t)(k + d), kmax)); }
};

template <typename T> struct Prob {
    int32_t nx, ny, steps, scheme;
    int32_t tri, coeff_mode, coeff_sets, bnd_sets;
    // general_heston_equation_coefficients scalars (coefficients.hpp:62-73) and theta, zeta
    T theta, zeta, alpha, beta, gamma, delta, ni, rho, h1;
    T *a, *b;                  // ping-pong state: step s reads (s even ? a : b), writes the other
    T *y1, *y2, *y3, *scr;     // intermediates and line-solver scratch (nx*ny each)
    T *y4;                     // HV's fourth stage: Y1's buffer (exact sweeps) or Y3's (partitioned)
    T *colA0, *colA1;          // v = 0 and v = v_max columns produced by boundary stage A (nx each)
    const uint32_t *idx1;      // index_of maps (lss_grid_config.hpp:110-118)
    const uint32_t *idx2;
    const T *coeff;            // coeff_sets sets of coeff_stride values
    // SEPARABLE: factor slot 2q + r of X (codes 0..11) and of Y (codes 12..23): -1 the vector is zero,
    // -2 it is one, k >= 0 it is stored as packed vector k of its block
    int32_t fcode[24];
    const T *bnd;              // bnd_sets sets of bnd_stride values: [v_upper(nx) | h_lower(ny) | h_upper*X'(1) (ny)]
    int64_t coeff_stride, bnd_stride;
    // heat source: null, or src_sets (1 or steps) sets of [F(t) | F(t_next)] (nx*ny each, row-major)
    const T *src;
    int32_t src_sets, pad_;
};

// the heat-source set of loop step `step` (null without a source)
template <typename T> __device__ __forceinline__ const T *src_set(const Prob<T> &p, int step)
{
    return p.src_sets == 0 ? nullptr
                           : p.src + (size_t)(p.src_sets == 1 ? 0 : step) * 2 * (size_t)p.nx * (size_t)p.ny;
}

// Workgroups are dispatched to the 8 XCDs round-robin (block b runs on XCD b % 8), and every XCD has
// its own L2.  xcd_tile maps block b so that each XCD takes a contiguous run of the g tiles:
// neighbouring tiles, which share cache lines and stencil halos, then meet in the same L2.
constexpr int NXCD = 8;
__device__ __forceinline__ int xcd_tile(int b, int g)
{
    const int q = g / NXCD, r = g % NXCD, x = b % NXCD, k = b / NXCD;
    return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + k;
}

struct Tile {
    int32_t prob;   // problem index
    int32_t first;  // first line / row / column of the tile
    int32_t count;  // number of lines in the tile (<= TILE)
    int32_t col;    // first column of a two-dimensional tile (0 otherwise)
};


// -------------------------------------------------------------------------------------------------
// Reverse mode (adjoint) of the time loop
// -------------------------------------------------------------------------------------------------
constexpr int NPAR = 6;    // model parameters per problem
constexpr int GTR = 8, GTC = 32;  // explicit step: GTR x GTC threads, each ERPT rows of one column
// gather tiles of the adjoint: AGR x AGC points; one thread per point of the tile and its halo ring
// (AGT = (AGR + 2) x (AGC + 2) = 512), so that every equation point is formed in one pass
constexpr int AGR = 14, AGC = 30, AGT = (AGR + 2) * (AGC + 2);
constexpr int ERPT = 8;           // rows per thread of the explicit step (its tiles: GTR * ERPT rows)
constexpr int ERES_T = 256, ERES_K = 6;  // threads and points per thread of the resident explicit loop

// v-lines per wavefront of the partitioned v-sweep for lines of N interior points and the plan's
// chunk bound M: 2 when the line fits 32 chunks of at most M <= 4 points, else 1 (the plan's v-line
// tiles and the kernel follow the same rule; larger chunk bounds keep one line per wavefront, whose
// register budget the second body would strain)
__host__ __device__ constexpr int vlines_per_wave(int N, int M) { return M <= 4 && N <= 32 * M ? 2 : 1; }
constexpr int NTERM = 12;  // derivative terms per problem

// Adjoint state of one problem.  The forward values a reverse step needs (U_s, U_{s+1} and the
// stage outputs of step s) are reached through the step's Prob table; Adj holds the lambdas.
//   la, lb     lambda of the time levels: reverse step s reads lambda_{s+1} from ((s + 1) & 1 ? lb : la)
//              and writes lambda_s to (s & 1 ? lb : la)
//   lv2, lx2   lambda of the equations of the last v-stage (DR: its only v-stage) and of the second
//   lv1, lx1   S-stage, of the first v-stage and of the first S-stage; written at equation points only
//   elo2, elo1 lambda of the lower (v = 0) Dirichlet end of each v-line, per row
//   inv*       inverse index maps: rows r with idx1[r] == i are inv1i[inv1o[i] .. inv1o[i + 1]) (same for idx2)
//   dfac       derivative terms: term t says d q_t(i, j) / d param[tp[t]] = X_t[i] * Y_t[j] for the
//              transformed coefficient q = a..f (tq[t] = 0..5).  Distinct factor vectors are stored
//              once: per set, nfx X vectors (nx each) then nfy Y vectors (ny each); X_t is vector fx[t]
//              (-2: all ones, not stored), Y_t is vector fy[t]
template <typename T> struct Adj {
    T *la, *lb;
    T *lv2, *lx2, *lv1, *lx1;
    T *ly2;                     // lambda(Y2) gathered from the X2 (and HV) equations (CS, MCS, HV)
    T *elo2, *elo1;
    const uint32_t *inv1o, *inv1i, *inv2o, *inv2i;
    const T *dfac;
    int64_t dfac_stride;
    int32_t dfac_sets, nterm, nfx, nfy;
    int32_t tp[NTERM], tq[NTERM], fx[NTERM], fy[NTERM];
};

// End conditions of a batched line solve, the same for every system of the batch, folded into the
// end rows as thomas_lu_solver_boundary does (lss_thomas_lu_boundary.hpp:114-205): Dirichlet (the end
// value is val; the unknowns run from 1 or to n - 2), Neumann (u' = -val, a ghost point folded into
// the first / last row) and Robin (u' + lin u + val = 0), with h2 = 2 h = 2 / (n - 1).  PLAIN keeps
// the rows as given (a[0] and c[n-1] ignored).
enum : int32_t { BC_DIRICHLET = 0, BC_NEUMANN = 1, BC_ROBIN = 2, BC_PLAIN = 3 };
template <typename T> struct LineBC {
    int32_t lo, hi;
    T lo_lin, lo_val, hi_lin, hi_val, h2;
    __host__ __device__ int start() const { return lo == BC_DIRICHLET ? 1 : 0; }
    __host__ __device__ int end(int n) const { return hi == BC_DIRICHLET ? n - 2 : n - 1; }
    // folds the first (row start) or last (row end) equation; operand order of the reference
    __device__ __forceinline__ void fold(bool first, bool last, T &a, T &b, T &c, T &r) const
    {
        if (first) {
            if (lo == BC_DIRICHLET) {
                r = r - a * lo_val;
            } else if (lo == BC_NEUMANN) {
                c = a + c;
                r = r - a * (h2 * lo_val);
            } else if (lo == BC_ROBIN) {
                b = b + a * (h2 * lo_lin);
                c = a + c;
                r = r - a * (h2 * lo_val);
            }
            a = (T)0;
        }
        if (last) {
            if (hi == BC_DIRICHLET) {
                r = r - c * hi_val;
            } else if (hi == BC_NEUMANN) {
                a = a + c;
                r = r + c * (h2 * hi_val);
            } else if (hi == BC_ROBIN) {
                a = a + c;
                b = b - c * (h2 * hi_lin);
                r = r + c * (h2 * hi_val);
            }
            c = (T)0;
        }
    }
};

// one device-to-device copy of n elements (checkpoint and lambda transfers)
struct CopyJob {
    const void *src;
    void *dst;
    int64_t n;
};

}  // namespace lssamd
