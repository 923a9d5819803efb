// lss_adi_capi.hip -- host side of liblss_amd.so: plans, time loop, hipGraph capture, C-ABI.
//
// A plan owns all device memory of its problems (state ping-pong, intermediates, line-solver
// scratch, sampled coefficients and boundaries, index maps) and a problem table.  Problems are
// ordered by step count, descending, so that at loop step s the problems still stepping form a
// prefix and each launch covers exactly the active tiles.  The whole loop is recorded once into a
// hipGraph (on a plan-owned stream) and replayed on the caller's stream.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#include "lss_adi_device.hpp"
#include "lss_adi_stencil.hpp"
#include "lss_amd.h"
#include "lss_capi_common.hpp"
#include "lss/lss_adi_sampler.hpp"

namespace lssamd {

template <typename T> __global__ void k_bnd_a(const Prob<T> *, const Tile *, int);
template <typename T> __global__ void k_bnd_b(const Prob<T> *, const Tile *, int);
template <typename T> __global__ void k_xsweep(const Prob<T> *, const Tile *, int, int);
template <typename T> __global__ void k_vsweep(const Prob<T> *, const Tile *, int, int);
template <typename T> __global__ void k_hv_y3(const Prob<T> *, const Tile *, int);
template <typename T, int PAT> __global__ void k_euler(const Prob<T> *, const Tile *, int);
template <typename T, int PAT> __global__ void k_euler_res(const Prob<T> *, const Tile *);
template <typename T>
__global__ void k_tridiag_lane(int64_t, int, int, const T *, const T *, const T *, const T *, T *, T *, LineBC<T>);
template <typename T, int M, int STAGE, int W, int PAT, int NC>
__global__ void k_xsweep_p(const Prob<T> *, const Tile *, int, int);
template <typename T, int M, int VR, int PAT> __global__ void k_vsweep_p(const Prob<T> *, const Tile *, int, int, int);
template <typename T, int M, int PAT>
__global__ void k_adj_v(const Prob<T> *, const Adj<T> *, const Tile *, int, int, double *);
template <typename T, int M, int W, int PAT, int STAGE, int NC>
__global__ void k_adj_x(const Prob<T> *, const Adj<T> *, const Tile *, int, double *);
template <typename T, int WHAT, int PAT>
__global__ void k_adj_g(const Prob<T> *, const Adj<T> *, const Tile *, int, double *);
template <typename T> __global__ void k_copy_jobs(const CopyJob *);
template <typename T, int M>
__global__ void k_tridiag_rows(int64_t, int, const T *, const T *, const T *, const T *, T *, LineBC<T>);
template <typename T, int M>
__global__ void k_tridiag_cols(int64_t, int, const T *, const T *, const T *, const T *, T *, LineBC<T>);
constexpr int P_NCH = 64, P_XTJ = 16, P_VROWS = 4;

}  // namespace lssamd

using namespace lssamd;

using namespace lssamd_host;

std::string &lssamd_host::last_error()
{
    thread_local std::string message;
    return message;
}

#define HIPCHK(expr) LSS_HIPCHK(expr)

namespace {

// Layout of one problem's device allocation, in elements of T (index maps appended as uint32).
struct ProbLayout {
    size_t off_a, off_b, off_y1, off_y2, off_y3, off_scr, off_colA0, off_colA1, off_coeff, off_bnd, off_src;
    size_t n_T;          // number of T elements
    size_t off_idx_bytes, bytes;
};

}  // namespace

struct lss_adi_plan {
    virtual ~lss_adi_plan() = default;
    int fp = LSS_FP64;
    int64_t count = 0;
    std::vector<int64_t> order;      // internal position -> caller index
    std::vector<int64_t> where;      // caller index -> internal position
    std::vector<int64_t> nx, ny, steps;
    std::vector<std::vector<char>> init_host;  // model plans: sampled initial surface per caller index
    std::vector<std::vector<char>> grid_host;  // model plans: physical S (nx) then v (ny) coordinates
    // parameter derivative terms per caller index (see Adj in lss_adi_device.hpp)
    struct Terms {
        int32_t nparam = 0, nterm = 0;
        int64_t sets = 0;
        int8_t tp[NTERM] = {}, tq[NTERM] = {};
        std::vector<char> fac;  // sets * nterm * (nx + ny) values of the plan's precision
    };
    std::vector<Terms> terms;
    virtual void forward_ckpt(const void *const *u_in, void *const *u_out, int ckpt, hipStream_t stream) = 0;
    virtual void backward(const void *const *lam_out, void *const *lam_in, double *dparams, hipStream_t stream) = 0;
    virtual void invalidate_adjoint() = 0;
    virtual void tune(hipStream_t stream, bool again, int *xw, int *vr, int *xcd) = 0;
    virtual void set_mapping(int xw, int vr, int xcd) = 0;
    virtual void set_overlap(int mode) = 0;
    virtual void kernel_config(int32_t *cfg) const = 0;
    virtual void run(const void *const *u_in, void *const *u_out, hipStream_t stream, bool use_graph, bool host) = 0;
    virtual void run_levels(const void *const *u_in, void *const *levels, hipStream_t stream, bool host) = 0;
    virtual void profile(const void *const *u_in, void *const *u_out, hipStream_t stream, double *ms, double *bytes,
                         int64_t *launches) = 0;
};

namespace {

// Per-launch event timing of a plan run (used by the benchmark for the roofline figure).
struct Profiler {
    std::vector<hipEvent_t> ev;
    std::vector<int> kind;
    std::vector<double> bytes;
    hipStream_t st = nullptr;
    ~Profiler()
    {
        for (auto e : ev)
            (void)hipEventDestroy(e);
    }
    void begin(int k, double b)
    {
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        HIPCHK(hipEventRecord(e0, st));
        ev.push_back(e0);
        ev.push_back(e1);
        kind.push_back(k);
        bytes.push_back(b);
    }
    void end() { HIPCHK(hipEventRecord(ev.back(), st)); }
};

template <typename T> struct PlanImpl final : lss_adi_plan {
    std::vector<Prob<T>> probs;        // host copy (device pointers inside)
    std::vector<ProbLayout> layouts;
    dev_buffer mem;                    // all problems
    dev_buffer probs_dev, tiles_dev;
    // tile tables: rows [0,nx), cols [0,ny), S-lines [1,ny-2], v-lines [1,nx-2], interior rows
    // adjoint: S-line tiles of P_XTJ, v-line tiles of P_VROWS, one tile per row (all rows)
    // forward partitioned sweeps: S-line tiles of 16 (TB_X) or 8 (TB_X8), v-line tiles of 4 (TB_V) or 16 (TB_V16)
    enum { TB_I = 0, TB_J, TB_X, TB_V, TB_R, TB_AX, TB_AV, TB_U, TB_X8, TB_V16, TB_X4, TB_E, TB_P, TB_N };
    std::vector<Tile> tiles[TB_N];
    size_t tile_off[TB_N] = {};
    std::vector<int> prefix[TB_N];     // prefix[tb][q] = tiles of internal problems [0, q)
    // Graph replays of plans of GRAPH_GROUP_MIN or more problems run as GROUPS concurrent groups
    // (internal problem q belongs to group q mod GROUPS, so every group keeps the steps-descending
    // order), one graph per group on a stream of the plan's own: the groups' launches share the CUs.
    // Measured: the 2048-problem Heston book 4.4 % faster with 2 groups than with one (3 and 4 groups
    // less), the 256-problem SABR book 3.5 % (f64) and 1.8 % (f32).
    static constexpr int GROUPS = 2;
    static constexpr int64_t GRAPH_GROUP_MIN = 2;
    int ngroups = 1;
    int set_ = 0;                      // tiles of the launches: 0 every problem, g + 1 group g
    std::vector<int> gprefix[GROUPS][TB_N];  // gprefix[g][tb][k] = tiles of group g's first k problems
    size_t gtile_off[GROUPS][TB_N] = {};
    hipStream_t gstream[GROUPS] = {};
    hipGraph_t ggraph[GROUPS] = {};
    hipGraphExec_t gexec[GROUPS] = {};
    hipEvent_t gfork = nullptr, gjoin[GROUPS] = {};

    // launch grid of tile table tb over the na problems still stepping, and the first tile, in the
    // current tile set
    int pre(int tb, int64_t na) const
    {
        if (set_ == 0)
            return prefix[tb][na];
        const int g = set_ - 1;
        const int64_t k = na > g ? (na - g + ngroups - 1) / ngroups : 0;
        return gprefix[g][tb][k];
    }
    size_t toff(int tb) const { return set_ == 0 ? tile_off[tb] : gtile_off[set_ - 1][tb]; }
    int max_steps = 0;
    int line_solver = LINE_EXACT;
    int pmx = 1, pmy = 1;              // points per chunk bound (template M) of S-lines and v-lines
    int xw = 16, vr = 4;               // S-lines per x-sweep workgroup, v-lines per v-sweep workgroup
    int swz = 0;                       // XCD-aware tile order (xcd_tile) in the sweeps
    int fpat = PAT_GENERIC;            // factor pattern shared by every problem (SEPARABLE), see FPat
    bool tuned = false;                // the autotuner has chosen xw and vr
    bool adj_ok = false;               // every line fits the partitioned solver (needed by the adjoint)
    bool euler = false;                // explicit Euler plan (no line sweeps)
    bool euler_res = false;            // ... whose problems all fit k_euler_res (the loop resident in LDS)
    hipStream_t cap_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;

    ~PlanImpl() override
    {
        if (graph_exec)
            (void)hipGraphExecDestroy(graph_exec);
        if (graph)
            (void)hipGraphDestroy(graph);
        if (cap_stream)
            (void)hipStreamDestroy(cap_stream);
        if (aux_stream)
            (void)hipStreamDestroy(aux_stream);
        for (int g = 0; g < GROUPS; ++g) {
            if (gexec[g])
                (void)hipGraphExecDestroy(gexec[g]);
            if (ggraph[g])
                (void)hipGraphDestroy(ggraph[g]);
            if (gstream[g])
                (void)hipStreamDestroy(gstream[g]);
            if (gjoin[g])
                (void)hipEventDestroy(gjoin[g]);
        }
        if (gfork)
            (void)hipEventDestroy(gfork);
        for (hipEvent_t e : {ev_fork, ev_ready[0], ev_ready[1], ev_free[0], ev_free[1]})
            if (e)
                (void)hipEventDestroy(e);
    }

    // schemes without the second pair of stages (no Y2, Y3)
    static bool one_stage(const lss_adi_desc &d)
    {
        return d.splitting == LSS_DOUGLAS_RACHFORD || d.splitting == LSS_EXPLICIT_EULER;
    }

    static void validate(const lss_adi_desc &d)
    {
        require(d.nx >= 4 && d.ny >= 4, "nx and ny must be >= 4");
        // kernels index a problem's arrays with 32-bit offsets (7 coefficient fields included)
        require(d.nx <= (1 << 20) && d.ny <= (1 << 20) && d.nx * d.ny <= (int64_t)1 << 28,
                "grid too large for one problem (nx * ny <= 2^28)");
        require(d.steps >= 1 && d.steps < (1 << 30), "steps must be >= 1");
        require(d.splitting >= 0 && d.splitting <= LSS_EXPLICIT_EULER, "unknown splitting method");
        require(d.tridiagonal == LSS_THOMAS_LU || d.tridiagonal == LSS_DOUBLE_SWEEP ||
                    d.tridiagonal == LSS_CUDA_SOLVER,
                "tridiagonal method must be ThomasLU, DoubleSweep or CUDASolver");
        require(d.coeff_mode == LSS_COEFF_FIELDS || d.coeff_mode == LSS_COEFF_SEPARABLE, "unknown coeff_mode");
        require(d.line_solver == LSS_LINE_EXACT || d.line_solver == LSS_LINE_PARTITIONED, "unknown line_solver");
        if (d.line_solver == LSS_LINE_PARTITIONED)
            require(d.nx - 2 <= P_NCH * 16 && d.ny - 2 <= P_NCH * 16,
                    "partitioned line solver needs nx - 2 and ny - 2 <= 1024");
        require(d.coeff_sets == 1 || d.coeff_sets == d.steps, "coeff_sets must be 1 or steps");
        require(d.bnd_sets == 1 || d.bnd_sets == d.steps, "bnd_sets must be 1 or steps");
        require(d.idx1 && d.idx2 && d.coeff && d.bnd, "null array in descriptor");
        require(d.src_sets == 0 || d.src_sets == 1 || d.src_sets == d.steps, "src_sets must be 0, 1 or steps");
        require((d.src_sets == 0) == (d.src == nullptr), "src must be null exactly when src_sets is 0");
        require(d.src_sets == 0 || d.splitting != LSS_EXPLICIT_EULER, "explicit Euler plans take no heat source");
        for (int64_t i = 0; i < d.nx; ++i)
            require(d.idx1[i] < (uint32_t)d.nx, "idx1 out of range");
        for (int64_t j = 0; j < d.ny; ++j)
            require(d.idx2[j] < (uint32_t)d.ny, "idx2 out of range");
    }

    static size_t coeff_stride(const lss_adi_desc &d)
    {
        return d.coeff_mode == LSS_COEFF_FIELDS ? (size_t)(7 * d.nx * d.ny + 3 * d.nx) : (size_t)(12 * (d.nx + d.ny));
    }

    // SEPARABLE factor packing: factor vectors that are zero (or one) in every coefficient set are not
    // stored; code[slot] says -1 (zero), -2 (one) or the packed position (see Prob::fcode)
    static void pack_separable(const lss_adi_desc &d, std::vector<T> &out, int32_t *code)
    {
        const size_t nx = (size_t)d.nx, ny = (size_t)d.ny, stride = 12 * (nx + ny);
        const T *src = (const T *)d.coeff;
        out.assign(stride * (size_t)d.coeff_sets, (T)0);
        for (int blk = 0; blk < 2; ++blk) {
            const size_t n = blk ? ny : nx, base = blk ? 12 * nx : 0;
            int packed = 0;
            for (int slot = 0; slot < 12; ++slot) {
                bool zero = true, one = true;
                for (int64_t k = 0; k < d.coeff_sets; ++k)
                    for (size_t i = 0; i < n; ++i) {
                        const T v = src[(size_t)k * stride + base + slot * n + i];
                        zero = zero && v == (T)0;
                        one = one && v == (T)1;
                    }
                int c = zero ? -1 : (one ? -2 : packed++);
                if (c >= 0)
                    for (int64_t k = 0; k < d.coeff_sets; ++k)
                        std::memcpy(out.data() + (size_t)k * stride + base + (size_t)c * n,
                                    src + (size_t)k * stride + base + slot * n, n * sizeof(T));
                code[blk * 12 + slot] = c;
            }
        }
    }

    void build(const lss_adi_desc *descs, int64_t n)
    {
        count = n;
        require(n >= 1, "count must be >= 1");
        for (int64_t p = 0; p < n; ++p) {
            require(descs[p].fp == (sizeof(T) == 8 ? LSS_FP64 : LSS_FP32), "all problems of a plan share fp");
            validate(descs[p]);
            require(descs[p].line_solver == descs[0].line_solver, "all problems of a plan share line_solver");
            require((descs[p].splitting == LSS_EXPLICIT_EULER) == (descs[0].splitting == LSS_EXPLICIT_EULER),
                    "a plan steps either explicit Euler problems or ADI problems, not both");
        }
        line_solver = descs[0].line_solver == LSS_LINE_PARTITIONED ? LINE_PARTITIONED : LINE_EXACT;
        euler = descs[0].splitting == LSS_EXPLICIT_EULER;
        euler_res = euler;
        for (int64_t q = 0; q < n; ++q)
            euler_res = euler_res && descs[q].nx >= 4 && descs[q].ny <= ERES_T &&
                        descs[q].nx * descs[q].ny <= (int64_t)ERES_T * ERES_K && descs[q].coeff_sets == 1 &&
                        descs[q].bnd_sets == 1;
        {
            int64_t mx = 1, my = 1;
            for (int64_t q = 0; q < n; ++q) {
                mx = std::max(mx, (descs[q].nx - 2 + P_NCH - 1) / P_NCH);
                my = std::max(my, (descs[q].ny - 2 + P_NCH - 1) / P_NCH);
            }
            while (pmx < mx)
                pmx *= 2;
            while (pmy < my)
                pmy *= 2;
            adj_ok = pmx <= 16 && pmy <= 16 && descs[0].splitting != LSS_EXPLICIT_EULER;
            if (pmx >= 16)
                xw = 8;  // long S-lines: at most 8 per workgroup
        }
        order.resize(n);
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(),
                         [&](int64_t x, int64_t y) { return descs[x].steps > descs[y].steps; });
        where.resize(n);
        for (int64_t q = 0; q < n; ++q)
            where[order[q]] = q;
        nx.resize(n);
        ny.resize(n);
        steps.resize(n);
        // layouts
        layouts.resize(n);
        size_t total = 0;
        for (int64_t q = 0; q < n; ++q) {
            const lss_adi_desc &d = descs[order[q]];
            const size_t nxy = (size_t)(d.nx * d.ny);
            ProbLayout L;
            size_t o = 0;
            auto take = [&](size_t k) {
                const size_t r = o;
                o += (k + 63) / 64 * 64;  // keep every array 256/512-byte aligned
                return r;
            };
            L.off_a = take(nxy);
            L.off_b = take(nxy);
            L.off_y1 = take(nxy);
            L.off_y2 = one_stage(d) ? 0 : take(nxy);
            L.off_y3 = one_stage(d) ? 0 : take(nxy);
            L.off_scr = take(nxy);
            L.off_colA0 = take((size_t)d.nx);
            L.off_colA1 = take((size_t)d.nx);
            L.off_coeff = take(coeff_stride(d) * (size_t)d.coeff_sets);
            L.off_bnd = take((size_t)(d.nx + 2 * d.ny) * (size_t)d.bnd_sets);
            L.off_src = d.src_sets ? take(2 * nxy * (size_t)d.src_sets) : 0;
            L.n_T = o;
            L.off_idx_bytes = o * sizeof(T);
            L.bytes = L.off_idx_bytes + ((size_t)(d.nx + d.ny) * 4 + 255) / 256 * 256;
            layouts[q] = L;
            total += L.bytes;
        }
        mem.alloc(total);
        HIPCHK(hipMemset(mem.p, 0, total));
        probs.resize(n);
        char *base = (char *)mem.p;
        for (int64_t q = 0; q < n; ++q) {
            const lss_adi_desc &d = descs[order[q]];
            const ProbLayout &L = layouts[q];
            T *t = (T *)base;
            Prob<T> &p = probs[q];
            std::memset(&p, 0, sizeof(p));
            p.nx = (int32_t)d.nx;
            p.ny = (int32_t)d.ny;
            p.steps = (int32_t)d.steps;
            p.scheme = d.splitting;
            p.tri = d.tridiagonal == LSS_DOUBLE_SWEEP ? TRI_DSWEEP : TRI_THOMAS;
            p.coeff_mode = d.coeff_mode == LSS_COEFF_FIELDS ? COEFF_FIELDS : COEFF_SEPARABLE;
            p.coeff_sets = (int32_t)d.coeff_sets;
            p.bnd_sets = (int32_t)d.bnd_sets;
            p.theta = (T)d.theta;
            p.zeta = (T)d.weighting;
            p.alpha = (T)d.alpha;
            p.beta = (T)d.beta;
            p.gamma = (T)d.gamma;
            p.delta = (T)d.delta;
            p.ni = (T)d.ni;
            p.rho = (T)d.rho;
            p.h1 = (T)d.h1;
            p.a = t + L.off_a;
            p.b = t + L.off_b;
            p.y1 = t + L.off_y1;
            p.y2 = one_stage(d) ? nullptr : t + L.off_y2;
            p.y3 = one_stage(d) ? nullptr : t + L.off_y3;
            p.y4 = line_solver == LINE_PARTITIONED ? p.y3 : p.y1;
            p.scr = t + L.off_scr;
            p.colA0 = t + L.off_colA0;
            p.colA1 = t + L.off_colA1;
            p.coeff = t + L.off_coeff;
            p.bnd = t + L.off_bnd;
            p.coeff_stride = (int64_t)coeff_stride(d);
            p.bnd_stride = d.nx + 2 * d.ny;
            max_nxy = std::max<int64_t>(max_nxy, (int64_t)d.nx * d.ny);
            uint32_t *ix = (uint32_t *)(base + L.off_idx_bytes);
            p.idx1 = ix;
            p.idx2 = ix + d.nx;
            if (p.coeff_mode == COEFF_SEPARABLE) {
                std::vector<T> packed;
                pack_separable(d, packed, p.fcode);
                HIPCHK(hipMemcpy((void *)p.coeff, packed.data(), sizeof(T) * packed.size(), hipMemcpyHostToDevice));
            } else {
                HIPCHK(hipMemcpy((void *)p.coeff, d.coeff, sizeof(T) * coeff_stride(d) * (size_t)d.coeff_sets,
                                 hipMemcpyHostToDevice));
            }
            HIPCHK(hipMemcpy((void *)p.bnd, d.bnd, sizeof(T) * (size_t)(d.nx + 2 * d.ny) * (size_t)d.bnd_sets,
                             hipMemcpyHostToDevice));
            if (d.src_sets) {
                p.src = t + L.off_src;
                p.src_sets = (int32_t)d.src_sets;
                HIPCHK(hipMemcpy((void *)p.src, d.src, sizeof(T) * 2 * (size_t)(d.nx * d.ny) * (size_t)d.src_sets,
                                 hipMemcpyHostToDevice));
            }
            HIPCHK(hipMemcpy(ix, d.idx1, sizeof(uint32_t) * (size_t)d.nx, hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(ix + d.nx, d.idx2, sizeof(uint32_t) * (size_t)d.ny, hipMemcpyHostToDevice));
            nx[order[q]] = d.nx;
            ny[order[q]] = d.ny;
            steps[order[q]] = d.steps;
            max_steps = std::max(max_steps, (int)d.steps);
            base += L.bytes;
        }
        // a factor pattern shared by every problem lets the sweeps use the specialized stencil
        fpat = PAT_GENERIC;
        for (int cand : {(int)PAT_HESTON, (int)PAT_SABR0}) {
            const int *c = cand == PAT_HESTON ? FPat<PAT_HESTON>::c : FPat<PAT_SABR0>::c;
            bool all = true;
            for (int64_t q = 0; q < n && all; ++q)
                all = probs[q].coeff_mode == COEFF_SEPARABLE && std::equal(c, c + 24, probs[q].fcode);
            if (all) {
                fpat = cand;
                break;
            }
        }
        // tile tables
        for (int tb = 0; tb < TB_N; ++tb) {
            prefix[tb].assign(n + 1, 0);
            for (int64_t q = 0; q < n; ++q) {
                const Prob<T> &p = probs[q];
                int lo = 0, hi = 0, step = TILE;
                switch (tb) {
                case TB_I: lo = 0, hi = p.nx; break;
                case TB_J: lo = 0, hi = p.ny; break;
                case TB_X: lo = 1, hi = p.ny - 1, step = line_solver == LINE_PARTITIONED ? P_XTJ : TILE; break;
                case TB_V:
                    lo = 1, hi = p.nx - 1;
                    step = line_solver == LINE_PARTITIONED ? P_VROWS * vlines_per_wave(p.ny - 2, pmy) : TILE;
                    break;
                case TB_AX: lo = 1, hi = p.ny - 1, step = P_XTJ; break;
                case TB_AV: lo = 1, hi = p.nx - 1, step = P_VROWS; break;
                case TB_U: lo = 0, hi = p.nx, step = 1; break;
                case TB_X8: lo = 1, hi = p.ny - 1, step = 8; break;
                case TB_X4: lo = 1, hi = p.ny - 1, step = 4; break;
                case TB_P: lo = 0, hi = 1, step = 1; break;  // one tile per problem
                case TB_V16: lo = 1, hi = p.nx - 1, step = 16 * vlines_per_wave(p.ny - 2, pmy); break;
                default: lo = 1, hi = p.nx - 1, step = 1; break;
                }
                if (tb == TB_U || tb == TB_E) {  // two-dimensional tiles (adjoint gathers, explicit step)
                    const int rows = tb == TB_U ? AGR : GTR * ERPT, cols = tb == TB_U ? AGC : GTC;
                    for (int r = 0; r < p.nx; r += rows)
                        for (int c = 0; c < p.ny; c += cols)
                            tiles[tb].push_back({(int32_t)q, r, std::min(rows, p.nx - r), c});
                } else {
                    for (int s = lo; s < hi; s += step) {
                        Tile t;
                        t.prob = (int32_t)q;
                        t.first = s;
                        t.count = std::min(step, hi - s);
                        t.col = 0;
                        tiles[tb].push_back(t);
                    }
                }
                prefix[tb][q + 1] = (int)tiles[tb].size();
            }
        }
        size_t ntiles = 0;
        for (int tb = 0; tb < TB_N; ++tb) {
            tile_off[tb] = ntiles;
            ntiles += tiles[tb].size();
        }
        // the groups' tile tables (copies of their problems' tiles)
        ngroups = n >= GRAPH_GROUP_MIN ? GROUPS : 1;
        std::vector<Tile> gtiles[GROUPS][TB_N];
        for (int g = 0; g < ngroups && ngroups > 1; ++g)
            for (int tb = 0; tb < TB_N; ++tb) {
                gprefix[g][tb].assign(1, 0);
                for (int64_t q = g; q < n; q += ngroups) {
                    gtiles[g][tb].insert(gtiles[g][tb].end(), tiles[tb].begin() + prefix[tb][q],
                                         tiles[tb].begin() + prefix[tb][q + 1]);
                    gprefix[g][tb].push_back((int)gtiles[g][tb].size());
                }
                gtile_off[g][tb] = ntiles;
                ntiles += gtiles[g][tb].size();
            }
        tiles_dev.alloc(ntiles * sizeof(Tile));
        for (int tb = 0; tb < TB_N; ++tb)
            HIPCHK(hipMemcpy((Tile *)tiles_dev.p + tile_off[tb], tiles[tb].data(), tiles[tb].size() * sizeof(Tile),
                             hipMemcpyHostToDevice));
        for (int g = 0; g < ngroups && ngroups > 1; ++g)
            for (int tb = 0; tb < TB_N; ++tb)
                HIPCHK(hipMemcpy((Tile *)tiles_dev.p + gtile_off[g][tb], gtiles[g][tb].data(),
                                 gtiles[g][tb].size() * sizeof(Tile), hipMemcpyHostToDevice));
        probs_dev.alloc(n * sizeof(Prob<T>));
        HIPCHK(hipMemcpy(probs_dev.p, probs.data(), n * sizeof(Prob<T>), hipMemcpyHostToDevice));
    }

    // number of internal problems still stepping at loop step s (problems sorted by steps desc)
    int64_t active(int s) const
    {
        int64_t lo = 0, hi = count;
        while (lo < hi) {
            const int64_t mid = (lo + hi) / 2;
            if (probs[mid].steps > s)
                lo = mid + 1;
            else
                hi = mid;
        }
        return lo;
    }

    // algorithmic HBM bytes of one launch of kernel kind k at step s: compulsory reads and writes of
    // the 2-D state arrays (coefficients and boundary data are O(nx + ny) and not counted)
    double algo_bytes(int k, int64_t na) const
    {
        double b = 0;
        for (int64_t q = 0; q < na; ++q) {
            const Prob<T> &p = probs[q];
            const double nxy = (double)p.nx * p.ny;
            const bool dr = p.scheme == SCHEME_DR, hv = p.scheme == SCHEME_HV;
            switch (k) {
            case LSS_K_BND_A: b += 5.0 * p.nx; break;
            case LSS_K_XSWEEP1: b += 2 * nxy + (line_solver == LINE_PARTITIONED ? 5.0 * p.nx : 0); break;
            case LSS_K_VSWEEP1: b += 3 * nxy; break;
            case LSS_K_HV_Y3: b += hv ? 4 * nxy : 0; break;
            case LSS_K_XSWEEP2: b += dr ? 0 : ((hv && line_solver != LINE_PARTITIONED) ? 3 * nxy : 4 * nxy); break;
            case LSS_K_VSWEEP2: b += dr ? 0 : 3 * nxy; break;
            case LSS_K_EULER: b += 2 * nxy; break;
            default: b += 4.0 * p.ny; break;
            }
        }
        return b * sizeof(T);
    }

    // partitioned sweeps with the chosen tile/lane mapping (xw, vr) and chunk bound (pmx, pmy)
    // Chunks per S-line.  Plans whose S-lines all have at most 512 points (M <= 8): 64 in stage 1 and
    // 32 (of at most 2M points) in stage 2 -- measured on the Heston book, stage 2 is 11 % faster with
    // 32, stage 1 faster with 64.  Plans with longer S-lines (M = 16): 128 chunks of at most 8 points
    // in both stages, so that a single large problem has enough wavefronts (the 1024 x 512 grid: four
    // per CU instead of two); such plans take 8 or 4 S-lines per workgroup.
    template <int M, int W, int PAT> void launch_x_p(int stage, int s, hipStream_t st, const Prob<T> *pd, int64_t na)
    {
        const int tb = W == 16 ? TB_X : (W == 8 ? TB_X8 : TB_X4);
        const dim3 g((unsigned)pre(tb, na));
        const Tile *tx = (const Tile *)tiles_dev.p + toff(tb);
        if constexpr (M >= 16) {
            if constexpr (W <= 8) {
                if (stage == 1)
                    hipLaunchKernelGGL((k_xsweep_p<T, 8, 1, W, PAT, 128>), g, dim3(W * 128), 0, st, pd, tx, s, swz);
                else
                    hipLaunchKernelGGL((k_xsweep_p<T, 8, 2, W, PAT, 128>), g, dim3(W * 128), 0, st, pd, tx, s, swz);
            } else {
                throw lss_error(LSS_ERR_INVALID, "S-lines of more than 512 points take 8 or 4 lines per workgroup");
            }
        } else if (stage == 1) {
            hipLaunchKernelGGL((k_xsweep_p<T, M, 1, W, PAT, 64>), g, dim3(W * 64), 0, st, pd, tx, s, swz);
        } else {
            hipLaunchKernelGGL((k_xsweep_p<T, 2 * M, 2, W, PAT, 32>), g, dim3(W * 32), 0, st, pd, tx, s, swz);
        }
    }
    template <int M, int W> void launch_x_m(int stage, int s, hipStream_t st, const Prob<T> *pd, int64_t na)
    {
        if (fpat == PAT_HESTON)
            return launch_x_p<M, W, PAT_HESTON>(stage, s, st, pd, na);
        if (fpat == PAT_SABR0)
            return launch_x_p<M, W, PAT_SABR0>(stage, s, st, pd, na);
        launch_x_p<M, W, PAT_GENERIC>(stage, s, st, pd, na);
    }
    template <int W> void launch_x_w(int stage, int s, hipStream_t st, const Prob<T> *pd, int64_t na)
    {
        switch (pmx) {
        case 1: launch_x_m<1, W>(stage, s, st, pd, na); break;
        case 2: launch_x_m<2, W>(stage, s, st, pd, na); break;
        case 4: launch_x_m<4, W>(stage, s, st, pd, na); break;
        case 8: launch_x_m<8, W>(stage, s, st, pd, na); break;
        default: launch_x_m<16, W>(stage, s, st, pd, na); break;
        }
    }
    void launch_x(int stage, int s, hipStream_t st, const Prob<T> *pd, int64_t na)
    {
        if (xw == 8)
            launch_x_w<8>(stage, s, st, pd, na);
        else if (xw == 4)
            launch_x_w<4>(stage, s, st, pd, na);
        else
            launch_x_w<16>(stage, s, st, pd, na);
    }
    template <int M, int VR, int PAT> void launch_v_p(int stage, int s, hipStream_t st, const Prob<T> *pd, int64_t na)
    {
        const int tb = VR == 4 ? TB_V : TB_V16;
        const dim3 g((unsigned)pre(tb, na)), b(64 * VR);
        hipLaunchKernelGGL((k_vsweep_p<T, M, VR, PAT>), g, b, 0, st, pd, (const Tile *)tiles_dev.p + toff(tb),
                           s, stage, swz);
    }
    template <int M, int VR> void launch_v_m(int stage, int s, hipStream_t st, const Prob<T> *pd, int64_t na)
    {
        if (fpat == PAT_HESTON)
            return launch_v_p<M, VR, PAT_HESTON>(stage, s, st, pd, na);
        if (fpat == PAT_SABR0)
            return launch_v_p<M, VR, PAT_SABR0>(stage, s, st, pd, na);
        launch_v_p<M, VR, PAT_GENERIC>(stage, s, st, pd, na);
    }
    template <int VR> void launch_v_r(int stage, int s, hipStream_t st, const Prob<T> *pd, int64_t na)
    {
        switch (pmy) {
        case 1: launch_v_m<1, VR>(stage, s, st, pd, na); break;
        case 2: launch_v_m<2, VR>(stage, s, st, pd, na); break;
        case 4: launch_v_m<4, VR>(stage, s, st, pd, na); break;
        case 8: launch_v_m<8, VR>(stage, s, st, pd, na); break;
        default: launch_v_m<16, VR>(stage, s, st, pd, na); break;
        }
    }
    void launch_v(int stage, int s, hipStream_t st, const Prob<T> *pd, int64_t na)
    {
        if (vr == 16)
            launch_v_r<16>(stage, s, st, pd, na);
        else
            launch_v_r<4>(stage, s, st, pd, na);
    }

    void drop_graph()
    {
        if (graph_exec) {  // the captured loop uses the previous mapping
            HIPCHK(hipGraphExecDestroy(graph_exec));
            graph_exec = nullptr;
            HIPCHK(hipGraphDestroy(graph));
            graph = nullptr;
        }
        for (int g = 0; g < GROUPS; ++g)
            if (gexec[g]) {
                HIPCHK(hipGraphExecDestroy(gexec[g]));
                gexec[g] = nullptr;
                HIPCHK(hipGraphDestroy(ggraph[g]));
                ggraph[g] = nullptr;
            }
    }

    // the time loop as ngroups graphs (captured once), run concurrently after the work queued on
    // `stream` and joined back into it
    void run_groups(hipStream_t stream)
    {
        if (!gexec[0]) {
            if (!gfork) {
                HIPCHK(hipEventCreateWithFlags(&gfork, hipEventDisableTiming));
                for (int g = 0; g < ngroups; ++g) {
                    HIPCHK(hipStreamCreateWithFlags(&gstream[g], hipStreamNonBlocking));
                    HIPCHK(hipEventCreateWithFlags(&gjoin[g], hipEventDisableTiming));
                }
            }
            for (int g = 0; g < ngroups; ++g) {
                set_ = g + 1;
                try {
                    HIPCHK(hipStreamBeginCapture(gstream[g], hipStreamCaptureModeThreadLocal));
                    for (int s = 0; s < probs[g].steps; ++s)  // group g's first problem has its most steps
                        launch_step(s, gstream[g]);
                } catch (...) {
                    set_ = 0;
                    hipGraph_t partial = nullptr;
                    if (hipStreamEndCapture(gstream[g], &partial) == hipSuccess && partial)
                        (void)hipGraphDestroy(partial);
                    drop_graph();
                    throw;
                }
                set_ = 0;
                HIPCHK(hipStreamEndCapture(gstream[g], &ggraph[g]));
                HIPCHK(hipGraphInstantiate(&gexec[g], ggraph[g], nullptr, nullptr, 0));
            }
        }
        HIPCHK(hipEventRecord(gfork, stream));
        for (int g = 0; g < ngroups; ++g) {
            HIPCHK(hipStreamWaitEvent(gstream[g], gfork, 0));
            HIPCHK(hipGraphLaunch(gexec[g], gstream[g]));
            HIPCHK(hipEventRecord(gjoin[g], gstream[g]));
        }
        for (int g = 0; g < ngroups; ++g)
            HIPCHK(hipStreamWaitEvent(stream, gjoin[g], 0));
    }

    void tune(hipStream_t stream, bool again, int *x, int *v, int *z) override
    {
        if (!tuned || again) {
            drop_graph();
            autotune(stream);
        }
        *x = xw;
        *v = vr;
        *z = swz;
    }

    void kernel_config(int32_t *cfg) const override
    {
        cfg[0] = line_solver;
        cfg[1] = xw;
        cfg[2] = vr;
        cfg[3] = swz;
        cfg[4] = fpat;
        cfg[5] = pmx;
        cfg[6] = pmy;
        cfg[7] = tuned ? 1 : 0;
        cfg[8] = ck > 0 ? apat : -1;
        cfg[9] = ck > 0 ? axw : 0;
        cfg[10] = ck > 0 ? (dbl ? 1 : 0) : -1;
        cfg[11] = ngroups;
        cfg[12] = euler_res ? 1 : 0;
    }

    void set_overlap(int mode) override
    {
        require(mode >= -1 && mode <= 1, "adjoint overlap mode must be -1, 0 or 1");
        overlap_mode = mode;
        invalidate_adjoint();
    }

    void set_mapping(int x, int v, int z) override
    {
        require((x == 16 || x == 8 || x == 4) && (v == 4 || v == 16) && (z == 0 || z == 1),
                "mapping must be (16, 8 or 4; 4 or 16; 0 or 1)");
        require(x != 16 || pmx < 16, "S-lines of more than 512 points take 8 or 4 lines per workgroup");
        drop_graph();
        xw = x;
        vr = v;
        swz = z;
        tuned = true;
    }

    // Autotuner: times every (xw, vr) mapping on the first steps of the loop (at most 8, from the
    // plan's current state) and keeps the fastest.  The mappings give identical results.
    void autotune(hipStream_t st)
    {
        tuned = true;
        if (line_solver != LINE_PARTITIONED || euler)
            return;
        const int nst = std::min(max_steps, 8);
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        float best = 1e30f;
        int bx = pmx >= 16 ? 8 : 16, bv = 4, bs = 0;
        for (int cz : {0, 1})
            for (int cx : {16, 8, 4})
                for (int cv : {4, 16}) {
                if (cx == 16 && pmx >= 16)
                    continue;  // long S-lines: at most 8 per workgroup
                xw = cx;
                vr = cv;
                swz = cz;
                for (int rep = 0; rep < 2; ++rep) {  // the first pass warms caches and code
                    HIPCHK(hipEventRecord(e0, st));
                    for (int s = 0; s < nst; ++s)
                        launch_step(s, st);
                    HIPCHK(hipEventRecord(e1, st));
                }
                HIPCHK(hipEventSynchronize(e1));
                float ms = 0;
                HIPCHK(hipEventElapsedTime(&ms, e0, e1));
                if (ms < best) {
                    best = ms;
                    bx = cx;
                    bv = cv;
                    bs = cz;
                }
            }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        xw = bx;
        vr = bv;
        swz = bs;
    }

    // one loop step of every active problem; `table` replaces the plan's problem table (the adjoint's
    // recomputation writes each step's stage outputs into their own buffers)
    void launch_step(int s, hipStream_t st, Profiler *prof = nullptr, const Prob<T> *table = nullptr)
    {
        const int64_t na = active(s);
        const Prob<T> *pd = table ? table : (const Prob<T> *)probs_dev.p;
        const Tile *td = (const Tile *)tiles_dev.p;
        auto grid = [&](int tb) { return (unsigned)pre(tb, na); };
        bool any_split2 = false, any_hv = false;
        for (int64_t q = 0; q < na; ++q) {
            any_split2 |= probs[q].scheme != SCHEME_DR;
            any_hv |= probs[q].scheme == SCHEME_HV;
        }
        auto pre = [&](int k) {
            if (prof)
                prof->begin(k, algo_bytes(k, na));
        };
        auto post = [&]() {
            if (prof)
                prof->end();
        };
        if (euler) {  // heston_euler_time_loop::run (lss_heston_euler_solver_scheme.hpp:63-67), one launch
            const dim3 g(grid(TB_E)), b(GTR * GTC);
            const Tile *tu = td + toff(TB_E);
            pre(LSS_K_EULER);
            if (fpat == PAT_HESTON)
                hipLaunchKernelGGL((k_euler<T, PAT_HESTON>), g, b, 0, st, pd, tu, s);
            else if (fpat == PAT_SABR0)
                hipLaunchKernelGGL((k_euler<T, PAT_SABR0>), g, b, 0, st, pd, tu, s);
            else
                hipLaunchKernelGGL((k_euler<T, PAT_GENERIC>), g, b, 0, st, pd, tu, s);
            post();
            return;
        }
        if (line_solver == LINE_EXACT) {  // (the partitioned stage-1 S-sweep evaluates stage A itself)
            pre(LSS_K_BND_A);
            hipLaunchKernelGGL(k_bnd_a<T>, dim3(grid(TB_I)), dim3(TILE), 0, st, pd, td + toff(TB_I), s);
            post();
        }
        auto xs = [&](int stage) {
            if (line_solver == LINE_EXACT) {
                hipLaunchKernelGGL(k_xsweep<T>, dim3(grid(TB_X)), dim3(TILE), 0, st, pd, td + toff(TB_X), s, stage);
                return;
            }
            launch_x(stage, s, st, pd, na);
        };
        auto vs = [&](int stage) {
            if (line_solver == LINE_EXACT) {
                hipLaunchKernelGGL(k_vsweep<T>, dim3(grid(TB_V)), dim3(TILE), 0, st, pd, td + toff(TB_V), s, stage);
                return;
            }
            launch_v(stage, s, st, pd, na);
        };
        pre(LSS_K_XSWEEP1);
        xs(1);
        post();
        pre(LSS_K_VSWEEP1);
        vs(1);
        post();
        if (any_hv && line_solver == LINE_EXACT) {
            pre(LSS_K_HV_Y3);
            hipLaunchKernelGGL(k_hv_y3<T>, dim3(grid(TB_R)), dim3(TILE), 0, st, pd, td + toff(TB_R), s);
            post();
        }
        if (any_split2) {
            pre(LSS_K_XSWEEP2);
            xs(2);
            post();
            pre(LSS_K_VSWEEP2);
            vs(2);
            post();
        }
        pre(LSS_K_BND_B);
        hipLaunchKernelGGL(k_bnd_b<T>, dim3(grid(TB_J)), dim3(TILE), 0, st, pd, td + toff(TB_J), s);
        post();
    }

    // ---------------------------------------------------------------------------------------------
    // Reverse mode: checkpointed forward, then segment by segment (last first) a recomputation that
    // keeps every stage output of the segment, followed by the adjoint steps of the segment.
    // ---------------------------------------------------------------------------------------------
    int ck = 0;                                   // checkpoint interval of the adjoint allocation
    bool fwd_done = false;
    dev_buffer adj_mem, adj_dev, seg_dev, slots_dev, jobs_dev, lam_jobs_dev;
    std::vector<Adj<T>> adjs;                     // host copy (device pointers inside)
    std::vector<T *> ckpt_base, seg_lv_base;      // per internal problem
    std::vector<size_t> fwd_job_off, bwd_job_off; // per checkpoint: first job in jobs_dev
    std::vector<int> fwd_job_n, bwd_job_n;        // per checkpoint: number of jobs
    int64_t max_nxy = 0;
    size_t slot_off[5] = {};                      // per adjoint kernel kind, in units of NPAR doubles
    size_t nslots = 0;
    int slot_tb[5] = {TB_AV, TB_AX, TB_AV, TB_AX, TB_U};
    int axw = 16;                                 // S-lines per adjoint x-sweep workgroup
    int apat = PAT_GENERIC;                       // compile-time factor and term tables of the adjoint
    // Two segment storage sets (dbl): the backward pass then recomputes segment c - 1 on aux_stream
    // while the adjoint steps of segment c run on the caller's stream.  overlap_mode -1 (automatic)
    // uses them when the second set takes at most SEG2_MAX_BYTES: plans of few problems, whose
    // kernels alone leave CUs idle.  (The 256-problem SABR book, which fills the GPU, gained 1 % from
    // a 35 GB second set.)
    static constexpr size_t SEG2_MAX_BYTES = (size_t)2 << 30;
    int overlap_mode = -1;
    bool dbl = false;
    hipStream_t aux_stream = nullptr;
    hipEvent_t ev_fork = nullptr, ev_ready[2] = {}, ev_free[2] = {};

    void invalidate_adjoint() override
    {
        ck = 0;
        fwd_done = false;
    }

    // problem table of step s0 + k of a segment stored in set `set`
    const Prob<T> *seg_table(int k, int set = 0) const
    {
        return (const Prob<T> *)seg_dev.p + ((size_t)set * ck_len + k) * count;
    }
    int ck_len = 0;  // steps per segment of the tables in seg_dev

    // Derivative terms on the device: terms with an all-zero factor are dropped, all-ones factors are
    // not stored (code -2) and identical factor vectors are stored once (see Adj::dfac).
    static void pack_terms(const Terms &tm, size_t nx, size_t ny, T *dst, Adj<T> &A)
    {
        const T *src = (const T *)tm.fac.data();
        const size_t L = nx + ny, per = (size_t)tm.nterm * L;
        const int64_t sets = tm.sets;
        auto vec = [&](int t, int64_t k, bool y) { return src + (size_t)k * per + (size_t)t * L + (y ? nx : 0); };
        auto all = [&](int t, bool y, T v) {
            const size_t n = y ? ny : nx;
            for (int64_t k = 0; k < sets; ++k)
                for (size_t i = 0; i < n; ++i)
                    if (vec(t, k, y)[i] != v)
                        return false;
            return true;
        };
        auto same = [&](int t, int u, bool y) {
            const size_t n = y ? ny : nx;
            for (int64_t k = 0; k < sets; ++k)
                if (std::memcmp(vec(t, k, y), vec(u, k, y), n * sizeof(T)) != 0)
                    return false;
            return true;
        };
        std::vector<int> ux, uy;  // representative term of each stored vector
        int kept = 0;
        int fx[NTERM], fy[NTERM], tp[NTERM], tq[NTERM];
        for (int t = 0; t < tm.nterm; ++t) {
            if (all(t, false, (T)0) || all(t, true, (T)0))
                continue;
            int codes[2];
            for (int y = 0; y < 2; ++y) {
                std::vector<int> &u = y ? uy : ux;
                int c = -2;
                if (!all(t, y, (T)1)) {
                    c = -1;
                    for (size_t z = 0; z < u.size() && c < 0; ++z)
                        if (same(t, u[z], y))
                            c = (int)z;
                    if (c < 0) {
                        c = (int)u.size();
                        u.push_back(t);
                    }
                }
                codes[y] = c;
            }
            fx[kept] = codes[0];
            fy[kept] = codes[1];
            tp[kept] = tm.tp[t];
            tq[kept] = tm.tq[t];
            ++kept;
        }
        const size_t stride = ux.size() * nx + uy.size() * ny;
        std::vector<T> packed(std::max<size_t>(stride * (size_t)sets, 1));
        for (int64_t k = 0; k < sets; ++k) {
            T *o = packed.data() + (size_t)k * stride;
            for (size_t z = 0; z < ux.size(); ++z)
                std::memcpy(o + z * nx, vec(ux[z], k, false), nx * sizeof(T));
            for (size_t z = 0; z < uy.size(); ++z)
                std::memcpy(o + ux.size() * nx + z * ny, vec(uy[z], k, true), ny * sizeof(T));
        }
        HIPCHK(hipMemcpy(dst, packed.data(), packed.size() * sizeof(T), hipMemcpyHostToDevice));
        A.dfac = dst;
        A.dfac_sets = (int32_t)sets;
        A.dfac_stride = (int64_t)stride;
        A.nterm = kept;
        A.nfx = (int32_t)ux.size();
        A.nfy = (int32_t)uy.size();
        for (int z = 0; z < NTERM; ++z) {
            A.tp[z] = z < kept ? tp[z] : 0;
            A.tq[z] = z < kept ? tq[z] : 0;
            A.fx[z] = z < kept ? fx[z] : -2;
            A.fy[z] = z < kept ? fy[z] : -2;
        }
    }

    void adj_setup(int K)
    {
        require(!euler, "explicit Euler plans have no adjoint");
        require(adj_ok, "the adjoint needs nx - 2 <= 1024 and ny - 2 <= 1024 for every problem");
        require(K >= 2 && K % 2 == 0 && K <= 4096, "ckpt_every must be even, >= 2 and <= 4096");
        ck = 0;
        fwd_done = false;
        const int64_t n = count;
        // sizes
        std::vector<size_t> off_T(n), n_T(n), off_U(n);
        size_t total = 0, seg_bytes = 0;
        max_nxy = 0;
        for (int64_t q = 0; q < n; ++q) {
            const Prob<T> &p = probs[q];
            const size_t nxy = ((size_t)p.nx * p.ny + 63) / 64 * 64;
            const bool dr = p.scheme == SCHEME_DR, hv = p.scheme == SCHEME_HV;
            const size_t nys = (dr ? 1 : 3) + ((hv && line_solver == LINE_EXACT) ? 1 : 0);
            seg_bytes += ((size_t)(K + 1) + (size_t)K * nys) * nxy * sizeof(T);
        }
        dbl = overlap_mode >= 0 ? overlap_mode == 1 : seg_bytes <= SEG2_MAX_BYTES;
        for (int64_t q = 0; q < n; ++q) {
            const Prob<T> &p = probs[q];
            const size_t nxy = (size_t)p.nx * p.ny, nx_ = (size_t)p.nx, ny_ = (size_t)p.ny;
            const bool dr = p.scheme == SCHEME_DR, hv = p.scheme == SCHEME_HV;
            const size_t nys = (dr ? 1 : 3) + ((hv && line_solver == LINE_EXACT) ? 1 : 0);
            const size_t nl = dr ? 4 : 7;
            const size_t nck = ((size_t)p.steps + K - 1) / K;
            const Terms *tm = terms.empty() ? nullptr : &terms[order[q]];
            const size_t nfac = tm ? tm->fac.size() / sizeof(T) : 0;
            auto al = [](size_t k) { return (k + 63) / 64 * 64; };
            size_t t = nl * al(nxy) + 2 * al(nx_) + nck * al(nxy) +
                       (dbl ? 2 : 1) * ((K + 1) * al(nxy) + (size_t)K * nys * al(nxy)) + al(nfac);
            off_T[q] = total;
            n_T[q] = t;
            total += t * sizeof(T);
            off_U[q] = total;
            total += ((2 * (nx_ + ny_) + 2) * 4 + 255) / 256 * 256;
            max_nxy = std::max<int64_t>(max_nxy, (int64_t)nxy);
        }
        adj_mem.~dev_buffer();
        new (&adj_mem) dev_buffer();
        adj_mem.alloc(total);
        HIPCHK(hipMemset(adj_mem.p, 0, total));
        char *base = (char *)adj_mem.p;
        adjs.assign(n, Adj<T>{});
        ckpt_base.assign(n, nullptr);
        seg_lv_base.assign(2 * n, nullptr);  // set 1 at q + n
        std::vector<T *> seg_y_base(2 * n, nullptr);
        for (int64_t q = 0; q < n; ++q) {
            const Prob<T> &p = probs[q];
            const size_t nxy = (size_t)p.nx * p.ny, nx_ = (size_t)p.nx, ny_ = (size_t)p.ny;
            const bool dr = p.scheme == SCHEME_DR, hv = p.scheme == SCHEME_HV;
            const size_t nys = (dr ? 1 : 3) + ((hv && line_solver == LINE_EXACT) ? 1 : 0);
            const size_t nck = ((size_t)p.steps + K - 1) / K;
            auto al = [](size_t k) { return (k + 63) / 64 * 64; };
            T *t = (T *)(base + off_T[q]);
            auto take = [&](size_t k) {
                T *r = t;
                t += al(k);
                return r;
            };
            Adj<T> &A = adjs[q];
            A.la = take(nxy);
            A.lb = take(nxy);
            A.lv2 = take(nxy);
            A.lx1 = take(nxy);
            A.lx2 = dr ? nullptr : take(nxy);
            A.lv1 = dr ? nullptr : take(nxy);
            A.ly2 = dr ? nullptr : take(nxy);
            A.elo2 = take(nx_);
            A.elo1 = take(nx_);
            ckpt_base[q] = take(nck * nxy);
            seg_lv_base[q] = take((K + 1) * nxy);
            seg_y_base[q] = take((size_t)K * nys * nxy);
            if (dbl) {
                seg_lv_base[q + n] = take((K + 1) * nxy);
                seg_y_base[q + n] = take((size_t)K * nys * nxy);
            }
            const Terms *tm = terms.empty() ? nullptr : &terms[order[q]];
            const size_t nfac = tm ? tm->fac.size() / sizeof(T) : 0;
            T *fac = take(nfac);
            if (tm && tm->nterm > 0)
                pack_terms(*tm, nx_, ny_, fac, A);
            // inverse index maps
            std::vector<uint32_t> h1, h2;
            h1.resize(probs[q].nx);
            h2.resize(probs[q].ny);
            HIPCHK(hipMemcpy(h1.data(), p.idx1, nx_ * 4, hipMemcpyDeviceToHost));
            HIPCHK(hipMemcpy(h2.data(), p.idx2, ny_ * 4, hipMemcpyDeviceToHost));
            auto inverse = [](const std::vector<uint32_t> &m, std::vector<uint32_t> &o, std::vector<uint32_t> &ix) {
                const size_t k = m.size();
                o.assign(k + 1, 0);
                for (uint32_t v : m)
                    o[v + 1] += 1;
                for (size_t z = 0; z < k; ++z)
                    o[z + 1] += o[z];
                ix.assign(k, 0);
                std::vector<uint32_t> fill(o.begin(), o.end() - 1);
                for (size_t z = 0; z < k; ++z)
                    ix[fill[m[z]]++] = (uint32_t)z;
            };
            std::vector<uint32_t> o1, i1, o2, i2;
            inverse(h1, o1, i1);
            inverse(h2, o2, i2);
            uint32_t *u = (uint32_t *)(base + off_U[q]);
            HIPCHK(hipMemcpy(u, o1.data(), o1.size() * 4, hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(u + o1.size(), i1.data(), i1.size() * 4, hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(u + o1.size() + i1.size(), o2.data(), o2.size() * 4, hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(u + o1.size() + i1.size() + o2.size(), i2.data(), i2.size() * 4, hipMemcpyHostToDevice));
            A.inv1o = u;
            A.inv1i = u + o1.size();
            A.inv2o = u + o1.size() + i1.size();
            A.inv2i = u + o1.size() + i1.size() + o2.size();
        }
        // the reference models' derivative-term tables (with their factor pattern) select the
        // specialized adjoint kernels
        apat = PAT_GENERIC;
        if (fpat == PAT_HESTON || fpat == PAT_SABR0) {
            bool all = true;
            for (int64_t q = 0; q < n && all; ++q) {
                const Adj<T> &A = adjs[q];
                const int nt = fpat == PAT_HESTON ? TPat<PAT_HESTON>::n : TPat<PAT_SABR0>::n;
                const int *tp = fpat == PAT_HESTON ? TPat<PAT_HESTON>::tp : TPat<PAT_SABR0>::tp;
                const int *tq = fpat == PAT_HESTON ? TPat<PAT_HESTON>::tq : TPat<PAT_SABR0>::tq;
                const int *fx = fpat == PAT_HESTON ? TPat<PAT_HESTON>::fx : TPat<PAT_SABR0>::fx;
                const int *fy = fpat == PAT_HESTON ? TPat<PAT_HESTON>::fy : TPat<PAT_SABR0>::fy;
                all = A.nterm == nt && std::equal(tp, tp + nt, A.tp) && std::equal(tq, tq + nt, A.tq) &&
                      std::equal(fx, fx + nt, A.fx) && std::equal(fy, fy + nt, A.fy);
            }
            if (all)
                apat = fpat;
        }
        adj_dev.~dev_buffer();
        new (&adj_dev) dev_buffer();
        adj_dev.alloc(n * sizeof(Adj<T>));
        HIPCHK(hipMemcpy(adj_dev.p, adjs.data(), n * sizeof(Adj<T>), hipMemcpyHostToDevice));
        // per-step problem tables of a segment: step s0 + k reads level slot k, writes slot k + 1
        const int nset = dbl ? 2 : 1;
        std::vector<Prob<T>> seg((size_t)nset * K * n);
        for (int set = 0; set < nset; ++set)
        for (int k = 0; k < K; ++k)
            for (int64_t q = 0; q < n; ++q) {
                Prob<T> t = probs[q];
                const size_t nxy = (size_t)t.nx * t.ny;
                T *lk = seg_lv_base[q + set * n] + (size_t)k * nxy, *lk1 = lk + nxy;
                if (k % 2 == 0) {
                    t.a = lk;
                    t.b = lk1;
                } else {
                    t.b = lk;
                    t.a = lk1;
                }
                const bool dr = t.scheme == SCHEME_DR, hv = t.scheme == SCHEME_HV;
                const size_t nys = (dr ? 1 : 3) + ((hv && line_solver == LINE_EXACT) ? 1 : 0);
                T *yb = seg_y_base[q + set * n] + (size_t)k * nys * nxy;
                t.y1 = yb;
                t.y2 = dr ? nullptr : yb + nxy;
                t.y3 = dr ? nullptr : yb + 2 * nxy;
                t.y4 = line_solver == LINE_PARTITIONED ? t.y3 : (hv ? yb + 3 * nxy : t.y1);
                seg[((size_t)set * K + k) * n + q] = t;
            }
        ck_len = K;
        seg_dev.~dev_buffer();
        new (&seg_dev) dev_buffer();
        seg_dev.alloc(seg.size() * sizeof(Prob<T>));
        HIPCHK(hipMemcpy(seg_dev.p, seg.data(), seg.size() * sizeof(Prob<T>), hipMemcpyHostToDevice));
        // checkpoint copy jobs
        const int nck_max = (max_steps + K - 1) / K;
        std::vector<CopyJob> jobs;
        fwd_job_off.assign(nck_max, 0);
        bwd_job_off.assign(nck_max, 0);
        fwd_job_n.assign(nck_max, 0);
        bwd_job_n.assign(nck_max, 0);
        for (int c = 0; c < nck_max; ++c) {
            fwd_job_off[c] = jobs.size();
            for (int64_t q = 0; q < n; ++q)
                if (probs[q].steps > c * K)
                    jobs.push_back({probs[q].a, ckpt_base[q] + (size_t)c * probs[q].nx * probs[q].ny,
                                    (int64_t)probs[q].nx * probs[q].ny});
            fwd_job_n[c] = (int)(jobs.size() - fwd_job_off[c]);
            bwd_job_off[c] = jobs.size();
            for (int64_t q = 0; q < n; ++q)
                if (probs[q].steps > c * K)
                    jobs.push_back({ckpt_base[q] + (size_t)c * probs[q].nx * probs[q].ny,
                                    seg_lv_base[q + (dbl ? (c & 1) : 0) * n], (int64_t)probs[q].nx * probs[q].ny});
            bwd_job_n[c] = (int)(jobs.size() - bwd_job_off[c]);
        }
        jobs_dev.~dev_buffer();
        new (&jobs_dev) dev_buffer();
        jobs_dev.alloc(jobs.size() * sizeof(CopyJob));
        HIPCHK(hipMemcpy(jobs_dev.p, jobs.data(), jobs.size() * sizeof(CopyJob), hipMemcpyHostToDevice));
        lam_jobs_dev.~dev_buffer();
        new (&lam_jobs_dev) dev_buffer();
        lam_jobs_dev.alloc(2 * n * sizeof(CopyJob));
        // parameter slots.  S-lines per adjoint S-sweep workgroup: 8 for S-lines of more than 512 points
        // (128 chunks per line), else 16 (measured faster than the forward's tuned width on the SABR book)
        axw = pmx >= 16 ? 8 : 16;
        slot_tb[1] = slot_tb[3] = axw == 8 ? TB_X8 : TB_AX;
        nslots = 0;
        for (int k = 0; k < 5; ++k) {
            slot_off[k] = nslots;
            nslots += tiles[slot_tb[k]].size();
        }
        slots_dev.~dev_buffer();
        new (&slots_dev) dev_buffer();
        slots_dev.alloc(nslots * NPAR * sizeof(double));
        ck = K;
    }

    void launch_jobs(const CopyJob *jd, int njobs, hipStream_t st)
    {
        if (njobs <= 0)
            return;
        const unsigned bx = (unsigned)std::max<int64_t>(1, std::min<int64_t>(64, (max_nxy + 255) / 256));
        for (int first = 0; first < njobs; first += 65535) {
            const int nj = std::min(65535, njobs - first);
            hipLaunchKernelGGL(k_copy_jobs<T>, dim3(bx, (unsigned)nj), dim3(256), 0, st, jd + first);
        }
    }

    template <int M, int PAT>
    void adj_v_p(int s, int last, const Prob<T> *table, double *slot, hipStream_t st, unsigned grid)
    {
        hipLaunchKernelGGL((k_adj_v<T, M, PAT>), dim3(grid), dim3(64 * P_VROWS), 0, st, table,
                           (const Adj<T> *)adj_dev.p, (const Tile *)tiles_dev.p + tile_off[TB_AV], s, last, slot);
    }
    template <int M> void adj_v(int s, int last, const Prob<T> *table, double *slot, hipStream_t st, unsigned grid)
    {
        if (apat == PAT_HESTON)
            adj_v_p<M, PAT_HESTON>(s, last, table, slot, st, grid);
        else if (apat == PAT_SABR0)
            adj_v_p<M, PAT_SABR0>(s, last, table, slot, st, grid);
        else
            adj_v_p<M, PAT_GENERIC>(s, last, table, slot, st, grid);
    }
    template <int M, int W, int PAT, int NC = P_NCH>
    void adj_x_w(int s, int stage, const Prob<T> *table, double *slot, hipStream_t st, unsigned grid)
    {
        const Adj<T> *ad = (const Adj<T> *)adj_dev.p;
        const Tile *tx = (const Tile *)tiles_dev.p + tile_off[W == 8 ? TB_X8 : (W == 4 ? TB_X4 : TB_AX)];
        if (stage == 1)
            hipLaunchKernelGGL((k_adj_x<T, M, W, PAT, 1, NC>), dim3(grid), dim3(W * NC), 0, st, table, ad, tx, s, slot);
        else
            hipLaunchKernelGGL((k_adj_x<T, M, W, PAT, 2, NC>), dim3(grid), dim3(W * NC), 0, st, table, ad, tx, s, slot);
    }
    template <int M, int PAT>
    void adj_x_p(int s, int stage, const Prob<T> *table, double *slot, hipStream_t st, unsigned grid)
    {
        // S-lines of more than 512 points (M = pmx = 16): 128 chunks of at most 8 points and 8 lines per
        // workgroup, so that a single large problem has twice the wavefronts (1024 x 512 MCS: backward
        // 142 -> 136 ms); shorter lines: 64 chunks and 16 lines per workgroup (= axw)
        if constexpr (M >= 16)
            adj_x_w<8, 8, PAT, 128>(s, stage, table, slot, st, grid);
        else
            adj_x_w<M, 16, PAT>(s, stage, table, slot, st, grid);
    }
    template <int M> void adj_x(int s, int stage, const Prob<T> *table, double *slot, hipStream_t st, unsigned grid)
    {
        if (apat == PAT_HESTON)
            adj_x_p<M, PAT_HESTON>(s, stage, table, slot, st, grid);
        else if (apat == PAT_SABR0)
            adj_x_p<M, PAT_SABR0>(s, stage, table, slot, st, grid);
        else
            adj_x_p<M, PAT_GENERIC>(s, stage, table, slot, st, grid);
    }
    template <int WHAT> void adj_g(int s, const Prob<T> *table, double *slot, hipStream_t st, unsigned grid)
    {
        const Adj<T> *ad = (const Adj<T> *)adj_dev.p;
        const Tile *tu = (const Tile *)tiles_dev.p + tile_off[TB_U];
        if (apat == PAT_HESTON)
            hipLaunchKernelGGL((k_adj_g<T, WHAT, PAT_HESTON>), dim3(grid), dim3(AGT), 0, st, table, ad, tu, s, slot);
        else if (apat == PAT_SABR0)
            hipLaunchKernelGGL((k_adj_g<T, WHAT, PAT_SABR0>), dim3(grid), dim3(AGT), 0, st, table, ad, tu, s, slot);
        else
            hipLaunchKernelGGL((k_adj_g<T, WHAT, PAT_GENERIC>), dim3(grid), dim3(AGT), 0, st, table, ad, tu, s,
                               slot);
    }
    void adj_v_m(int s, int last, const Prob<T> *table, double *slot, hipStream_t st, unsigned grid)
    {
        switch (pmy) {
        case 1: adj_v<1>(s, last, table, slot, st, grid); break;
        case 2: adj_v<2>(s, last, table, slot, st, grid); break;
        case 4: adj_v<4>(s, last, table, slot, st, grid); break;
        case 8: adj_v<8>(s, last, table, slot, st, grid); break;
        default: adj_v<16>(s, last, table, slot, st, grid); break;
        }
    }
    void adj_x_m(int s, int stage, const Prob<T> *table, double *slot, hipStream_t st, unsigned grid)
    {
        switch (pmx) {
        case 1: adj_x<1>(s, stage, table, slot, st, grid); break;
        case 2: adj_x<2>(s, stage, table, slot, st, grid); break;
        case 4: adj_x<4>(s, stage, table, slot, st, grid); break;
        case 8: adj_x<8>(s, stage, table, slot, st, grid); break;
        default: adj_x<16>(s, stage, table, slot, st, grid); break;
        }
    }

    void launch_adj_step(int s, const Prob<T> *table, hipStream_t st)
    {
        const int64_t na = active(s);
        auto grid = [&](int tb) { return (unsigned)prefix[tb][na]; };
        bool any_split2 = false;
        for (int64_t q = 0; q < na; ++q)
            any_split2 |= probs[q].scheme != SCHEME_DR;
        double *sl = (double *)slots_dev.p;
        adj_v_m(s, 1, table, sl + slot_off[0] * NPAR, st, grid(TB_AV));
        if (any_split2) {
            adj_x_m(s, 2, table, sl + slot_off[1] * NPAR, st, grid(slot_tb[1]));
            adj_g<1>(s, table, sl, st, grid(TB_U));
            adj_v_m(s, 0, table, sl + slot_off[2] * NPAR, st, grid(TB_AV));
        }
        adj_x_m(s, 1, table, sl + slot_off[3] * NPAR, st, grid(slot_tb[3]));
        adj_g<0>(s, table, sl + slot_off[4] * NPAR, st, grid(TB_U));
    }

    void forward_ckpt(const void *const *u_in, void *const *u_out, int K, hipStream_t stream) override
    {
        require(u_in && u_out, "null pointer array");
        for (int64_t c = 0; c < count; ++c)
            require(u_in[c] && u_out[c], "null surface pointer");
        if (ck != K)
            adj_setup(K);
        if (!tuned)
            autotune(stream);
        fwd_done = false;
        for (int64_t q = 0; q < count; ++q) {
            const Prob<T> &p = probs[q];
            HIPCHK(hipMemcpyAsync(p.a, u_in[order[q]], sizeof(T) * (size_t)p.nx * p.ny, hipMemcpyDeviceToDevice,
                                  stream));
        }
        const CopyJob *jd = (const CopyJob *)jobs_dev.p;
        for (int s = 0; s < max_steps; ++s) {
            if (s % K == 0)
                launch_jobs(jd + fwd_job_off[s / K], fwd_job_n[s / K], stream);
            launch_step(s, stream);
        }
        HIPCHK(hipGetLastError());
        for (int64_t q = 0; q < count; ++q) {
            const Prob<T> &p = probs[q];
            const T *res = (p.steps & 1) ? p.b : p.a;
            HIPCHK(hipMemcpyAsync(u_out[order[q]], res, sizeof(T) * (size_t)p.nx * p.ny, hipMemcpyDeviceToDevice,
                                  stream));
        }
        fwd_done = true;
    }

    void backward(const void *const *lam_out, void *const *lam_in, double *dparams, hipStream_t stream) override
    {
        require(fwd_done && ck > 0, "lss_adi_backward needs a preceding lss_adi_forward on this plan");
        require(lam_out && dparams, "null lam_out or dparams");
        for (int64_t c = 0; c < count; ++c)
            require(lam_out[c], "null lambda pointer");
        const int K = ck;
        HIPCHK(hipStreamSynchronize(stream));
        HIPCHK(hipMemsetAsync(slots_dev.p, 0, nslots * NPAR * sizeof(double), stream));
        std::vector<CopyJob> lj;
        for (int64_t q = 0; q < count; ++q) {
            const Prob<T> &p = probs[q];
            lj.push_back({lam_out[order[q]], (p.steps & 1) ? adjs[q].lb : adjs[q].la, (int64_t)p.nx * p.ny});
        }
        const size_t nin = lj.size();
        if (lam_in)
            for (int64_t q = 0; q < count; ++q)
                if (lam_in[order[q]])
                    lj.push_back({adjs[q].la, lam_in[order[q]], (int64_t)probs[q].nx * probs[q].ny});
        HIPCHK(hipMemcpy(lam_jobs_dev.p, lj.data(), lj.size() * sizeof(CopyJob), hipMemcpyHostToDevice));
        const CopyJob *ljd = (const CopyJob *)lam_jobs_dev.p;
        launch_jobs(ljd, (int)nin, stream);
        const CopyJob *jd = (const CopyJob *)jobs_dev.p;
        const int nseg = (max_steps + K - 1) / K;
        // recomputation of segment c from its checkpoint into storage set `set`, keeping every stage
        auto recompute = [&](int c, int set, hipStream_t st) {
            const int s0 = c * K, s1 = std::min(max_steps, s0 + K);
            launch_jobs(jd + bwd_job_off[c], bwd_job_n[c], st);
            for (int s = s0; s < s1; ++s)
                launch_step(s, st, nullptr, seg_table(s - s0, set));
        };
        auto adjoint = [&](int c, int set) {
            const int s0 = c * K, s1 = std::min(max_steps, s0 + K);
            for (int s = s1 - 1; s >= s0; --s)
                launch_adj_step(s, seg_table(s - s0, set), stream);
        };
        if (!dbl) {
            for (int c = nseg - 1; c >= 0; --c) {
                recompute(c, 0, stream);
                adjoint(c, 0);
            }
        } else {
            // segment c lives in set c & 1: aux_stream recomputes segment c - 1 (after the adjoint of
            // segment c + 1 has released that set) while this stream runs the adjoint of segment c
            if (!aux_stream) {
                HIPCHK(hipStreamCreateWithFlags(&aux_stream, hipStreamNonBlocking));
                HIPCHK(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
                for (int k = 0; k < 2; ++k) {
                    HIPCHK(hipEventCreateWithFlags(&ev_ready[k], hipEventDisableTiming));
                    HIPCHK(hipEventCreateWithFlags(&ev_free[k], hipEventDisableTiming));
                }
            }
            HIPCHK(hipEventRecord(ev_fork, stream));
            HIPCHK(hipStreamWaitEvent(aux_stream, ev_fork, 0));
            recompute(nseg - 1, (nseg - 1) & 1, aux_stream);
            HIPCHK(hipEventRecord(ev_ready[(nseg - 1) & 1], aux_stream));
            for (int c = nseg - 1; c >= 0; --c) {
                if (c >= 1) {
                    const int set = (c - 1) & 1;
                    if (c + 1 <= nseg - 1)  // the set held segment c + 1
                        HIPCHK(hipStreamWaitEvent(aux_stream, ev_free[set], 0));
                    recompute(c - 1, set, aux_stream);
                    HIPCHK(hipEventRecord(ev_ready[set], aux_stream));
                }
                HIPCHK(hipStreamWaitEvent(stream, ev_ready[c & 1], 0));
                adjoint(c, c & 1);
                HIPCHK(hipEventRecord(ev_free[c & 1], stream));
            }
        }
        launch_jobs(ljd + nin, (int)(lj.size() - nin), stream);
        HIPCHK(hipGetLastError());
        std::vector<double> sl(nslots * NPAR);
        HIPCHK(hipMemcpyAsync(sl.data(), slots_dev.p, sl.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        std::fill(dparams, dparams + count * NPAR, 0.0);
        for (int k = 0; k < 5; ++k) {
            const auto &tt = tiles[slot_tb[k]];
            for (size_t t = 0; t < tt.size(); ++t) {
                const int64_t c = order[tt[t].prob];
                for (int z = 0; z < NPAR; ++z)
                    dparams[c * NPAR + z] += sl[(slot_off[k] + t) * NPAR + z];
            }
        }
    }

    // every step of every problem of a plan of small explicit Euler problems (euler_res), the levels
    // resident in LDS; the result is where the step-by-step loop leaves it
    void launch_euler_res(hipStream_t st)
    {
        const Prob<T> *pd = (const Prob<T> *)probs_dev.p;
        const Tile *tp = (const Tile *)tiles_dev.p + tile_off[TB_P];
        const dim3 g((unsigned)count), b(ERES_T);
        if (fpat == PAT_HESTON)
            hipLaunchKernelGGL((k_euler_res<T, PAT_HESTON>), g, b, 0, st, pd, tp);
        else if (fpat == PAT_SABR0)
            hipLaunchKernelGGL((k_euler_res<T, PAT_SABR0>), g, b, 0, st, pd, tp);
        else
            hipLaunchKernelGGL((k_euler_res<T, PAT_GENERIC>), g, b, 0, st, pd, tp);
        HIPCHK(hipGetLastError());
    }

    void launch_loop(hipStream_t st)
    {
        for (int s = 0; s < max_steps; ++s)
            launch_step(s, st);
        HIPCHK(hipGetLastError());
    }

    void profile(const void *const *u_in, void *const *u_out, hipStream_t stream, double *ms, double *bytes,
                 int64_t *launches) override
    {
        Profiler prof;
        prof.st = stream;
        for (int64_t q = 0; q < count; ++q) {
            const Prob<T> &p = probs[q];
            HIPCHK(hipMemcpyAsync(p.a, u_in[order[q]], sizeof(T) * (size_t)p.nx * p.ny, hipMemcpyDeviceToDevice,
                                  stream));
        }
        for (int s = 0; s < max_steps; ++s)
            launch_step(s, stream, &prof);
        HIPCHK(hipGetLastError());
        for (int64_t q = 0; q < count; ++q) {
            const Prob<T> &p = probs[q];
            const T *res = (p.steps & 1) ? p.b : p.a;
            HIPCHK(hipMemcpyAsync(u_out[order[q]], res, sizeof(T) * (size_t)p.nx * p.ny, hipMemcpyDeviceToDevice,
                                  stream));
        }
        HIPCHK(hipStreamSynchronize(stream));
        for (int k = 0; k < LSS_K_COUNT; ++k) {
            ms[k] = 0;
            bytes[k] = 0;
            launches[k] = 0;
        }
        for (size_t l = 0; l < prof.kind.size(); ++l) {
            float t = 0;
            HIPCHK(hipEventElapsedTime(&t, prof.ev[2 * l], prof.ev[2 * l + 1]));
            ms[prof.kind[l]] += t;
            bytes[prof.kind[l]] += prof.bytes[l];
            launches[prof.kind[l]] += 1;
        }
    }

    void run(const void *const *u_in, void *const *u_out, hipStream_t stream, bool use_graph, bool host) override
    {
        require(u_in && u_out, "null pointer array");
        for (int64_t c = 0; c < count; ++c)
            require(u_in[c] && u_out[c], "null surface pointer");
        const hipMemcpyKind kin = host ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
        const hipMemcpyKind kout = host ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
        if (!tuned)
            autotune(stream);
        // device surfaces of a plan of many problems: two batched copy launches instead of one copy
        // per problem and direction
        const bool batched_io = !host && count >= IO_BATCH_MIN;
        if (batched_io) {
            io_jobs(u_in, u_out, stream);
            launch_jobs((const CopyJob *)io_jobs_dev.p, (int)count, stream);
        } else {
            for (int64_t q = 0; q < count; ++q) {
                const Prob<T> &p = probs[q];
                HIPCHK(hipMemcpyAsync(p.a, u_in[order[q]], sizeof(T) * (size_t)p.nx * p.ny, kin, stream));
            }
        }
        if (euler_res) {  // the whole loop in one launch, one workgroup per problem
            launch_euler_res(stream);
        } else if (use_graph && ngroups > 1) {
            run_groups(stream);
        } else if (use_graph) {
            if (!graph_exec) {
                HIPCHK(hipStreamCreateWithFlags(&cap_stream, hipStreamNonBlocking));
                HIPCHK(hipStreamBeginCapture(cap_stream, hipStreamCaptureModeThreadLocal));
                for (int s = 0; s < max_steps; ++s)
                    launch_step(s, cap_stream);
                HIPCHK(hipStreamEndCapture(cap_stream, &graph));
                HIPCHK(hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0));
            }
            HIPCHK(hipGraphLaunch(graph_exec, stream));
        } else {
            launch_loop(stream);
        }
        if (batched_io) {
            launch_jobs((const CopyJob *)io_jobs_dev.p + count, (int)count, stream);
        } else {
            for (int64_t q = 0; q < count; ++q) {
                const Prob<T> &p = probs[q];
                const T *res = (p.steps & 1) ? p.b : p.a;
                HIPCHK(hipMemcpyAsync(u_out[order[q]], res, sizeof(T) * (size_t)p.nx * p.ny, kout, stream));
            }
        }
        if (host)
            HIPCHK(hipStreamSynchronize(stream));
    }

    // Every time level: levels[p] receives steps + 1 surfaces of nx * ny values, row s = the surface after s
    // steps (row 0 = u_in[p]).  One batched copy launch per step moves the new level of every problem still
    // stepping; the job table (sum over problems of steps + 1 jobs) is rebuilt when the pointers change.
    dev_buffer lv_jobs_dev;
    std::vector<const void *> lv_last;
    std::vector<size_t> lv_first;  // first job of each step s = 0..max_steps (s = 0: the inputs)
    void level_jobs(const void *const *u_in, void *const *levels, hipStream_t stream)
    {
        std::vector<const void *> key(u_in, u_in + count);
        key.insert(key.end(), levels, levels + count);
        if (key == lv_last)
            return;
        std::vector<CopyJob> jobs;
        lv_first.assign(max_steps + 2, 0);
        for (int s = 0; s <= max_steps; ++s) {
            lv_first[s] = jobs.size();
            const int64_t na = s == 0 ? count : active(s - 1);
            for (int64_t q = 0; q < na; ++q) {
                const Prob<T> &p = probs[q];
                const size_t nxy = (size_t)p.nx * p.ny;
                T *dst = (T *)levels[order[q]] + (size_t)s * nxy;
                // step s - 1 wrote ((s - 1) & 1 ? a : b)
                const T *src = s == 0 ? (const T *)u_in[order[q]] : (((s - 1) & 1) ? p.a : p.b);
                jobs.push_back({src, dst, (int64_t)nxy});
            }
        }
        lv_first[max_steps + 1] = jobs.size();
        HIPCHK(hipStreamSynchronize(stream));  // an earlier run may still read the old table
        lv_jobs_dev.~dev_buffer();
        new (&lv_jobs_dev) dev_buffer();
        lv_jobs_dev.alloc(jobs.size() * sizeof(CopyJob));
        HIPCHK(hipMemcpy(lv_jobs_dev.p, jobs.data(), jobs.size() * sizeof(CopyJob), hipMemcpyHostToDevice));
        lv_last = std::move(key);
    }

    void run_levels(const void *const *u_in, void *const *levels, hipStream_t stream, bool host) override
    {
        require(u_in && levels, "null pointer array");
        for (int64_t c = 0; c < count; ++c)
            require(u_in[c] && levels[c], "null surface pointer");
        if (!tuned)
            autotune(stream);
        if (host) {  // stage inputs and levels through one device allocation
            std::vector<size_t> off(count);
            size_t tot = 0;
            for (int64_t c = 0; c < count; ++c) {
                off[c] = tot;
                tot += (size_t)nx[c] * ny[c] * (size_t)(steps[c] + 2);
            }
            dev_buffer stage;
            stage.alloc(tot * sizeof(T));
            std::vector<const void *> di(count);
            std::vector<void *> dl(count);
            for (int64_t c = 0; c < count; ++c) {
                const size_t nxy = (size_t)nx[c] * ny[c];
                T *base = (T *)stage.p + off[c];
                HIPCHK(hipMemcpy(base, u_in[c], nxy * sizeof(T), hipMemcpyHostToDevice));
                di[c] = base;
                dl[c] = base + nxy;
            }
            run_levels(di.data(), dl.data(), nullptr, false);
            HIPCHK(hipDeviceSynchronize());
            lv_last.clear();  // the staging buffer goes away
            for (int64_t c = 0; c < count; ++c) {
                const size_t nxy = (size_t)nx[c] * ny[c];
                HIPCHK(hipMemcpy(levels[c], (T *)stage.p + off[c] + nxy, nxy * (size_t)(steps[c] + 1) * sizeof(T),
                                 hipMemcpyDeviceToHost));
            }
            return;
        }
        level_jobs(u_in, levels, stream);
        const CopyJob *jd = (const CopyJob *)lv_jobs_dev.p;
        for (int64_t q = 0; q < count; ++q) {
            const Prob<T> &p = probs[q];
            HIPCHK(hipMemcpyAsync(p.a, u_in[order[q]], sizeof(T) * (size_t)p.nx * p.ny, hipMemcpyDeviceToDevice,
                                  stream));
        }
        launch_jobs(jd, (int)(lv_first[1] - lv_first[0]), stream);
        for (int s = 0; s < max_steps; ++s) {
            launch_step(s, stream);
            launch_jobs(jd + lv_first[s + 1], (int)(lv_first[s + 2] - lv_first[s + 1]), stream);
        }
        HIPCHK(hipGetLastError());
    }

    // copy jobs of the surfaces of a device run (count inputs, then count outputs), rebuilt when the
    // caller's pointers change
    static constexpr int64_t IO_BATCH_MIN = 8;
    dev_buffer io_jobs_dev;
    std::vector<const void *> io_last;
    void io_jobs(const void *const *u_in, void *const *u_out, hipStream_t stream)
    {
        std::vector<const void *> key(u_in, u_in + count);
        key.insert(key.end(), u_out, u_out + count);
        if (key == io_last)
            return;
        std::vector<CopyJob> jobs(2 * count);
        for (int64_t q = 0; q < count; ++q) {
            const Prob<T> &p = probs[q];
            const int64_t n = (int64_t)p.nx * p.ny;
            jobs[q] = {u_in[order[q]], p.a, n};
            jobs[count + q] = {(p.steps & 1) ? p.b : p.a, u_out[order[q]], n};
        }
        if (!io_jobs_dev.p)
            io_jobs_dev.alloc(jobs.size() * sizeof(CopyJob));
        HIPCHK(hipStreamSynchronize(stream));  // an earlier run may still read the old table
        HIPCHK(hipMemcpy(io_jobs_dev.p, jobs.data(), jobs.size() * sizeof(CopyJob), hipMemcpyHostToDevice));
        io_last = std::move(key);
    }
};

lss_adi_plan *make_plan(int fp, const lss_adi_desc *descs, int64_t n)
{
    if (fp == LSS_FP64) {
        auto p = std::make_unique<PlanImpl<double>>();
        p->fp = fp;
        p->build(descs, n);
        return p.release();
    }
    if (fp == LSS_FP32) {
        auto p = std::make_unique<PlanImpl<float>>();
        p->fp = fp;
        p->build(descs, n);
        return p.release();
    }
    throw lss_error(LSS_ERR_INVALID, "fp must be LSS_FP64 or LSS_FP32");
}

// ------------------------------------------------------------------------------------------------
// Reference models (the data of the reference's Heston and SABR tests)
// ------------------------------------------------------------------------------------------------
template <typename T> lss_amd_detail::grid_spec<T> spec_of(const lss_model &m)
{
    lss_amd_detail::grid_spec<T> g;
    g.x_lo = (T)m.x_lo;
    g.x_hi = (T)m.x_hi;
    g.y_lo = (T)m.y_lo;
    g.y_hi = (T)m.y_hi;
    g.t_lo = (T)m.t_lo;
    g.t_hi = (T)m.t_hi;
    g.nx = (size_t)m.nx;
    g.ny = (size_t)m.ny;
    g.nt = (size_t)m.nt;
    g.accumulation_point = (T)m.accumulation_point;
    g.alpha_scale = (T)m.alpha_scale;
    g.beta_scale = (T)m.beta_scale;
    g.nonuniform = m.grid == LSS_GRID_NONUNIFORM;
    g.splitting = m.splitting;
    g.weighting = (T)m.weighting;
    g.implicit_scheme = m.implicit_scheme;
    g.tridiagonal = m.tridiagonal;
    g.direction = m.direction;
    return g;
}

// The reference's test lambdas capture double constants, so each expression is evaluated in
// double and rounded to T on return (lss_print_t.hpp:2156-2186, 2039-2050).
template <typename T> lss_amd_detail::user_data<T> closures_of(const lss_model &m)
{
    lss_amd_detail::user_data<T> u;
    const double K = m.strike, r = m.rate, sig = m.sig_sig, kap = m.sig_kappa, th = m.sig_theta, rho = m.rho;
    const double be = m.beta, mat = m.maturity, lo = m.bnd_low, hi = m.bnd_high;
    if (m.kind == LSS_MODEL_HESTON) {
        u.A = [](T, T s, T v) { return (T)(0.5 * v * s * s); };
        u.B = [=](T, T, T v) { return (T)(0.5 * sig * sig * v); };
        u.C = [=](T, T s, T v) { return (T)(rho * sig * v * s); };
        u.D = [=](T, T s, T) { return (T)(r * s); };
        u.E = [=](T, T, T v) { return (T)(kap * (th - v)); };
        u.F = [=](T, T, T) { return (T)(-r); };
    } else {
        auto Df = [=](T t) { return std::exp(-r * (mat - t)); };
        u.A = [=](T t, T s, T al) {
            return (T)(0.5 * al * al * std::pow((double)s, 2.0 * be) * std::pow(Df(t), 2.0 * (1.0 - be)));
        };
        u.B = [=](T, T, T al) { return (T)(0.5 * sig * sig * al * al); };
        u.C = [=](T t, T s, T al) {
            return (T)(rho * sig * al * al * std::pow((double)s, be) * std::pow(Df(t), (1.0 - be)));
        };
        u.D = [=](T, T s, T) { return (T)(r * s); };
        u.E = [](T, T, T) { return (T)0.0; };
        u.F = [=](T, T, T) { return (T)(-r); };
    }
    u.init = [=](T s, T) { return std::max<T>((T)0.0, (T)(s - K)); };
    u.h_lower = [=](T, T) { return (T)lo; };
    u.h_upper = [=](T, T) { return (T)hi; };
    u.v_upper = [](T, T s) { return s; };
    return u;
}

// Separable factors of the transformed coefficients for the two models (coefficients evaluated
// in double from the T-rounded grid data, then rounded to T).
template <typename T>
void separable_set(const lss_model &m, const lss_amd_detail::grid_spec<T> &g,
                   const lss_amd_detail::sampled<T> &s, T t, T *out)
{
    const lss_amd_detail::grid_map<T> gm(g);
    const size_t nx = g.nx, ny = g.ny;
    T *X = out, *Y = out + 12 * nx;
    std::fill(out, out + 12 * (nx + ny), (T)0);
    auto x = [&](int q, int r, size_t i) -> T & { return X[(2 * q + r) * nx + i]; };
    auto y = [&](int q, int r, size_t j) -> T & { return Y[(2 * q + r) * ny + j]; };
    const T h_1 = T(1) / static_cast<T>(nx - 1), h_2 = T(1) / static_cast<T>(ny - 1);
    const double r = m.rate, sig = m.sig_sig, rho = m.rho;
    const double Dt = m.kind == LSS_MODEL_SABR ? std::exp(-r * (m.maturity - (double)t)) : 1.0;
    for (size_t i = 0; i < nx; ++i) {
        const T z = h_1 * static_cast<T>(i);
        const double S = s.xs[i], dX = gm.dx(z), ddX = gm.ddx(z);
        double aX, cX;
        if (m.kind == LSS_MODEL_HESTON) {
            aX = 0.5 * S * S;
            cX = S;
        } else {
            aX = 0.5 * std::pow(S, 2.0 * m.beta) * std::pow(Dt, 2.0 * (1.0 - m.beta));
            cX = std::pow(S, m.beta) * std::pow(Dt, 1.0 - m.beta);
        }
        x(0, 0, i) = (T)(aX / (dX * dX));
        x(1, 0, i) = (T)1;
        x(2, 0, i) = (T)(cX / dX);
        x(3, 0, i) = (T)(r * S / dX);
        x(3, 1, i) = (T)(-aX * ddX / (dX * dX * dX));
        x(4, 0, i) = (T)1;
        x(5, 0, i) = (T)1;
    }
    for (size_t j = 0; j < ny; ++j) {
        const T e = h_2 * static_cast<T>(j);
        const double V = s.ys[j], dY = gm.dy(e), ddY = gm.ddy(e);
        double aY, bY, cY, eY;
        if (m.kind == LSS_MODEL_HESTON) {
            aY = V;
            bY = 0.5 * sig * sig * V;
            cY = rho * sig * V;
            eY = m.sig_kappa * (m.sig_theta - V);
        } else {
            aY = V * V;
            bY = 0.5 * sig * sig * V * V;
            cY = rho * sig * V * V;
            eY = 0.0;
        }
        y(0, 0, j) = (T)aY;
        y(1, 0, j) = (T)(bY / (dY * dY));
        y(2, 0, j) = (T)(cY / dY);
        y(3, 0, j) = (T)1;
        y(3, 1, j) = (T)aY;
        y(4, 0, j) = (T)(eY / dY - bY * ddY / (dY * dY * dY));
        y(5, 0, j) = (T)(-r);
    }
}

// Derivative terms of the transformed coefficients with respect to the model parameters
// (transform of lss_heat_data_transform.hpp:174-214 applied to the derivatives of the closures):
//   Heston (kappa, theta_v, sigma, rho, r):
//     e' = (theta_v - v) / Y', kappa / Y';  b' = sigma v / Y'^2, c' = rho v S / (X' Y'),
//     e' = -sigma v Y'' / Y'^3;  c' = sigma v S / (X' Y');  d' = S / X', f' = -1
//   SABR (nu, rho, beta, lam), D = exp(-r (T - t)):
//     b' = nu al^2 / Y'^2, c' = rho al^2 cX / (X' Y'), e' = -nu al^2 Y'' / Y'^3;  c' = nu al^2 cX / (X' Y');
//     a' = aX_b al^2 / X'^2, c' = rho nu al^2 cX_b / (X' Y'), d' = -aX_b al^2 X'' / X'^3
//     with aX = F^(2 beta) D^(2 (1 - beta)) / 2, cX = F^beta D^(1 - beta), aX_b = 2 aX ln(F / D),
//     cX_b = cX ln(F / D) (0 at F = 0);
//     lam scales the alpha axis about y_lo (Yh = al - y_lo): a' = aX 2 al Yh / X'^2,
//     b' = nu^2 al (Yh - al) / Y'^2, c' = rho nu cX al (2 Yh - al) / (X' Y'), d' = -aX X'' 2 al Yh / X'^3,
//     e' = -nu^2 al Y'' (Yh - al) / Y'^3.
template <typename T>
void param_terms(const lss_model &m, const lss_amd_detail::grid_spec<T> &g, const lss_amd_detail::sampled<T> &s,
                 T t, lss_adi_plan::Terms &tm, T *out)
{
    const lss_amd_detail::grid_map<T> gm(g);
    const size_t nx = g.nx, ny = g.ny, L = nx + ny;
    const T h_1 = T(1) / static_cast<T>(nx - 1), h_2 = T(1) / static_cast<T>(ny - 1);
    const double r = m.rate, sig = m.sig_sig, rho = m.rho;
    std::vector<double> S(nx), dX(nx), ddX(nx), V(ny), dY(ny), ddY(ny);
    for (size_t i = 0; i < nx; ++i) {
        const T z = h_1 * static_cast<T>(i);
        S[i] = s.xs[i];
        dX[i] = gm.dx(z);
        ddX[i] = gm.ddx(z);
    }
    for (size_t j = 0; j < ny; ++j) {
        const T e = h_2 * static_cast<T>(j);
        V[j] = s.ys[j];
        dY[j] = gm.dy(e);
        ddY[j] = gm.ddy(e);
    }
    int nt = 0;
    auto term = [&](int par, int q, auto fx, auto fy) {
        tm.tp[nt] = (int8_t)par;
        tm.tq[nt] = (int8_t)q;
        T *x = out + (size_t)nt * L, *y = x + nx;
        for (size_t i = 0; i < nx; ++i)
            x[i] = (T)fx(i);
        for (size_t j = 0; j < ny; ++j)
            y[j] = (T)fy(j);
        ++nt;
    };
    auto one = [](size_t) { return 1.0; };
    if (m.kind == LSS_MODEL_HESTON) {
        const double kap = m.sig_kappa, th = m.sig_theta;
        auto sx = [&](size_t i) { return S[i] / dX[i]; };
        term(0, 4, one, [&](size_t j) { return (th - V[j]) / dY[j]; });
        term(1, 4, one, [&](size_t j) { return kap / dY[j]; });
        term(2, 1, one, [&](size_t j) { return sig * V[j] / (dY[j] * dY[j]); });
        term(2, 2, sx, [&](size_t j) { return rho * V[j] / dY[j]; });
        term(2, 4, one, [&](size_t j) { return -sig * V[j] * ddY[j] / (dY[j] * dY[j] * dY[j]); });
        term(3, 2, sx, [&](size_t j) { return sig * V[j] / dY[j]; });
        term(4, 3, sx, one);
        term(4, 5, one, [](size_t) { return -1.0; });
        tm.nparam = 5;
    } else {
        const double be = m.beta, lnD = -r * (m.maturity - (double)t);
        const double Dt = std::exp(lnD);
        auto aX = [&](size_t i) { return 0.5 * std::pow(S[i], 2.0 * be) * std::pow(Dt, 2.0 * (1.0 - be)); };
        auto cX = [&](size_t i) { return std::pow(S[i], be) * std::pow(Dt, 1.0 - be); };
        auto lnFD = [&](size_t i) { return S[i] > 0.0 ? std::log(S[i]) - lnD : 0.0; };
        auto a2 = [&](size_t j) { return V[j] * V[j]; };
        term(0, 1, one, [&](size_t j) { return sig * a2(j) / (dY[j] * dY[j]); });
        term(0, 2, [&](size_t i) { return cX(i) / dX[i]; }, [&](size_t j) { return rho * a2(j) / dY[j]; });
        term(0, 4, one, [&](size_t j) { return -sig * a2(j) * ddY[j] / (dY[j] * dY[j] * dY[j]); });
        term(1, 2, [&](size_t i) { return cX(i) / dX[i]; }, [&](size_t j) { return sig * a2(j) / dY[j]; });
        term(2, 0, [&](size_t i) { return 2.0 * aX(i) * lnFD(i) / (dX[i] * dX[i]); }, a2);
        term(2, 2, [&](size_t i) { return cX(i) * lnFD(i) / dX[i]; }, [&](size_t j) { return rho * sig * a2(j) / dY[j]; });
        term(2, 3, [&](size_t i) { return -2.0 * aX(i) * lnFD(i) * ddX[i] / (dX[i] * dX[i] * dX[i]); }, a2);
        // alpha axis scaled by lam about y_lo (grid, and with it the price point at a fixed eta):
        // d Y / d lam = Y - y_lo = Yh, d Y' / d lam = Y', d Y'' / d lam = Y''; with y_lo = 0 only the a, c
        // and d terms are non-zero (b and e are invariant), and dPrice / d alpha0 = dPrice / d lam / alpha0
        // when y_hi is a multiple of alpha0
        const double ylo = (double)g.y_lo;
        auto yh = [&](size_t j) { return V[j] - ylo; };
        term(3, 0, [&](size_t i) { return aX(i) / (dX[i] * dX[i]); }, [&](size_t j) { return 2.0 * V[j] * yh(j); });
        term(3, 1, one, [&](size_t j) { return sig * sig * V[j] * (yh(j) - V[j]) / (dY[j] * dY[j]); });
        term(3, 2, [&](size_t i) { return cX(i) / dX[i]; },
             [&](size_t j) { return rho * sig * (V[j] * (2.0 * yh(j) - V[j])) / dY[j]; });
        term(3, 3, [&](size_t i) { return -aX(i) * ddX[i] / (dX[i] * dX[i] * dX[i]); },
             [&](size_t j) { return 2.0 * V[j] * yh(j); });
        term(3, 4, one, [&](size_t j) { return -sig * sig * V[j] * ddY[j] * (yh(j) - V[j]) / (dY[j] * dY[j] * dY[j]); });
        tm.nparam = 4;
    }
    tm.nterm = nt;
}

template <typename T> struct model_samples {
    std::vector<lss_amd_detail::sampled<T>> s;
    std::vector<lss_adi_desc> d;
};

template <typename T> lss_adi_plan *model_plan(const lss_model *models, int64_t n)
{
    model_samples<T> ms;
    ms.s.resize(n);
    ms.d.resize(n);
    for (int64_t q = 0; q < n; ++q) {
        const lss_model &m = models[q];
        require(m.kind == LSS_MODEL_HESTON || m.kind == LSS_MODEL_SABR, "unknown model kind");
        require(m.nx >= 4 && m.ny >= 4 && m.nt >= 2, "model grid too small");
        require(m.coeff_mode == LSS_COEFF_FIELDS || m.coeff_mode == LSS_COEFF_SEPARABLE, "unknown coeff_mode");
        const auto g = spec_of<T>(m);
        const auto u = closures_of<T>(m);
        auto &s = ms.s[q];
        const bool tdep = m.kind == LSS_MODEL_SABR && m.rate != 0.0;
        const auto cmode = tdep ? lss_amd_detail::TIME_DEPENDENT : lss_amd_detail::TIME_INDEPENDENT;
        if (m.coeff_mode == LSS_COEFF_FIELDS) {
            lss_amd_detail::sample_fields(u, g, s, cmode, lss_amd_detail::TIME_INDEPENDENT);
        } else {
            // grid, initial data and boundaries as for fields; coefficients as separable factors
            lss_amd_detail::sample_grid(g, s);
            s.init.resize(g.nx * g.ny);
            for (size_t i = 0; i < g.nx; ++i)
                for (size_t j = 0; j < g.ny; ++j)
                    s.init[i * g.ny + j] = u.init(s.xs[i], s.ys[j]);
            s.bnd_sets = 1;
            s.bnd.assign(g.nx + 2 * g.ny, (T)0);
            lss_amd_detail::sample_bnd_set(u, g, s, s.times[0], s.bnd.data());
            const size_t stride = 12 * (g.nx + g.ny);
            s.coeff_sets = tdep ? s.steps : 1;
            s.coeff.assign(stride * s.coeff_sets, (T)0);
            for (size_t k = 0; k < s.coeff_sets; ++k)
                separable_set(m, g, s, s.times[k], s.coeff.data() + k * stride);
        }
        ms.d[q] = s.desc(m.coeff_mode, m.line_solver);
    }
    lss_adi_plan *plan = make_plan(sizeof(T) == 8 ? LSS_FP64 : LSS_FP32, ms.d.data(), n);
    plan->terms.resize(n);
    for (int64_t q = 0; q < n; ++q) {
        const lss_model &m = models[q];
        const auto g = spec_of<T>(m);
        const auto &sm = ms.s[q];
        auto &tm = plan->terms[q];
        const bool tdep = m.kind == LSS_MODEL_SABR && m.rate != 0.0;
        tm.sets = tdep ? sm.steps : 1;
        const size_t L = (size_t)(m.nx + m.ny), per = (size_t)NTERM * L;
        std::vector<T> fac(per * tm.sets, (T)0);
        for (int64_t k = 0; k < tm.sets; ++k)
            param_terms(m, g, sm, sm.times[k], tm, fac.data() + k * per);
        // keep only the used terms of each set
        std::vector<T> used((size_t)tm.nterm * L * tm.sets);
        for (int64_t k = 0; k < tm.sets; ++k)
            std::copy(fac.begin() + k * per, fac.begin() + k * per + (size_t)tm.nterm * L,
                      used.begin() + k * (size_t)tm.nterm * L);
        tm.fac.resize(used.size() * sizeof(T));
        std::memcpy(tm.fac.data(), used.data(), tm.fac.size());
    }
    plan->init_host.resize(n);
    plan->grid_host.resize(n);
    for (int64_t q = 0; q < n; ++q) {
        const auto &v = ms.s[q].init;
        plan->init_host[q].resize(v.size() * sizeof(T));
        std::memcpy(plan->init_host[q].data(), v.data(), v.size() * sizeof(T));
        const auto &xs = ms.s[q].xs, &ys = ms.s[q].ys;
        plan->grid_host[q].resize((xs.size() + ys.size()) * sizeof(T));
        std::memcpy(plan->grid_host[q].data(), xs.data(), xs.size() * sizeof(T));
        std::memcpy(plan->grid_host[q].data() + xs.size() * sizeof(T), ys.data(), ys.size() * sizeof(T));
    }
    return plan;
}

constexpr int TRI_WAVES_H = 4;  // systems per workgroup of k_tridiag_rows

template <typename T, int M>
void tridiag_part_m(int layout, int64_t B, int n, const void *a, const void *b, const void *c, const void *d, void *x,
                    size_t lds_rows, const LineBC<T> &bc, hipStream_t st)
{
    const T *pa = (const T *)a, *pb = (const T *)b, *pc = (const T *)c, *pd = (const T *)d;
    T *px = (T *)x;
    if (layout == LSS_LAYOUT_CONTIGUOUS) {
        const unsigned blocks = (unsigned)((B + TRI_WAVES_H - 1) / TRI_WAVES_H);
        hipLaunchKernelGGL((k_tridiag_rows<T, M>), dim3(blocks), dim3(64 * TRI_WAVES_H), (unsigned)lds_rows, st, B, n,
                           pa, pb, pc, pd, px, bc);
    } else {
        const unsigned blocks = (unsigned)((B + P_XTJ - 1) / P_XTJ);
        hipLaunchKernelGGL((k_tridiag_cols<T, M>), dim3(blocks), dim3(P_XTJ * P_NCH), 0, st, B, n, pa, pb, pc, pd,
                           px, bc);
    }
}

template <typename T>
void tridiag_part(int M, int layout, int64_t B, int n, const void *a, const void *b, const void *c, const void *d,
                  void *x, size_t lds_rows, const LineBC<T> &bc, hipStream_t st)
{
    switch (M) {
    case 1: tridiag_part_m<T, 1>(layout, B, n, a, b, c, d, x, lds_rows, bc, st); break;
    case 2: tridiag_part_m<T, 2>(layout, B, n, a, b, c, d, x, lds_rows, bc, st); break;
    case 4: tridiag_part_m<T, 4>(layout, B, n, a, b, c, d, x, lds_rows, bc, st); break;
    case 8: tridiag_part_m<T, 8>(layout, B, n, a, b, c, d, x, lds_rows, bc, st); break;
    default: tridiag_part_m<T, 16>(layout, B, n, a, b, c, d, x, lds_rows, bc, st); break;
    }
}

}  // namespace

// Scratch of lss_tridiag_batched_noalloc: one array per device, held until the process ends.  Host
// code never passes `tridiag_scratch` itself to a kernel (on the host it is only a stand-in for the
// symbol); it uses the device address that hipGetSymbolAddress returns for the current device.
static constexpr size_t TRIDIAG_SCRATCH_BYTES = (size_t)LSS_TRIDIAG_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char tridiag_scratch[TRIDIAG_SCRATCH_BYTES];

namespace {

std::mutex g_scratch_mutex;
std::vector<void *> g_scratch_ptr;  // per device: address of tridiag_scratch (null until first use)

void *static_scratch()
{
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_scratch_mutex);
    if ((int)g_scratch_ptr.size() <= dev)
        g_scratch_ptr.resize(dev + 1, nullptr);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(tridiag_scratch)));
        require(ptr != nullptr, "hipGetSymbolAddress returned a null scratch address");
        g_scratch_ptr[dev] = ptr;
    }
    return g_scratch_ptr[dev];
}

template <typename T> LineBC<T> line_bc(const lss_line_bc *lo, const lss_line_bc *hi, int64_t n)
{
    LineBC<T> bc;
    bc.lo = lo ? lo->type : BC_PLAIN;
    bc.hi = hi ? hi->type : BC_PLAIN;
    bc.lo_lin = lo ? (T)lo->lin : (T)0;
    bc.lo_val = lo ? (T)lo->val : (T)0;
    bc.hi_lin = hi ? (T)hi->lin : (T)0;
    bc.hi_val = hi ? (T)hi->val : (T)0;
    const T step = (T)1 / (T)(n - 1);  // the solver's h (thomas_lu_boundary.hpp: two * step * value)
    bc.h2 = (T)2 * step;
    return bc;
}

// lss_tridiag_batched(_bc, _noalloc): lower/upper null = plain rows.  The one-lane-per-system kernel
// needs B * n values of scratch: with noalloc it uses tridiag_scratch (sizes beyond it are rejected),
// otherwise a stream-ordered allocation that is freed on `st` after the kernel, so that calls on
// different streams never share scratch memory.
void tridiag_run(int32_t fp, int32_t layout, int64_t B, int64_t n, const void *a, const void *b, const void *c,
                 const void *d, const lss_line_bc *lo, const lss_line_bc *hi, void *x, hipStream_t st,
                 bool noalloc = false)
{
    require(fp == LSS_FP64 || fp == LSS_FP32, "fp must be LSS_FP64 or LSS_FP32");
    require(layout == LSS_LAYOUT_CONTIGUOUS || layout == LSS_LAYOUT_STRIDED, "unknown layout");
    require(B >= 0 && n >= 4 && n < (1 << 30), "need B >= 0 and n >= 4");
    require(a && b && c && d && x, "null array");
    for (const lss_line_bc *e : {lo, hi})
        require(!e || e->type == LSS_BC_DIRICHLET || e->type == LSS_BC_NEUMANN || e->type == LSS_BC_ROBIN,
                "end condition type must be LSS_BC_DIRICHLET, LSS_BC_NEUMANN or LSS_BC_ROBIN");
    if (B == 0)
        return;
    const size_t es = fp == LSS_FP64 ? 8 : 4;
    // unknowns of the folded system
    const int64_t nf = n - ((lo && lo->type == LSS_BC_DIRICHLET) ? 1 : 0) - ((hi && hi->type == LSS_BC_DIRICHLET) ? 1 : 0);
    // partitioned solvers for lines of up to NCH * 16 unknowns (rows: while the staged systems fit
    // 64 KiB of LDS), otherwise one lane per system with a scratch workspace
    int M = 1;
    while (M * P_NCH < nf && M < 32)
        M *= 2;
    const size_t lds_rows = (size_t)TRI_WAVES_H * 4 * (size_t)n * es;
    const bool part = M <= 16 && (layout == LSS_LAYOUT_STRIDED || lds_rows <= 65536);
    if (part) {
        if (fp == LSS_FP64)
            tridiag_part<double>(M, layout, B, (int)n, a, b, c, d, x, lds_rows, line_bc<double>(lo, hi, n), st);
        else
            tridiag_part<float>(M, layout, B, (int)n, a, b, c, d, x, lds_rows, line_bc<float>(lo, hi, n), st);
    } else {
        const size_t bytes = (size_t)B * (size_t)n * es;
        void *scr = nullptr;
        if (noalloc) {
            require(bytes <= TRIDIAG_SCRATCH_BYTES,
                    "lss_tridiag_batched_noalloc: B * n * sizeof(fp) = " + std::to_string(bytes) +
                        " bytes of scratch exceeds its static maximum of " + std::to_string(TRIDIAG_SCRATCH_BYTES) +
                        " bytes (systems of more than 1024 unknowns, or contiguous fp64 systems of more than 512, "
                        "need B * n values of scratch)");
            scr = static_scratch();
        } else {
            hipError_t e = hipMallocAsync(&scr, bytes, st);
            if (e != hipSuccess)
                throw lss_error(LSS_ERR_ALLOC, std::string("scratch hipMallocAsync failed: ") + hipGetErrorString(e));
        }
        const unsigned blocks = (unsigned)((B + 255) / 256);
        if (fp == LSS_FP64)
            hipLaunchKernelGGL(k_tridiag_lane<double>, dim3(blocks), dim3(256), 0, st, B, (int)n, (int)layout,
                               (const double *)a, (const double *)b, (const double *)c, (const double *)d,
                               (double *)x, (double *)scr, line_bc<double>(lo, hi, n));
        else
            hipLaunchKernelGGL(k_tridiag_lane<float>, dim3(blocks), dim3(256), 0, st, B, (int)n, (int)layout,
                               (const float *)a, (const float *)b, (const float *)c, (const float *)d,
                               (float *)x, (float *)scr, line_bc<float>(lo, hi, n));
        const hipError_t le = hipGetLastError();
        if (!noalloc)
            HIPCHK(hipFreeAsync(scr, st));
        HIPCHK(le);
        return;
    }
    HIPCHK(hipGetLastError());
}

}  // namespace

extern "C" {

const char *lss_last_error(void) { return last_error().c_str(); }

int32_t lss_version(void) { return 1; }

int lss_adi_plan_create(const lss_adi_desc *descs, int64_t count, lss_adi_plan **plan)
{
    return guarded([&] {
        require(descs && plan && count >= 1, "null descriptor/plan pointer or count < 1");
        *plan = nullptr;
        *plan = make_plan(descs[0].fp, descs, count);
    });
}

int lss_adi_plan_destroy(lss_adi_plan *plan)
{
    return guarded([&] { delete plan; });
}

int lss_adi_run_dev(lss_adi_plan *plan, const void *const *u_in, void *const *u_out, void *stream, int32_t use_graph)
{
    return guarded([&] {
        require(plan != nullptr, "null plan");
        plan->run(u_in, u_out, (hipStream_t)stream, use_graph != 0, false);
    });
}

int lss_adi_run_profiled(lss_adi_plan *plan, const void *const *u_in, void *const *u_out, void *stream, double *ms,
                         double *bytes, int64_t *launches)
{
    return guarded([&] {
        require(plan && u_in && u_out && ms && bytes && launches, "null pointer");
        plan->profile(u_in, u_out, (hipStream_t)stream, ms, bytes, launches);
    });
}

int lss_adi_run_host(lss_adi_plan *plan, const void *const *u_in, void *const *u_out)
{
    return guarded([&] {
        require(plan != nullptr, "null plan");
        plan->run(u_in, u_out, nullptr, false, true);
    });
}

int lss_adi_run_levels(lss_adi_plan *plan, const void *const *u_in, void *const *levels, void *stream, int32_t host)
{
    return guarded([&] {
        require(plan != nullptr, "null plan");
        plan->run_levels(u_in, levels, (hipStream_t)stream, host != 0);
    });
}

int lss_model_plan_create(const lss_model *models, int64_t count, int32_t fp, lss_adi_plan **plan)
{
    return guarded([&] {
        require(models && plan && count >= 1, "null model/plan pointer or count < 1");
        *plan = nullptr;
        if (fp == LSS_FP64)
            *plan = model_plan<double>(models, count);
        else if (fp == LSS_FP32)
            *plan = model_plan<float>(models, count);
        else
            throw lss_error(LSS_ERR_INVALID, "fp must be LSS_FP64 or LSS_FP32");
    });
}

int lss_model_initial(const lss_adi_plan *plan, int64_t p, void *u0_host)
{
    return guarded([&] {
        require(plan && u0_host, "null pointer");
        require(p >= 0 && p < plan->count, "problem index out of range");
        require((int64_t)plan->init_host.size() == plan->count, "plan was not created from models");
        std::memcpy(u0_host, plan->init_host[p].data(), plan->init_host[p].size());
    });
}

int lss_model_grid(const lss_adi_plan *plan, int64_t p, void *xs_host, void *ys_host)
{
    return guarded([&] {
        require(plan && xs_host && ys_host, "null pointer");
        require(p >= 0 && p < plan->count, "problem index out of range");
        require(!plan->grid_host.empty(), "not a model plan");
        const size_t es = plan->fp == LSS_FP64 ? 8 : 4, nxb = (size_t)plan->nx[p] * es;
        std::memcpy(xs_host, plan->grid_host[p].data(), nxb);
        std::memcpy(ys_host, plan->grid_host[p].data() + nxb, (size_t)plan->ny[p] * es);
    });
}

int lss_plan_problem_shape(const lss_adi_plan *plan, int64_t p, int64_t *nx, int64_t *ny, int64_t *steps)
{
    return guarded([&] {
        require(plan != nullptr, "null plan");
        require(p >= 0 && p < plan->count, "problem index out of range");
        if (nx)
            *nx = plan->nx[p];
        if (ny)
            *ny = plan->ny[p];
        if (steps)
            *steps = plan->steps[p];
    });
}

int lss_adi_set_param_terms(lss_adi_plan *plan, int64_t p, int32_t nparam, int32_t nterm, const int32_t *term_param,
                            const int32_t *term_coeff, int64_t sets, const void *factors)
{
    return guarded([&] {
        require(plan != nullptr, "null plan");
        require(p >= 0 && p < plan->count, "problem index out of range");
        require(nparam >= 0 && nparam <= LSS_MAX_PARAMS, "nparam must be in [0, LSS_MAX_PARAMS]");
        require(nterm >= 0 && nterm <= LSS_MAX_TERMS, "nterm must be in [0, LSS_MAX_TERMS]");
        require(sets == 1 || sets == plan->steps[p], "sets must be 1 or the problem's step count");
        require(nterm == 0 || (term_param && term_coeff && factors), "null term array");
        if (plan->terms.empty())
            plan->terms.resize(plan->count);
        lss_adi_plan::Terms tm;
        tm.nparam = nparam;
        tm.nterm = nterm;
        tm.sets = sets;
        for (int t = 0; t < nterm; ++t) {
            require(term_param[t] >= 0 && term_param[t] < nparam, "term parameter out of range");
            require(term_coeff[t] >= LSS_Q_A && term_coeff[t] <= LSS_Q_F, "term coefficient out of range");
            tm.tp[t] = (int8_t)term_param[t];
            tm.tq[t] = (int8_t)term_coeff[t];
        }
        const size_t es = plan->fp == LSS_FP64 ? 8 : 4;
        tm.fac.resize((size_t)sets * nterm * (size_t)(plan->nx[p] + plan->ny[p]) * es);
        if (!tm.fac.empty())
            std::memcpy(tm.fac.data(), factors, tm.fac.size());
        plan->terms[p] = std::move(tm);
        plan->invalidate_adjoint();
    });
}

int lss_adi_autotune(lss_adi_plan *plan, void *stream, int32_t again, int32_t *s_lines_per_group,
                     int32_t *v_lines_per_group, int32_t *xcd_order)
{
    return guarded([&] {
        require(plan && s_lines_per_group && v_lines_per_group && xcd_order, "null pointer");
        int x = 0, v = 0, z = 0;
        plan->tune((hipStream_t)stream, again != 0, &x, &v, &z);
        *s_lines_per_group = x;
        *v_lines_per_group = v;
        *xcd_order = z;
    });
}

int lss_adi_set_mapping(lss_adi_plan *plan, int32_t s_lines_per_group, int32_t v_lines_per_group,
                        int32_t xcd_order)
{
    return guarded([&] {
        require(plan != nullptr, "null plan");
        plan->set_mapping(s_lines_per_group, v_lines_per_group, xcd_order);
    });
}

int lss_adi_set_adjoint_overlap(lss_adi_plan *plan, int32_t mode)
{
    return guarded([&] {
        require(plan != nullptr, "null plan");
        plan->set_overlap(mode);
    });
}

int lss_plan_kernel_config(const lss_adi_plan *plan, int32_t *config)
{
    return guarded([&] {
        require(plan && config, "null pointer");
        plan->kernel_config(config);
    });
}

int lss_plan_param_count(const lss_adi_plan *plan, int64_t p, int32_t *nparam)
{
    return guarded([&] {
        require(plan && nparam, "null pointer");
        require(p >= 0 && p < plan->count, "problem index out of range");
        *nparam = plan->terms.empty() ? 0 : plan->terms[p].nparam;
    });
}

int lss_adi_forward(lss_adi_plan *plan, const void *const *u_in, void *const *u_out, int32_t ckpt_every, void *stream)
{
    return guarded([&] {
        require(plan != nullptr, "null plan");
        plan->forward_ckpt(u_in, u_out, ckpt_every, (hipStream_t)stream);
    });
}

int lss_adi_backward(lss_adi_plan *plan, const void *const *lam_out, void *const *lam_in, double *dparams,
                     void *stream)
{
    return guarded([&] {
        require(plan != nullptr, "null plan");
        plan->backward(lam_out, lam_in, dparams, (hipStream_t)stream);
    });
}

// host-pointer forms of the adjoint pair: the surfaces are staged through device memory; the plan keeps
// its checkpoints between the two calls as with the device forms
int lss_adi_forward_host(lss_adi_plan *plan, const void *const *u_in, void *const *u_out, int32_t ckpt_every)
{
    return guarded([&] {
        require(plan && u_in && u_out, "null pointer");
        const size_t es = plan->fp == LSS_FP64 ? 8 : 4;
        std::vector<size_t> off(plan->count);
        size_t tot = 0;
        for (int64_t c = 0; c < plan->count; ++c) {
            require(u_in[c] && u_out[c], "null surface pointer");
            off[c] = tot;
            tot += 2 * (size_t)(plan->nx[c] * plan->ny[c]) * es;
        }
        dev_buffer stage;
        stage.alloc(tot);
        std::vector<const void *> di(plan->count);
        std::vector<void *> dout(plan->count);
        for (int64_t c = 0; c < plan->count; ++c) {
            const size_t bytes = (size_t)(plan->nx[c] * plan->ny[c]) * es;
            char *base = (char *)stage.p + off[c];
            HIPCHK(hipMemcpy(base, u_in[c], bytes, hipMemcpyHostToDevice));
            di[c] = base;
            dout[c] = base + bytes;
        }
        plan->forward_ckpt(di.data(), dout.data(), ckpt_every, nullptr);
        HIPCHK(hipDeviceSynchronize());
        for (int64_t c = 0; c < plan->count; ++c)
            HIPCHK(hipMemcpy(u_out[c], dout[c], (size_t)(plan->nx[c] * plan->ny[c]) * es, hipMemcpyDeviceToHost));
    });
}

int lss_adi_backward_host(lss_adi_plan *plan, const void *const *lam_out, void *const *lam_in, double *dparams)
{
    return guarded([&] {
        require(plan && lam_out && dparams, "null pointer");
        const size_t es = plan->fp == LSS_FP64 ? 8 : 4;
        std::vector<size_t> off(plan->count);
        size_t tot = 0;
        for (int64_t c = 0; c < plan->count; ++c) {
            require(lam_out[c] != nullptr, "null adjoint seed");
            off[c] = tot;
            tot += 2 * (size_t)(plan->nx[c] * plan->ny[c]) * es;
        }
        dev_buffer stage;
        stage.alloc(tot);
        std::vector<const void *> dl(plan->count);
        std::vector<void *> dg(plan->count, nullptr);
        for (int64_t c = 0; c < plan->count; ++c) {
            const size_t bytes = (size_t)(plan->nx[c] * plan->ny[c]) * es;
            char *base = (char *)stage.p + off[c];
            HIPCHK(hipMemcpy(base, lam_out[c], bytes, hipMemcpyHostToDevice));
            dl[c] = base;
            if (lam_in && lam_in[c])
                dg[c] = base + bytes;
        }
        plan->backward(dl.data(), dg.data(), dparams, nullptr);
        HIPCHK(hipDeviceSynchronize());
        for (int64_t c = 0; c < plan->count; ++c)
            if (dg[c])
                HIPCHK(hipMemcpy(lam_in[c], dg[c], (size_t)(plan->nx[c] * plan->ny[c]) * es, hipMemcpyDeviceToHost));
    });
}

int lss_tridiag_batched(int32_t fp, int32_t layout, int64_t B, int64_t n, const void *a, const void *b,
                        const void *c, const void *d, void *x, void *stream)
{
    return guarded([&] { tridiag_run(fp, layout, B, n, a, b, c, d, nullptr, nullptr, x, (hipStream_t)stream); });
}

int lss_tridiag_batched_bc(int32_t fp, int32_t layout, int64_t B, int64_t n, const void *a, const void *b,
                           const void *c, const void *d, const lss_line_bc *lower, const lss_line_bc *upper, void *x,
                           void *stream)
{
    return guarded([&] {
        require(lower && upper, "null end condition");
        tridiag_run(fp, layout, B, n, a, b, c, d, lower, upper, x, (hipStream_t)stream);
    });
}

int lss_tridiag_batched_noalloc(int32_t fp, int32_t layout, int64_t B, int64_t n, const void *a, const void *b,
                                const void *c, const void *d, const lss_line_bc *lower, const lss_line_bc *upper,
                                void *x, void *stream)
{
    return guarded([&] {
        require((lower == nullptr) == (upper == nullptr), "lower and upper end conditions: both or neither");
        tridiag_run(fp, layout, B, n, a, b, c, d, lower, upper, x, (hipStream_t)stream, true);
    });
}

}  // extern "C"
