// ppe.hip -- pressure Poisson equation (PPE) of the semi-implicit MPS step, assembled and
// solved on the device.
//
// The reference declares this stage (CalcPoisson, Source/DeviceProcess.h:50-53;
// CalculatePressureCG / NJPCG, Source/CGPUCellSet.h:218-225) and keeps its coefficients
// (CoefficientCIterate = 2d/(lambda n0), PressureCoefficient = lambda n0 rho / (2 d dt^2),
// radix/SPVSOLVER/MultiGPUSolver/MultiGPUHandler.cpp:1303-1314) but never defines it.  Its live
// step uses an equation of state instead (DeviceProcess_Kernel.h:783-787).
//
// System, one row per sorted particle t:
//   sum_j w_ij (p_t - p_j) = alpha max(n*_t - n0s, 0) / n0s,   alpha = lambda_lap n0_lap rho / (2 d dt^2)
// with w_ij = GetWeight(re_lap, max(r_ij, re_lap 1e-3)) from the build-precision distance.  n0s
// is the PND of the ideal lattice under this build's distance and weight (zero source for a
// perfect lattice); lambda_lap and n0_lap are the exact-sqrt lattice values at re_lap.  Rows of
// particles that are out of the zone, not binned, or at the free surface (n* <= beta n0) are
// Dirichlet p = 0; the walls contribute no Laplacian term (Neumann), only their PND to n*.  Only
// compression is a source: rows just below the free surface (beta n0 < n* < n0s) would otherwise
// be strong negative sources whose suction cancels the hydrostatic pressure until the fluid has
// compressed for many steps, and the pressure then switches on abruptly (particles collapse at
// l0 = 2e-3 m).  A is an M-matrix, so a non-negative source gives p >= 0.  The pressure handed to
// the corrector is still max(p, 0) (an inexact solve can dip below 0): the corrector is the
// reference's symmetric (P_i + P_j) gradient, which would turn a suction pressure into an
// attraction.
//
// Storage: SELL-64/4 -- slices of 64 row positions (one wave).  Inside every window of 256 rows
// the rows take positions by decreasing entry bound (SELL-C-sigma, sigma = 256); entry k of
// position P sits at slice_off[P/64] + 256 (k/4) + 4 (P%64) + k%4, so a lane loads four
// consecutive entries in one 16 B access and a wave's group of four is 1 KB contiguous.  Every
// reader stops at its row's entry count rounded up to four; only that last group is padded.  The
// diagonal is a separate FP64 vector.  Rows keep the canonical neighbour order, so the assembled
// matrix equals the CPU restatement's bit for bit.
//
// Assembly: the PND pass of the PPE path (k_pnd with PPE, particles.hip) gives every row its
// active flag, a bound on its entries (all Laplacian-radius neighbours), its FP64 diagonal and its
// Dirichlet flag; k_sell_order places the rows and sizes the slices from the bounds; k_ppe_fill
// writes the entries (Dirichlet neighbours dropped), the real counts and the source.
//
// Solvers:
//   FP64   Jacobi-preconditioned CG, FP64 values, vectors and dots.
//   MIXED  iterative refinement: FP64 residual r = b - A x (FP32 values, FP64 accumulation),
//          inner PCG with FP32 values, FP32 vectors and SpMV, FP64 dot products, then x += e in
//          FP64.  Repeats until the FP64 residual meets the tolerance.
//   Either with the Jacobi or the two-level preconditioner (aggregates of cells, dense coarse
//   inverse).  A two-level iteration is seven launches (SpMV + p.q, alpha, restriction, coarse
//   product, update + z + dots, beta, direction); every kernel reads a device-side "done" flag and
//   returns at once after convergence, so the host queues fixed chunks of iterations and looks at
//   the flag once per chunk.
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>
#include <algorithm>
#include "mpsr_internal.h"
#include "neighbors.h"
#include <utility>
#include "ppe.h"
#include "mpnn.h"

namespace mpsr {

static constexpr int PB = 256;          // block of the vector kernels (4 waves = 4 slices)
static constexpr int CHUNK = 32;        // PCG iterations queued between two flag reads
static constexpr int CHUNK_TAIL = 4;    // ... after a first chunk sized from the previous solve
static constexpr int HINT_PASSES = 4;   // refinement passes whose iteration counts are remembered
static constexpr int SELL_SIGMA = 256;  // rows ordered by stored length inside windows of this many
static constexpr int NB_WIDTH = 192;    // Laplacian-radius neighbours listed per row by the PND pass
static constexpr int WIN_BYTES = 32768; // LDS a workgroup of the FP32 SpMV may stage p in
static constexpr int RESID_WIN_BYTES = 49152;  // ... and of the residual, x in FP64
static_assert(SELL_SIGMA == PB, "one SpMV workgroup covers one SELL window");
// MPSR_PPE_COLUMNS=32 keeps the SpMV on the 32-bit columns (for comparing the two forms)
static bool c16_off() {
    static const int off = [] {
        const char* e = std::getenv("MPSR_PPE_COLUMNS");
        return (e && std::strcmp(e, "32") == 0) ? 1 : 0;
    }();
    return off != 0;
}
// MPSR_PPE_WINDOW=0 keeps the FP32 SpMV's gathers of p in global memory (for comparing the two forms)
static bool win_off() {
    static const int off = [] {
        const char* e = std::getenv("MPSR_PPE_WINDOW");
        return (e && std::strcmp(e, "0") == 0) ? 1 : 0;
    }();
    return off != 0;
}
// S_SKIP and S_OK serve the batched solve: its start found the residual already within the
// tolerance; its final residual passed the checks that let the pressure be written back
enum Scal { S_RZ, S_PQ, S_ALPHA, S_BETA, S_RR, S_TARGET, S_DONE, S_ITERS, S_MAXIT, S_BB, S_RR0, S_SKIP, S_OK, S_COUNT };

template <typename R> struct PpeSolver {
    int cap = 0;
    int nslices = 0;
    double alpha = 0, n0 = 0, n0_lap = 0, lambda_lap = 0;
    int* active = nullptr;
    int* dir = nullptr;
    int* cnt = nullptr;
    int hint[HINT_PASSES] = {};  // iterations of each refinement pass of the last solve (0: none)
    int* nbq = nullptr;          // [rows * NB_WIDTH] neighbour list of the PND pass (sorted index)
    int* nbn = nullptr;          // [rows] its neighbour count per row (-1: row skipped)
    int* slice_w = nullptr;
    int* pos = nullptr;          // [positions] SELL position of each row (rows ordered by length per window)
    int* rowp = nullptr;         // [positions] row at each SELL position, -1 when empty
    int* slice_off = nullptr;
    int* total = nullptr;
    int* cols = nullptr;
    uint16_t* cols16 = nullptr;  // the columns as 16-bit codes for the PCG SpMV (k_ppe_fill explains them)
    int* first_ge = nullptr;     // [ncell + 1] first sorted row whose cell key is >= the index
    int* wbox = nullptr;         // [windows * 8] per SELL window: the p ranges its rows read (k_win_box)
    int wcap = 0;                // p values a window stages in LDS in this solve's SpMV (0: none)
    int c16 = 0;                 // this solve's SpMV reads the 16-bit codes (every column fitted one)
    int cols_valid = 0;          // the 32-bit columns of the current matrix are written
    int cols_wide = 0;           // a matrix had columns beyond the codes: the fill writes 32-bit columns
    const uint32_t* skeys = nullptr;  // this solve's sorted cell keys
    int gz = 0, gygz = 0;        // cell-key strides of one grid row and one x-slab
    double* vals64 = nullptr;
    float* vals32 = nullptr;
    long long nnz_cap = 0;
    double* diag = nullptr;
    double* rhs = nullptr;
    double* x = nullptr;
    double* prev = nullptr;      // last solution, slot order
    int have_prev = 0;
    double *r = nullptr, *z = nullptr, *p0 = nullptr, *p1 = nullptr, *q = nullptr;
    float *rf = nullptr, *zf = nullptr, *p0f = nullptr, *p1f = nullptr, *qf = nullptr, *ef = nullptr;
    double* part = nullptr;      // [2 * nblocks]
    double* scal = nullptr;      // device scalars
    double* scal_h = nullptr;    // pinned host mirror
    int* cnt_h = nullptr;        // pinned host mirror of nnz_dev
    int mixed = 0;
    SortScratch scan{};
    // last solve
    int iters = 0, passes = 0, status = 0, fallbacks = 0;
    double relres = 0, relres0 = 0;
    char why[256] = {0};         // what went wrong in the last solve
    int* nnz_dev = nullptr;      // [4]: off-diagonal entries, non-Dirichlet rows, wide columns, largest window
    long long nnz_real = 0;
    long long nnz_stored = 0;     // entries kept, SELL padding included
    int rows_real = 0;
    // SpMV timing (profiling)
    int profile = 0;             // 1: events around every SpMV launch; 2: around those of every 4th chunk
    unsigned chunk_tick = 0;
    std::vector<hipEvent_t> ev;
    size_t ev_used = 0;
    // two-level preconditioner: aggregates of a x a x a cells, dense inverse of Ac = P^T A P
    int precond = 0;
    int agg_edge = 4;
    int cgx = 0, cgy = 0, cgz = 0, nfull = 0;
    int* agg_flag = nullptr;     // [nfull]
    int* agg_map = nullptr;      // [nfull] compact index (valid where flagged)
    int* agg_full = nullptr;     // [nfull] compact -> full
    int* agg_of_row = nullptr;   // [rows] compact aggregate of a non-Dirichlet row, else -1
    // cell level of the V-cycle preconditioner (MPSR_PRECOND_CELL_VCYCLE), over the grid cells
    int vcycle = 0;              // this solve uses it
    int ncell = 0;
    int cell_age = -1;           // solves since Ac1 was built (-1: never)
    int cell_rebuilt = 0;        // this solve rebuilt Ac1
    int iters_at_cell = 0;       // fewest iterations of a solve since Ac1 was built
    int* cxyz = nullptr;         // [rows] packed cell coordinates of every row
    int* cell_of_row = nullptr;  // [rows] cell of a non-Dirichlet row, else -1
    int* c_agg = nullptr;        // [ncell] compact aggregate of the cell, else -1
    float* c_op = nullptr;       // [ncell * 27] Ac1 = P1^T A P1 as a 27-point stencil on the cell grid
    float* c_dinv = nullptr;     // [ncell] omega / Ac1 diagonal (0 for a cell without rows)
    double* c_rc = nullptr;      // [ncell] restricted residual
    double* c_e = nullptr;       // [ncell] pre-smoothed correction
    double* c_ef = nullptr;      // [ncell] correction after the coarse correction and post-smoothing
    int* nc_dev = nullptr;
    int nc = 0, nc_cap = 0;
    double* Ac = nullptr;        // [nc_cap^2] coarse matrix, then its inverse (column-major)
    float* Af = nullptr;         // [nc_cap^2] the inverse in FP32 (the FP32 inner solve reads it)
    int2* agg_rng = nullptr;     // [nc_cap * a^2] row range of every cell stack of an aggregate
    double* cpart = nullptr;     // [nc_cap * a^3 * 27] per-cell slot sums of the coarse assembly
    long long cpart_cap = 0;
    int* rxyz = nullptr;         // [rows] packed aggregate coordinates of each row's aggregate
    double* rc = nullptr;        // [nc_cap] restricted residual
    double* cc = nullptr;        // [nc_cap] coarse correction
    int* info_dev = nullptr;
    SortScratch agg_scan{};
    int two_level_ok = 0;        // this solve applies a coarse correction
    int coarse_valid = 0;        // Ac holds the inverse built for the frozen aggregate map
    int coarse_age = 0;          // solves since the last rebuild
    int coarse_rebuilt = 0;      // this solve rebuilt it
    int iters_at_build = 0;      // iterations of the first solve after the last rebuild
    int iters_last = 0;
    // A rebuild for iteration growth that did not cut the iterations (the system got harder, the
    // operator was not stale) turns that trigger off until the next scheduled rebuild
    int growth_off = 0, cell_growth_off = 0;
    int growth_check = 0, cell_growth_check = 0;   // this solve rebuilt for growth: compare after it
    int iters_before_growth = 0, iters_before_cell_growth = 0;
    hipEvent_t setup_ev[2] = {nullptr, nullptr};
    double setup_ms = 0.0;
    // learned warm start and preconditioner weights
    const double* pdiag = nullptr; // diagonal used by the preconditioner (diag, or diag / omega)
    double* pdiag_buf = nullptr;   // [rows]
    double* omega = nullptr;       // [rows]
    double* xprev_s = nullptr;     // [rows] previous pressure in sorted order (network input)
    double* xl = nullptr;          // [rows] learned warm start kept across the calibration probe
    Mpnn* net = nullptr;
    int learned = 0, learned_fallback = 0;
    int calib_age = 0;             // solves until the next plain 5-iteration calibration
    double plain5 = -1.0;          // ||r||/||b|| of the plain solver after 5 iterations
    int probe_at = 0;              // next PCG: check r.r after this many iterations (0 = no check)
    double probe_rr = 0.0;         // ... and return -2 when it exceeds this
    int spmv_launches = 0;
    double spmv_ms = 0.0;
    double spmv_bytes = 0.0;
};

template <int NW> __device__ inline double block_sum(double v, double* lds) {
    v = wave_sum(v);
    const int lane = lane_id(), w = (int)(threadIdx.x >> 6);
    if (lane == 0) lds[w] = v;
    __syncthreads();
    double r = 0;
    if (w == 0) {
        r = lane < NW ? lds[lane] : 0.0;
        r = wave_sum(r);
    }
    __syncthreads();
    return r;
}

// ---------------------------------------------------------------------------------------------
// assembly
// ---------------------------------------------------------------------------------------------
// SELL-C-sigma order: inside every window of SELL_SIGMA rows the rows take positions by stored
// length (descending, ties by row), so the 64 rows of a slice have similar lengths and little
// padding.  One workgroup per window sorts the words ((2^20 - 2 - length) << 12 | row in window)
// with a bitonic network in LDS (rows past N carry the largest word and sort last); the slice
// width is the length of the slice's first (longest) row rounded up to 4.  Positions past the
// window's last row are empty (row -1) and their slices have width 0.
static constexpr int SELL_BLOCK = SELL_SIGMA < 1024 ? SELL_SIGMA : 1024;
__device__ inline void sell_order_body(int blk, int N, int nslices, const int* __restrict__ cnt,
                                       const int* __restrict__ dir, int* pos, int* rowp,
                                       int* slice_w) {
    __shared__ uint32_t v[SELL_SIGMA];
    const int w0 = blk * SELL_SIGMA;
    const int pc = min(SELL_SIGMA, nslices * 64 - w0);
    constexpr uint32_t KTOP = (1u << 20) - 2u;
    for (int i = threadIdx.x; i < SELL_SIGMA; i += SELL_BLOCK) {
        const int t = w0 + i;
        const int k = t < N ? (dir[t] ? 0 : min(cnt[t], (int)KTOP)) : -1;
        v[i] = ((KTOP - (uint32_t)k) << 12) | (uint32_t)i;   // k = -1 gives the largest key
    }
    __syncthreads();
    for (int kk = 2; kk <= SELL_SIGMA; kk <<= 1)
        for (int j = kk >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < SELL_SIGMA; i += SELL_BLOCK) {
                const int ixj = i ^ j;
                if (ixj > i) {
                    const uint32_t a = v[i], b = v[ixj];
                    const bool up = (i & kk) == 0;
                    if ((a > b) == up) {
                        v[i] = b;
                        v[ixj] = a;
                    }
                }
            }
            __syncthreads();
        }
    for (int p = threadIdx.x; p < pc; p += SELL_BLOCK) {
        const uint32_t w = v[p];
        const int t = w0 + (int)(w & 0xfffu);
        const int P = w0 + p;
        if (t < N) {
            const int k = (int)(KTOP - (w >> 12));
            pos[t] = P;
            rowp[P] = t;
            if ((p & 63) == 0) slice_w[P >> 6] = ((k + 3) & ~3) * 64;
        } else {
            rowp[P] = -1;
            if ((p & 63) == 0) slice_w[P >> 6] = 0;
        }
    }
}
__global__ __launch_bounds__(SELL_BLOCK) void k_sell_order(int N, int nslices, const int* __restrict__ cnt,
                                                           const int* __restrict__ dir, int* pos, int* rowp,
                                                           int* slice_w) {
    sell_order_body(blockIdx.x, N, nslices, cnt, dir, pos, rowp, slice_w);
}

// first_ge[K] = the first sorted row whose cell key is >= K (a lower bound per cell, empty cells
// included), K = 0 .. ncell
__device__ inline void first_ge_body(int blk, int N, int ncell, const uint32_t* __restrict__ skeys, int* fg) {
    const int K = blk * blockDim.x + threadIdx.x;
    if (K > ncell) return;
    int lo = 0, hi = N;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if ((int)skeys[mid] < K) lo = mid + 1;
        else hi = mid;
    }
    fg[K] = lo;
}
__global__ __launch_bounds__(PB) void k_first_ge(int N, int ncell, const uint32_t* __restrict__ skeys, int* fg) {
    first_ge_body(blockIdx.x, N, ncell, skeys, fg);
}

// 16-bit column codes.  The cells are numbered x-slowest, z-fastest, so the neighbours of a row in
// cell (i, j, k) that lie in the x-slab i + dx (cells (i + dx, j - 1 .. j + 1, k - 1 .. k + 1))
// fill the sorted range [first_ge[K + dx gy gz - gz - 1], first_ge[K + dx gy gz + gz + 2]), K the
// row's key.  A column is coded as (dx + 1) << 14 | (its offset in that range); the SpMV rebuilds
// it from the row's three range starts.  A column outside its ranges or 16384 or more into one
// (a cell-id wrap at the grid edge, or very dense cells) marks the matrix wide, and the SpMV then
// reads the 32-bit columns.
struct C16Row {
    int b0, b1, b2, e0, e1, e2;
};
__device__ inline bool c16_row(const uint32_t* skeys, const int* fg, int t, int gz, int gygz, int ncell, C16Row& r) {
    const long long key = (long long)skeys[t];
    const long long lo = key - gygz - gz - 1, hi = key + gygz + gz + 2;
    if (lo < 0 || hi > ncell) return false;
    r.b0 = fg[lo];
    r.e0 = fg[key - gygz + gz + 2];
    r.b1 = fg[key - gz - 1];
    r.e1 = fg[key + gz + 2];
    r.b2 = fg[key + gygz - gz - 1];
    r.e2 = fg[hi];
    return true;
}
__device__ inline int c16_code(const C16Row& r, int q, bool& wide) {
    int g, off;
    if (q < r.e0) { g = 0; off = q - r.b0; }
    else if (q < r.e1) { g = 1; off = q - r.b1; }
    else { g = 2; off = q - r.b2; if (q >= r.e2) wide = true; }
    if (off < 0 || off >= (1 << 14)) wide = true;
    return (g << 14) | (off & 0x3fff);
}

// The p values a SELL window's rows read.  Window w holds the sorted rows t0 .. t1 (keys K0 <= K1),
// so every column of its rows in the x-slab i + dx lies in [first_ge[K0 + dx gy gz - gz - 1],
// first_ge[K1 + dx gy gz + gz + 2]) (bounds clamped to the grid), whatever the rows' cells.
// wbox[8 w ..] = the three range starts, the three ends and their total length; *maxw = the
// largest total.
__global__ __launch_bounds__(PB) void k_win_box(int N, int nwin, int ncell, int gz, int gygz,
                                                const uint32_t* __restrict__ skeys, const int* __restrict__ fg,
                                                int* wbox, int* maxw) {
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nwin) return;
    const int t0 = w * SELL_SIGMA, t1 = min(N, t0 + SELL_SIGMA) - 1;
    int* o = wbox + 8 * w;
    int tot = 0;
    for (int g = 0; g < 3; ++g) {
        int b = 0, e = 0;
        if (t0 <= t1) {
            const long long sh = (long long)(g - 1) * gygz;
            const long long lo = min(max((long long)skeys[t0] + sh - gz - 1, 0LL), (long long)ncell);
            const long long hi = min(max((long long)skeys[t1] + sh + gz + 2, 0LL), (long long)ncell);
            b = fg[lo];
            e = max(fg[hi], b);
        }
        o[g] = b;
        o[3 + g] = e;
        tot += e - b;
    }
    o[6] = tot;
    o[7] = 0;
    atomicMax(maxw, tot);
}

template <typename R, typename VA>
__device__ inline void ppe_fill_body(int blk, int N, int nslices, Consts<R> c, const int* __restrict__ active,
                                     const int* __restrict__ dir, const R4<R>* __restrict__ spos,
                                     const int* __restrict__ cs, const int* __restrict__ ce,
                                     const int* __restrict__ slice_off, const int* __restrict__ slice_w,
                                     const int* __restrict__ perm, PState<R> out, double alpha, double n0,
                                     const int* __restrict__ rowp, int* cols, VA* vals, double* rhs,
                                     int* cnt, int* nnz, const int* __restrict__ nbq,
                                     int nb_width, uint16_t* cols16,
                                     const uint32_t* __restrict__ skeys, const int* __restrict__ fg, int wc = 1) {
    const int P = blk * blockDim.x + threadIdx.x;
    const int slice = P >> 6, lane = lane_id();
    if (slice >= nslices) return;
    const int t = rowp[P];   // -1: empty position
    const int base = slice_off[slice], width = slice_w[slice] >> 6;
    int k = 0;
    int c0 = 0, c1 = 0, c2 = 0, c3 = 0;
    int h0 = 0, h1 = 0, h2 = 0, h3 = 0;
    VA v0 = 0, v1 = 0, v2 = 0, v3 = 0;
    C16Row cr{};
    bool wide = false;
    auto push = [&](int col, VA val) {
        c0 = c1; c1 = c2; c2 = c3; c3 = col;
        h0 = h1; h1 = h2; h2 = h3; h3 = c16_code(cr, col, wide);
        v0 = v1; v1 = v2; v2 = v3; v3 = val;
        ++k;
        if ((k & 3) == 0) {
            const int e = sell_at(base, lane, k - 4);
            if (wc) quad_store(cols + e, c0, c1, c2, c3);
            quad_store(cols16 + e, h0, h1, h2, h3);
            quad_store(vals + e, v0, v1, v2, v3);
        }
    };
    if (t >= 0 && !dir[t] && !c16_row(skeys, fg, t, c.gz, c.gy * c.gz, c.ncell, cr)) wide = true;
    if (t >= 0 && !dir[t]) {
        const int bound = cnt[t];   // Laplacian-radius neighbours, from the PND pass
        if (bound <= nb_width) {
            // the PND pass listed them with their weights, in the canonical order
            // (the weights are formed again from the positions, as the PND pass formed them; a
            // group's 4 active flags and positions are loaded together with the next group's
            // entries and waited for once; the last group is padded with a valid index)
            const int lbase = (t >> 6) * 64 * nb_width, ll = t & 63;
            const R4<R> X = spos[t];
            const R rl = c.re_lap, mindl = c.mind_lap;
            int q[4] = {t, t, t, t};
            if (bound > 0) quad_load(nbq + sell_at(lbase, ll, 0), q);
            for (int j = 0; j < bound; j += 4) {
                int act[4], nq[4];
                R4<R> Q[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) act[u] = active[q[u]];
#pragma unroll
                for (int u = 0; u < 4; ++u) Q[u] = spos[q[u]];
                quad_load(nbq + sell_at(lbase, ll, j + 4 < bound ? j + 4 : j), nq);
#pragma unroll
                for (int u = 0; u < 4; ++u) asm volatile("" ::"v"(act[u]), "v"(Q[u].x), "v"(Q[u].y), "v"(Q[u].z));
                const int m = bound - j < 4 ? bound - j : 4;
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    if (u < m && act[u] && k < width) {
                        R d = dist3(X.x, X.y, X.z, Q[u].x, Q[u].y, Q[u].z);
                        if (d < mindl) d = mindl;
                        push(q[u], (VA)(-(double)weight(rl, d)));
                    }
#pragma unroll
                for (int u = 0; u < 4; ++u) q[u] = nq[u];
            }
        } else {
            const R4<R> X = spos[t];
            const R re = c.re_lap, mind = c.mind_lap;
            for_candidates(c, cs, ce, X.x, X.y, X.z, t, [&](int q) {
                const R4<R> Q = spos[q];
                R d = dist3(X.x, X.y, X.z, Q.x, Q.y, Q.z);
                if (d < re) {
                    if (d < mind) d = mind;
                    const R w = weight(re, d);
                    if (active[q] && k < width) push(q, (VA)(-(double)w));
                }
            });
        }
    }
    // the row's entries end at k (its Dirichlet neighbours dropped); readers stop at k rounded up
    // to a group of 4, so only the last group is padded
    const int real = k;
    const int self = t >= 0 ? t : 0;
    while (k & 3) push(self, (VA)0);
    if (t >= 0) {
        cnt[t] = real;
        rhs[t] = dir[t] ? 0.0 : alpha * fmax((double)out.n[perm[t]] - n0, 0.0) / n0;
    }
    const int tot = wave_sum(real);
    const int rows = wave_sum((t >= 0 && !dir[t]) ? 1 : 0);
    const int wides = wave_sum(wide ? 1 : 0);
    if (lane == 0) {
        atomicAdd(&nnz[0], tot);
        atomicAdd(&nnz[1], rows);
        if (wides) atomicAdd(&nnz[2], wides);
    }
    (void)N;
}
template <typename R, typename VA>
__global__ __launch_bounds__(PB) void k_ppe_fill(int N, int nslices, Consts<R> c, const int* __restrict__ active,
                                                  const int* __restrict__ dir, const R4<R>* __restrict__ spos,
                                                  const int* __restrict__ cs, const int* __restrict__ ce,
                                                  const int* __restrict__ slice_off, const int* __restrict__ slice_w,
                                                  const int* __restrict__ perm, PState<R> out, double alpha, double n0,
                                                  const int* __restrict__ rowp, int* cols, VA* vals, double* rhs,
                                                  int* cnt, int* nnz, const int* __restrict__ nbq,
                                                  int nb_width, uint16_t* cols16,
                                                  const uint32_t* __restrict__ skeys, const int* __restrict__ fg, int wc) {
    ppe_fill_body<R, VA>(blockIdx.x, N, nslices, c, active, dir, spos, cs, ce, slice_off, slice_w, perm, out, alpha,
                         n0, rowp, cols, vals, rhs, cnt, nnz, nbq, nb_width, cols16, skeys, fg, wc);
}
// The 32-bit columns of a matrix whose columns all fitted the 16-bit codes, from the codes, for the
// readers that take that form (the coarse and cell operator assembly, the network, the matrix
// getter): the mixed solver's fill leaves them out.  Entries up to each row's count rounded up to
// a group of 4, as the fill writes them.
__global__ __launch_bounds__(PB) void k_decode_cols(int nslices, const int* __restrict__ slice_off,
                                                     const int* __restrict__ rowp, const int* __restrict__ cnt,
                                                     const int* __restrict__ dir, const uint16_t* __restrict__ codes,
                                                     const uint32_t* __restrict__ skeys, const int* __restrict__ fg,
                                                     int gz, int gygz, int ncell, int* cols) {
    const int P = blockIdx.x * blockDim.x + threadIdx.x;
    const int slice = P >> 6, lane = lane_id();
    if (slice >= nslices) return;
    const int t = rowp[P];
    if (t < 0 || dir[t]) return;
    C16Row cr;
    if (!c16_row(skeys, fg, t, gz, gygz, ncell, cr)) return;
    const int base = slice_off[slice];
    const int n4 = (cnt[t] + 3) & ~3;
    for (int k = 0; k < n4; k += 4) {
        const int e = sell_at(base, lane, k);
        int h[4], o[4];
        quad_load(codes + e, h);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int g = h[u] >> 14, off = h[u] & 0x3fff;
            o[u] = (g == 0 ? cr.b0 : g == 1 ? cr.b1 : cr.b2) + off;
        }
        quad_store(cols + e, o[0], o[1], o[2], o[3]);
    }
}
// the 32-bit columns of the current matrix, decoded when its fill left them out
template <typename R> static void ensure_cols(PpeSolver<R>* s, const Dev<R>& d, hipStream_t st) {
    if (s->cols_valid || s->nslices <= 0) return;
    hipLaunchKernelGGL(k_decode_cols, dim3((s->nslices * 64 + PB - 1) / PB), dim3(PB), 0, st, s->nslices,
                       s->slice_off, s->rowp, s->cnt, s->dir, s->cols16, d.skeys, s->first_ge, d.c.gz,
                       d.c.gy * d.c.gz, d.c.ncell, s->cols);
    s->cols_valid = 1;
}

// ---------------------------------------------------------------------------------------------
// solver kernels
// ---------------------------------------------------------------------------------------------
// start of a PCG on A e = r0 (vectors V): r = r0, z = r / diag, p = z, partial r.z and r.r
template <typename V>
__device__ inline void pcg_start_body(int blk, int N, const double* __restrict__ r0, const double* __restrict__ diag,
                                      V* r, V* z, V* p, V* e, int zero_e, double* part) {
    __shared__ double lds[PB / 64];
    const int t = blk * blockDim.x + threadIdx.x;
    double rz = 0.0, rr = 0.0;
    if (t < N) {
        const V rv = (V)r0[t];
        const V zv = rv / (V)diag[t];
        r[t] = rv;
        z[t] = zv;
        p[t] = zv;
        if (zero_e) e[t] = (V)0;
        rz = (double)rv * (double)zv;
        rr = (double)rv * (double)rv;
    }
    rz = block_sum<PB / 64>(rz, lds);
    rr = block_sum<PB / 64>(rr, lds);
    if (threadIdx.x == 0) {
        part[2 * blk] = rz;
        part[2 * blk + 1] = rr;
    }
}
template <typename V>
__global__ __launch_bounds__(PB) void k_pcg_start(int N, const double* __restrict__ r0, const double* __restrict__ diag,
                                                   V* r, V* z, V* p, V* e, int zero_e, double* part) {
    pcg_start_body<V>(blockIdx.x, N, r0, diag, r, z, p, e, zero_e, part);
}

// q = A p (values VA, arithmetic V), partial p.q.  One lane per row of a 64-row slice; the entry
// loop is unrolled by 8 so every lane keeps 8 column/value loads and then 8 gathers of p in
// flight (the loop is latency-bound otherwise).  Matrix streams are loaded non-temporally: they
// are read once per SpMV and would only evict p from L2.  With C16 the columns come from the
// 16-bit codes (2 bytes an entry instead of 4) and the row's three x-slab range starts.  With WIN
// (FP32 vectors, 16-bit codes) a workgroup first copies the p values its window's rows can read
// (the three slab ranges of k_win_box, a few thousand values) into LDS with coalesced loads, and
// the gathers then read LDS: the range starts become LDS offsets, the codes are unchanged.  A
// window larger than wcap gathers from global memory.  Both forms add the same products in the
// same order.
template <bool C16> struct ColSource {
    const int* cols;
    const uint16_t* codes;
    int b0, b1, b2;
    __device__ inline void load4(int e, int* c) const {
        if constexpr (C16) {
            quad_load(codes + e, c);
            // the range start by arithmetic on the slab number (a select between the three members
            // becomes an indexed load, and the object then lives in LDS or scratch)
            const int d1 = b1 - b0, d2 = b2 - b0;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int g = c[u] >> 14;
                c[u] = b0 + (d1 & -(int)(g == 1)) + (d2 & -(int)(g == 2)) + (c[u] & 0x3fff);
            }
        } else {
            quad_load(cols + e, c);
        }
    }
};

template <typename V> struct GatherGlobal {
    const V* __restrict__ p;
    __device__ inline V operator()(int c) const { return p[c]; }
};
template <typename V> struct GatherLds {
    const V* w;
    __device__ inline V operator()(int c) const { return w[c]; }
};

// a += the products of a row's entries [0, width) (width a multiple of 4) in entry order
template <typename VA, typename V, bool C16, typename G>
__device__ inline V spmv_row(V a, int base, int lane, int width, const ColSource<C16>& src,
                             const VA* __restrict__ vals, G gather) {
    int k = 0;
    // FP32 matrix and vectors: 32 and then 16 entries per round (16 or 8 vector loads, then 32
    // or 16 gathers of p in flight); FP32 values with FP64 vectors: rounds of 16; with FP64
    // values the registers of such rounds cut the occupancy, so they keep rounds of 8
    if constexpr (sizeof(VA) == 4 && sizeof(V) == 4) {
        for (; k + 32 <= width; k += 32) {
            int c[32];
            VA v[32];
            const int e = sell_at(base, lane, k);
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                src.load4(e + 256 * u, c + 4 * u);
                quad_load(vals + e + 256 * u, v + 4 * u);
            }
            V g[32];
#pragma unroll
            for (int u = 0; u < 32; ++u) g[u] = gather(c[u]);
#pragma unroll
            for (int u = 0; u < 32; ++u) a += (V)v[u] * g[u];
        }
    }
    // FP32 values (FP32 or FP64 vectors): then rounds of 16
    if constexpr (sizeof(VA) == 4) {
        for (; k + 16 <= width; k += 16) {
            int c[16];
            VA v[16];
            const int e = sell_at(base, lane, k);
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                src.load4(e + 256 * u, c + 4 * u);
                quad_load(vals + e + 256 * u, v + 4 * u);
            }
            V g[16];
#pragma unroll
            for (int u = 0; u < 16; ++u) g[u] = gather(c[u]);
#pragma unroll
            for (int u = 0; u < 16; ++u) a += (V)v[u] * g[u];
        }
    }
    for (; k + 8 <= width; k += 8) {
        int c[8];
        VA v[8];
        const int e = sell_at(base, lane, k);
        src.load4(e, c);
        src.load4(e + 256, c + 4);
        quad_load(vals + e, v);
        quad_load(vals + e + 256, v + 4);
        V g[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) g[u] = gather(c[u]);
#pragma unroll
        for (int u = 0; u < 8; ++u) a += (V)v[u] * g[u];
    }
    if (k < width) {   // one group of 4 left
        int c[4];
        VA v[4];
        const int e = sell_at(base, lane, k);
        src.load4(e, c);
        quad_load(vals + e, v);
        V g[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) g[u] = gather(c[u]);
#pragma unroll
        for (int u = 0; u < 4; ++u) a += (V)v[u] * g[u];
    }
    return a;
}

// Copies the three slab ranges of v that the rows of SELL window blk read (k_win_box)
// into LDS and sets lb[g] = LDS offset minus first index of range g; false (nothing staged) when
// the window holds more than wcap values.  The same for every thread of the workgroup.
template <typename V>
__device__ inline bool stage_window(int blk, const int* __restrict__ wbox, int wcap, const V* __restrict__ v, V* win,
                                    int& lb0, int& lb1, int& lb2) {
    const int* o = wbox + 8 * blk;
    if (o[6] > wcap) return false;
    const int B0 = o[0], B1 = o[1], B2 = o[2], E0 = o[3], E1 = o[4], E2 = o[5];
    const int S1 = E0 - B0, S2 = S1 + (E1 - B1);
    for (int i = threadIdx.x; i < E0 - B0; i += PB) win[i] = v[B0 + i];
    for (int i = threadIdx.x; i < E1 - B1; i += PB) win[S1 + i] = v[B1 + i];
    for (int i = threadIdx.x; i < E2 - B2; i += PB) win[S2 + i] = v[B2 + i];
    lb0 = -B0;
    lb1 = S1 - B1;
    lb2 = S2 - B2;
    __syncthreads();
    return true;
}

template <typename VA, typename V, bool C16, bool WIN>
__device__ inline void pcg_spmv_body(int blk, int nslices, const int* __restrict__ rowp,
                                     const int* __restrict__ cnt, const int* __restrict__ slice_off,
                                     const int* __restrict__ slice_w, const int* __restrict__ cols,
                                     const VA* __restrict__ vals, const double* __restrict__ diag,
                                     const V* __restrict__ p, V* q, const double* __restrict__ scal,
                                     double* part, const uint16_t* __restrict__ codes,
                                     const uint32_t* __restrict__ skeys, const int* __restrict__ fg,
                                     int gz, int gygz, const int* __restrict__ wbox, int wcap) {
    static_assert(!WIN || C16, "the staged window needs the 16-bit codes");
    if (scal[S_DONE] != 0.0) return;
    __shared__ double lds[PB / 64];
    extern __shared__ __align__(16) unsigned char spmv_dyn[];
    V* win = reinterpret_cast<V*>(spmv_dyn);
    // LDS offset minus first p index of each slab range (0 when gathering from global memory)
    int lb0 = 0, lb1 = 0, lb2 = 0;
    bool staged = false;
    if constexpr (WIN) staged = stage_window(blk, wbox, wcap, p, win, lb0, lb1, lb2);
    const int P = blk * blockDim.x + threadIdx.x;   // SELL position
    const int slice = P >> 6, lane = lane_id();
    const int t = slice < nslices ? rowp[P] : -1;
    double pq = 0.0;
    if (t >= 0) {
        const V pt = p[t];
        V a = (V)diag[t] * pt;
        // the row's own entries (rounded up to a group of 4): a lane stops fetching at its row's end,
        // so the slice padding beyond it is not read
        const int base = slice_off[slice], width = min(slice_w[slice] >> 6, (cnt[t] + 3) & ~3);
        ColSource<C16> src{cols, codes, 0, 0, 0};
        if constexpr (C16) {
            if (width > 0) {   // a row with entries lies inside the grid (k_ppe_fill checked its ranges)
                const int key = (int)skeys[t];
                src.b0 = fg[key - gygz - gz - 1] + lb0;
                src.b1 = fg[key - gz - 1] + lb1;
                src.b2 = fg[key + gygz - gz - 1] + lb2;
            }
        }
        if (staged)
            a = spmv_row<VA, V, C16>(a, base, lane, width, src, vals, GatherLds<V>{win});
        else
            a = spmv_row<VA, V, C16>(a, base, lane, width, src, vals, GatherGlobal<V>{p});
        q[t] = a;
        pq = (double)pt * (double)a;
    }
    pq = block_sum<PB / 64>(pq, lds);
    if (threadIdx.x == 0) part[blk] = pq;
}
template <typename VA, typename V, bool C16, bool WIN>
__global__ __launch_bounds__(PB) void k_pcg_spmv(int nslices, const int* __restrict__ rowp,
                                                  const int* __restrict__ cnt, const int* __restrict__ slice_off,
                                                  const int* __restrict__ slice_w, const int* __restrict__ cols,
                                                  const VA* __restrict__ vals, const double* __restrict__ diag,
                                                  const V* __restrict__ p, V* q, const double* __restrict__ scal,
                                                  double* part, const uint16_t* __restrict__ codes,
                                                  const uint32_t* __restrict__ skeys, const int* __restrict__ fg,
                                                  int gz, int gygz, const int* __restrict__ wbox, int wcap) {
    pcg_spmv_body<VA, V, C16, WIN>(blockIdx.x, nslices, rowp, cnt, slice_off, slice_w, cols, vals, diag, p, q, scal,
                                   part, codes, skeys, fg, gz, gygz, wbox, wcap);
}

// r = b - A x with FP64 accumulation (x, b, r FP64; values VA); partial sums of r.r and b.b.
// Columns and the staged window as in k_pcg_spmv (the window holds FP64 values of x); the sum runs
// entry by entry in either form.
template <typename VA, bool C16, bool WIN>
__device__ __attribute__((always_inline)) inline void resid_body(int blk, int N, int nslices,
                                                                  const int* __restrict__ slice_off,
                                  const int* __restrict__ slice_w, const int* __restrict__ cols,
                                  const VA* __restrict__ vals, const double* __restrict__ diag,
                                  const double* __restrict__ x, const double* __restrict__ b,
                                  const int* __restrict__ rowp, const int* __restrict__ cnt, double* r,
                                  double* part, const uint16_t* __restrict__ codes,
                                  const uint32_t* __restrict__ skeys, const int* __restrict__ fg, int gz,
                                  int gygz, const int* __restrict__ wbox, int wcap) {
    static_assert(!WIN || C16, "the staged window needs the 16-bit codes");
    __shared__ double lds[PB / 64];
    extern __shared__ __align__(16) unsigned char resid_dyn[];
    double* win = reinterpret_cast<double*>(resid_dyn);
    int lb0 = 0, lb1 = 0, lb2 = 0;
    bool staged = false;
    if constexpr (WIN) staged = stage_window(blk, wbox, wcap, x, win, lb0, lb1, lb2);
    const int P = blk * blockDim.x + threadIdx.x;   // SELL position
    const int slice = P >> 6, lane = lane_id();
    const int t = slice < nslices ? rowp[P] : -1;
    double rr = 0.0, bb = 0.0;
    if (t >= 0) {
        const int base = slice_off[slice], width = min(slice_w[slice] >> 6, (cnt[t] + 3) & ~3);
        double a = diag[t] * x[t];
        ColSource<C16> src{cols, codes, 0, 0, 0};
        if constexpr (C16) {
            if (width > 0) {
                const int key = (int)skeys[t];
                src.b0 = fg[key - gygz - gz - 1] + lb0;
                src.b1 = fg[key - gz - 1] + lb1;
                src.b2 = fg[key + gygz - gz - 1] + lb2;
            }
        }
        if (staged)
            a = spmv_row<VA, double, C16>(a, base, lane, width, src, vals, GatherLds<double>{win});
        else
            a = spmv_row<VA, double, C16>(a, base, lane, width, src, vals, GatherGlobal<double>{x});
        const double rv = b[t] - a;
        r[t] = rv;
        rr = rv * rv;
        bb = b[t] * b[t];
    }
    rr = block_sum<PB / 64>(rr, lds);
    bb = block_sum<PB / 64>(bb, lds);
    if (threadIdx.x == 0) {
        part[2 * blk] = rr;
        part[2 * blk + 1] = bb;
    }
    (void)N;
}
template <typename VA, bool C16, bool WIN>
__global__ __launch_bounds__(PB) void k_resid(int N, int nslices, const int* __restrict__ slice_off,
                                               const int* __restrict__ slice_w, const int* __restrict__ cols,
                                               const VA* __restrict__ vals, const double* __restrict__ diag,
                                               const double* __restrict__ x, const double* __restrict__ b,
                                               const int* __restrict__ rowp, const int* __restrict__ cnt, double* r,
                                               double* part, const uint16_t* __restrict__ codes,
                                               const uint32_t* __restrict__ skeys, const int* __restrict__ fg, int gz,
                                               int gygz, const int* __restrict__ wbox, int wcap) {
    resid_body<VA, C16, WIN>(blockIdx.x, N, nslices, slice_off, slice_w, cols, vals, diag, x, b, rowp, cnt, r, part,
                             codes, skeys, fg, gz, gygz, wbox, wcap);
}

// The first residual of a solve, launched before the host has read the fill's counts: the 16-bit
// codes and the staged FP64 window when the fill coded every column (nnz[2] == 0, read here on the
// device), else the 32-bit columns.  Either form adds the same products in the same order.
template <typename VA>
__global__ __launch_bounds__(PB) void k_resid_first(int N, int nslices, const int* __restrict__ slice_off,
                                                     const int* __restrict__ slice_w, const int* __restrict__ cols,
                                                     const VA* __restrict__ vals, const double* __restrict__ diag,
                                                     const double* __restrict__ x, const double* __restrict__ b,
                                                     const int* __restrict__ rowp, const int* __restrict__ cnt,
                                                     double* r, double* part, const uint16_t* __restrict__ codes,
                                                     const uint32_t* __restrict__ skeys, const int* __restrict__ fg,
                                                     int gz, int gygz, const int* __restrict__ wbox, int wcap,
                                                     const int* __restrict__ nnz, int have_cols) {
    if (nnz[2] == 0)
        resid_body<VA, true, true>(blockIdx.x, N, nslices, slice_off, slice_w, cols, vals, diag, x, b, rowp, cnt, r,
                                   part, codes, skeys, fg, gz, gygz, wbox, wcap);
    else if (have_cols)   // without them the solve fills the matrix again with its 32-bit columns
        resid_body<VA, false, false>(blockIdx.x, N, nslices, slice_off, slice_w, cols, vals, diag, x, b, rowp, cnt, r,
                                     part, nullptr, nullptr, nullptr, 0, 0, nullptr, 0);
}

// x += alpha p; r -= alpha q; z = r / diag; partial r.z and r.r
template <typename V>
__device__ inline void pcg_update_body(int blk, int N, V* x, V* r, V* z, const V* __restrict__ p,
                                       const V* __restrict__ q, const double* __restrict__ diag,
                                       const double* __restrict__ scal, double* part) {
    if (scal[S_DONE] != 0.0) return;
    __shared__ double lds[PB / 64];
    const int t = blk * blockDim.x + threadIdx.x;
    const V alpha = (V)scal[S_ALPHA];
    double rz = 0.0, rr = 0.0;
    if (t < N) {
        x[t] = x[t] + alpha * p[t];
        const V rv = r[t] - alpha * q[t];
        const V zv = rv / (V)diag[t];
        r[t] = rv;
        z[t] = zv;
        rz = (double)rv * (double)zv;
        rr = (double)rv * (double)rv;
    }
    rz = block_sum<PB / 64>(rz, lds);
    rr = block_sum<PB / 64>(rr, lds);
    if (threadIdx.x == 0) {
        part[2 * blk] = rz;
        part[2 * blk + 1] = rr;
    }
}
template <typename V>
__global__ __launch_bounds__(PB) void k_pcg_update(int N, V* x, V* r, V* z, const V* __restrict__ p,
                                                    const V* __restrict__ q, const double* __restrict__ diag,
                                                    const double* __restrict__ scal, double* part) {
    pcg_update_body<V>(blockIdx.x, N, x, r, z, p, q, diag, scal, part);
}

// p = z + beta p (after the beta reduction; skipped once the solve has converged)
template <typename V>
__device__ inline void pcg_direction_body(int blk, int N, const V* __restrict__ z, V* p,
                                          const double* __restrict__ scal) {
    if (scal[S_DONE] != 0.0) return;
    const int t = blk * blockDim.x + threadIdx.x;
    if (t < N) p[t] = z[t] + (V)scal[S_BETA] * p[t];
}
template <typename V>
__global__ __launch_bounds__(PB) void k_pcg_direction(int N, const V* __restrict__ z, V* p,
                                                       const double* __restrict__ scal) {
    pcg_direction_body<V>(blockIdx.x, N, z, p, scal);
}

enum { RED_START = 0, RED_ALPHA = 1, RED_BETA = 2, RED_RESID = 3 };

// one workgroup: sums the block partials (1 or 2 interleaved sums) and advances the scalars
__device__ inline void pcg_reduce_body(int blk, const double* __restrict__ part, int nb, int nsum, double* scal,
                                       int mode) {
    if (mode == RED_ALPHA || mode == RED_BETA)
        if (scal[S_DONE] != 0.0) return;
    __shared__ double lds[16];
    double s0 = 0.0, s1 = 0.0;
    for (int b = threadIdx.x; b < nb; b += 1024) {
        if (nsum == 2) {
            s0 += part[2 * b];
            s1 += part[2 * b + 1];
        } else {
            s0 += part[b];
        }
    }
    s0 = block_sum<16>(s0, lds);
    s1 = block_sum<16>(s1, lds);
    if (threadIdx.x != 0) return;
    if (mode == RED_START) {
        scal[S_RZ] = s0;
        scal[S_RR] = s1;
        scal[S_RR0] = s1;
        scal[S_BETA] = 0.0;
        scal[S_ITERS] = 0.0;
        scal[S_DONE] = (s1 <= scal[S_TARGET]) ? 1.0 : 0.0;
    } else if (mode == RED_ALPHA) {
        if (!(s0 > 0.0) || !isfinite(s0)) {
            scal[S_DONE] = 3.0;
            return;
        }
        scal[S_PQ] = s0;
        scal[S_ALPHA] = scal[S_RZ] / s0;
    } else if (mode == RED_BETA) {
        scal[S_ITERS] += 1.0;
        scal[S_BETA] = s0 / scal[S_RZ];
        scal[S_RZ] = s0;
        scal[S_RR] = s1;
        if (!isfinite(s1)) scal[S_DONE] = 3.0;
        else if (s1 <= scal[S_TARGET]) scal[S_DONE] = 1.0;
        else if (scal[S_ITERS] >= scal[S_MAXIT]) scal[S_DONE] = 2.0;
    } else {  // RED_RESID: residual norm and rhs norm of an FP64 residual pass
        scal[S_RR] = s0;
        scal[S_BB] = s1;
    }
}
__global__ __launch_bounds__(1024) void k_pcg_reduce(const double* __restrict__ part, int nb, int nsum, double* scal,
                                                      int mode) {
    pcg_reduce_body(blockIdx.x, part, nb, nsum, scal, mode);
}

// previous pressure in sorted order (0 on Dirichlet rows or before the first solve)
__global__ __launch_bounds__(PB) void k_xprev(int N, const int* __restrict__ perm, const int* __restrict__ dir,
                                              const double* __restrict__ prev, int have, double* xs) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < N) xs[t] = (have && !dir[t]) ? prev[perm[t]] : 0.0;
}
__global__ __launch_bounds__(PB) void k_div_diag(int N, const double* __restrict__ diag, const double* __restrict__ om,
                                                 double* out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < N) out[t] = diag[t] / om[t];
}
__global__ __launch_bounds__(PB) void k_copy_d(int N, const double* __restrict__ a, double* b) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < N) b[t] = a[t];
}

__global__ __launch_bounds__(PB) void k_add_f2d(int N, double* x, const float* e) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < N) x[t] = x[t] + (double)e[t];
}

// x0 from the previous solution (slot order) or zero; Dirichlet rows are 0
__device__ inline void warm_body(int blk, int N, const int* __restrict__ perm, const int* __restrict__ dir,
                                 const double* __restrict__ prev, int use_prev, double* x) {
    const int t = blk * blockDim.x + threadIdx.x;
    if (t >= N) return;
    x[t] = (use_prev && !dir[t]) ? prev[perm[t]] : 0.0;
}
__global__ __launch_bounds__(PB) void k_warm(int N, const int* __restrict__ perm, const int* __restrict__ dir,
                                              const double* __restrict__ prev, int use_prev, double* x) {
    warm_body(blockIdx.x, N, perm, dir, prev, use_prev, x);
}

// pressure to the particle state (the EOS kernel's writes, Kern :796-800) and to the sorted
// pressure array; the solution is kept in slot order for the next warm start
template <typename R>
__device__ inline void ppe_writeback_body(int blk, int N, Consts<R> c, const int* __restrict__ perm,
                                          const int* __restrict__ dir, const double* __restrict__ x,
                                          PState<R> in, PState<R> out, const R4<R>* __restrict__ spos,
                                          R* spres, double* prev) {
    const int t = blk * blockDim.x + threadIdx.x;
    if (t >= N) return;
    const int s = perm[t];
    const R4<R> P = spos[t];
    prev[s] = dir[t] ? 0.0 : x[t];
    if (!include(c, P.x, P.y, P.z)) {
        in.p[s] = 0;
        spres[t] = out.p[s];
        return;
    }
    R pr = dir[t] ? R(0) : (R)x[t];
    if (pr < 0) pr = 0;
    out.p[s] = pr;
    in.p[s] = pr;
    spres[t] = pr;
}
template <typename R>
__global__ __launch_bounds__(PB) void k_ppe_writeback(int N, Consts<R> c, const int* __restrict__ perm,
                                                       const int* __restrict__ dir, const double* __restrict__ x,
                                                       PState<R> in, PState<R> out, const R4<R>* __restrict__ spos,
                                                       R* spres, double* prev) {
    ppe_writeback_body<R>(blockIdx.x, N, c, perm, dir, x, in, out, spos, spres, prev);
}

// ---------------------------------------------------------------------------------------------
// two-level preconditioner: z = D^-1 r + P Ac^-1 P^T r with P the indicator of aggregates of
// a x a x a cells (non-Dirichlet rows only) and Ac = P^T A P.  Every aggregate is handled by one
// wave whose lanes take its cells, so restriction and coarse assembly are fixed-order sums.
// ---------------------------------------------------------------------------------------------
struct AggGeom {
    int gx, gy, gz;      // cell grid
    int a;               // aggregate edge in cells
    int cgx, cgy, cgz;   // aggregate grid
};

__device__ inline int agg_full_of_cell(const AggGeom& g, int cell) {
    const int ci = cell / (g.gy * g.gz), cj = (cell / g.gz) % g.gy, ck = cell % g.gz;
    return ((ci / g.a) * g.cgy + cj / g.a) * g.cgz + ck / g.a;
}

__global__ __launch_bounds__(PB) void k_agg_flag(int N, AggGeom g, const uint32_t* __restrict__ skeys,
                                                  const int* __restrict__ dir, int* flag) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N || dir[t]) return;
    flag[agg_full_of_cell(g, (int)skeys[t])] = 1;
}
// rows of aggregates outside the (possibly frozen) map get no coarse correction
__global__ __launch_bounds__(PB) void k_agg_rows(int N, AggGeom g, const uint32_t* __restrict__ skeys,
                                                  const int* __restrict__ dir, const int* __restrict__ flag,
                                                  const int* __restrict__ map, int* agg_of_row) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N) return;
    if (dir[t]) {
        agg_of_row[t] = -1;
        return;
    }
    const int f = agg_full_of_cell(g, (int)skeys[t]);
    agg_of_row[t] = flag[f] ? map[f] : -1;
}
__global__ __launch_bounds__(PB) void k_agg_list(int nfull, const int* __restrict__ flag, const int* __restrict__ map,
                                                  int* full_of) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f < nfull && flag[f]) full_of[map[f]] = f;
}

// Aggregate coordinates of every row's aggregate, packed (ax << 20 | ay << 10 | az), or -1 for
// rows outside the aggregates
__global__ __launch_bounds__(PB) void k_row_aggxyz(int N, AggGeom g, const int* __restrict__ agg_of_row,
                                                    const int* __restrict__ full_of, int* xyz) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N) return;
    const int I = agg_of_row[t];
    int v = -1;
    if (I >= 0) {
        const int f = full_of[I];
        const int ax = f / (g.cgy * g.cgz), ay = (f / g.cgz) % g.cgy, az = f % g.cgz;
        v = (ax << 20) | (ay << 10) | az;
    }
    xyz[t] = v;
}

// Ac = P^T A P, part 1: one wave per (aggregate I, cell c of its a^3 cells), lanes over the cell's
// rows.  A row's entries come in stencil order, so consecutive entries mostly fall in the same
// neighbour slot (dx, dy, dz) in {-1,0,1}^3: each lane sums a run of one slot in a register and
// adds it to its LDS column when the slot changes.  The 27 slot sums of the cell go to
// part[(I * a^3 + c) * 27 + slot].
template <typename VA>
__global__ __launch_bounds__(64) void k_coarse_cells(int nc, AggGeom g, const int* __restrict__ full_of,
                                                      const int* __restrict__ cs, const int* __restrict__ ce,
                                                      const int* __restrict__ dir, const int* __restrict__ cnt,
                                                      const double* __restrict__ diag,
                                                      const int* __restrict__ slice_off, const int* __restrict__ pos,
                                                      const int* __restrict__ cols, const VA* __restrict__ vals,
                                                      const int* __restrict__ xyz, double* part) {
    const int cells = g.a * g.a * g.a;
    const int I = blockIdx.x / cells, c = blockIdx.x % cells;
    if (I >= nc) return;
    const int f = full_of[I];
    const int ax = f / (g.cgy * g.cgz), ay = (f / g.cgz) % g.cgy, az = f % g.cgz;
    __shared__ double acc[27][64];
    const int lane = lane_id();
    for (int q = 0; q < 27; ++q) acc[q][lane] = 0.0;
    const int i = ax * g.a + c / (g.a * g.a), j = ay * g.a + (c / g.a) % g.a, k = az * g.a + c % g.a;
    if (i < g.gx && j < g.gy && k < g.gz) {
        const int cell = (i * g.gy + j) * g.gz + k;
        const int a0 = cs[cell];
        const int a1 = a0 < 0 ? a0 : ce[cell];
        for (int t = a0 + lane; t < a1; t += 64) {
            if (dir[t]) continue;
            acc[13][lane] += diag[t];
            const int P = pos[t], sb = slice_off[P >> 6], n = cnt[t];
            int cur = -1;
            double run = 0.0;
            for (int e4 = 0; e4 < n; ++e4) {
                const int e = sell_at(sb, P & 63, e4);
                const int v = xyz[cols[e]];
                int slot = -1;
                if (v >= 0) {
                    const int dx = (v >> 20) - ax, dy = ((v >> 10) & 1023) - ay, dz = (v & 1023) - az;
                    // entries reach at most one aggregate away; anything else came from a cell-id
                    // wrap at the grid edge and is left out of the preconditioner
                    if (dx >= -1 && dx <= 1 && dy >= -1 && dy <= 1 && dz >= -1 && dz <= 1)
                        slot = (dx + 1) * 9 + (dy + 1) * 3 + (dz + 1);
                }
                if (slot != cur) {
                    if (cur >= 0) acc[cur][lane] += run;
                    cur = slot;
                    run = 0.0;
                }
                if (slot >= 0) run += (double)vals[e];
            }
            if (cur >= 0) acc[cur][lane] += run;
        }
    }
    __syncthreads();
    if (lane < 27) {
        double sum = 0.0;
        for (int l = 0; l < 64; ++l) sum += acc[lane][l];
        part[((size_t)I * cells + c) * 27 + lane] = sum;
    }
}

// Ac, part 2: entry (neighbour of I in slot q, I) = sum over the a^3 cells of I in cell order
__global__ __launch_bounds__(PB) void k_coarse_gather(int nc, AggGeom g, const int* __restrict__ full_of,
                                                       const int* __restrict__ flag, const int* __restrict__ map,
                                                       const double* __restrict__ part, double* Ac) {
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= nc * 27) return;
    const int I = id / 27, q = id % 27;
    const int cells = g.a * g.a * g.a;
    double sum = 0.0;
    for (int c = 0; c < cells; ++c) sum += part[((size_t)I * cells + c) * 27 + q];
    const int f = full_of[I];
    const int ax = f / (g.cgy * g.cgz), ay = (f / g.cgz) % g.cgy, az = f % g.cgz;
    const int bx = ax + q / 9 - 1, by = ay + (q / 3) % 3 - 1, bz = az + q % 3 - 1;
    if (bx >= 0 && by >= 0 && bz >= 0 && bx < g.cgx && by < g.cgy && bz < g.cgz) {
        const int fb = (bx * g.cgy + by) * g.cgz + bz;
        if (flag[fb]) Ac[(size_t)map[fb] * nc + I] = sum;
    }
}

template <typename T> __global__ __launch_bounds__(PB) void k_symmetrize(int nc, T* A) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)nc * nc) return;
    const int j = (int)(idx / nc), i = (int)(idx % nc);   // element (i, j) at A[j * nc + i]
    if (i < j) A[(size_t)j * nc + i] = A[(size_t)i * nc + j];
}

// ---------------------------------------------------------------------------------------------
// Dense inverse of the SPD coarse operator in FP64: blocked in-place Gauss-Jordan without pivoting
// (the sweep of every diagonal block in turn).  For pivot block p with D = A_pp^-1:
//   A_pp <- D;  A_pj <- D A_pj (j != p);  A_ij <- A_ij - A_ip A_pj (i, j != p);  A_ip <- -A_ip D (i != p)
// After the last block A holds A^-1.  A non-positive or non-finite pivot (the matrix is not SPD)
// sets *fail.  Column-major, leading dimension n; blocks of GJ_B rows and columns.
// ---------------------------------------------------------------------------------------------
static constexpr int GJ_B = 64;
static constexpr int GJ_LD = GJ_B + 1;   // LDS row pitch (odd: column reads spread over banks)

// the diagonal block: scalar Gauss-Jordan by one workgroup of 256 threads, each holding a 4 x 4
// piece of the 64 x 64 block in registers (rows 4 (t % 16) .., columns 4 (t / 16) ..).  Step K
// needs the block's column K and row K as they stand after step K - 1: their owners leave them in
// LDS (two buffers, alternating), so one barrier separates the steps.  With p the pivot and inv =
// 1 / p, every element takes the same operations as in a column-by-column sweep:
//   M_Kj <- M_Kj inv, M_ij <- M_ij - M_iK (M_Kj inv) (fma), M_iK <- -M_iK inv, M_KK <- inv.
// A partial last block is padded with the identity, which the sweep leaves as it is.
__global__ __launch_bounds__(256) void k_gj_pivot(int n, int p0, double* A, int* fail) {
    __shared__ double colk[2][GJ_B], rowk[2][GJ_B];
    const int bp = min(GJ_B, n - p0);
    const int r0 = (threadIdx.x % 16) * 4, c0 = (threadIdx.x / 16) * 4;
    double m[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int i = r0 + a, j = c0 + b;
            m[a][b] = (i < bp && j < bp) ? A[(size_t)(p0 + j) * n + p0 + i] : (i == j ? 1.0 : 0.0);
        }
    // column 0 and row 0 for the first step
    if (c0 == 0)
#pragma unroll
        for (int a = 0; a < 4; ++a) colk[0][r0 + a] = m[a][0];
    if (r0 == 0)
#pragma unroll
        for (int b = 0; b < 4; ++b) rowk[0][c0 + b] = m[0][b];
    __syncthreads();
    for (int K = 0; K < GJ_B; ++K) {
        const int buf = K & 1;
        const double piv = rowk[buf][K];
        if (threadIdx.x == 0 && (!(piv > 0.0) || !isfinite(piv))) *fail = 1;
        const double inv = 1.0 / piv;
        double ck[4], rk[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) ck[a] = colk[buf][r0 + a];
#pragma unroll
        for (int b = 0; b < 4; ++b) rk[b] = rowk[buf][c0 + b] * inv;
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const int i = r0 + a, j = c0 + b;
                if (j != K) m[a][b] = i != K ? fma(-ck[a], rk[b], m[a][b]) : rk[b];
                else m[a][b] = i != K ? -m[a][b] * inv : inv;
            }
        if (K + 1 < GJ_B) {
            const int nb = buf ^ 1, k1 = K + 1;
            if (k1 >= c0 && k1 < c0 + 4)
#pragma unroll
                for (int a = 0; a < 4; ++a)
#pragma unroll
                    for (int b = 0; b < 4; ++b)
                        if (c0 + b == k1) colk[nb][r0 + a] = m[a][b];
            if (k1 >= r0 && k1 < r0 + 4)
#pragma unroll
                for (int a = 0; a < 4; ++a)
#pragma unroll
                    for (int b = 0; b < 4; ++b)
                        if (r0 + a == k1) rowk[nb][c0 + b] = m[a][b];
        }
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int i = r0 + a, j = c0 + b;
            if (i < bp && j < bp) A[(size_t)(p0 + j) * n + p0 + i] = m[a][b];
        }
}

// loads block (r0.., c0..) of A (rows x cols, at most GJ_B each) into t[r][c], zero-padded
__device__ inline void gj_load(const double* A, int n, int r0, int c0, int rows, int cols, double* t) {
    for (int e = threadIdx.x; e < GJ_B * GJ_B; e += blockDim.x) {
        const int r = e % GJ_B, c = e / GJ_B;
        t[r * GJ_LD + c] = (r < rows && c < cols) ? A[(size_t)(c0 + c) * n + r0 + r] : 0.0;
    }
}
// each of the 256 threads forms a 4 x 4 piece of the 64 x 64 product X Y (X, Y in LDS, K = GJ_B)
__device__ inline void gj_mul(const double* X, const double* Y, double acc[4][4]) {
    const int tr = (threadIdx.x % 16) * 4, tc = (threadIdx.x / 16) * 4;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
    for (int k = 0; k < GJ_B; ++k) {
        double x[4], y[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) x[a] = X[(tr + a) * GJ_LD + k];
#pragma unroll
        for (int b = 0; b < 4; ++b) y[b] = Y[k * GJ_LD + tc + b];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] = fma(x[a], y[b], acc[a][b]);
    }
}

// row panel: A_pj <- D A_pj for every column block j != p (blockIdx.x = j)
__global__ __launch_bounds__(256) void k_gj_row(int n, int p0, double* A) {
    __shared__ double D[GJ_B * GJ_LD], Y[GJ_B * GJ_LD];
    const int j0 = blockIdx.x * GJ_B;
    if (j0 == p0) return;
    const int bp = min(GJ_B, n - p0), bj = min(GJ_B, n - j0);
    gj_load(A, n, p0, p0, bp, bp, D);
    gj_load(A, n, p0, j0, bp, bj, Y);
    __syncthreads();
    double acc[4][4];
    gj_mul(D, Y, acc);
    const int tr = (threadIdx.x % 16) * 4, tc = (threadIdx.x / 16) * 4;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
            if (tr + a < bp && tc + b < bj) A[(size_t)(j0 + tc + b) * n + p0 + tr + a] = acc[a][b];
}

// trailing update: A_ij <- A_ij - A_ip A_pj for i, j != p, on the FP64 matrix cores
// (v_mfma_f64_16x16x4: a wave multiplies a 16 x 4 piece of A_ip by a 4 x 16 piece of A_pj into a
// 16 x 16 piece of the product; lane l brings A_ip(l % 16, l / 16) and A_pj(l / 16, l % 16) and
// holds rows l / 16 + 4 v, v = 0 .. 3, of column l % 16 of the result).  A workgroup takes 64 rows
// (blockIdx.x, one row block) and 128 columns (blockIdx.y); each of its 4 waves takes 32 of the
// columns, as 4 x 2 pieces of 16 x 16, over K = GJ_B in steps of 4.  The pieces are read straight
// from memory (A is column-major: the A_ip reads of a step are 4 runs of 16 rows, the A_pj reads
// 16 runs of 4).  The columns of the pivot block inside a tile are left alone.
static constexpr int GJ_TN = 128;
typedef double gj_d4 __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void k_gj_update(int n, int p0, double* A) {
    const int i0 = blockIdx.x * GJ_B, j0 = blockIdx.y * GJ_TN;
    if (i0 == p0) return;
    const int bp = min(GJ_B, n - p0);
    if (j0 >= p0 && min(j0 + GJ_TN, n) <= p0 + bp) return;   // the tile is all pivot columns
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lr = lane & 15, lk = lane >> 4;
    const int jw = j0 + wave * 32;
    gj_d4 acc[4][2];
#pragma unroll
    for (int rb = 0; rb < 4; ++rb)
#pragma unroll
        for (int cb = 0; cb < 2; ++cb) acc[rb][cb] = gj_d4{0.0, 0.0, 0.0, 0.0};
    bool rok[4], cok[2];
#pragma unroll
    for (int rb = 0; rb < 4; ++rb) rok[rb] = i0 + rb * 16 + lr < n;
#pragma unroll
    for (int cb = 0; cb < 2; ++cb) cok[cb] = jw + cb * 16 + lr < n;
#pragma unroll 4
    for (int k = 0; k < GJ_B; k += 4) {
        const int kk = k + lk;
        const bool kok = kk < bp;
        double x[4], y[2];
#pragma unroll
        for (int rb = 0; rb < 4; ++rb)
            x[rb] = (kok && rok[rb]) ? A[(size_t)(p0 + kk) * n + i0 + rb * 16 + lr] : 0.0;
#pragma unroll
        for (int cb = 0; cb < 2; ++cb)
            y[cb] = (kok && cok[cb]) ? A[(size_t)(jw + cb * 16 + lr) * n + p0 + kk] : 0.0;
#pragma unroll
        for (int rb = 0; rb < 4; ++rb)
#pragma unroll
            for (int cb = 0; cb < 2; ++cb)
                acc[rb][cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(x[rb], y[cb], acc[rb][cb], 0, 0, 0);
    }
#pragma unroll
    for (int cb = 0; cb < 2; ++cb) {
        const int col = jw + cb * 16 + lr;
        if (col >= n || (col >= p0 && col < p0 + bp)) continue;
#pragma unroll
        for (int rb = 0; rb < 4; ++rb)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = i0 + rb * 16 + lk + 4 * v;
                if (row < n) {
                    double* q = A + (size_t)col * n + row;
                    *q = *q - acc[rb][cb][v];
                }
            }
    }
}

// column panel: A_ip <- -A_ip D for every row block i != p (blockIdx.x = i)
__global__ __launch_bounds__(256) void k_gj_col(int n, int p0, double* A) {
    __shared__ double X[GJ_B * GJ_LD], D[GJ_B * GJ_LD];
    const int i0 = blockIdx.x * GJ_B;
    if (i0 == p0) return;
    const int bp = min(GJ_B, n - p0), bi = min(GJ_B, n - i0);
    gj_load(A, n, i0, p0, bi, bp, X);
    gj_load(A, n, p0, p0, bp, bp, D);
    __syncthreads();
    double acc[4][4];
    gj_mul(X, D, acc);
    const int tr = (threadIdx.x % 16) * 4, tc = (threadIdx.x / 16) * 4;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
            if (tr + a < bi && tc + b < bp) A[(size_t)(p0 + tc + b) * n + i0 + tr + a] = -acc[a][b];
}

// A (n x n, SPD, column-major) <- A^-1 on stream st; *fail (device) is set when a pivot is not
// positive
static void spd_inverse(int n, double* A, int* fail, hipStream_t st) {
    const int nb = (n + GJ_B - 1) / GJ_B;
    for (int p = 0; p < nb; ++p) {
        const int p0 = p * GJ_B;
        hipLaunchKernelGGL(k_gj_pivot, dim3(1), dim3(256), 0, st, n, p0, A, fail);
        hipLaunchKernelGGL(k_gj_row, dim3(nb), dim3(256), 0, st, n, p0, A);
        hipLaunchKernelGGL(k_gj_update, dim3(nb, (n + GJ_TN - 1) / GJ_TN), dim3(256), 0, st, n, p0, A);
        hipLaunchKernelGGL(k_gj_col, dim3(nb), dim3(256), 0, st, n, p0, A);
    }
}

// the inverse of one SPD matrix given in host memory, on a stream of its own (the C entry point
// mpsr_spd_inverse below)
static int spd_inverse_host(int n, const double* a, double* inv) {
    double* d = nullptr;
    int* fail = nullptr;
    hipStream_t st = nullptr;
    int rc = 0;
    const size_t bytes = (size_t)n * n * sizeof(double);
    if (hipMalloc(&d, bytes) != hipSuccess || hipMalloc(&fail, sizeof(int)) != hipSuccess ||
        hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
        rc = 2;
    } else {
        int h = 0;
        if (hipMemcpyAsync(d, a, bytes, hipMemcpyHostToDevice, st) != hipSuccess ||
            hipMemcpyAsync(fail, &h, sizeof(int), hipMemcpyHostToDevice, st) != hipSuccess) {
            rc = 3;
        } else {
            spd_inverse(n, d, fail, st);
            if (hipMemcpyAsync(inv, d, bytes, hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipMemcpyAsync(&h, fail, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipStreamSynchronize(st) != hipSuccess)
                rc = 3;
            else if (h)
                rc = 1;
        }
    }
    if (st) (void)hipStreamDestroy(st);
    if (fail) (void)hipFree(fail);
    if (d) (void)hipFree(d);
    return rc;
}

// FP32 copy of the coarse operator's inverse, the one the mixed solver's V-cycle applies
__global__ __launch_bounds__(PB) void k_d2f(long long n, const double* __restrict__ a, float* b) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}
// FP64 copy of the symmetric FP32 inverse, for the FP64 solver paths

__global__ __launch_bounds__(PB) void k_fill_d(double* a, long long n, double v) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

// Row ranges of the aggregates' cell stacks for this solve's binning.  Cells are ordered z
// fastest, so the a cells (i, j, k0 .. k0+a-1) of aggregate I hold one contiguous range of sorted
// rows: rng[I * a^2 + s] = [first row, end) of stack s (empty when its cells are).
__global__ __launch_bounds__(PB) void k_agg_ranges(int nc, AggGeom g, const int* __restrict__ full_of,
                                                    const int* __restrict__ cs, const int* __restrict__ ce, int2* rng) {
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    const int stacks = g.a * g.a;
    if (id >= nc * stacks) return;
    const int I = id / stacks, st = id % stacks;
    const int f = full_of[I];
    const int ax = f / (g.cgy * g.cgz), ay = (f / g.cgz) % g.cgy, az = f % g.cgz;
    const int i = ax * g.a + st / g.a, j = ay * g.a + st % g.a;
    int lo = -1, hi = 0;
    if (i < g.gx && j < g.gy) {
        for (int k = az * g.a; k < az * g.a + g.a && k < g.gz; ++k) {
            const int cell = (i * g.gy + j) * g.gz + k;
            const int a0 = cs[cell];
            if (a0 < 0) continue;
            if (lo < 0) lo = a0;
            hi = ce[cell];
        }
    }
    if (lo < 0) lo = hi = 0;
    rng[id] = make_int2(lo, hi);
}

// rc[I] = sum of r over the non-Dirichlet rows of aggregate I: one wave per aggregate, 64 / a^2
// lanes per cell stack, each lane with 4 rows in flight
// With upd the residual is the updated one, r - alpha q (the PCG update that k_zcomp then
// writes; the same expression, so both kernels see the same values).
template <typename V>
__global__ __launch_bounds__(64) void k_restrict(int nc, int stacks, const int2* __restrict__ rng,
                                                  const int* __restrict__ dir, const V* __restrict__ r,
                                                  const V* __restrict__ q, int upd, double* rc,
                                                  const double* __restrict__ scal, int check_done) {
    if (check_done && scal[S_DONE] != 0.0) return;
    const int I = blockIdx.x;
    if (I >= nc) return;
    const int lane = lane_id();
    const V alpha = upd ? (V)scal[S_ALPHA] : (V)0;
    auto rv = [&](int t) -> V { return upd ? (V)(r[t] - alpha * q[t]) : r[t]; };
    const int lps = 64 / stacks;
    const int st = lane / lps, sub = lane % lps;
    double acc = 0.0;
    if (st < stacks) {
        const int2 b = rng[(size_t)I * stacks + st];
        int t = b.x + sub;
        for (; t + 3 * lps < b.y; t += 4 * lps) {
            const int t1 = t + lps, t2 = t + 2 * lps, t3 = t + 3 * lps;
            const int d0 = dir[t], d1 = dir[t1], d2 = dir[t2], d3 = dir[t3];
            const V r0 = rv(t), r1 = rv(t1), r2 = rv(t2), r3 = rv(t3);
            if (!d0) acc += (double)r0;
            if (!d1) acc += (double)r1;
            if (!d2) acc += (double)r2;
            if (!d3) acc += (double)r3;
        }
        for (; t < b.y; t += lps)
            if (!dir[t]) acc += (double)rv(t);
    }
    acc = wave_sum(acc);
    if (lane == 0) rc[I] = acc;
}

// ---- cell level of the V-cycle preconditioner -------------------------------------------------
// z = D^-1 r + P1 V(P1^T r): P1 sums the non-Dirichlet rows of each grid cell (the rows of a cell
// are one sorted range), and V is one symmetric V-cycle on Ac1 = P1^T A P1: weighted Jacobi
// (omega / diag Ac1), the dense coarse correction on the aggregates of cells, weighted Jacobi
// again.  Ac1 couples a cell to its 26 neighbours only (the Laplacian radius is one cell), so it
// is kept as a 27-point stencil.  Every step is linear and symmetric, and the pre- and
// post-smoothers are the same, so the preconditioner is SPD.
static constexpr double VC_OMEGA = 0.9;
static constexpr int GROWTH_MIN_AGE = 10;  // the iteration-growth rebuild waits this many solves after a rebuild
static constexpr int PASS_CAP_MIN = 100;    // an inner refinement pass stops after max(this, 4 x its hint) iterations
static constexpr int PASS_CAP_COLD = 500;   // ... or after this many when the pass has no hint yet
static constexpr int CELL_REBUILD = 50;  // Ac1 is rebuilt every this many solves and with the aggregates,
static constexpr int CELL_MIN_AGE = 5;   // ... and from this many solves on when the last solve took more
                                         // than 1.2 x the fewest iterations of any solve since it was built

__device__ inline int stencil_cell(const AggGeom& g, int ci, int cj, int ck, int b) {
    const int x = ci + b / 9 - 1, y = cj + (b / 3) % 3 - 1, z = ck + b % 3 - 1;
    if (x < 0 || y < 0 || z < 0 || x >= g.gx || y >= g.gy || z >= g.gz) return -1;
    return (x * g.gy + y) * g.gz + z;
}

// per row: packed cell coordinates; the cell of a non-Dirichlet row
__global__ __launch_bounds__(PB) void k_row_cell(int N, AggGeom g, const uint32_t* __restrict__ skeys,
                                                  const int* __restrict__ dir, int* cxyz, int* cell_of_row) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N) return;
    const int key = (int)skeys[t];
    const int ci = key / (g.gy * g.gz), cj = (key / g.gz) % g.gy, ck = key % g.gz;
    cxyz[t] = (ci << 20) | (cj << 10) | ck;
    cell_of_row[t] = dir[t] ? -1 : key;
}
// per cell: its compact aggregate in the (possibly frozen) map
__global__ __launch_bounds__(PB) void k_cell_agg_map(int ncell, AggGeom g, const int* __restrict__ flag,
                                                      const int* __restrict__ map, int* c_agg) {
    const int I = blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= ncell) return;
    const int f = agg_full_of_cell(g, I);
    c_agg[I] = flag[f] ? map[f] : -1;
}
// Ac1 row of cell I: one wave, lanes over the cell's rows, each lane summing its rows' entries per
// neighbour slot in its own LDS column; the slots are then summed over the lanes in lane order
template <typename VA>
__global__ __launch_bounds__(64) void k_cell_op(int ncell, AggGeom g, const int* __restrict__ cs,
                                                 const int* __restrict__ ce, const int* __restrict__ dir,
                                                 const int* __restrict__ cnt, const int* __restrict__ pos,
                                                 const int* __restrict__ slice_off, const int* __restrict__ cols,
                                                 const VA* __restrict__ vals, const double* __restrict__ diag,
                                                 const int* __restrict__ cxyz, float* op, float* dinv) {
    __shared__ double lds[27 * 64];
    const int I = blockIdx.x;
    if (I >= ncell) return;
    const int lane = lane_id();
    for (int b = 0; b < 27; ++b) lds[b * 64 + lane] = 0.0;
    const int a0 = cs[I];
    if (a0 >= 0) {
        const int b0 = ce[I];
        const int ci = I / (g.gy * g.gz), cj = (I / g.gz) % g.gy, ck = I % g.gz;
        for (int t = a0 + lane; t < b0; t += 64) {
            if (dir[t]) continue;
            lds[13 * 64 + lane] += diag[t];
            const int P = pos[t];
            const int base = slice_off[P >> 6], l = P & 63, n = cnt[t];
            for (int k = 0; k < n; ++k) {
                const int e = sell_at(base, l, k);
                const int x = cxyz[cols[e]];
                const int dx = (x >> 20) - ci, dy = ((x >> 10) & 1023) - cj, dz = (x & 1023) - ck;
                // a neighbour within the Laplacian radius lies in an adjacent cell
                if (dx < -1 || dx > 1 || dy < -1 || dy > 1 || dz < -1 || dz > 1) continue;
                lds[((dx + 1) * 9 + (dy + 1) * 3 + (dz + 1)) * 64 + lane] += (double)vals[e];
            }
        }
    }
    __syncthreads();
    if (lane < 27) {
        double sum = 0.0;
        for (int l = 0; l < 64; ++l) sum += lds[lane * 64 + l];
        op[(size_t)I * 27 + lane] = (float)sum;
        if (lane == 13) dinv[I] = sum > 0.0 ? (float)(VC_OMEGA / sum) : 0.0f;
    }
}
// k_cell_op sums each row of Ac1 in its own order, so the mirrored entries op[I][b] and
// op[J][26 - b] (J the neighbour of I in slot b) can differ in the last bit.  Both take the mean of
// the two, rounded once, so Ac1 and with it the preconditioner are exactly symmetric.  Every
// mirrored pair is handled by one thread (the one holding its slot below the centre).
__global__ __launch_bounds__(PB) void k_cell_sym(int ncell, AggGeom g, float* op) {
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (long long)ncell * 27) return;
    const int I = (int)(id / 27), b = (int)(id % 27);
    if (b >= 13) return;
    const int ci = I / (g.gy * g.gz), cj = (I / g.gz) % g.gy, ck = I % g.gz;
    const int J = stencil_cell(g, ci, cj, ck, b);
    if (J < 0) return;
    float* a = op + (size_t)I * 27 + b;
    float* c = op + (size_t)J * 27 + (26 - b);
    const float m = (float)(0.5 * ((double)*a + (double)*c));
    *a = m;
    *c = m;
}
// rc1 = P1^T r (r - alpha q with upd, as k_zcomp then writes it) and e = omega rc1 / diag Ac1:
// eight lanes per cell striding over its contiguous rows, summed in a fixed lane order
template <typename V>
__global__ __launch_bounds__(PB) void k_cell_pre(int ncell, const int* __restrict__ cs, const int* __restrict__ ce,
                                                  const int* __restrict__ dir, const V* __restrict__ r,
                                                  const V* __restrict__ q, int upd, const float* __restrict__ dinv,
                                                  double* rc, double* e, const double* __restrict__ scal,
                                                  int check_done) {
    if (check_done && scal[S_DONE] != 0.0) return;
    const int g8 = blockIdx.x * blockDim.x + threadIdx.x;
    const int I = g8 >> 3, sub = g8 & 7;
    const V alpha = upd ? (V)scal[S_ALPHA] : (V)0;
    double acc = 0.0;
    const int a0 = I < ncell ? cs[I] : -1;
    if (a0 >= 0) {
        const int b0 = ce[I];
        for (int t = a0 + sub; t < b0; t += 8)
            if (!dir[t]) acc += (double)(upd ? (V)(r[t] - alpha * q[t]) : r[t]);
    }
    acc += __shfl_xor(acc, 1, 64);
    acc += __shfl_xor(acc, 2, 64);
    acc += __shfl_xor(acc, 4, 64);
    if (I < ncell && sub == 0) {
        rc[I] = acc;
        e[I] = (double)dinv[I] * acc;
    }
}
// rc2[A] = sum over the cells I of aggregate A of (rc1 - Ac1 e)[I]: one wave per aggregate
__global__ __launch_bounds__(64) void k_cell_agg(int nc, AggGeom g, const int* __restrict__ full_of,
                                                  const float* __restrict__ op, const double* __restrict__ rc1,
                                                  const double* __restrict__ e, double* rc2,
                                                  const double* __restrict__ scal, int check_done) {
    if (check_done && scal[S_DONE] != 0.0) return;
    const int A = blockIdx.x;
    if (A >= nc) return;
    const int lane = lane_id();
    const int f = full_of[A];
    const int ax = f / (g.cgy * g.cgz), ay = (f / g.cgz) % g.cgy, az = f % g.cgz;
    const int ncells = g.a * g.a * g.a;
    double acc = 0.0;
    // eight lanes per cell: lane `sub` takes the cell's own residual term (sub 0) and the stencil
    // slots sub, sub + 8, ...; a cell's coordinates are decoded once
    for (int it = lane; it < ncells * 8; it += 64) {
        const int c = it >> 3, sub = it & 7;
        const int ci = ax * g.a + c / (g.a * g.a), cj = ay * g.a + (c / g.a) % g.a, ck = az * g.a + c % g.a;
        if (ci >= g.gx || cj >= g.gy || ck >= g.gz) continue;
        const int I = (ci * g.gy + cj) * g.gz + ck;
        if (sub == 0) acc += rc1[I];
        const float* o = op + (size_t)I * 27;
        for (int b = sub; b < 27; b += 8) {
            const int J = stencil_cell(g, ci, cj, ck, b);
            if (J >= 0) acc -= (double)o[b] * e[J];
        }
    }
    acc = wave_sum(acc);
    if (lane == 0) rc2[A] = acc;
}
// ef = e1 + omega (rc1 - Ac1 e1) / diag Ac1 with e1 = e + the coarse correction of the cell's aggregate:
// eight lanes per cell, lane `sub` over the stencil slots sub, sub + 8, ..., summed in a fixed lane order
__global__ __launch_bounds__(PB) void k_cell_post(int ncell, AggGeom g, const int* __restrict__ c_agg,
                                                   const float* __restrict__ op, const float* __restrict__ dinv,
                                                   const double* __restrict__ rc1, const double* __restrict__ e,
                                                   const double* __restrict__ cc, double* ef,
                                                   const double* __restrict__ scal, int check_done) {
    if (check_done && scal[S_DONE] != 0.0) return;
    const int g8 = blockIdx.x * blockDim.x + threadIdx.x;
    const int I = g8 >> 3, sub = g8 & 7;
    auto e1 = [&](int J) {
        const int A = c_agg[J];
        return e[J] + (A >= 0 ? cc[A] : 0.0);
    };
    const bool live = I < ncell && dinv[I] != 0.0f;
    double acc = 0.0;
    if (live) {
        const int ci = I / (g.gy * g.gz), cj = (I / g.gz) % g.gy, ck = I % g.gz;
        const float* o = op + (size_t)I * 27;
        for (int b = sub; b < 27; b += 8) {
            const int J = stencil_cell(g, ci, cj, ck, b);
            if (J >= 0) acc += (double)o[b] * e1(J);
        }
    }
    acc += __shfl_xor(acc, 1, 64);
    acc += __shfl_xor(acc, 2, 64);
    acc += __shfl_xor(acc, 4, 64);
    if (I < ncell && sub == 0) ef[I] = live ? e1(I) + (double)dinv[I] * (rc1[I] - acc) : 0.0;
}

// cc = Ainv rc (one wave per row; Ainv symmetric so row I = column I, contiguous)
template <typename M>
__global__ __launch_bounds__(64) void k_coarse_mv(int nc, const M* __restrict__ Ainv, const double* __restrict__ rc,
                                                   double* cc, const double* __restrict__ scal, int check_done) {
    if (check_done && scal[S_DONE] != 0.0) return;
    const int I = blockIdx.x;
    if (I >= nc) return;
    const M* row = Ainv + (size_t)I * nc;
    const int l = lane_id();
    double acc = 0.0;
    if constexpr (std::is_same<M, float>::value) {
        // the FP32 row in 16-byte loads: a scalar head up to the first 16-byte boundary (the
        // buffer itself is 256-byte aligned), float4 loads, a scalar tail; four independent sums
        const int h = min((int)((4 - (((size_t)I * nc) & 3)) & 3), nc);
        const int nv = (nc - h) >> 2;
        const float4* row4 = reinterpret_cast<const float4*>(row + h);
        const double* rc4 = rc + h;
        double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
        if (l < h) a0 = (double)row[l] * rc[l];
#pragma unroll 4
        for (int v = l; v < nv; v += 64) {
            const float4 a = row4[v];
            const double* rv = rc4 + 4 * v;
            a0 += (double)a.x * rv[0];
            a1 += (double)a.y * rv[1];
            a2 += (double)a.z * rv[2];
            a3 += (double)a.w * rv[3];
        }
        for (int J = h + 4 * nv + l; J < nc; J += 64) a0 += (double)row[J] * rc[J];
        acc = (a0 + a1) + (a2 + a3);
    } else {
        for (int J = l; J < nc; J += 64) acc += (double)row[J] * rc[J];
    }
    acc = wave_sum(acc);
    if (l == 0) cc[I] = acc;
}

// z = r / diag + cc[agg]; optionally p = z; partial r.z and r.r
// With upd first the PCG update x += alpha p, r -= alpha q (written back).
template <typename V>
__global__ __launch_bounds__(PB) void k_zcomp(int N, V* r, const double* __restrict__ diag,
                                               const int* __restrict__ agg_of_row, const double* __restrict__ cc,
                                               V* z, V* p, int set_p, V* x, const V* __restrict__ q, int upd,
                                               double* part, const double* __restrict__ scal, int check_done) {
    if (check_done && scal[S_DONE] != 0.0) return;
    __shared__ double lds[PB / 64];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    double rz = 0.0, rr = 0.0;
    if (t < N) {
        V rv = r[t];
        if (upd) {
            const V alpha = (V)scal[S_ALPHA];
            x[t] = x[t] + alpha * p[t];
            rv = rv - alpha * q[t];
            r[t] = rv;
        }
        const int I = agg_of_row[t];
        const V zv = (V)((double)(rv / (V)diag[t]) + (I >= 0 ? cc[I] : 0.0));
        z[t] = zv;
        if (set_p) p[t] = zv;
        rz = (double)rv * (double)zv;
        rr = (double)rv * (double)rv;
    }
    rz = block_sum<PB / 64>(rz, lds);
    rr = block_sum<PB / 64>(rr, lds);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = rz;
        part[2 * blockIdx.x + 1] = rr;
    }
}

// r = r0 (V) and, where asked, e = 0
template <typename V>
__global__ __launch_bounds__(PB) void k_pcg_start2(int N, const double* __restrict__ r0, V* r, V* e, int zero_e) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N) return;
    r[t] = (V)r0[t];
    if (zero_e) e[t] = (V)0;
}

// ---------------------------------------------------------------------------------------------
template <typename T> static bool dev_alloc(T** p, size_t n) {
    return hipMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(T)) == hipSuccess;
}

template <typename R> static double ppe_alpha(const mpsr_params& P, double n0_lap, double lambda_lap) {
    return lambda_lap * n0_lap * (double)(R)P.Density / (2.0 * P.Dimension * (double)(R)P.Dt * (double)(R)P.Dt);
}

template <typename R>
PpeSolver<R>* ppe_create(const Dev<R>& d, const mpsr_params& P, const mpsr_ppe_options& opt, double n0_lap,
                         double lambda_lap, double n0_src) {
    auto* s = new PpeSolver<R>();
    s->cap = d.cap;
    s->nslices = (d.cap + 63) / 64;
    s->n0_lap = n0_lap;
    s->lambda_lap = lambda_lap;
    s->alpha = ppe_alpha<R>(P, n0_lap, lambda_lap);
    s->n0 = n0_src;
    const int rows = s->nslices * 64;
    const int nb = (rows + PB - 1) / PB;
    bool ok = true;
    ok &= dev_alloc(&s->active, rows) && dev_alloc(&s->dir, rows) && dev_alloc(&s->cnt, rows);
    ok &= dev_alloc(&s->nbq, (size_t)rows * NB_WIDTH) && dev_alloc(&s->nbn, rows) && dev_alloc(&s->first_ge, (size_t)d.c.ncell + 1);
    ok &= dev_alloc(&s->slice_w, s->nslices + 1) && dev_alloc(&s->slice_off, s->nslices + 1) && dev_alloc(&s->total, 1);
    const int positions = ((d.cap + SELL_SIGMA - 1) / SELL_SIGMA) * SELL_SIGMA;
    ok &= dev_alloc(&s->pos, positions) && dev_alloc(&s->rowp, positions);
    ok &= dev_alloc(&s->diag, rows) && dev_alloc(&s->rhs, rows) && dev_alloc(&s->x, rows) && dev_alloc(&s->prev, rows);
    ok &= dev_alloc(&s->r, rows) && dev_alloc(&s->z, rows) && dev_alloc(&s->p0, rows) && dev_alloc(&s->p1, rows) &&
          dev_alloc(&s->q, rows);
    ok &= dev_alloc(&s->rf, rows) && dev_alloc(&s->zf, rows) && dev_alloc(&s->p0f, rows) && dev_alloc(&s->p1f, rows) &&
          dev_alloc(&s->qf, rows) && dev_alloc(&s->ef, rows);
    ok &= dev_alloc(&s->part, 2 * nb + 2) && dev_alloc(&s->scal, S_COUNT);
    s->scan.scan_tmp_cap = (int)scan_tmp_ints(s->nslices + 1);
    ok &= dev_alloc(&s->scan.scan_tmp, s->scan.scan_tmp_cap);
    ok &= hipHostMalloc((void**)&s->scal_h, sizeof(double) * S_COUNT) == hipSuccess;
    ok &= hipHostMalloc((void**)&s->cnt_h, sizeof(int) * 4) == hipSuccess;
    ok &= dev_alloc(&s->nnz_dev, 4) && dev_alloc(&s->wbox, (size_t)(positions / SELL_SIGMA) * 8);
    ok &= dev_alloc(&s->pdiag_buf, rows) && dev_alloc(&s->omega, rows) && dev_alloc(&s->xprev_s, rows) &&
          dev_alloc(&s->xl, rows);
    if (ok && opt.warm_start == MPSR_WARM_LEARNED) {
        s->net = mpnn_create(rows);
        ok = s->net != nullptr;
    }
    if (!ok) {
        ppe_destroy(s);
        return nullptr;
    }
    (void)hipMemset(s->prev, 0, sizeof(double) * rows);
    (void)hipMemset(s->scal, 0, sizeof(double) * S_COUNT);
    s->mixed = opt.solver == MPSR_PCG_MIXED;
    // two-level preconditioner geometry and buffers
    s->agg_edge = opt.coarse_cells > 0 ? opt.coarse_cells : 2;
    s->cgx = (d.c.gx + s->agg_edge - 1) / s->agg_edge;
    s->cgy = (d.c.gy + s->agg_edge - 1) / s->agg_edge;
    s->cgz = (d.c.gz + s->agg_edge - 1) / s->agg_edge;
    s->nfull = s->cgx * s->cgy * s->cgz;
    ok &= dev_alloc(&s->agg_flag, s->nfull) && dev_alloc(&s->agg_map, s->nfull) && dev_alloc(&s->agg_full, s->nfull);
    ok &= dev_alloc(&s->agg_of_row, rows) && dev_alloc(&s->nc_dev, 1) && dev_alloc(&s->info_dev, 2);
    ok &= dev_alloc(&s->rxyz, rows);
    s->agg_scan.scan_tmp_cap = (int)scan_tmp_ints(s->nfull + 1);
    ok &= dev_alloc(&s->agg_scan.scan_tmp, s->agg_scan.scan_tmp_cap);
    if (!ok) {
        ppe_destroy(s);
        return nullptr;
    }
    return s;
}

static constexpr int NC_MAX = 8192;   // largest coarse system factorised densely

// Builds the aggregates of this solve's non-Dirichlet rows, Ac = P^T A P and its inverse.  When
// the aggregates outnumber what is factorised densely, the aggregate edge grows by one cell at a
// time (up to 8; the flag and map arrays were sized for the configured, smaller edge) and the
// larger edge is kept for later solves.  Returns false (and the solve uses the Jacobi
// preconditioner) when there are no aggregates, still too many, or the factorisation fails.
template <typename R, typename VA>
static bool setup_two_level(PpeSolver<R>* s, const Dev<R>& d, const VA* vals, hipStream_t st) {
    const int N = d.N;
    int nc = 0;
    for (;;) {
        const AggGeom g0{d.c.gx, d.c.gy, d.c.gz, s->agg_edge, s->cgx, s->cgy, s->cgz};
        fill_int(s->agg_flag, 0, s->nfull, st);
        hipLaunchKernelGGL(k_agg_flag, dim3(grid_for(N, PB)), dim3(PB), 0, st, N, g0, d.skeys, s->dir, s->agg_flag);
        exclusive_scan(s->agg_flag, s->agg_map, nullptr, s->nfull, s->nfull, s->nc_dev, s->agg_scan, st);
        if (hipMemcpyAsync(&s->scal_h[S_COUNT - 1], s->nc_dev, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess)
            return false;
        if (hipStreamSynchronize(st) != hipSuccess) return false;
        std::memcpy(&nc, &s->scal_h[S_COUNT - 1], sizeof(int));
        if (nc <= NC_MAX || s->agg_edge >= 8) break;
        s->agg_edge += 1;
        s->cgx = (d.c.gx + s->agg_edge - 1) / s->agg_edge;
        s->cgy = (d.c.gy + s->agg_edge - 1) / s->agg_edge;
        s->cgz = (d.c.gz + s->agg_edge - 1) / s->agg_edge;
        s->nfull = s->cgx * s->cgy * s->cgz;
    }
    const AggGeom g{d.c.gx, d.c.gy, d.c.gz, s->agg_edge, s->cgx, s->cgy, s->cgz};
    hipLaunchKernelGGL(k_agg_list, dim3(grid_for(s->nfull, PB)), dim3(PB), 0, st, s->nfull, s->agg_flag, s->agg_map,
                       s->agg_full);
    s->nc = nc;
    if (nc <= 0 || nc > NC_MAX) return false;
    if (g.cgx >= 1024 || g.cgy >= 1024 || g.cgz >= 1024) return false;   // packed coordinates hold 10 bits
    if (nc > s->nc_cap) {
        if (s->Ac) (void)hipFree(s->Ac);
        if (s->Af) (void)hipFree(s->Af);
        if (s->agg_rng) (void)hipFree(s->agg_rng);
        if (s->rc) (void)hipFree(s->rc);
        if (s->cc) (void)hipFree(s->cc);
        s->Ac = nullptr; s->Af = nullptr; s->agg_rng = nullptr; s->rc = nullptr; s->cc = nullptr;
        s->nc_cap = std::min(NC_MAX, nc + nc / 4 + 64);
        if (!dev_alloc(&s->Ac, (size_t)s->nc_cap * s->nc_cap) || !dev_alloc(&s->Af, (size_t)s->nc_cap * s->nc_cap) ||
            !dev_alloc(&s->agg_rng, (size_t)s->nc_cap * 64) || !dev_alloc(&s->rc, s->nc_cap) ||
            !dev_alloc(&s->cc, s->nc_cap)) {
            s->nc_cap = 0;
            return false;
        }
    }
    hipLaunchKernelGGL(k_agg_rows, dim3(grid_for(N, PB)), dim3(PB), 0, st, N, g, d.skeys, s->dir, s->agg_flag,
                       s->agg_map, s->agg_of_row);
    const long long n2 = (long long)nc * nc;
    hipLaunchKernelGGL(k_fill_d, dim3(grid_for(n2, PB)), dim3(PB), 0, st, s->Ac, n2, 0.0);
    const int cells = g.a * g.a * g.a;
    const long long need = (long long)nc * cells * 27;
    if (need > s->cpart_cap) {
        if (s->cpart) (void)hipFree(s->cpart);
        s->cpart = nullptr;
        s->cpart_cap = need + need / 4;
        if (!dev_alloc(&s->cpart, (size_t)s->cpart_cap)) {
            s->cpart_cap = 0;
            return false;
        }
    }
    ensure_cols(s, d, st);
    hipLaunchKernelGGL(k_row_aggxyz, dim3(grid_for(N, PB)), dim3(PB), 0, st, N, g, s->agg_of_row, s->agg_full, s->rxyz);
    hipLaunchKernelGGL((k_coarse_cells<VA>), dim3((unsigned)(nc * cells)), dim3(64), 0, st, nc, g, s->agg_full,
                       d.cell_start, d.cell_end, s->dir, s->cnt, s->diag, s->slice_off, s->pos, s->cols, vals, s->rxyz,
                       s->cpart);
    hipLaunchKernelGGL(k_coarse_gather, dim3(grid_for((long long)nc * 27, PB)), dim3(PB), 0, st, nc, g, s->agg_full,
                       s->agg_flag, s->agg_map, s->cpart, s->Ac);
    // the dense inverse in FP64 (blocked Gauss-Jordan), made exactly symmetric, and its FP32 copy
    fill_int(s->info_dev, 0, 1, st);
    spd_inverse(nc, s->Ac, s->info_dev, st);
    int info = 0;
    if (hipMemcpyAsync(&info, s->info_dev, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess) return false;
    if (hipStreamSynchronize(st) != hipSuccess) return false;
    if (info != 0) return false;
    hipLaunchKernelGGL(k_symmetrize<double>, dim3(grid_for(n2, PB)), dim3(PB), 0, st, nc, s->Ac);
    hipLaunchKernelGGL(k_d2f, dim3(grid_for(n2, PB)), dim3(PB), 0, st, n2, s->Ac, s->Af);
    return true;
}

// PCG on A e = r0 (V = FP32, e from 0) or A x = b (V = FP64, x = warm start, r0 = b - A x) with
// the two-level preconditioner.  Returns iterations or -1 on breakdown.
// The cell level of the V-cycle for this solve: every row's cell and every cell's aggregate in the
// (possibly frozen) aggregate map; Ac1 = P1^T A P1 from this solve's matrix every CELL_REBUILD
// solves and whenever the aggregates were rebuilt (in between the last Ac1 is applied: a fixed
// SPD operator, so the preconditioner stays SPD).  Returns 1; 0 when
// the cell grid is too large for the packed coordinates (the solve keeps the two-level
// preconditioner); -1 when the cell arrays cannot be allocated.
template <typename R, typename VA>
static int setup_cells(PpeSolver<R>* s, const Dev<R>& d, const VA* vals, hipStream_t st) {
    const int N = d.N;
    const int ncell = d.c.ncell;
    if (d.c.gx >= 1024 || d.c.gy >= 1024 || d.c.gz >= 1024) return 0;   // packed coordinates hold 10 bits
    if (!s->c_op) {
        // per-row arrays cover every slot the handle can hold (N grows with bucket inflow and
        // mpsr_set_particles), as the row buffers of ppe_create do
        const size_t rows_cap = (size_t)((s->cap + 63) / 64) * 64;
        bool ok = dev_alloc(&s->cxyz, rows_cap) && dev_alloc(&s->cell_of_row, rows_cap) &&
                  dev_alloc(&s->c_agg, ncell) && dev_alloc(&s->c_op, (size_t)ncell * 27) && dev_alloc(&s->c_dinv, ncell) &&
                  dev_alloc(&s->c_rc, ncell) && dev_alloc(&s->c_e, ncell) && dev_alloc(&s->c_ef, ncell);
        if (!ok) return -1;
    }
    s->ncell = ncell;
    const AggGeom g{d.c.gx, d.c.gy, d.c.gz, s->agg_edge, s->cgx, s->cgy, s->cgz};
    hipLaunchKernelGGL(k_row_cell, dim3(grid_for(N, PB)), dim3(PB), 0, st, N, g, d.skeys, s->dir, s->cxyz,
                       s->cell_of_row);
    hipLaunchKernelGGL(k_cell_agg_map, dim3(grid_for(ncell, PB)), dim3(PB), 0, st, ncell, g, s->agg_flag, s->agg_map,
                       s->c_agg);
    const bool cell_due = s->cell_age < 0 || s->cell_age + 1 >= CELL_REBUILD || s->coarse_rebuilt;
    const bool cell_grew = !s->cell_growth_off && s->iters_at_cell > 0 && s->cell_age + 1 >= CELL_MIN_AGE &&
                           5 * s->iters_last > 6 * s->iters_at_cell;
    s->cell_rebuilt = cell_due || cell_grew;
    if (cell_due) s->cell_growth_off = 0;
    s->cell_growth_check = !cell_due && cell_grew;
    if (s->cell_growth_check) s->iters_before_cell_growth = s->iters_last;
    if (s->cell_rebuilt) {
        ensure_cols(s, d, st);
        hipLaunchKernelGGL((k_cell_op<VA>), dim3(ncell), dim3(64), 0, st, ncell, g, d.cell_start, d.cell_end, s->dir,
                           s->cnt, s->pos, s->slice_off, s->cols, vals, s->diag, s->cxyz, s->c_op, s->c_dinv);
        hipLaunchKernelGGL(k_cell_sym, dim3(grid_for((long long)ncell * 27, PB)), dim3(PB), 0, st, ncell, g, s->c_op);
        s->cell_age = 0;
    } else {
        s->cell_age += 1;
    }
    return 1;
}

// q = A p and the partial p.q, with the 16-bit column codes when every column of the matrix fitted
// one, else with the 32-bit columns; FP32 vectors with the codes stage each window's p in LDS
template <typename R, typename VA, typename V>
static void launch_spmv(PpeSolver<R>* s, const VA* vals, const V* p, V* q, hipStream_t st) {
    const int nb = (s->nslices * 64 + PB - 1) / PB;
    if constexpr (sizeof(V) == 4) {
        if (s->c16 && s->wcap > 0) {
            hipLaunchKernelGGL((k_pcg_spmv<VA, V, true, true>), dim3(nb), dim3(PB), sizeof(V) * (size_t)s->wcap, st,
                               s->nslices, s->rowp, s->cnt, s->slice_off, s->slice_w, s->cols, vals, s->diag, p, q,
                               s->scal, s->part, s->cols16, s->skeys, s->first_ge, s->gz, s->gygz, s->wbox, s->wcap);
            return;
        }
    }
    if (s->c16)
        hipLaunchKernelGGL((k_pcg_spmv<VA, V, true, false>), dim3(nb), dim3(PB), 0, st, s->nslices, s->rowp, s->cnt,
                           s->slice_off, s->slice_w, s->cols, vals, s->diag, p, q, s->scal, s->part, s->cols16,
                           s->skeys, s->first_ge, s->gz, s->gygz, (const int*)nullptr, 0);
    else
        hipLaunchKernelGGL((k_pcg_spmv<VA, V, false, false>), dim3(nb), dim3(PB), 0, st, s->nslices, s->rowp, s->cnt,
                           s->slice_off, s->slice_w, s->cols, vals, s->diag, p, q, s->scal, s->part,
                           (const uint16_t*)nullptr, (const uint32_t*)nullptr, (const int*)nullptr, 0, 0,
                           (const int*)nullptr, 0);
}

// Iterations queued before the next flag read.  Every launch queued after convergence returns at
// once but still costs its dispatch, so with a hint (the iterations the same pass took in the last
// solve) the first chunk stops one short of it and the rest come in chunks of CHUNK_TAIL.
static int chunk_len(int it, int hint, int probe) {
    if (probe > 0 && it == 0) return std::min(probe, CHUNK);
    if (hint <= 0) return CHUNK;
    if (it == 0) return std::max(CHUNK_TAIL, hint - 1);
    return CHUNK_TAIL;
}

template <typename R, typename VA, typename V>
static int run_pcg2(PpeSolver<R>* s, const Dev<R>& d, int N, const VA* vals, V* x, V* r, V* z, V* p, V* q,
                    const double* r0, double target, int maxit, int zero_x, hipStream_t st, int hint = 0) {
    const int rows = s->nslices * 64;
    const int nb = (rows + PB - 1) / PB;
    const int nc = s->nc;
    const AggGeom g{d.c.gx, d.c.gy, d.c.gz, s->agg_edge, s->cgx, s->cgy, s->cgz};
    s->scal_h[S_TARGET] = target;
    s->scal_h[S_MAXIT] = (double)maxit;
    if (hipMemcpyAsync(s->scal + S_TARGET, s->scal_h + S_TARGET, sizeof(double), hipMemcpyHostToDevice, st) != hipSuccess)
        return -1;
    if (hipMemcpyAsync(s->scal + S_MAXIT, s->scal_h + S_MAXIT, sizeof(double), hipMemcpyHostToDevice, st) != hipSuccess)
        return -1;
    // the FP32 inner solve applies the FP32 copy of the coarse inverse
    const bool f32 = std::is_same<V, float>::value;
    const int stacks = g.a * g.a;
    hipLaunchKernelGGL(k_agg_ranges, dim3(grid_for((long long)nc * stacks, PB)), dim3(PB), 0, st, nc, g, s->agg_full,
                       d.cell_start, d.cell_end, s->agg_rng);
    hipLaunchKernelGGL(k_pcg_start2<V>, dim3(nb), dim3(PB), 0, st, N, r0, r, x, zero_x);
    // the coarse part of the preconditioner on r (upd = 0) or on r - alpha q (upd = 1): the aggregate
    // correction cc, or with the V-cycle the cell correction c_ef
    const bool vc = s->vcycle != 0;
    const int ncell = s->ncell;
    auto coarse = [&](int upd, int check) {
        if (vc) {
            hipLaunchKernelGGL(k_cell_pre<V>, dim3(grid_for((long long)ncell * 8, PB)), dim3(PB), 0, st, ncell,
                               d.cell_start, d.cell_end, s->dir, r, q, upd, s->c_dinv, s->c_rc, s->c_e, s->scal, check);
            hipLaunchKernelGGL(k_cell_agg, dim3(nc), dim3(64), 0, st, nc, g, s->agg_full, s->c_op, s->c_rc, s->c_e,
                               s->rc, s->scal, check);
        } else {
            hipLaunchKernelGGL(k_restrict<V>, dim3(nc), dim3(64), 0, st, nc, stacks, s->agg_rng, s->dir, r, q, upd,
                               s->rc, s->scal, check);
        }
        if (f32) hipLaunchKernelGGL(k_coarse_mv<float>, dim3(nc), dim3(64), 0, st, nc, s->Af, s->rc, s->cc, s->scal, check);
        else hipLaunchKernelGGL(k_coarse_mv<double>, dim3(nc), dim3(64), 0, st, nc, s->Ac, s->rc, s->cc, s->scal, check);
        if (vc)
            hipLaunchKernelGGL(k_cell_post, dim3(grid_for((long long)ncell * 8, PB)), dim3(PB), 0, st, ncell, g, s->c_agg, s->c_op,
                               s->c_dinv, s->c_rc, s->c_e, s->cc, s->c_ef, s->scal, check);
    };
    const int* zrow = vc ? s->cell_of_row : s->agg_of_row;
    const double* zcor = vc ? s->c_ef : s->cc;
    coarse(0, 0);
    hipLaunchKernelGGL(k_zcomp<V>, dim3(nb), dim3(PB), 0, st, N, r, s->pdiag, zrow, zcor, z, p, 1, x, q, 0, s->part,
                       s->scal, 0);
    hipLaunchKernelGGL(k_pcg_reduce, dim3(1), dim3(1024), 0, st, s->part, nb, 2, s->scal, (int)RED_START);
    int done_iters = 0;
    const int probe = s->probe_at;
    s->probe_at = 0;
    for (int it = 0; it < maxit;) {
        const int len = chunk_len(it, hint, probe);
        const size_t chunk_ev0 = s->ev_used;
        const bool timed = s->profile && (s->profile == 1 || (s->chunk_tick++ & 3) == 0) &&
                           s->ev_used + 2 * (size_t)len <= s->ev.size();
        for (int k = 0; k < len; ++k) {
            if (timed) (void)hipEventRecord(s->ev[s->ev_used], st);
            launch_spmv<R, VA, V>(s, vals, p, q, st);
            if (timed) {
                (void)hipEventRecord(s->ev[s->ev_used + 1], st);
                s->ev_used += 2;
            }
            hipLaunchKernelGGL(k_pcg_reduce, dim3(1), dim3(1024), 0, st, s->part, nb, 1, s->scal, (int)RED_ALPHA);
            coarse(1, 1);
            hipLaunchKernelGGL(k_zcomp<V>, dim3(nb), dim3(PB), 0, st, N, r, s->pdiag, zrow, zcor, z, p, 0, x, q, 1,
                               s->part, s->scal, 1);
            hipLaunchKernelGGL(k_pcg_reduce, dim3(1), dim3(1024), 0, st, s->part, nb, 2, s->scal, (int)RED_BETA);
            hipLaunchKernelGGL(k_pcg_direction<V>, dim3(nb), dim3(PB), 0, st, N, z, p, s->scal);
        }
        if (read_scal(s, st) != hipSuccess) {
            std::snprintf(s->why, sizeof(s->why), "device error while reading the PCG scalars");
            return -1;
        }
        const int live = std::max(0, std::min(len, (int)s->scal_h[S_ITERS] - done_iters));
        if (timed) s->ev_used = chunk_ev0 + 2 * (size_t)live;
        done_iters = (int)s->scal_h[S_ITERS];
        it += len;
        if (s->scal_h[S_DONE] != 0.0) break;
        if (probe > 0 && it == probe && s->scal_h[S_RR] > s->probe_rr) return -2;
    }
    if (s->scal_h[S_DONE] == 3.0) {
        std::snprintf(s->why, sizeof(s->why),
                      "two-level PCG breakdown after %d iterations (p.Ap = %.3e, r.z = %.3e, r.r = %.3e)", done_iters,
                      s->scal_h[S_PQ], s->scal_h[S_RZ], s->scal_h[S_RR]);
        return -1;
    }
    return done_iters;
}

template <typename R> void ppe_destroy(PpeSolver<R>* s) {
    if (!s) return;
    void* ps[] = {s->active, s->dir, s->cnt, s->slice_w, s->slice_off, s->pos, s->rowp, s->total, s->cols, s->vals64, s->vals32,
                  s->cols16, s->first_ge, s->wbox,
                  s->diag, s->rhs, s->x, s->prev, s->r, s->z, s->p0, s->p1, s->q, s->rf, s->zf, s->p0f,
                  s->p1f, s->qf, s->ef, s->part, s->scal, s->scan.scan_tmp, s->nnz_dev, s->agg_flag,
                  s->agg_map, s->agg_full, s->agg_of_row, s->nc_dev, s->info_dev, s->Ac, s->Af, s->agg_rng, s->cpart, s->rxyz, s->rc, s->cc,
                  s->agg_scan.scan_tmp, s->pdiag_buf, s->omega, s->xprev_s, s->xl, s->nbq, s->nbn, s->cxyz,
                  s->cell_of_row, s->c_agg, s->c_op, s->c_dinv, s->c_rc, s->c_e, s->c_ef};
    if (s->net) mpnn_destroy(s->net);
    for (void* p : ps)
        if (p) (void)hipFree(p);
    if (s->scal_h) (void)hipHostFree(s->scal_h);
    if (s->cnt_h) (void)hipHostFree(s->cnt_h);
    for (hipEvent_t e : s->ev) (void)hipEventDestroy(e);
    for (hipEvent_t e : s->setup_ev)
        if (e) (void)hipEventDestroy(e);
    delete s;
}

template <typename R> void ppe_update_params(PpeSolver<R>* s, const mpsr_params& P) {
    s->alpha = ppe_alpha<R>(P, s->n0_lap, s->lambda_lap);
}

// read the device scalars (synchronises the stream)
template <typename R> static hipError_t read_scal(PpeSolver<R>* s, hipStream_t st) {
    hipError_t e = hipMemcpyAsync(s->scal_h, s->scal, sizeof(double) * S_COUNT, hipMemcpyDeviceToHost, st);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(st);
}

// Runs one PCG on A e = r0 from e = 0 (V = FP32) or on A x = b from the x in s->x (V = FP64)
// until r.r <= target.  Returns iterations, or -1 on breakdown.
template <typename R, typename VA, typename V>
static int run_pcg(PpeSolver<R>* s, int N, const VA* vals, V* x, V* r, V* z, V* pa, V* pb, V* q,
                   const double* r0, double target, int maxit, hipStream_t st, int hint = 0) {
    const int rows = s->nslices * 64;
    const int nb = (rows + PB - 1) / PB;
    s->scal_h[S_TARGET] = target;
    s->scal_h[S_MAXIT] = (double)maxit;
    if (hipMemcpyAsync(s->scal + S_TARGET, s->scal_h + S_TARGET, sizeof(double), hipMemcpyHostToDevice, st) != hipSuccess)
        return -1;
    if (hipMemcpyAsync(s->scal + S_MAXIT, s->scal_h + S_MAXIT, sizeof(double), hipMemcpyHostToDevice, st) != hipSuccess)
        return -1;
    // the PCG start zeroes x only on the refinement path (x is the correction e there)
    const int zero_x = std::is_same<V, float>::value ? 1 : 0;
    hipLaunchKernelGGL(k_pcg_start<V>, dim3(nb), dim3(PB), 0, st, N, r0, s->pdiag, r, z, pa, x, zero_x, s->part);
    hipLaunchKernelGGL(k_pcg_reduce, dim3(1), dim3(1024), 0, st, s->part, nb, 2, s->scal, (int)RED_START);
    int done_iters = 0;
    const int probe = s->probe_at;
    s->probe_at = 0;
    for (int it = 0; it < maxit;) {
        // SpMV launches are bracketed by events while the pool lasts; after the chunk, the
        // events of launches that found the solve already converged (and returned at once) are
        // dropped, so the mean covers working launches only.  With a probe, the first chunk is
        // `probe` iterations long and r.r is checked after it.
        const int len = chunk_len(it, hint, probe);
        const size_t chunk_ev0 = s->ev_used;
        const bool timed = s->profile && (s->profile == 1 || (s->chunk_tick++ & 3) == 0) &&
                           s->ev_used + 2 * (size_t)len <= s->ev.size();
        for (int k = 0; k < len; ++k) {
            if (timed) (void)hipEventRecord(s->ev[s->ev_used], st);
            launch_spmv<R, VA, V>(s, vals, pa, q, st);
            if (timed) {
                (void)hipEventRecord(s->ev[s->ev_used + 1], st);
                s->ev_used += 2;
            }
            hipLaunchKernelGGL(k_pcg_reduce, dim3(1), dim3(1024), 0, st, s->part, nb, 1, s->scal, (int)RED_ALPHA);
            hipLaunchKernelGGL(k_pcg_update<V>, dim3(nb), dim3(PB), 0, st, N, x, r, z, pa, q, s->pdiag, s->scal,
                               s->part);
            hipLaunchKernelGGL(k_pcg_reduce, dim3(1), dim3(1024), 0, st, s->part, nb, 2, s->scal, (int)RED_BETA);
            hipLaunchKernelGGL(k_pcg_direction<V>, dim3(nb), dim3(PB), 0, st, N, z, pa, s->scal);
        }
        if (read_scal(s, st) != hipSuccess) {
            std::snprintf(s->why, sizeof(s->why), "device error while reading the PCG scalars");
            return -1;
        }
        const int live = std::max(0, std::min(len, (int)s->scal_h[S_ITERS] - done_iters));
        if (timed) s->ev_used = chunk_ev0 + 2 * (size_t)live;
        done_iters = (int)s->scal_h[S_ITERS];
        it += len;
        if (s->scal_h[S_DONE] != 0.0) break;
        if (probe > 0 && it == probe && s->scal_h[S_RR] > s->probe_rr) return -2;
    }
    if (s->scal_h[S_DONE] == 3.0) {
        std::snprintf(s->why, sizeof(s->why),
                      "PCG breakdown after %d iterations (p.Ap = %.3e, r.z = %.3e, r.r = %.3e, target %.3e)",
                      done_iters, s->scal_h[S_PQ], s->scal_h[S_RZ], s->scal_h[S_RR], s->scal_h[S_TARGET]);
        return -1;
    }
    return done_iters;
}

// r = b - A x: the first residual of a solve decides between the 16-bit codes and the 32-bit
// columns on the device (the FP32 matrix) or takes the 32-bit columns; later ones use the codes and,
// for the FP32 matrix, the staged window when the solve found every column coded
template <typename R, typename VA>
static void launch_resid(PpeSolver<R>* s, int N, int nb, const VA* vals, hipStream_t st, bool first) {
    if (first && s->mixed && sizeof(VA) == 4 && !c16_off() && !win_off()) {
        hipLaunchKernelGGL(k_resid_first<VA>, dim3(nb), dim3(PB), (size_t)RESID_WIN_BYTES, st, N, s->nslices,
                           s->slice_off, s->slice_w, s->cols, vals, s->diag, s->x, s->rhs, s->rowp, s->cnt, s->r,
                           s->part, s->cols16, s->skeys, s->first_ge, s->gz, s->gygz, s->wbox,
                           RESID_WIN_BYTES / (int)sizeof(double), s->nnz_dev, s->cols_valid);
        return;
    }
    const int wcap = s->wcap > 0 ? std::min(s->wcap, RESID_WIN_BYTES / (int)sizeof(double)) : 0;
    if (s->c16 && wcap > 0 && sizeof(VA) == 4)
        hipLaunchKernelGGL((k_resid<VA, true, true>), dim3(nb), dim3(PB), sizeof(double) * (size_t)wcap, st, N,
                           s->nslices, s->slice_off, s->slice_w, s->cols, vals, s->diag, s->x, s->rhs, s->rowp, s->cnt,
                           s->r, s->part, s->cols16, s->skeys, s->first_ge, s->gz, s->gygz, s->wbox, wcap);
    else if (s->c16)
        hipLaunchKernelGGL((k_resid<VA, true, false>), dim3(nb), dim3(PB), 0, st, N, s->nslices, s->slice_off,
                           s->slice_w, s->cols, vals, s->diag, s->x, s->rhs, s->rowp, s->cnt, s->r, s->part, s->cols16,
                           s->skeys, s->first_ge, s->gz, s->gygz, (const int*)nullptr, 0);
    else
        hipLaunchKernelGGL((k_resid<VA, false, false>), dim3(nb), dim3(PB), 0, st, N, s->nslices, s->slice_off,
                           s->slice_w, s->cols, vals, s->diag, s->x, s->rhs, s->rowp, s->cnt, s->r, s->part,
                           (const uint16_t*)nullptr, (const uint32_t*)nullptr, (const int*)nullptr, 0, 0,
                           (const int*)nullptr, 0);
}

template <typename R>
static double resid_pass(PpeSolver<R>* s, int N, hipStream_t st, double* bb_out, bool first = false) {
    const int rows = s->nslices * 64;
    const int nb = (rows + PB - 1) / PB;
    if (s->mixed)
        launch_resid(s, N, nb, s->vals32, st, first);
    else
        launch_resid(s, N, nb, s->vals64, st, first);
    hipLaunchKernelGGL(k_pcg_reduce, dim3(1), dim3(1024), 0, st, s->part, nb, 2, s->scal, (int)RED_RESID);
    if (read_scal(s, st) != hipSuccess) return -1.0;
    *bb_out = s->scal_h[S_BB];
    return s->scal_h[S_RR];
}

// the assembled system and the particle geometry as the network sees them
template <typename R> static MpnnSys mpnn_sys(const PpeSolver<R>* s, const Dev<R>& d) {
    MpnnSys y;
    y.N = d.N;
    y.nslices = s->nslices;
    y.rows_active = s->rows_real;
    y.slice_off = s->slice_off;
    y.pos = s->pos;
    y.slice_w = s->slice_w;
    y.cols = s->cols;
    y.cnt = s->cnt;
    y.dir = s->dir;
    y.v32 = s->mixed ? s->vals32 : nullptr;
    y.v64 = s->mixed ? nullptr : s->vals64;
    y.diag = s->diag;
    y.rhs = s->rhs;
    y.xprev = s->xprev_s;
    y.xsol = s->x;
    return y;
}
template <typename R> static MpnnGeo<R> mpnn_geo(const Dev<R>& d) {
    MpnnGeo<R> g;
    g.spos = d.spos;
    g.svel = d.svel;
    g.perm = d.perm;
    g.stl_mag = d.stl_mag;
    g.l0 = (double)d.c.l0;
    g.dt = (double)d.c.dt;
    g.re = (double)d.c.re_lap;
    return g;
}

template <typename R> void ppe_start(PpeSolver<R>* s, const Dev<R>& d, hipStream_t st) {
    launch_pnd_ppe(d, s->active, s->cnt, s->diag, s->dir, s->nbq, s->nbn, NB_WIDTH, st);
}
template <typename R> NbList ppe_neighbors(const PpeSolver<R>* s) {
    NbList nb;
    nb.q = s->nbq;
    nb.n = s->nbn;
    nb.width = NB_WIDTH;
    return nb;
}

template <typename R>
mpsr_status ppe_solve(PpeSolver<R>* s, Dev<R>& d, const mpsr_ppe_options& opt, int profile, hipStream_t st) {
    const int N = d.N;
    s->profile = profile;
    if (profile && s->ev.empty()) {
        s->ev.resize(2 * 4096);
        for (auto& e : s->ev)
            if (hipEventCreate(&e) != hipSuccess) return MPSR_GPUERR;
        for (auto& e : s->setup_ev)
            if (hipEventCreate(&e) != hipSuccess) return MPSR_GPUERR;
    }
    const int nsl = (N + 63) / 64;
    const int rows = nsl * 64;
    const int nb = (rows + PB - 1) / PB;
    s->mixed = opt.solver == MPSR_PCG_MIXED;
    s->status = 0;
    s->why[0] = 0;
    // the column form and the window are decided anew for this matrix after its first residual
    s->c16 = 0;
    s->wcap = 0;
    // ---- assembly ----
    // active flags, row bounds, diagonal and Dirichlet flags come from the PND pass (ppe_start)
    fill_int(s->nnz_dev, 0, 4, st);
    s->skeys = d.skeys;
    s->gz = d.c.gz;
    s->gygz = d.c.gy * d.c.gz;
    if (N > 0)
        hipLaunchKernelGGL(k_first_ge, dim3(grid_for((long long)d.c.ncell + 1, PB)), dim3(PB), 0, st, N, d.c.ncell,
                           d.skeys, s->first_ge);
    if (N > 0 && s->mixed)
        hipLaunchKernelGGL(k_win_box, dim3(grid_for(nb, PB)), dim3(PB), 0, st, N, nb, d.c.ncell, s->gz, s->gygz,
                           d.skeys, s->first_ge, s->wbox, s->nnz_dev + 3);
    if (N > 0)
        hipLaunchKernelGGL(k_sell_order, dim3(grid_for(N, SELL_SIGMA)), dim3(SELL_BLOCK), 0, st, N, nsl, s->cnt, s->dir,
                           s->pos, s->rowp, s->slice_w);
    exclusive_scan(s->slice_w, s->slice_off, nullptr, nsl, nsl, s->total, s->scan, st);
    int total = 0;
    if (hipMemcpyAsync(&s->scal_h[0], s->total, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess)
        return MPSR_GPUERR;
    if (hipStreamSynchronize(st) != hipSuccess) return MPSR_GPUERR;
    std::memcpy(&total, &s->scal_h[0], sizeof(int));
    s->nnz_stored = total;
    if ((long long)total > s->nnz_cap) {
        if (s->cols) (void)hipFree(s->cols);
        if (s->cols16) (void)hipFree(s->cols16);
        if (s->vals64) (void)hipFree(s->vals64);
        if (s->vals32) (void)hipFree(s->vals32);
        s->cols = nullptr; s->cols16 = nullptr; s->vals64 = nullptr; s->vals32 = nullptr;
        s->nnz_cap = (long long)(total * 1.25) + 4096;
        if (!dev_alloc(&s->cols, (size_t)s->nnz_cap)) return MPSR_MEMERR;
        if (!dev_alloc(&s->cols16, (size_t)s->nnz_cap)) return MPSR_MEMERR;
        if (!dev_alloc(&s->vals64, (size_t)s->nnz_cap)) return MPSR_MEMERR;
        if (!dev_alloc(&s->vals32, (size_t)s->nnz_cap)) return MPSR_MEMERR;
    }
    // the source's share of the density error corrected per step, 0.1 unless set (a double product,
    // evaluated alike by the oracle)
    const double alpha_src = s->alpha * (opt.source_relaxation > 0 ? opt.source_relaxation : 0.1);
    // the mixed solver's SpMV and residuals read the 16-bit codes, so its fill leaves the 32-bit
    // columns out (a tenth of the fill's writes at 1M) unless a matrix of this solver had columns
    // beyond the codes, or the codes or the window are switched off; the readers of the 32-bit
    // form decode them first (ensure_cols)
    const int wc = (!s->mixed || s->cols_wide || c16_off() || win_off()) ? 1 : 0;
    if (s->mixed)
        hipLaunchKernelGGL((k_ppe_fill<R, float>), dim3(nb), dim3(PB), 0, st, N, nsl, d.c, s->active, s->dir, d.spos,
                           d.cell_start, d.cell_end, s->slice_off, s->slice_w, d.perm, d.out, alpha_src, s->n0,
                           s->rowp, s->cols, s->vals32, s->rhs, s->cnt, s->nnz_dev, s->nbq, NB_WIDTH,
                           s->cols16, d.skeys, s->first_ge, wc);
    else
        hipLaunchKernelGGL((k_ppe_fill<R, double>), dim3(nb), dim3(PB), 0, st, N, nsl, d.c, s->active, s->dir, d.spos,
                           d.cell_start, d.cell_end, s->slice_off, s->slice_w, d.perm, d.out, alpha_src, s->n0,
                           s->rowp, s->cols, s->vals64, s->rhs, s->cnt, s->nnz_dev, s->nbq, NB_WIDTH,
                           s->cols16, d.skeys, s->first_ge, 1);
    s->cols_valid = wc;
    // entry and row counts of the filled matrix, read after the first residual pass synchronises
    if (hipMemcpyAsync(s->cnt_h, s->nnz_dev, 4 * sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess)
        return MPSR_GPUERR;
    s->nslices = nsl;
    // ---- solve ----
    const double tol = opt.rel_tol > 0 ? opt.rel_tol : 1e-8;
    const int maxit = opt.max_iter > 0 ? opt.max_iter : 20000;
    hipLaunchKernelGGL(k_xprev, dim3(grid_for(N, PB)), dim3(PB), 0, st, N, d.perm, s->dir, s->prev, s->have_prev,
                       s->xprev_s);
    s->pdiag = s->diag;
    s->learned = opt.warm_start == MPSR_WARM_LEARNED && s->net && mpnn_has_weights(s->net) && N > 0;
    s->learned_fallback = 0;
    // the plain warm start: the previous pressure (learned and previous modes) or zero
    const int use_prev = (opt.warm_start != MPSR_WARM_ZERO && s->have_prev) ? 1 : 0;
    if (s->learned) {
        ensure_cols(s, d, st);
        const MpnnSys sys = mpnn_sys(s, d);
        mpnn_features<R>(s->net, sys, mpnn_geo(d), st);
        mpnn_forward(s->net, sys, st);
        mpnn_apply(s->net, sys, s->x, opt.learned_omega ? s->omega : nullptr, st);
        if (opt.learned_omega) {
            hipLaunchKernelGGL(k_div_diag, dim3(grid_for(N, PB)), dim3(PB), 0, st, N, s->diag, s->omega,
                               s->pdiag_buf);
            s->pdiag = s->pdiag_buf;
        }
    } else {
        hipLaunchKernelGGL(k_warm, dim3(grid_for(N, PB)), dim3(PB), 0, st, N, d.perm, s->dir, s->prev, use_prev, s->x);
    }
    double bb = 0.0;
    double rr = resid_pass(s, N, st, &bb, true);
    if (rr < 0) return MPSR_GPUERR;
    if (s->cnt_h[2] != 0 && !s->cols_valid) {
        // columns beyond the codes and no 32-bit columns: the first residual pass skipped its work;
        // this solver fills them from now on, and the solve starts again from the PND pass
        s->cols_wide = 1;
        ppe_start(s, d, st);
        return ppe_solve(s, d, opt, profile, st);
    }
    s->nnz_real = s->cnt_h[0];
    s->rows_real = s->cnt_h[1];
    s->c16 = (s->cnt_h[2] == 0 && !c16_off()) ? 1 : 0;
    // LDS per SpMV workgroup: the largest window, at most WIN_BYTES (larger windows gather from
    // global memory)
    s->wcap = (s->mixed && s->c16 && s->cnt_h[3] > 0 && !win_off())
                  ? std::min((s->cnt_h[3] + 63) / 64 * 64, WIN_BYTES / (int)sizeof(float)) : 0;
    s->relres0 = bb > 0 ? std::sqrt(rr / bb) : 0.0;
    const double target = tol * tol * bb;
    s->iters = 0;
    s->passes = 0;
    s->fallbacks = 0;
    s->two_level_ok = 0;
    s->coarse_rebuilt = 0;
    s->setup_ms = 0.0;
    const bool coarse_pre = opt.precond == MPSR_PRECOND_TWO_LEVEL || opt.precond == MPSR_PRECOND_CELL_VCYCLE;
    if (coarse_pre && bb > 0 && rr > target) {
        const int every = opt.coarse_rebuild > 0 ? opt.coarse_rebuild : 200;
        const bool due = !s->coarse_valid || s->coarse_age + 1 >= every;
        const bool grew = !s->growth_off && s->iters_at_build > 0 && s->coarse_age + 1 >= GROWTH_MIN_AGE &&
                          4 * s->iters_last > 5 * s->iters_at_build;
        const bool rebuild = due || grew;
        if (due) s->growth_off = 0;
        s->growth_check = !due && grew;
        if (s->growth_check) s->iters_before_growth = s->iters_last;
        if (rebuild) {
            if (profile) (void)hipEventRecord(s->setup_ev[0], st);
            s->two_level_ok = s->mixed ? setup_two_level<R, float>(s, d, s->vals32, st)
                                       : setup_two_level<R, double>(s, d, s->vals64, st);
            if (profile) (void)hipEventRecord(s->setup_ev[1], st);
            s->coarse_valid = s->two_level_ok;
            s->coarse_age = 0;
            s->coarse_rebuilt = 1;
        } else {
            // the frozen aggregate map and coarse inverse, applied to this solve's rows
            const AggGeom g{d.c.gx, d.c.gy, d.c.gz, s->agg_edge, s->cgx, s->cgy, s->cgz};
            hipLaunchKernelGGL(k_agg_rows, dim3(grid_for(N, PB)), dim3(PB), 0, st, N, g, d.skeys, s->dir,
                               s->agg_flag, s->agg_map, s->agg_of_row);
            s->two_level_ok = 1;
            s->coarse_age += 1;
        }
    }
    s->vcycle = 0;
    if (s->two_level_ok && opt.precond == MPSR_PRECOND_CELL_VCYCLE) {
        const int ok = s->mixed ? setup_cells<R, float>(s, d, s->vals32, st) : setup_cells<R, double>(s, d, s->vals64, st);
        if (ok < 0) return MPSR_MEMERR;
        s->vcycle = ok;
    }
    const bool two = s->two_level_ok != 0;
    if (s->learned && bb > 0 && rr > target) {
        // Calibration, every 16 solves: five PCG iterations from the plain start with the plain
        // preconditioner give the residual the network's start has to beat after five iterations.
        auto five = [&]() -> int {
            if (s->mixed)
                return two ? run_pcg2<R, float, double>(s, d, N, s->vals32, s->x, s->r, s->z, s->p0, s->q, s->r, 0.0, 5,
                                                        0, st)
                           : run_pcg<R, float, double>(s, N, s->vals32, s->x, s->r, s->z, s->p0, s->p1, s->q, s->r,
                                                       0.0, 5, st);
            return two ? run_pcg2<R, double, double>(s, d, N, s->vals64, s->x, s->r, s->z, s->p0, s->q, s->r, 0.0, 5, 0,
                                                     st)
                       : run_pcg<R, double, double>(s, N, s->vals64, s->x, s->r, s->z, s->p0, s->p1, s->q, s->r, 0.0, 5,
                                                    st);
        };
        if (s->calib_age <= 0 || s->plain5 < 0) {
            hipLaunchKernelGGL(k_copy_d, dim3(grid_for(N, PB)), dim3(PB), 0, st, N, s->x, s->xl);
            hipLaunchKernelGGL(k_warm, dim3(grid_for(N, PB)), dim3(PB), 0, st, N, d.perm, s->dir, s->prev, use_prev,
                               s->x);
            s->pdiag = s->diag;
            double bbp = 0.0;
            if (resid_pass(s, N, st, &bbp) < 0) return MPSR_GPUERR;
            const int itp = five();
            const double rrp = resid_pass(s, N, st, &bbp);
            if (rrp < 0) return MPSR_GPUERR;
            s->plain5 = (itp >= 0 && bbp > 0) ? std::sqrt(rrp / bbp) : 1e300;
            s->calib_age = 16;
            hipLaunchKernelGGL(k_copy_d, dim3(grid_for(N, PB)), dim3(PB), 0, st, N, s->xl, s->x);
            s->pdiag = opt.learned_omega ? s->pdiag_buf : s->diag;
            rr = resid_pass(s, N, st, &bb);
            if (rr < 0) return MPSR_GPUERR;
        }
        s->calib_age -= 1;
    }
    // with the network, the first PCG checks its residual after 5 iterations against the plain
    // solver's (same preconditioner family, plain start) and returns -2 when it is larger
    auto arm_probe = [&]() {
        if (s->learned && s->plain5 >= 0) {
            s->probe_at = 5;
            s->probe_rr = s->plain5 * s->plain5 * bb;
        }
    };
    auto fall_back = [&]() -> bool {
        s->learned_fallback = 1;
        s->iters += 5;
        hipLaunchKernelGGL(k_warm, dim3(grid_for(N, PB)), dim3(PB), 0, st, N, d.perm, s->dir, s->prev, use_prev, s->x);
        s->pdiag = s->diag;
        rr = resid_pass(s, N, st, &bb);
        return rr >= 0;
    };
    if (!s->mixed) {
        if (bb > 0 && rr > target) {
            // FP64 PCG from the warm start; r already holds b - A x
            auto pcg = [&]() {
                return two ? run_pcg2<R, double, double>(s, d, N, s->vals64, s->x, s->r, s->z, s->p0, s->q, s->r,
                                                         target, maxit, 0, st, s->hint[0])
                           : run_pcg<R, double, double>(s, N, s->vals64, s->x, s->r, s->z, s->p0, s->p1, s->q, s->r,
                                                        target, maxit, st, s->hint[0]);
            };
            arm_probe();
            int it = pcg();
            if (it == -2) {
                if (!fall_back()) return MPSR_GPUERR;
                it = pcg();
            }
            if (it < 0) s->status = 1;
            else s->iters += it;
            s->hint[0] = it > 0 ? it : 0;
            s->passes = 1;
            rr = resid_pass(s, N, st, &bb);
            if (rr < 0) return MPSR_GPUERR;
        }
    } else {
        // iterative refinement: FP64 residual, FP32 inner PCG with FP64 dots, FP64 update
        bool fallback = false;
        arm_probe();
        while (bb > 0 && rr > target && s->passes < 16 && s->iters < maxit) {
            const double inner_rel = std::max(1e-6, 0.5 * tol * std::sqrt(bb / rr));
            const int hp = std::min(s->passes, HINT_PASSES - 1);
            // a pass that stalls in FP32 ends at its cap; the next pass restarts from the FP64 residual
            const int cap = std::min(maxit - s->iters, s->hint[hp] > 0 ? std::max(PASS_CAP_MIN, 4 * s->hint[hp])
                                                                       : PASS_CAP_COLD);
            const int it =
                two ? run_pcg2<R, float, float>(s, d, N, s->vals32, s->ef, s->rf, s->zf, s->p0f, s->qf, s->r,
                                                inner_rel * inner_rel * rr, cap, 1, st, s->hint[hp])
                    : run_pcg<R, float, float>(s, N, s->vals32, s->ef, s->rf, s->zf, s->p0f, s->p1f, s->qf, s->r,
                                               inner_rel * inner_rel * rr, cap, st, s->hint[hp]);
            if (it >= 0 && it < cap) s->hint[hp] = it;
            if (it == -2) {
                if (!fall_back()) return MPSR_GPUERR;
                continue;
            }
            if (it < 0) {
                fallback = true;
                break;
            }
            hipLaunchKernelGGL(k_add_f2d, dim3(grid_for(N, PB)), dim3(PB), 0, st, N, s->x, s->ef);
            s->iters += it;
            s->passes += 1;
            const double prev_rr = rr;
            rr = resid_pass(s, N, st, &bb);
            if (rr < 0) return MPSR_GPUERR;
            if (it == 0 || !(rr < 0.5 * prev_rr)) {
                // the FP32 inner solve no longer reduces the FP64 residual
                fallback = rr > target;
                break;
            }
        }
        if (fallback && bb > 0 && rr > target) {
            // plain PCG in FP64 vectors on the same (FP32-stored) matrix from the current x;
            // r holds b - A x from the last residual pass
            s->fallbacks += 1;
            const int it = two ? run_pcg2<R, float, double>(s, d, N, s->vals32, s->x, s->r, s->z, s->p0, s->q, s->r,
                                                            target, std::max(1, maxit - s->iters), 0, st)
                               : run_pcg<R, float, double>(s, N, s->vals32, s->x, s->r, s->z, s->p0, s->p1, s->q,
                                                           s->r, target, std::max(1, maxit - s->iters), st);
            if (it < 0) s->status = 1;
            else s->iters += it;
            s->passes += 1;
            rr = resid_pass(s, N, st, &bb);
            if (rr < 0) return MPSR_GPUERR;
        }
    }
    s->relres = bb > 0 ? std::sqrt(rr / bb) : 0.0;
    if (s->coarse_rebuilt && s->two_level_ok) s->iters_at_build = s->iters;
    if (s->growth_check && 10 * s->iters >= 9 * s->iters_before_growth) s->growth_off = 1;
    if (s->cell_growth_check && 10 * s->iters >= 9 * s->iters_before_cell_growth) s->cell_growth_off = 1;
    s->growth_check = s->cell_growth_check = 0;
    if (s->vcycle) s->iters_at_cell = s->cell_rebuilt ? s->iters : std::min(s->iters_at_cell, s->iters);
    s->iters_last = s->iters;
    if (profile && s->coarse_rebuilt) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, s->setup_ev[0], s->setup_ev[1]) == hipSuccess) s->setup_ms = ms;
    }
    if (!std::isfinite(s->relres)) {
        s->status = 1;
        std::snprintf(s->why, sizeof(s->why), "non-finite residual (r.r = %.3e, b.b = %.3e)", rr, bb);
    } else if (s->status == 0 && bb > 0 && rr > target) {
        // the iteration or refinement cap ended the solve above the tolerance
        s->status = 2;
        std::snprintf(s->why, sizeof(s->why),
                      "not converged: relative residual %.3e above %.3e after %d iterations (max_iter %d)",
                      s->relres, tol, s->iters, maxit);
    }
    {
        // algorithmic bytes of one SpMV launch: every stored entry's value and column once, and
        // per row p and diag read and q written; gathers of p at the columns not counted
        const double sv = s->mixed ? 4.0 : 8.0;
        // 16-bit codes: 2 bytes a column, and per row its cell key
        s->spmv_bytes = s->c16 ? (double)s->nnz_real * (sv + 2.0) + (double)N * (2.0 * sv + 8.0 + 4.0)
                               : (double)s->nnz_real * (sv + 4.0) + (double)N * (2.0 * sv + 8.0);
    }
    if (s->status != 0) {
        // a failed solve leaves the state as it was: no pressure is written back, and the caller's
        // step stops here with the reason in ppe_why()
        if (!s->why[0]) std::snprintf(s->why, sizeof(s->why), "PCG breakdown");
        return MPSR_CALCERR;
    }
    hipLaunchKernelGGL(k_ppe_writeback<R>, dim3(grid_for(N, PB)), dim3(PB), 0, st, N, d.c, d.perm, s->dir, s->x, d.in,
                       d.out, d.spos, d.spres, s->prev);
    s->have_prev = 1;
    return MPSR_NOERR;
}

template <typename R> mpsr_status ppe_check(PpeSolver<R>* s, mpsr_timings* tm) {
    // SpMV launch statistics over the events of the call (complete: the caller synchronised)
    s->spmv_launches = 0;
    s->spmv_ms = 0.0;
    if (s->profile) {
        double tot_ms = 0.0;
        int cnt = 0;
        for (size_t e = 0; e + 1 < s->ev_used; e += 2) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, s->ev[e], s->ev[e + 1]) == hipSuccess) {
                tot_ms += ms;
                ++cnt;
            }
        }
        s->spmv_launches = cnt;
        s->spmv_ms = cnt ? tot_ms / cnt : 0.0;
    }
    tm->spmv_ms = s->spmv_ms;
    tm->spmv_launches = s->spmv_launches;
    tm->spmv_bytes = s->spmv_bytes;
    tm->ppe_nnz = s->nnz_real;
    tm->ppe_stored = s->nnz_stored;
    tm->ppe_rows = s->rows_real;
    tm->ppe_iterations = s->iters;
    tm->ppe_refinements = s->passes;
    tm->ppe_rel_residual = s->relres;
    tm->ppe_initial_rel_residual = s->relres0;
    tm->ppe_fallbacks = s->fallbacks;
    tm->ppe_coarse_rows = s->two_level_ok ? s->nc : 0;
    tm->ppe_coarse_rebuilt = s->coarse_rebuilt;
    tm->ppe_setup_ms = s->setup_ms;
    tm->ppe_learned = s->learned;
    tm->ppe_learned_fallback = s->learned_fallback;
    return s->status == 0 ? MPSR_NOERR : MPSR_CALCERR;
}

template <typename R> const char* ppe_why(PpeSolver<R>* s) { return s->why; }
// the SpMV launch events of a step or step_many call start from an empty pool
template <typename R> void ppe_begin_call(PpeSolver<R>* s) { s->ev_used = 0; }
template <typename R> void ppe_last(const PpeSolver<R>* s, int* iters, double* relres, int* rebuilt) {
    *iters = s->iters;
    *relres = s->relres;
    *rebuilt = s->coarse_rebuilt && s->two_level_ok;
}

template <typename R> mpsr_status ppe_mpnn_set(PpeSolver<R>* s, const float* w, hipStream_t st) {
    if (!s->net) {
        s->net = mpnn_create(((s->cap + 63) / 64) * 64);
        if (!s->net) return MPSR_MEMERR;
    }
    return mpnn_set_weights(s->net, w, st);
}
template <typename R> mpsr_status ppe_mpnn_get(PpeSolver<R>* s, float* w, hipStream_t st) {
    if (!s->net || !mpnn_has_weights(s->net)) return MPSR_PARAMERR;
    return mpnn_get_weights(s->net, w, st);
}
// One training sample: the system of the last solve with its solution as the target.  Computes
// the loss and all gradients; with update set, one Adam step of rate lr follows.
template <typename R>
mpsr_status ppe_mpnn_train(PpeSolver<R>* s, const Dev<R>& d, double lr, double beta, int update, float* grads,
                           double* parts, hipStream_t st) {
    if (!s->net || !mpnn_has_weights(s->net)) return MPSR_PARAMERR;
    if (s->nslices <= 0 || d.N <= 0) return MPSR_PARAMERR;
    ensure_cols(s, d, st);
    const MpnnSys sys = mpnn_sys(s, d);
    mpnn_features<R>(s->net, sys, mpnn_geo(d), st);
    mpnn_forward(s->net, sys, st);
    mpsr_status r = mpnn_loss_grad(s->net, sys, beta, parts, st);
    if (r != MPSR_NOERR) return r;
    if (grads) {
        r = mpnn_get_grads(s->net, grads, st);
        if (r != MPSR_NOERR) return r;
    }
    if (update) mpnn_adam(s->net, (float)lr, st);
    return hipStreamSynchronize(st) == hipSuccess ? MPSR_NOERR : MPSR_GPUERR;
}
// Features and outputs of the network on the last system, slot order (rows of slots that are not
// binned are 0), and the pressure scale s.
template <typename R>
mpsr_status ppe_mpnn_rows(PpeSolver<R>* s, const Dev<R>& d, float* X, float* Y, double* scale, hipStream_t st) {
    if (!s->net || !mpnn_has_weights(s->net)) return MPSR_PARAMERR;
    ensure_cols(s, d, st);
    const MpnnSys sys = mpnn_sys(s, d);
    mpnn_features<R>(s->net, sys, mpnn_geo(d), st);
    mpnn_forward(s->net, sys, st);
    const int N = d.N;
    std::vector<float> xs((size_t)std::max(N, 1) * NN_F), ys((size_t)std::max(N, 1) * 2);
    std::vector<int> perm(std::max(N, 1));
    mpsr_status r = mpnn_get_rows(s->net, N, xs.data(), ys.data(), scale, st);
    if (r != MPSR_NOERR) return r;
    if (N > 0 && hipMemcpy(perm.data(), d.perm, sizeof(int) * N, hipMemcpyDeviceToHost) != hipSuccess)
        return MPSR_GPUERR;
    if (X) std::fill(X, X + (size_t)N * NN_F, 0.f);
    if (Y) std::fill(Y, Y + (size_t)N * 2, 0.f);
    for (int t = 0; t < N; ++t) {
        const int slot = perm[t];
        if (X) std::copy(&xs[(size_t)t * NN_F], &xs[(size_t)t * NN_F] + NN_F, X + (size_t)slot * NN_F);
        if (Y) {
            Y[2 * (size_t)slot] = ys[2 * (size_t)t];
            Y[2 * (size_t)slot + 1] = ys[2 * (size_t)t + 1];
        }
    }
    return MPSR_NOERR;
}

template <typename R> mpsr_status ppe_get(PpeSolver<R>* s, const Dev<R>& d, double* rhs, double* x, int* dir) {
    const int N = d.N;
    std::vector<double> b(N), xx(N);
    std::vector<int> dd(N), pp(N);
    if (hipMemcpy(b.data(), s->rhs, sizeof(double) * N, hipMemcpyDeviceToHost) != hipSuccess) return MPSR_GPUERR;
    if (hipMemcpy(xx.data(), s->x, sizeof(double) * N, hipMemcpyDeviceToHost) != hipSuccess) return MPSR_GPUERR;
    if (hipMemcpy(dd.data(), s->dir, sizeof(int) * N, hipMemcpyDeviceToHost) != hipSuccess) return MPSR_GPUERR;
    if (hipMemcpy(pp.data(), d.perm, sizeof(int) * N, hipMemcpyDeviceToHost) != hipSuccess) return MPSR_GPUERR;
    for (int t = 0; t < N; ++t) {
        const int sl = pp[t];
        if (rhs) rhs[sl] = b[t];
        if (x) x[sl] = xx[t];
        if (dir) dir[sl] = dd[t];
    }
    return MPSR_NOERR;
}

template <typename R>
mpsr_status ppe_get_matrix(PpeSolver<R>* s, const Dev<R>& d, int* ptr, int* col, double* val, double* diag, int cap,
                           int* nnz) {
    const int N = d.N;
    const int nsl = (N + 63) / 64;
    if (!s->cols_valid) {
        ensure_cols(s, d, nullptr);
        if (hipDeviceSynchronize() != hipSuccess) return MPSR_GPUERR;
    }
    std::vector<int> sw(nsl), so(nsl), cn(N), pp(N), dd(N), po(N);
    std::vector<double> dg(N);
    if (hipMemcpy(sw.data(), s->slice_w, sizeof(int) * nsl, hipMemcpyDeviceToHost) != hipSuccess) return MPSR_GPUERR;
    if (hipMemcpy(so.data(), s->slice_off, sizeof(int) * nsl, hipMemcpyDeviceToHost) != hipSuccess) return MPSR_GPUERR;
    if (hipMemcpy(cn.data(), s->cnt, sizeof(int) * N, hipMemcpyDeviceToHost) != hipSuccess) return MPSR_GPUERR;
    if (hipMemcpy(po.data(), s->pos, sizeof(int) * N, hipMemcpyDeviceToHost) != hipSuccess) return MPSR_GPUERR;
    if (hipMemcpy(pp.data(), d.perm, sizeof(int) * N, hipMemcpyDeviceToHost) != hipSuccess) return MPSR_GPUERR;
    if (hipMemcpy(dd.data(), s->dir, sizeof(int) * N, hipMemcpyDeviceToHost) != hipSuccess) return MPSR_GPUERR;
    if (hipMemcpy(dg.data(), s->diag, sizeof(double) * N, hipMemcpyDeviceToHost) != hipSuccess) return MPSR_GPUERR;
    const long long total = nsl ? (long long)so[nsl - 1] + sw[nsl - 1] : 0;
    std::vector<int> cc(total);
    std::vector<double> vv(total);
    if (total) {
        if (hipMemcpy(cc.data(), s->cols, sizeof(int) * total, hipMemcpyDeviceToHost) != hipSuccess) return MPSR_GPUERR;
        if (s->mixed) {
            std::vector<float> vf(total);
            if (hipMemcpy(vf.data(), s->vals32, sizeof(float) * total, hipMemcpyDeviceToHost) != hipSuccess)
                return MPSR_GPUERR;
            for (long long e = 0; e < total; ++e) vv[e] = vf[e];
        } else if (hipMemcpy(vv.data(), s->vals64, sizeof(double) * total, hipMemcpyDeviceToHost) != hipSuccess) {
            return MPSR_GPUERR;
        }
    }
    std::vector<int> row_of(N);
    for (int t = 0; t < N; ++t) row_of[pp[t]] = t;
    int e = 0;
    ptr[0] = 0;
    for (int sl = 0; sl < N; ++sl) {
        const int t = row_of[sl];
        if (diag) diag[sl] = dg[t];
        const int n = dd[t] ? 0 : cn[t];
        for (int k = 0; k < n; ++k, ++e) {
            const long long idx = sell_at(so[po[t] >> 6], po[t] & 63, k);
            if (e < cap) {
                col[e] = pp[cc[idx]];
                val[e] = vv[idx];
            }
        }
        ptr[sl + 1] = e;
    }
    if (nnz) *nnz = e;
    return MPSR_NOERR;
}

}  // namespace mpsr

extern "C" mpsr_status mpsr_spd_inverse(int n, const double* a, double* inv) {
    if (n <= 0 || n > 16384 || !a || !inv) return MPSR_PARAMERR;
    switch (mpsr::spd_inverse_host(n, a, inv)) {
        case 0: return MPSR_NOERR;
        case 1: return MPSR_CALCERR;
        case 2: return MPSR_MEMERR;
        default: return MPSR_GPUERR;
    }
}

namespace mpsr {
// ---------------------------------------------------------------------------------------------
// The PPE solves of a batched step: the PPE buckets of one precision, every kernel one launch over
// (bucket, block) work items.  Each bucket solves with the FP64 Jacobi-PCG of ppe_solve (its matrix,
// residuals, reductions and stopping rules), so its pressures equal a single-domain run's with
// that solver.  The host decisions of ppe_solve become device steps: the start is skipped when the
// first residual is already within the tolerance (S_SKIP), the iterations run in fixed chunks
// that return at once after convergence, and the final residual decides on the device whether the
// pressure is written back (S_OK).  The matrix holds NB_WIDTH entries a row; a bucket whose
// matrix would need more is reported after the step and left unsolved.
// ---------------------------------------------------------------------------------------------
template <typename R> struct PpeView {
    int dev = 0;                 // index of the bucket in the group's device table
    int N = 0, nslices = 0, nb = 0, use_prev = 0, maxit = 0;
    double tol = 1e-8, alpha_src = 0.0, n0 = 0.0;
    long long nnz_cap = 0;
    int *active, *dir, *cnt, *nbq, *nbn;
    int *slice_w, *slice_off, *pos, *rowp, *total, *first_ge, *cols, *nnz;
    uint16_t* cols16;
    double *vals, *diag, *rhs, *x, *prev, *r, *z, *p, *q, *part, *scal;
    int* flags;                  // [0] a previous pressure exists, [1] the matrix overflowed
};

template <typename R> struct OpPFirstGe {
    const Dev<R>* dv;
    const PpeView<R>* pv;
    __device__ void operator()(int j, int b) const {
        const Dev<R>& d = dv[pv[j].dev];
        first_ge_body(b, d.N, (int)d.c.ncell, d.skeys, pv[j].first_ge);
    }
};
template <typename R> struct OpPSell {
    const PpeView<R>* pv;
    __device__ void operator()(int j, int b) const {
        const PpeView<R>& v = pv[j];
        sell_order_body(b, v.N, v.nslices, v.cnt, v.dir, v.pos, v.rowp, v.slice_w);
    }
};
template <typename R> struct OpPFill {
    const Dev<R>* dv;
    const PpeView<R>* pv;
    __device__ void operator()(int j, int b) const {
        const PpeView<R>& v = pv[j];
        const Dev<R>& d = dv[v.dev];
        if ((long long)*v.total > v.nnz_cap) {   // the same for every workgroup of the bucket
            if (b == 0 && threadIdx.x == 0) v.flags[1] = 1;
            return;
        }
        ppe_fill_body<R, double>(b, d.N, v.nslices, d.c, v.active, v.dir, d.spos, d.cell_start, d.cell_end,
                                 v.slice_off, v.slice_w, d.perm, d.out, v.alpha_src, v.n0, v.rowp, v.cols, v.vals,
                                 v.rhs, v.cnt, v.nnz, v.nbq, NB_WIDTH, v.cols16, d.skeys, v.first_ge);
    }
};
template <typename R> struct OpPWarm {
    const Dev<R>* dv;
    const PpeView<R>* pv;
    __device__ void operator()(int j, int b) const {
        const PpeView<R>& v = pv[j];
        warm_body(b, v.N, dv[v.dev].perm, v.dir, v.prev, v.use_prev && v.flags[0], v.x);
    }
};
template <typename R> struct OpPResid {
    const PpeView<R>* pv;
    __device__ void operator()(int j, int b) const {
        const PpeView<R>& v = pv[j];
        if (v.flags[1]) return;
        resid_body<double, false, false>(b, v.N, v.nslices, v.slice_off, v.slice_w, v.cols, v.vals, v.diag, v.x, v.rhs,
                                         v.rowp, v.cnt, v.r, v.part, nullptr, nullptr, nullptr, 0, 0, nullptr, 0);
    }
};
template <typename R> struct OpPReduce {
    const PpeView<R>* pv;
    int mode;
    __device__ void operator()(int j, int b) const {
        const PpeView<R>& v = pv[j];
        if (mode == RED_START && v.scal[S_SKIP] != 0.0) {
            if (threadIdx.x == 0) {
                v.scal[S_ITERS] = 0.0;
                v.scal[S_DONE] = 1.0;
            }
            return;
        }
        pcg_reduce_body(b, v.part, v.nb, mode == RED_ALPHA ? 1 : 2, v.scal, mode);
    }
};
// the decisions ppe_solve takes on the host after the first residual: the target, the iteration
// cap, and whether the PCG runs at all
template <typename R> struct OpPGate {
    const PpeView<R>* pv;
    __device__ void operator()(int j, int) const {
        const PpeView<R>& v = pv[j];
        if (threadIdx.x != 0) return;
        const double rr = v.scal[S_RR], bb = v.scal[S_BB];
        const double target = v.tol * v.tol * bb;
        v.scal[S_TARGET] = target;
        v.scal[S_MAXIT] = (double)v.maxit;
        v.scal[S_SKIP] = (v.flags[1] || !(bb > 0 && rr > target)) ? 1.0 : 0.0;
        v.scal[S_DONE] = 0.0;
        v.scal[S_OK] = 0.0;
    }
};
template <typename R> struct OpPStart {
    const PpeView<R>* pv;
    __device__ void operator()(int j, int b) const {
        const PpeView<R>& v = pv[j];
        if (v.scal[S_SKIP] != 0.0) return;
        pcg_start_body<double>(b, v.N, v.r, v.diag, v.r, v.z, v.p, v.x, 0, v.part);
    }
};
template <typename R> struct OpPSpmv {
    const PpeView<R>* pv;
    __device__ void operator()(int j, int b) const {
        const PpeView<R>& v = pv[j];
        pcg_spmv_body<double, double, false, false>(b, v.nslices, v.rowp, v.cnt, v.slice_off, v.slice_w, v.cols,
                                                    v.vals, v.diag, v.p, v.q, v.scal, v.part, nullptr, nullptr,
                                                    nullptr, 0, 0, nullptr, 0);
    }
};
template <typename R> struct OpPUpdate {
    const PpeView<R>* pv;
    __device__ void operator()(int j, int b) const {
        const PpeView<R>& v = pv[j];
        pcg_update_body<double>(b, v.N, v.x, v.r, v.z, v.p, v.q, v.diag, v.scal, v.part);
    }
};
template <typename R> struct OpPDirection {
    const PpeView<R>* pv;
    __device__ void operator()(int j, int b) const {
        const PpeView<R>& v = pv[j];
        pcg_direction_body<double>(b, v.N, v.z, v.p, v.scal);
    }
};
// ppe_solve's checks of the final residual: no breakdown, a finite relative residual, and the
// residual within the tolerance
template <typename R> struct OpPOutcome {
    const PpeView<R>* pv;
    __device__ void operator()(int j, int) const {
        const PpeView<R>& v = pv[j];
        if (threadIdx.x != 0) return;
        const double rr = v.scal[S_RR], bb = v.scal[S_BB];
        const double rel = bb > 0 ? sqrt(rr / bb) : 0.0;
        const bool ok = !v.flags[1] && v.scal[S_DONE] != 3.0 && isfinite(rel) && !(bb > 0 && rr > v.scal[S_TARGET]);
        v.scal[S_OK] = ok ? 1.0 : 0.0;
    }
};
template <typename R> struct OpPWriteback {
    const Dev<R>* dv;
    const PpeView<R>* pv;
    __device__ void operator()(int j, int b) const {
        const PpeView<R>& v = pv[j];
        if (v.scal[S_OK] == 0.0) return;
        const Dev<R>& d = dv[v.dev];
        ppe_writeback_body<R>(b, v.N, d.c, d.perm, v.dir, v.x, d.in, d.out, d.spos, d.spres, v.prev);
        if (b == 0 && threadIdx.x == 0) v.flags[0] = 1;
    }
};
template <typename R> struct OpPFlags {
    const PpeView<R>* pv;
    int* done;
    __device__ void operator()(int j, int) const {
        if (threadIdx.x == 0) done[j] = pv[j].scal[S_DONE] != 0.0 ? 1 : 0;
    }
};

template <typename R> struct PpeBatch {
    std::vector<PpeSolver<R>*> s;
    std::vector<PpeView<R>> hv;
    PpeView<R>* views = nullptr;
    PndRows<R>* pnd = nullptr;
    int* flags = nullptr;      // [n * 2]
    int* done = nullptr;       // [n] device
    int* done_h = nullptr;     // [n] pinned host
    char* tab = nullptr;
    size_t tab_bytes = 0;
    WorkList wpnd, wcells, wwin, wvec, wrow, wone, wfill, wscan_t, wscan_1;
    const FillJob* fills = nullptr;
    const ScanJob* scans = nullptr;
    int n = 0;
};

template <typename R> PpeBatch<R>* ppe_batch_create() { return new PpeBatch<R>(); }
template <typename R> void ppe_batch_destroy(PpeBatch<R>* b) {
    if (!b) return;
    for (void* p : {(void*)b->views, (void*)b->pnd, (void*)b->flags, (void*)b->done, (void*)b->tab})
        if (p) (void)hipFree(p);
    if (b->done_h) (void)hipHostFree(b->done_h);
    delete b;
}

template <typename R>
mpsr_status ppe_batch_build(PpeBatch<R>* b, PpeSolver<R>* const* solvers, const int* dev, const Dev<R>* host_devs,
                            const mpsr_ppe_options* opts, int n, int block) {
    const bool first = b->views == nullptr;
    if (!first && n != b->n) return MPSR_PARAMERR;
    b->n = n;
    b->s.assign(solvers, solvers + n);
    b->hv.assign(n, PpeView<R>{});
    if (first) {
        if (hipMalloc((void**)&b->views, sizeof(PpeView<R>) * std::max(1, n)) != hipSuccess ||
            hipMalloc((void**)&b->pnd, sizeof(PndRows<R>) * std::max(1, n)) != hipSuccess ||
            hipMalloc((void**)&b->flags, sizeof(int) * 2 * std::max(1, n)) != hipSuccess ||
            hipMalloc((void**)&b->done, sizeof(int) * std::max(1, n)) != hipSuccess ||
            hipHostMalloc((void**)&b->done_h, sizeof(int) * std::max(1, n)) != hipSuccess)
            return MPSR_MEMERR;
        std::vector<int> f(2 * std::max(1, n), 0);
        for (int j = 0; j < n; ++j) f[2 * j] = solvers[j]->have_prev;
        if (hipMemcpy(b->flags, f.data(), sizeof(int) * f.size(), hipMemcpyHostToDevice) != hipSuccess)
            return MPSR_GPUERR;
    }
    std::vector<PndRows<R>> rows(std::max(1, n));
    std::vector<int2> pnd, cells, win, vec, row, one, fill, scan_t, scan_1;
    std::vector<FillJob> fj;
    std::vector<ScanJob> sj;
    for (int j = 0; j < n; ++j) {
        PpeSolver<R>* s = solvers[j];
        const Dev<R>& d = host_devs[dev[j]];
        const mpsr_ppe_options& o = opts[j];
        const int N = d.N;
        const int nsl = (N + 63) / 64;
        // capacity for NB_WIDTH entries a row, the matrix arrays of the FP64 solver
        const long long need = (long long)nsl * 64 * NB_WIDTH;
        if (need > s->nnz_cap) {
            for (void* p : {(void*)s->cols, (void*)s->cols16, (void*)s->vals64, (void*)s->vals32})
                if (p) (void)hipFree(p);
            s->cols = nullptr; s->cols16 = nullptr; s->vals64 = nullptr; s->vals32 = nullptr;
            s->nnz_cap = need;
            if (!dev_alloc(&s->cols, (size_t)need) || !dev_alloc(&s->cols16, (size_t)need) ||
                !dev_alloc(&s->vals64, (size_t)need) || !dev_alloc(&s->vals32, (size_t)need))
                return MPSR_MEMERR;
        }
        s->nslices = nsl;
        s->mixed = 0;
        s->c16 = 0;
        s->wcap = 0;
        s->skeys = d.skeys;
        PpeView<R> v{};
        v.dev = dev[j];
        v.N = N;
        v.nslices = nsl;
        v.nb = (nsl * 64 + PB - 1) / PB;
        v.use_prev = o.warm_start != MPSR_WARM_ZERO ? 1 : 0;
        v.maxit = o.max_iter > 0 ? o.max_iter : 20000;
        v.tol = o.rel_tol > 0 ? o.rel_tol : 1e-8;
        v.alpha_src = s->alpha * (o.source_relaxation > 0 ? o.source_relaxation : 0.1);
        v.n0 = s->n0;
        v.nnz_cap = s->nnz_cap;
        v.active = s->active; v.dir = s->dir; v.cnt = s->cnt; v.nbq = s->nbq; v.nbn = s->nbn;
        v.slice_w = s->slice_w; v.slice_off = s->slice_off; v.pos = s->pos; v.rowp = s->rowp; v.total = s->total;
        v.first_ge = s->first_ge; v.cols = s->cols; v.nnz = s->nnz_dev; v.cols16 = s->cols16; v.vals = s->vals64;
        v.diag = s->diag; v.rhs = s->rhs; v.x = s->x; v.prev = s->prev; v.r = s->r; v.z = s->z; v.p = s->p0;
        v.q = s->q; v.part = s->part; v.scal = s->scal;
        v.flags = b->flags + 2 * j;
        b->hv[j] = v;
        rows[j] = PndRows<R>{dev[j], s->active, s->cnt, s->diag, s->dir, s->nbq, s->nbn, NB_WIDTH};
        for (int k = 0; k < (int)grid_for(N, block); ++k) pnd.push_back(make_int2(j, k));
        for (int k = 0; k < (int)grid_for((long long)d.c.ncell + 1, PB); ++k) cells.push_back(make_int2(j, k));
        for (int k = 0; k < (int)grid_for(N, SELL_SIGMA); ++k) win.push_back(make_int2(j, k));
        for (int k = 0; k < v.nb; ++k) vec.push_back(make_int2(j, k));
        for (int k = 0; k < (int)grid_for(N, PB); ++k) row.push_back(make_int2(j, k));
        one.push_back(make_int2(j, 0));
        const int jf = (int)fj.size();
        fj.push_back(FillJob{s->nnz_dev, 0, 4});
        fj.push_back(FillJob{b->flags + 2 * j + 1, 0, 1});   // the overflow flag of the last step
        fill.push_back(make_int2(jf, 0));
        fill.push_back(make_int2(jf + 1, 0));
        const int js = (int)sj.size();
        sj.push_back(ScanJob{s->slice_w, s->slice_off, nullptr, nsl, nsl, s->total, s->scan.scan_tmp});
        for (int k = 0; k < (int)grid_for(nsl, scan_tile()); ++k) scan_t.push_back(make_int2(js, k));
        scan_1.push_back(make_int2(js, 0));
    }
    if (hipMemcpy(b->views, b->hv.data(), sizeof(PpeView<R>) * std::max(1, n), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->pnd, rows.data(), sizeof(PndRows<R>) * rows.size(), hipMemcpyHostToDevice) != hipSuccess)
        return MPSR_GPUERR;
    std::vector<char> img;
    auto put = [&](const void* p, size_t bytes) {
        const size_t off = (img.size() + 15) & ~(size_t)15;
        img.resize(off + bytes);
        if (bytes) std::memcpy(img.data() + off, p, bytes);
        return off;
    };
    std::vector<std::pair<size_t, int>> refs;
    for (const auto* v : {&pnd, &cells, &win, &vec, &row, &one, &fill, &scan_t, &scan_1})
        refs.push_back({put(v->data(), v->size() * sizeof(int2)), (int)v->size()});
    const size_t ofj = put(fj.data(), fj.size() * sizeof(FillJob)), osj = put(sj.data(), sj.size() * sizeof(ScanJob));
    if (b->tab && img.size() > b->tab_bytes) {
        (void)hipFree(b->tab);
        b->tab = nullptr;
    }
    if (!b->tab) {
        b->tab_bytes = std::max<size_t>(16, img.size());
        if (hipMalloc((void**)&b->tab, b->tab_bytes) != hipSuccess) return MPSR_MEMERR;
    }
    if (hipMemcpy(b->tab, img.data(), img.size(), hipMemcpyHostToDevice) != hipSuccess) return MPSR_GPUERR;
    auto wl = [&](int i) { return WorkList{reinterpret_cast<const int2*>(b->tab + refs[i].first), refs[i].second}; };
    b->wpnd = wl(0); b->wcells = wl(1); b->wwin = wl(2); b->wvec = wl(3); b->wrow = wl(4); b->wone = wl(5);
    b->wfill = wl(6); b->wscan_t = wl(7); b->wscan_1 = wl(8);
    b->fills = reinterpret_cast<const FillJob*>(b->tab + ofj);
    b->scans = reinterpret_cast<const ScanJob*>(b->tab + osj);
    return MPSR_NOERR;
}

// the PND pass of the PPE buckets (active rows, bounds, diagonals, neighbour lists)
template <typename R> void ppe_batch_pnd(PpeBatch<R>* b, const BatchLaunch<R>& bl, hipStream_t st) {
    launch_pnd_ppe_b(bl, b->pnd, b->wpnd, st);
}
template <typename R> void ppe_batch_iterate(PpeBatch<R>* b, int iters, hipStream_t st) {
    for (int k = 0; k < iters; ++k) {
        launch_batched<PB>(OpPSpmv<R>{b->views}, b->wvec, PB, st);
        launch_batched<1024>(OpPReduce<R>{b->views, RED_ALPHA}, b->wone, 1024, st);
        launch_batched<PB>(OpPUpdate<R>{b->views}, b->wvec, PB, st);
        launch_batched<1024>(OpPReduce<R>{b->views, RED_BETA}, b->wone, 1024, st);
        launch_batched<PB>(OpPDirection<R>{b->views}, b->wvec, PB, st);
    }
}
// assembly, warm start, first residual, the decisions on it, the PCG start and `iters` iterations
template <typename R> void ppe_batch_head(PpeBatch<R>* b, const BatchLaunch<R>& bl, int iters, hipStream_t st) {
    fill_int_batched(b->fills, b->wfill, st);
    launch_batched<PB>(OpPFirstGe<R>{bl.devs, b->views}, b->wcells, PB, st);
    launch_batched<SELL_BLOCK>(OpPSell<R>{b->views}, b->wwin, SELL_BLOCK, st);
    exclusive_scan_batched(b->scans, b->wscan_t, b->wscan_1, st);
    launch_batched<PB>(OpPFill<R>{bl.devs, b->views}, b->wvec, PB, st);
    launch_batched<PB>(OpPWarm<R>{bl.devs, b->views}, b->wrow, PB, st);
    launch_batched<PB>(OpPResid<R>{b->views}, b->wvec, PB, st);
    launch_batched<1024>(OpPReduce<R>{b->views, RED_RESID}, b->wone, 1024, st);
    launch_batched<64>(OpPGate<R>{b->views}, b->wone, 64, st);
    launch_batched<PB>(OpPStart<R>{b->views}, b->wvec, PB, st);
    launch_batched<1024>(OpPReduce<R>{b->views, RED_START}, b->wone, 1024, st);
    ppe_batch_iterate(b, iters, st);
}
// whether every bucket's PCG has stopped (synchronises the stream)
template <typename R> mpsr_status ppe_batch_converged(PpeBatch<R>* b, hipStream_t st, bool* all) {
    launch_batched<64>(OpPFlags<R>{b->views, b->done}, b->wone, 64, st);
    if (hipMemcpyAsync(b->done_h, b->done, sizeof(int) * b->n, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return MPSR_GPUERR;
    *all = true;
    for (int j = 0; j < b->n; ++j) *all = *all && b->done_h[j] != 0;
    return MPSR_NOERR;
}
// final residual, its checks, and the pressure written back where they passed
template <typename R> void ppe_batch_tail(PpeBatch<R>* b, const BatchLaunch<R>& bl, hipStream_t st) {
    launch_batched<PB>(OpPResid<R>{b->views}, b->wvec, PB, st);
    launch_batched<1024>(OpPReduce<R>{b->views, RED_RESID}, b->wone, 1024, st);
    launch_batched<64>(OpPOutcome<R>{b->views}, b->wone, 64, st);
    launch_batched<PB>(OpPWriteback<R>{bl.devs, b->views}, b->wrow, PB, st);
}
// after a batched call: the solvers' host state, and the first bucket whose last solve failed
template <typename R> mpsr_status ppe_batch_finish(PpeBatch<R>* b, int* failed, std::string* why) {
    *failed = -1;
    std::vector<int> f(2 * std::max(1, b->n));
    if (hipMemcpy(f.data(), b->flags, sizeof(int) * f.size(), hipMemcpyDeviceToHost) != hipSuccess) return MPSR_GPUERR;
    for (int j = 0; j < b->n; ++j) {
        PpeSolver<R>* s = b->s[j];
        s->cols_valid = 1;   // the batched fill writes the 32-bit columns
        double sc[S_COUNT];
        if (hipMemcpy(sc, s->scal, sizeof(sc), hipMemcpyDeviceToHost) != hipSuccess) return MPSR_GPUERR;
        s->have_prev = f[2 * j];
        s->iters = (int)sc[S_ITERS];
        s->passes = 1;
        s->relres = sc[S_BB] > 0 ? std::sqrt(sc[S_RR] / sc[S_BB]) : 0.0;
        s->status = (f[2 * j + 1] || sc[S_OK] == 0.0) ? 1 : 0;
        if (*failed < 0 && s->status) {
            *failed = j;
            char buf[256];
            if (f[2 * j + 1])
                std::snprintf(buf, sizeof(buf), "PPE matrix needs more than %d entries a row", NB_WIDTH);
            else if (sc[S_DONE] == 3.0)
                std::snprintf(buf, sizeof(buf), "PCG breakdown after %d iterations", (int)sc[S_ITERS]);
            else
                std::snprintf(buf, sizeof(buf), "not converged: relative residual %.3e after %d iterations",
                              s->relres, (int)sc[S_ITERS]);
            *why = buf;
        }
    }
    return MPSR_NOERR;
}

#define MPSR_INSTANTIATE_P(R)                                                                                        \
    template mpsr_status ppe_get_matrix<R>(PpeSolver<R>*, const Dev<R>&, int*, int*, double*, double*, int, int*);   \
    template PpeSolver<R>* ppe_create<R>(const Dev<R>&, const mpsr_params&, const mpsr_ppe_options&, double, double, \
                                         double);                                                                    \
    template void ppe_destroy<R>(PpeSolver<R>*);                                                                     \
    template void ppe_start<R>(PpeSolver<R>*, const Dev<R>&, hipStream_t);                                          \
    template mpsr_status ppe_solve<R>(PpeSolver<R>*, Dev<R>&, const mpsr_ppe_options&, int, hipStream_t);            \
    template mpsr_status ppe_check<R>(PpeSolver<R>*, mpsr_timings*);                                                 \
    template const char* ppe_why<R>(PpeSolver<R>*);                                                                 \
    template mpsr_status ppe_get<R>(PpeSolver<R>*, const Dev<R>&, double*, double*, int*);                           \
    template void ppe_update_params<R>(PpeSolver<R>*, const mpsr_params&);                                         \
    template NbList ppe_neighbors<R>(const PpeSolver<R>*);                                                          \
    template void ppe_begin_call<R>(PpeSolver<R>*);                                                                 \
    template void ppe_last<R>(const PpeSolver<R>*, int*, double*, int*);                                            \
    template mpsr_status ppe_mpnn_set<R>(PpeSolver<R>*, const float*, hipStream_t);                                   \
    template mpsr_status ppe_mpnn_get<R>(PpeSolver<R>*, float*, hipStream_t);                                         \
    template mpsr_status ppe_mpnn_train<R>(PpeSolver<R>*, const Dev<R>&, double, double, int, float*, double*,       \
                                           hipStream_t);                                                              \
    template mpsr_status ppe_mpnn_rows<R>(PpeSolver<R>*, const Dev<R>&, float*, float*, double*, hipStream_t);     \
    template PpeBatch<R>* ppe_batch_create<R>();                                                                    \
    template void ppe_batch_destroy<R>(PpeBatch<R>*);                                                              \
    template mpsr_status ppe_batch_build<R>(PpeBatch<R>*, PpeSolver<R>* const*, const int*, const Dev<R>*,          \
                                            const mpsr_ppe_options*, int, int);                                    \
    template void ppe_batch_pnd<R>(PpeBatch<R>*, const BatchLaunch<R>&, hipStream_t);                              \
    template void ppe_batch_head<R>(PpeBatch<R>*, const BatchLaunch<R>&, int, hipStream_t);                        \
    template void ppe_batch_iterate<R>(PpeBatch<R>*, int, hipStream_t);                                            \
    template mpsr_status ppe_batch_converged<R>(PpeBatch<R>*, hipStream_t, bool*);                                 \
    template void ppe_batch_tail<R>(PpeBatch<R>*, const BatchLaunch<R>&, hipStream_t);                             \
    template mpsr_status ppe_batch_finish<R>(PpeBatch<R>*, int*, std::string*);
MPSR_INSTANTIATE_P(float)
MPSR_INSTANTIATE_P(double)
}  // namespace mpsr
