// walls.hip -- STL wall kernels.
//
//   triangle binning   RegisterTriangleTopology_Kernel / RegisterTriangle / RegisterLine
//                      (Source/DeviceProcess_Kernel.h:443-460, Utility_Inline.h:1747-1857).
//                      The reference appends ids with atomicCAS/atomicInc in arrival order and
//                      wraps at 100 per cell.  Here: every triangle marches its edges and its
//                      shrunken copies twice (count, then emit (cell, triangle) pairs), the pairs
//                      are radix-sorted by cell (stable, so ids ascend inside a cell), duplicates
//                      dropped, and each cell lists its first 100 ids.  Deterministic.
//   k_stl_distance     CalcSTLDistance_Kernel (Kern :635-699), nearest non-drag triangle
//   k_drag             CalcDragEffect_Kernel (Kern :106-281), nearest drag triangle
//   k_translate        UpdateTrianglePosition_Kernel (Kern :463-482)
//   k_rotate           RotateTrianglePosition_Kernel (Kern :930-998); cos/sin of the step's
//                      angle come from the host so the rotation is bit-reproducible
#include "mpsr_internal.h"
#include "walls_geom.h"

namespace mpsr {

static constexpr int kCellMaxTriangles = 100;
static constexpr int kMaxShrink = 4096;
static constexpr int kMaxLineSteps = 1 << 20;


// ---------------------------------------------------------------------------------------------
// triangle binning
// ---------------------------------------------------------------------------------------------
// One RegisterLine march: sample V + TAB*D for TAB = 0, cell, 2*cell, ... while TAB <= Mag.
// Emits a cell only when it differs from the previous emission of this triangle.
template <typename R, typename F>
__device__ inline void march_line(const Consts<R>& c, P3<R> v, P3<R> d, R m, int& last, F&& emit) {
    R tab = 0;
    int steps = 0;
    do {
        const R tx = v.x + tab * d.x, ty = v.y + tab * d.y, tz = v.z + tab * d.z;
        int i, j, k;
        if (cell_coords(c, tx, ty, tz, i, j, k)) {
            const long long cid = cell_of(c, i, j, k);
            if (cid >= 0 && cid < c.ncell && (int)cid != last) {
                last = (int)cid;
                emit((int)cid);
            }
        }
        tab += c.cell;
    } while (tab <= m && ++steps < kMaxLineSteps);
}

template <typename R, typename F> __device__ inline void march_triangle(const Consts<R>& c, const TriD<R>& tr, F&& emit) {
    P3<R> A = p3(tr.v[0], tr.v[1], tr.v[2]), B = p3(tr.v[3], tr.v[4], tr.v[5]), C = p3(tr.v[6], tr.v[7], tr.v[8]);
    for (int q = 0; q < 9; ++q)
        if (!isfinite(tr.v[q])) return;
    const R cs = c.cell;
    R mBC, mAB, mCA;
    int iter = 0, last = -1;
    do {
        const P3<R> AB = psub(B, A), BA = psub(A, B), BC = psub(C, B), CB = psub(B, C), CA = psub(A, C), AC = psub(C, A);
        mBC = pmag(BC); mAB = pmag(AB); mCA = pmag(CA);
        if (mAB > 0.0) march_line(c, A, p3(AB.x / mAB, AB.y / mAB, AB.z / mAB), mAB, last, emit);
        if (mBC > 0.0) march_line(c, B, p3(BC.x / mBC, BC.y / mBC, BC.z / mBC), mBC, last, emit);
        if (mCA > 0.0) march_line(c, C, p3(CA.x / mCA, CA.y / mCA, CA.z / mCA), mCA, last, emit);
        P3<R> AM = padd(AB, AC), BM = padd(BA, BC), CM = padd(CA, CB);
        const R mAM = pmag(AM), mBM = pmag(BM), mCM = pmag(CM);
        if (mAM > 0.0) { AM.x /= mAM; AM.y /= mAM; AM.z /= mAM; A.x += AM.x * cs; A.y += AM.y * cs; A.z += AM.z * cs; }
        if (mBM > 0.0) { BM.x /= mBM; BM.y /= mBM; BM.z /= mBM; B.x += BM.x * cs; B.y += BM.y * cs; B.z += BM.z * cs; }
        if (mCM > 0.0) { CM.x /= mCM; CM.y /= mCM; CM.z /= mCM; C.x += CM.x * cs; C.y += CM.y * cs; C.z += CM.z * cs; }
    } while (mBC >= cs && mAB >= cs && mCA >= cs && ++iter < kMaxShrink);
}

template <typename R>
__device__ inline void tri_count_body(int blk, int T, Consts<R> c, const TriD<R>* __restrict__ tris, int* emit) {
    const int t = blk * blockDim.x + threadIdx.x;
    if (t >= T) return;
    const TriD<R> tr = tris[t];
    int n = 0;
    march_triangle(c, tr, [&](int) { ++n; });
    emit[t] = n;
}
template <typename R>
__global__ __launch_bounds__(64) void k_tri_count(int T, Consts<R> c, const TriD<R>* __restrict__ tris, int* emit) {
    tri_count_body<R>(blockIdx.x, T, c, tris, emit);
}

template <typename R>
__device__ inline void tri_fill_body(int blk, int T, Consts<R> c, const TriD<R>* __restrict__ tris,
                                     const int* __restrict__ off, const int* __restrict__ pair_n,
                                     int pair_cap, uint32_t* pair_cell, int* pair_tri, int* overflow) {
    const int t = blk * blockDim.x + threadIdx.x;
    if (t >= T) return;
    if (*pair_n > pair_cap) {
        if (t == 0) *overflow = 1;
        return;
    }
    const TriD<R> tr = tris[t];
    int e = off[t];
    march_triangle(c, tr, [&](int cid) {
        pair_cell[e] = (uint32_t)cid;
        pair_tri[e] = t;
        ++e;
    });
}
template <typename R>
__global__ __launch_bounds__(64) void k_tri_fill(int T, Consts<R> c, const TriD<R>* __restrict__ tris,
                                                 const int* __restrict__ off, const int* __restrict__ pair_n,
                                                 int pair_cap, uint32_t* pair_cell, int* pair_tri, int* overflow) {
    tri_fill_body<R>(blockIdx.x, T, c, tris, off, pair_n, pair_cap, pair_cell, pair_tri, overflow);
}

// 1 where a sorted pair differs from its predecessor
__device__ inline void pair_flags_body(int blk, const int* pair_n, int pair_cap, const uint32_t* __restrict__ cell,
                                       const int* __restrict__ tri, int* flag) {
    const int e = blk * blockDim.x + threadIdx.x;
    const int n = min(*pair_n, pair_cap);
    if (e >= n) return;
    flag[e] = (e == 0 || cell[e] != cell[e - 1] || tri[e] != tri[e - 1]) ? 1 : 0;
}
__global__ __launch_bounds__(256) void k_pair_flags(const int* pair_n, int pair_cap, const uint32_t* __restrict__ cell,
                                                     const int* __restrict__ tri, int* flag) {
    pair_flags_body(blockIdx.x, pair_n, pair_cap, cell, tri, flag);
}
__device__ inline void pair_compact_body(int blk, const int* pair_n, int pair_cap, const uint32_t* __restrict__ cell,
                                         const int* __restrict__ tri, const int* __restrict__ flag,
                                         const int* __restrict__ pos, uint32_t* ucell, int* utri) {
    const int e = blk * blockDim.x + threadIdx.x;
    const int n = min(*pair_n, pair_cap);
    if (e >= n) return;
    if (flag[e]) {
        ucell[pos[e]] = cell[e];
        utri[pos[e]] = tri[e];
    }
}
__global__ __launch_bounds__(256) void k_pair_compact(const int* pair_n, int pair_cap, const uint32_t* __restrict__ cell,
                                                       const int* __restrict__ tri, const int* __restrict__ flag,
                                                       const int* __restrict__ pos, uint32_t* ucell, int* utri) {
    pair_compact_body(blockIdx.x, pair_n, pair_cap, cell, tri, flag, pos, ucell, utri);
}
// first index and (capped) length of every cell's run in the unique list
__device__ inline void pair_ranges_body(int blk, const int* uniq_n, const uint32_t* __restrict__ ucell, int* cstart,
                                        int* ccount) {
    const int e = blk * blockDim.x + threadIdx.x;
    const int n = *uniq_n;
    if (e >= n) return;
    const uint32_t k = ucell[e];
    if (e == 0 || ucell[e - 1] != k) {
        int f = e + 1;
        while (f < n && ucell[f] == k && f - e < kCellMaxTriangles) ++f;
        cstart[k] = e;
        ccount[k] = f - e;
    }
}
__global__ __launch_bounds__(256) void k_pair_ranges(const int* uniq_n, const uint32_t* __restrict__ ucell, int* cstart,
                                                      int* ccount) {
    pair_ranges_body(blockIdx.x, uniq_n, ucell, cstart, ccount);
}

template <typename R> void launch_tri_count_total(Dev<R>& d, hipStream_t st) {
    if (d.ntri <= 0) {
        fill_int(d.pair_n, 0, 1, st);
        return;
    }
    hipLaunchKernelGGL(k_tri_count<R>, dim3(grid_for(d.ntri, 64)), dim3(64), 0, st, d.ntri, d.c, d.tris, d.tri_emit);
    exclusive_scan(d.tri_emit, d.tri_off, nullptr, d.ntri, d.ntri, d.pair_n, d.tsort, st);
}

// The walk of every cell that holds particles, for nearest_triangle_listed, in two passes, into a
// pool shared by the cells (head[cell] = (offset, count); count -1: no walk, the full one is taken).
// First, per cell: a cell without particles gets -1, one whose stencil lists are all empty gets 0;
// the others reserve min(total, kStencilTris) pool entries (total: the stencil lists' lengths,
// repeats included) and go to a work list (in any order: each cell's walk depends only on its own
// lists); a cell that finds the pool full gets -1 and its work item is skipped.  One workgroup makes
// one reservation in *pool_ctr and one append to *work_ctr.  Every thread of the workgroup calls it;
// it returns the cell's work-list slot, or -1.
__device__ inline int cell_walk_needed(long long cell, int ncell, int gy, int gz, const int* __restrict__ cs,
                                       const int* __restrict__ tcount, int2* head, int* pool_ctr, int cap,
                                       int* work_ctr) {
    int need = 0;
    if (cell < ncell) {
        if (cs[cell] < 0) {
            head[cell] = make_int2(-1, -1);
        } else {
            const long long gyz = (long long)gy * gz;
            const int ci = (int)(cell / gyz), cj = (int)((cell / gz) % gy), ck = (int)(cell % gz);
            int total = 0;
#pragma unroll
            for (int s = 0; s < 27; ++s) {
                const long long m =
                    gyz * (ci + s % 3 - 1) + (long long)gz * (cj + (s / 3) % 3 - 1) + (ck + s / 9 - 1);
                total += (m >= 0 && m < ncell) ? tcount[m] : 0;
            }
            if (total == 0) head[cell] = make_int2(0, 0);
            else need = total < kStencilTris ? total : kStencilTris;
        }
    }
    // the workgroup's reservation and append: wave prefix sums, then one atomic pair per workgroup
    __shared__ int wsum[kMaxBlock / 64], wcnt[kMaxBlock / 64], wbase2[2];
    const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int inc = wave_inclusive_scan(need);
    const unsigned long long mask = __ballot(need > 0);
    if (lane_id() == 63) {
        wsum[w] = inc;
        wcnt[w] = __popcll(mask);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int ps = 0, pc = 0;
        for (int i = 0; i < nw; ++i) {
            const int a = wsum[i], b = wcnt[i];
            wsum[i] = ps;
            wcnt[i] = pc;
            ps += a;
            pc += b;
        }
        wbase2[0] = ps > 0 ? atomicAdd(pool_ctr, ps) : 0;
        wbase2[1] = pc > 0 ? atomicAdd(work_ctr, pc) : 0;
    }
    __syncthreads();
    const int pbase = wbase2[0] + wsum[w], wbase = wbase2[1] + wcnt[w];
    __syncthreads();   // the next call of the workgroup reuses the arrays
    if (need == 0) return -1;
    const int off = pbase + inc - need;
    head[cell] = off > cap - need ? make_int2(-1, -1) : make_int2(off, -1);
    return wbase + __popcll(mask & ((1ull << lane_id()) - 1ull));
}
__global__ __launch_bounds__(256) void k_cell_walk_cells(int ncell, int gy, int gz, const int* __restrict__ cs,
                                                         const int* __restrict__ tcount, int* work, int* ctr,
                                                         int2* head, int cap) {
    const long long cell = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int slot = cell_walk_needed(cell, ncell, gy, gz, cs, tcount, head, &ctr[1], cap, &ctr[0]);
    if (slot >= 0) work[slot] = (int)cell;
}
// Then one wave per listed cell (a workgroup of 64).  Lane s < 27 reads stencil cell s's list (z
// outer, x inner); the ids are then taken 64 at a time in walk order, one per lane, and an id is
// kept when neither the ids kept so far (LDS, read 4 at a time) nor an earlier lane of the same 64
// holds it: its first meeting.  More than kStencilTris kept ids give -1 (the full walk).  The kept
// ids are then ordered by tri_box_bound over the cell's box (bitonic sort in LDS), each with its
// position in the walk, and written at the cell's pool offset for nearest_triangle_listed.
struct WalkLds {
    int off[28], beg[28];
    __align__(16) int kept[kStencilTris];
    unsigned key[kStencilTris];
};
template <typename R>
__device__ inline void cell_walk_wave(long long cell, const Consts<R>& c, const int* __restrict__ tstart,
                                      const int* __restrict__ tcount, const int* __restrict__ tids,
                                      const TriD<R>* __restrict__ tris, int2* ulist, int2* head, WalkLds& L) {
    if (head[cell].x < 0) return;   // the pool was full: the cell keeps the full walk
    const int lane = threadIdx.x;
    const int ncell = c.ncell, gy = c.gy, gz = c.gz;
    const long long gyz = (long long)gy * gz;
    const int ci = (int)(cell / gyz), cj = (int)((cell / gz) % gy), ck = (int)(cell % gz);
    int cnt = 0, b0 = 0;
    if (lane < 27) {
        const long long m = gyz * (ci + lane % 3 - 1) + (long long)gz * (cj + (lane / 3) % 3 - 1) + (ck + lane / 9 - 1);
        if (m >= 0 && m < ncell) {
            cnt = tcount[m];
            b0 = tstart[m];
        }
    }
    const int inc = wave_inclusive_scan(cnt);
    const int total = __shfl(inc, 63, 64);
    if (lane < 27) {
        L.off[lane] = inc - cnt;
        L.beg[lane] = b0;
    }
    if (lane == 0) L.off[27] = total;
    for (int q = lane; q < kStencilTris; q += 64) L.kept[q] = -1;
    __syncthreads();
    int n = 0;
    bool over = false;
    for (int base = 0; base < total && !over; base += 64) {
        const int f = base + lane;
        int tid = -1;
        if (f < total) {
            int s = 0;   // the stencil cell of walk position f: off[s] <= f < off[s + 1]
            for (int h = 16; h > 0; h >>= 1)
                if (s + h < 27 && L.off[s + h] <= f) s += h;
            tid = tids[L.beg[s] + f - L.off[s]];
        }
        bool first = f < total;
        for (int q = 0; q < n; q += 4) {   // unused slots hold -1, never an id
            const int4 k4 = *reinterpret_cast<const int4*>(&L.kept[q]);
            first = first && k4.x != tid && k4.y != tid && k4.z != tid && k4.w != tid;
        }
        for (int j = 0; j < 63; ++j) {
            const int tj = __shfl(tid, j, 64);
            if (j < lane && tj == tid) first = false;
        }
        const unsigned long long mask = __ballot(first);
        const int rank = __popcll(mask & ((1ull << lane) - 1ull));
        const int add = __popcll(mask);
        __syncthreads();
        if (first && n + rank < kStencilTris) L.kept[n + rank] = tid;
        __syncthreads();
        if (n + add > kStencilTris) over = true;
        n = min(n + add, kStencilTris);
    }
    if (over) {
        if (lane == 0) head[cell].y = -1;
        __syncthreads();
        return;
    }
    // sort keys: the bound's float bits with the walk position in the low 10 bits (clearing them
    // lowers a non-negative bound), padding after the n ids up to a power of two
    int n2 = 1;
    while (n2 < n) n2 <<= 1;
    const R lox = c.bminx + (R)ci * c.cell, loy = c.bminy + (R)cj * c.cell, loz = c.bminz + (R)ck * c.cell;
    for (int q = lane; q < n2; q += 64) {
        unsigned key = 0xffffffffu;
        if (q < n) key = (__float_as_uint(tri_box_bound(tris[L.kept[q]], lox, loy, loz, c.cell)) & ~kWalkPosMask) | (unsigned)q;
        L.key[q] = key;
    }
    __syncthreads();
    for (int size = 2; size <= n2; size <<= 1)
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int i = lane; i < (n2 >> 1); i += 64) {
                const int lo = 2 * stride * (i / stride) + (i % stride), hi = lo + stride;
                const unsigned ka = L.key[lo], kb = L.key[hi];
                if ((ka > kb) == ((lo & size) == 0)) {
                    L.key[lo] = kb;
                    L.key[hi] = ka;
                    const int t = L.kept[lo];
                    L.kept[lo] = L.kept[hi];
                    L.kept[hi] = t;
                }
            }
            __syncthreads();
        }
    const int off = head[cell].x;
    for (int q = lane; q < n; q += 64) ulist[off + q] = make_int2(L.kept[q], (int)L.key[q]);
    if (lane == 0) head[cell].y = n;
    __syncthreads();
}
template <typename R>
__global__ __launch_bounds__(64) void k_cell_walks(Consts<R> c, const int* __restrict__ work,
                                                   const int* __restrict__ nwork, const int* __restrict__ tstart,
                                                   const int* __restrict__ tcount, const int* __restrict__ tids,
                                                   const TriD<R>* __restrict__ tris, int2* ulist, int2* head) {
    __shared__ WalkLds L;
    const int nw = *nwork;
    for (int w = blockIdx.x; w < nw; w += gridDim.x)
        cell_walk_wave<R>(work[w], c, tstart, tcount, tids, tris, ulist, head, L);
}

template <typename R> static void launch_cell_stencil_tris(const Dev<R>& d, hipStream_t st) {
    fill_int(d.cell_work_n, 0, 2, st);
    hipLaunchKernelGGL(k_cell_walk_cells, dim3(grid_for(d.c.ncell, 256)), dim3(256), 0, st, (int)d.c.ncell, d.c.gy,
                       d.c.gz, d.cell_start, d.cell_tri_count, d.cell_work, d.cell_work_n, d.cell_uhead, d.utri_cap);
    hipLaunchKernelGGL(k_cell_walks<R>, dim3(2048), dim3(64), 0, st, d.c, d.cell_work, d.cell_work_n,
                       d.cell_tri_start, d.cell_tri_count, d.uniq_tri, d.tris, d.cell_utri, d.cell_uhead);
}

template <typename R> void launch_tri_bin(Dev<R>& d, hipStream_t st) {
    fill_int(d.cell_tri_count, 0, d.c.ncell, st);
    if (d.ntri <= 0) {
        fill_int(d.uniq_n, 0, 1, st);
        launch_cell_stencil_tris(d, st);
        return;
    }
    const int T = d.ntri;
    hipLaunchKernelGGL(k_tri_count<R>, dim3(grid_for(T, 64)), dim3(64), 0, st, T, d.c, d.tris, d.tri_emit);
    exclusive_scan(d.tri_emit, d.tri_off, nullptr, T, T, d.pair_n, d.tsort, st);
    hipLaunchKernelGGL(k_tri_fill<R>, dim3(grid_for(T, 64)), dim3(64), 0, st, T, d.c, d.tris, d.tri_off, d.pair_n,
                       d.pair_cap, d.pair_cell, d.pair_tri, d.overflow);
    const int bits = tri_pair_bits(d.c.ncell);
    const int which = radix_sort_pairs(d.pair_cell, d.pair_tri, d.pair_n, 0, d.pair_cap, bits, d.tsort, st);
    const uint32_t* sc = which ? d.tsort.k1 : d.pair_cell;
    const int* stri = which ? d.tsort.v1 : d.pair_tri;
    const unsigned g = grid_for(d.pair_cap, 256);
    hipLaunchKernelGGL(k_pair_flags, dim3(g), dim3(256), 0, st, d.pair_n, d.pair_cap, sc, stri, d.pair_flag);
    exclusive_scan(d.pair_flag, d.pair_pos, d.pair_n, 0, d.pair_cap, d.uniq_n, d.tsort, st);
    hipLaunchKernelGGL(k_pair_compact, dim3(g), dim3(256), 0, st, d.pair_n, d.pair_cap, sc, stri, d.pair_flag,
                       d.pair_pos, d.uniq_cell, d.uniq_tri);
    hipLaunchKernelGGL(k_pair_ranges, dim3(g), dim3(256), 0, st, d.uniq_n, d.uniq_cell, d.cell_tri_start,
                       d.cell_tri_count);
    launch_cell_stencil_tris(d, st);
}

template <typename R> __device__ inline R4<R> mk4w(R x, R y, R z);
template <> __device__ inline float4 mk4w<float>(float x, float y, float z) { return make_float4(x, y, z, 0.f); }
template <> __device__ inline double4 mk4w<double>(double x, double y, double z) { return make_double4(x, y, z, 0.0); }

template <typename R>
__device__ inline void stl_distance_body(int blk, int N, Consts<R> c, const R4<R>* __restrict__ spos,
                                         const int* __restrict__ tstart, const int* __restrict__ tcount,
                                         const int* __restrict__ tids, const TriD<R>* __restrict__ tris,
                                         R* stl_mag, R4<R>* stl_dir, int* stl_id, const int2* __restrict__ ulist,
                                         const int2* __restrict__ uhead) {
    const int t = blk * blockDim.x + threadIdx.x;
    if (t >= N) return;
    const R4<R> P = spos[t];
    R best;
    P3<R> dir;
    const int bid = nearest_triangle_listed(c, P.x, P.y, P.z, ulist, uhead, tstart, tcount, tids, tris, 0, &best,
                                            &dir);
    stl_mag[t] = best;
    stl_dir[t] = mk4w<R>(dir.x, dir.y, dir.z);
    stl_id[t] = bid;
}
template <typename R>
__global__ __launch_bounds__(kMaxBlock) void k_stl_distance(int N, Consts<R> c, const R4<R>* __restrict__ spos,
                                                       const int* __restrict__ tstart, const int* __restrict__ tcount,
                                                       const int* __restrict__ tids, const TriD<R>* __restrict__ tris,
                                                       R* stl_mag, R4<R>* stl_dir, int* stl_id,
                                                       const int2* __restrict__ ulist, const int2* __restrict__ uhead) {
    stl_distance_body<R>(blockIdx.x, N, c, spos, tstart, tcount, tids, tris, stl_mag, stl_dir, stl_id, ulist, uhead);
}


template <typename R>
__device__ inline void drag_body(int blk, int N, Consts<R> c, const int* __restrict__ perm, PState<R> in,
                                 PState<R> out, R4<R>* __restrict__ spos, R4<R>* __restrict__ svel,
                                 const int* __restrict__ tstart, const int* __restrict__ tcount,
                                 const int* __restrict__ tids, const TriD<R>* __restrict__ tris,
                                 const mpsr_tri_params* __restrict__ models,
                                 const mpsr_drag_params* __restrict__ drag, R4<R>* drag_acc,
                                 R* drag_temp, const int2* __restrict__ ulist, const int2* __restrict__ uhead) {
    const int t = blk * blockDim.x + threadIdx.x;
    if (t >= N) return;
    const int s = perm[t];
    if (out.id[s] < 0) return;
    const R4<R> P = spos[t];
    R best;
    P3<R> dir;
    const int bid = nearest_triangle_listed(c, P.x, P.y, P.z, ulist, uhead, tstart, tcount, tids, tris, 1, &best,
                                            &dir);
    drag_acc[t] = mk4w<R>(0, 0, 0);
    drag_temp[t] = 0;
    out.type[s] = 0;
    if (!(bid != -1 && best != (R)1e8)) return;
    const mpsr_drag_params dp = drag[models[tris[bid].model].DragTriangleID];
    if (!(best <= (R)dp.CoefficientofInflueneRegion * c.l0)) return;
    out.type[s] = 2;
    const R4<R> V = svel[t];
    if (dp.DragToken == 'E') {
        const char k1 = dp.InputVelocityToken1;
        if (k1 == 'V' || k1 == 'B') {
            const R sv = build_sqrt(V.x * V.x + V.y * V.y + V.z * V.z);
            R pr = 0;
            pr += (R)dp.VelocityIndex * powd(sv, (R)dp.VelocityExponential);
            const TriD<R>& tr = tris[bid];
            const R kp = (R)dp.VelocityDragPermiabilityConstant;
            if (kp != 0) drag_acc[t] = mk4w<R>(pr * tr.nrm[0] / kp, pr * tr.nrm[1] / kp, pr * tr.nrm[2] / kp);
        }
        if (k1 == 'T' || k1 == 'B') {
            const R sv = build_sqrt(V.x * V.x + V.y * V.y + V.z * V.z);
            const R en = (R)dp.TempIndex * powd(sv, (R)dp.TempExponential);
            const R den = build_sqrt(V.x * V.x + V.y * V.y + V.z * V.z) * c.density * c.cp;
            R dtmp = 0;
            if (den != 0) dtmp = (en * (R)dp.TempDragPermiabilityConstant * c.dt * c.thermal_step) / den;
            drag_temp[t] = dtmp;
            out.solid[s] = dtmp;
        }
    } else if (dp.DragToken == 'C') {
        if (dp.InputVelocityToken1 == 'V' || dp.InputVelocityToken1 == 'v') {
            const R cx = (R)dp.ConstantVx, cy = (R)dp.ConstantVy, cz = (R)dp.ConstantVz;
            out.vx[s] = cx; out.vy[s] = cy; out.vz[s] = cz;
            in.vx[s] = cx; in.vy[s] = cy; in.vz[s] = cz;
            svel[t] = mk4w<R>(cx, cy, cz);
        }
    } else if (dp.DragToken == 'F') {
        if (dp.InputVelocityToken1 == 'V' || dp.InputVelocityToken1 == 'v') {
            const R f = (R)dp.VelocityMagnitueFactor;
            const R nx = V.x * f, ny = V.y * f, nz = V.z * f;
            in.vx[s] = nx; in.vy[s] = ny; in.vz[s] = nz;
            out.vx[s] = out.vx[s] * f; out.vy[s] = out.vy[s] * f; out.vz[s] = out.vz[s] * f;
            svel[t] = mk4w<R>(nx, ny, nz);
        }
    }
}
template <typename R>
__global__ __launch_bounds__(kMaxBlock) void k_drag(int N, Consts<R> c, const int* __restrict__ perm, PState<R> in,
                                               PState<R> out, R4<R>* __restrict__ spos, R4<R>* __restrict__ svel,
                                               const int* __restrict__ tstart, const int* __restrict__ tcount,
                                               const int* __restrict__ tids, const TriD<R>* __restrict__ tris,
                                               const mpsr_tri_params* __restrict__ models,
                                               const mpsr_drag_params* __restrict__ drag, R4<R>* drag_acc,
                                               R* drag_temp, const int2* __restrict__ ulist,
                                               const int2* __restrict__ uhead) {
    drag_body<R>(blockIdx.x, N, c, perm, in, out, spos, svel, tstart, tcount, tids, tris, models, drag, drag_acc,
                 drag_temp, ulist, uhead);
}

template <typename R>
__device__ inline void translate_body(int blk, int T, R dt, TriD<R>* tris, const mpsr_tri_params* __restrict__ models) {
    const int t = blk * blockDim.x + threadIdx.x;
    if (t >= T) return;
    TriD<R> tr = tris[t];
    const mpsr_float3 mv = models[tr.model].Velocity;
    R vx = (R)mv.x, vy = (R)mv.y, vz = (R)mv.z;
    vx = vx * dt; vy = vy * dt; vz = vz * dt;
    for (int k = 0; k < 3; ++k) {
        tr.v[3 * k] = tr.v[3 * k] + vx;
        tr.v[3 * k + 1] = tr.v[3 * k + 1] + vy;
        tr.v[3 * k + 2] = tr.v[3 * k + 2] + vz;
    }
    tris[t] = tr;
}
template <typename R>
__global__ __launch_bounds__(256) void k_translate(int T, R dt, TriD<R>* tris, const mpsr_tri_params* __restrict__ models) {
    translate_body<R>(blockIdx.x, T, dt, tris, models);
}

template <typename R>
__device__ inline void rotate_body(int blk, int T, TriD<R>* tris, const RotCoef<R>* __restrict__ rot,
                                   const int* __restrict__ step_ctr, int nmodel) {
    const int t = blk * blockDim.x + threadIdx.x;
    if (t >= T) return;
    TriD<R> tr = tris[t];
    const RotCoef<R> r = rot[(step_ctr ? (size_t)(*step_ctr) * nmodel : 0) + tr.model];
    if (!r.active) return;
    const R a = r.a, b = r.b, cc = r.c, u = r.u, v = r.v, w = r.w, cs = r.cs, sn = r.sn, omc = r.omc;
    for (int k = 0; k < 3; ++k) {
        const R x = tr.v[3 * k], y = tr.v[3 * k + 1], z = tr.v[3 * k + 2];
        tr.v[3 * k] = (a * (v * v + w * w) - (u * (b * v + cc * w - u * x - v * y - w * z))) * omc + (x * cs) +
                      (((-cc * v) + b * w - w * y + v * z) * sn);
        tr.v[3 * k + 1] = (b * (u * u + w * w) - (v * (a * u + cc * w - u * x - v * y - w * z))) * omc + (y * cs) +
                          (((cc * u) - a * w + w * x - u * z) * sn);
        tr.v[3 * k + 2] = (cc * (u * u + v * v) - (w * (a * u + b * v - u * x - v * y - w * z))) * omc + (z * cs) +
                          (((-b * u) + a * v - v * x + u * y) * sn);
    }
    tris[t] = tr;
}
template <typename R>
__global__ __launch_bounds__(256) void k_rotate(int T, TriD<R>* tris, const RotCoef<R>* __restrict__ rot,
                                                const int* __restrict__ step_ctr, int nmodel) {
    rotate_body<R>(blockIdx.x, T, tris, rot, step_ctr, nmodel);
}

template <typename R> void launch_stl_distance(const Dev<R>& d, hipStream_t st) {
    if (d.N <= 0) return;
    hipLaunchKernelGGL(k_stl_distance<R>, dim3(grid_for(d.N, d.block)), dim3(d.block), 0, st, d.N, d.c, d.spos,
                       d.cell_tri_start, d.cell_tri_count, d.uniq_tri, d.tris, d.stl_mag, d.stl_dir, d.stl_id,
                       d.cell_utri, d.cell_uhead);
}
template <typename R> void launch_drag(const Dev<R>& d, hipStream_t st) {
    if (d.N <= 0) return;
    hipLaunchKernelGGL(k_drag<R>, dim3(grid_for(d.N, d.block)), dim3(d.block), 0, st, d.N, d.c, d.perm, d.in, d.out, d.spos,
                       d.svel, d.cell_tri_start, d.cell_tri_count, d.uniq_tri, d.tris, d.models, d.drag, d.drag_acc,
                       d.drag_temp, d.cell_utri, d.cell_uhead);
}
template <typename R> void launch_translate(const Dev<R>& d, hipStream_t st) {
    if (d.ntri <= 0) return;
    hipLaunchKernelGGL(k_translate<R>, dim3(grid_for(d.ntri, 256)), dim3(256), 0, st, d.ntri, d.c.dt, d.tris,
                       d.models);
}
template <typename R> void launch_rotate(const Dev<R>& d, hipStream_t st) {
    if (d.ntri <= 0) return;
    hipLaunchKernelGGL(k_rotate<R>, dim3(grid_for(d.ntri, 256)), dim3(256), 0, st, d.ntri, d.tris, d.rot, d.step_ctr,
                       d.nmodel > 0 ? d.nmodel : 1);
}

// advances the step counter of graph replays (one thread, a vector store)
__global__ void k_step_advance(int* ctr) {
    if (blockIdx.x == 0 && threadIdx.x == 0) ctr[0] = ctr[0] + 1;
}
void launch_step_advance(int* ctr, hipStream_t st) { hipLaunchKernelGGL(k_step_advance, dim3(1), dim3(64), 0, st, ctr); }

// ---- batched forms: one work item = one block of one domain's particles or triangles ---------
template <typename R> struct OpTriCount {
    const Dev<R>* dv;
    __device__ void operator()(int k, int b) const {
        const Dev<R>& d = dv[k];
        tri_count_body<R>(b, d.ntri, d.c, d.tris, d.tri_emit);
    }
};
template <typename R> struct OpTriFill {
    const Dev<R>* dv;
    __device__ void operator()(int k, int b) const {
        const Dev<R>& d = dv[k];
        tri_fill_body<R>(b, d.ntri, d.c, d.tris, d.tri_off, d.pair_n, d.pair_cap, d.pair_cell, d.pair_tri, d.overflow);
    }
};
// the sorted pairs are where the pair sort's passes left them (radix_sort_pairs' parity)
template <typename R> struct OpPairs {
    const Dev<R>* dv;
    int pass;
    __device__ void operator()(int k, int b) const {
        const Dev<R>& d = dv[k];
        const bool odd = radix_passes_dev(tri_pair_bits(d.c.ncell)) & 1;
        const uint32_t* sc = odd ? d.tsort.k1 : d.pair_cell;
        const int* stri = odd ? d.tsort.v1 : d.pair_tri;
        if (pass == 0) pair_flags_body(b, d.pair_n, d.pair_cap, sc, stri, d.pair_flag);
        else if (pass == 1) pair_compact_body(b, d.pair_n, d.pair_cap, sc, stri, d.pair_flag, d.pair_pos, d.uniq_cell,
                                              d.uniq_tri);
        else pair_ranges_body(b, d.uniq_n, d.uniq_cell, d.cell_tri_start, d.cell_tri_count);
    }
};
template <typename R> struct OpStl {
    const Dev<R>* dv;
    __device__ void operator()(int k, int b) const {
        const Dev<R>& d = dv[k];
        stl_distance_body<R>(b, d.N, d.c, d.spos, d.cell_tri_start, d.cell_tri_count, d.uniq_tri, d.tris, d.stl_mag,
                             d.stl_dir, d.stl_id, d.cell_utri, d.cell_uhead);
    }
};
template <typename R> struct OpDrag {
    const Dev<R>* dv;
    __device__ void operator()(int k, int b) const {
        const Dev<R>& d = dv[k];
        drag_body<R>(b, d.N, d.c, d.perm, d.in, d.out, d.spos, d.svel, d.cell_tri_start, d.cell_tri_count, d.uniq_tri,
                     d.tris, d.models, d.drag, d.drag_acc, d.drag_temp, d.cell_utri, d.cell_uhead);
    }
};
// the walks of the batched step: (bucket, block of 256 cells) items sort the cells as
// k_cell_walk_cells does and append (bucket, cell) to one list; one wave per listed cell then
template <typename R> struct OpCellWalkCells {
    const Dev<R>* dv;
    int2* work;
    int* nwork;
    __device__ void operator()(int k, int b) const {
        const Dev<R>& d = dv[k];
        const long long cell = (long long)b * 256 + threadIdx.x;
        const int slot = cell_walk_needed(cell, (int)d.c.ncell, d.c.gy, d.c.gz, d.cell_start, d.cell_tri_count,
                                          d.cell_uhead, &d.cell_work_n[1], d.utri_cap, nwork);
        if (slot >= 0) work[slot] = make_int2(k, (int)cell);
    }
};
template <typename R>
__global__ __launch_bounds__(64) void k_cell_walks_b(const Dev<R>* __restrict__ dv, const int2* __restrict__ work,
                                                     const int* __restrict__ nwork) {
    __shared__ WalkLds L;
    const int nw = *nwork;
    for (int w = blockIdx.x; w < nw; w += gridDim.x) {
        const int2 it = work[w];
        const Dev<R>& d = dv[it.x];
        cell_walk_wave<R>(it.y, d.c, d.cell_tri_start, d.cell_tri_count, d.uniq_tri, d.tris, d.cell_utri,
                          d.cell_uhead, L);
    }
}
template <typename R> struct OpTranslate {
    const Dev<R>* dv;
    __device__ void operator()(int k, int b) const {
        const Dev<R>& d = dv[k];
        translate_body<R>(b, d.ntri, d.c.dt, d.tris, d.models);
    }
};
template <typename R> struct OpRotate {
    const Dev<R>* dv;
    __device__ void operator()(int k, int b) const {
        const Dev<R>& d = dv[k];
        rotate_body<R>(b, d.ntri, d.tris, d.rot, d.step_ctr, d.nmodel > 0 ? d.nmodel : 1);
    }
};
template <typename R> struct OpStepAdvance {
    const Dev<R>* dv;
    __device__ void operator()(int k, int) const {
        int* ctr = const_cast<int*>(dv[k].step_ctr);
        if (threadIdx.x == 0 && ctr) ctr[0] = ctr[0] + 1;
    }
};
template <typename R> void launch_tri_count_b(const BatchLaunch<R>& b, WorkList t64, hipStream_t st) {
    launch_batched<64>(OpTriCount<R>{b.devs}, t64, 64, st);
}
template <typename R> void launch_tri_fill_b(const BatchLaunch<R>& b, WorkList t64, hipStream_t st) {
    launch_batched<64>(OpTriFill<R>{b.devs}, t64, 64, st);
}
template <typename R> void launch_pairs_b(const BatchLaunch<R>& b, int pass, WorkList pairs, hipStream_t st) {
    launch_batched<256>(OpPairs<R>{b.devs, pass}, pairs, 256, st);
}
template <typename R>
void launch_cell_walks_b(const BatchLaunch<R>& b, WorkList c256, int2* work, int* nwork, hipStream_t st) {
    fill_int(nwork, 0, 1, st);
    launch_batched<256>(OpCellWalkCells<R>{b.devs, work, nwork}, c256, 256, st);
    hipLaunchKernelGGL(k_cell_walks_b<R>, dim3(2048), dim3(64), 0, st, b.devs, work, nwork);
}
template <typename R> void launch_stl_distance_b(const BatchLaunch<R>& b, WorkList p, hipStream_t st) {
    launch_batched<kMaxBlock>(OpStl<R>{b.devs}, p, b.block, st);
}
template <typename R> void launch_drag_b(const BatchLaunch<R>& b, WorkList p, hipStream_t st) {
    launch_batched<kMaxBlock>(OpDrag<R>{b.devs}, p, b.block, st);
}
template <typename R> void launch_translate_b(const BatchLaunch<R>& b, WorkList t256, hipStream_t st) {
    launch_batched<256>(OpTranslate<R>{b.devs}, t256, 256, st);
}
template <typename R> void launch_rotate_b(const BatchLaunch<R>& b, WorkList t256, hipStream_t st) {
    launch_batched<256>(OpRotate<R>{b.devs}, t256, 256, st);
}
template <typename R> void launch_step_advance_b(const BatchLaunch<R>& b, WorkList one, hipStream_t st) {
    launch_batched<64>(OpStepAdvance<R>{b.devs}, one, 64, st);
}

#define MPSR_INSTANTIATE_W(R)                                               \
    template void launch_tri_bin<R>(Dev<R>&, hipStream_t);                  \
    template void launch_tri_count_total<R>(Dev<R>&, hipStream_t);          \
    template void launch_stl_distance<R>(const Dev<R>&, hipStream_t);       \
    template void launch_drag<R>(const Dev<R>&, hipStream_t);               \
    template void launch_translate<R>(const Dev<R>&, hipStream_t);          \
    template void launch_rotate<R>(const Dev<R>&, hipStream_t);                                  \
    template void launch_tri_count_b<R>(const BatchLaunch<R>&, WorkList, hipStream_t);           \
    template void launch_tri_fill_b<R>(const BatchLaunch<R>&, WorkList, hipStream_t);            \
    template void launch_pairs_b<R>(const BatchLaunch<R>&, int, WorkList, hipStream_t);          \
    template void launch_cell_walks_b<R>(const BatchLaunch<R>&, WorkList, int2*, int*, hipStream_t); \
    template void launch_stl_distance_b<R>(const BatchLaunch<R>&, WorkList, hipStream_t);        \
    template void launch_drag_b<R>(const BatchLaunch<R>&, WorkList, hipStream_t);                \
    template void launch_translate_b<R>(const BatchLaunch<R>&, WorkList, hipStream_t);           \
    template void launch_rotate_b<R>(const BatchLaunch<R>&, WorkList, hipStream_t);              \
    template void launch_step_advance_b<R>(const BatchLaunch<R>&, WorkList, hipStream_t);
MPSR_INSTANTIATE_W(float)
MPSR_INSTANTIATE_W(double)

}  // namespace mpsr
