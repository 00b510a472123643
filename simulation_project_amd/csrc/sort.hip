// sort.hip -- LDS-staged stable LSD radix sort of (key, value) pairs and a device-wide
// exclusive scan, both with an optional device-resident element count.
//
// Replaces thrust::sort_by_key of the reference's sorted-binning variant
// (radix/SPVSOLVER/MultiGPUSolver/DeviceProcess.cu:207-220).  The sort is stable, so sorting
// (cell, slot) pairs produced in slot order yields cells in ascending order and, inside a cell,
// ascending slots: the canonical permutation.
//
// One pass over a digit of <= 8 bits:
//   upsweep   per 2048-key tile, a digit histogram in LDS -> hist[digit * ntiles + tile]
//   scan      exclusive scan of hist (digit-major), one workgroup
//   downsweep per tile: wave64 ballot matching ranks every key among equal digits of its
//             wave, an LDS table turns that into a rank inside the tile, the tile is staged in
//             LDS in digit order and written out so that equal digits go to consecutive
//             addresses.
#include "mpsr_internal.h"

namespace mpsr {

static constexpr int RS_BLOCK = 256;  // 4 waves
static constexpr int RS_WAVES = RS_BLOCK / 64;
// rounds of 64 keys per wave (tile = 256 * items keys); any of 4, 8, 16 sorts the same way,
// the choice only trades histogram size against LDS and registers per tile
static constexpr int RS_ITEMS_MIN = 4;
static constexpr int RS_TILE_MIN = RS_BLOCK * RS_ITEMS_MIN;
static constexpr int SCAN_BLOCK = 1024;
static constexpr int SCAN_ITEMS = 4;
static constexpr int SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;

size_t radix_hist_ints(int n_cap) { return (size_t)256 * ((n_cap + RS_TILE_MIN - 1) / RS_TILE_MIN + 1); }
bool radix_items_valid(int items) { return items == 4 || items == 8 || items == 16; }
size_t scan_tmp_ints(int n_cap) { return (size_t)((n_cap + SCAN_TILE - 1) / SCAN_TILE + 2) + 64; }

// element count: the device-resident count clamped to the capacity, or the host count
__device__ inline int dev_n(const int* d_n, int n_host, int n_cap) { return d_n ? min(*d_n, n_cap) : n_host; }

__device__ inline void fill_int_body(int blk, int* p, int value, long long n) {
    long long i = (long long)blk * blockDim.x + threadIdx.x;
    if (i < n) p[i] = value;
}
__global__ void k_fill_int(int* p, int value, long long n) { fill_int_body(blockIdx.x, p, value, n); }
void fill_int(int* p, int value, long long n, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill_int, dim3(grid_for(n, 256)), dim3(256), 0, st, p, value, n);
}

// ---------------------------------------------------------------------------------------------
// generic exclusive scan: per-tile sums -> scan of tile sums in one workgroup -> tile scans
// ---------------------------------------------------------------------------------------------
__device__ inline void scan_tile_sums_body(int blk, const int* in, const int* d_n, int n_host, int n_cap, int* sums) {
    const int n = dev_n(d_n, n_host, n_cap);
    const int base = blk * SCAN_TILE;
    if (base >= n) return;
    __shared__ int lds[SCAN_BLOCK / 64 + 1];
    int s = 0;
#pragma unroll
    for (int r = 0; r < SCAN_ITEMS; ++r) {
        int i = base + r * SCAN_BLOCK + threadIdx.x;
        if (i < n) s += in[i];
    }
    s = wave_sum(s);
    if (lane_id() == 0) lds[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < SCAN_BLOCK / 64; ++w) t += lds[w];
        sums[blk] = t;
    }
}
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_tile_sums(const int* in, const int* d_n, int n_host, int n_cap,
                                                               int* sums) {
    scan_tile_sums_body(blockIdx.x, in, d_n, n_host, n_cap, sums);
}

// In-place exclusive scan of m = ceil(n / tile) * mult ints in one workgroup; the total to
// *d_total.  Used for tile sums (tile = SCAN_TILE) and for radix histograms (tile = the sort tile,
// mult = the radix).  Each thread owns one run of k consecutive ints (k = m / SCAN_BLOCK rounded
// up to 4): the runs' sums (16-byte loads, issued together), one block scan of them, then every run
// rewritten from its start.  Integer sums, so the result does not depend on the split.
__device__ inline void scan_single_body(int* a, const int* d_n, int n_host, int n_cap, int tile, int mult, int* d_total) {
    const int n = dev_n(d_n, n_host, n_cap);
    const int m = ((n + tile - 1) / tile) * mult;
    __shared__ int lds[SCAN_BLOCK / 64 + 1];
    const int k = ((m + SCAN_BLOCK - 1) / SCAN_BLOCK + 3) & ~3;
    const int i0 = threadIdx.x * k;
    const bool vec = (reinterpret_cast<uintptr_t>(a) & 15) == 0;
    int local = 0;
#pragma unroll 8
    for (int r = 0; r < k; r += 4) {
        const int i = i0 + r;
        if (vec && i + 3 < m) {
            const int4 v = *reinterpret_cast<const int4*>(a + i);
            local += v.x + v.y + v.z + v.w;
        } else {
            for (int u = 0; u < 4; ++u)
                if (i + u < m) local += a[i + u];
        }
    }
    int total;
    int run = block_exclusive_scan<SCAN_BLOCK / 64>(local, lds, &total);
#pragma unroll 8
    for (int r = 0; r < k; r += 4) {
        const int i = i0 + r;
        if (vec && i + 3 < m) {
            const int4 v = *reinterpret_cast<const int4*>(a + i);
            int4 o;
            o.x = run;
            o.y = run + v.x;
            o.z = o.y + v.y;
            o.w = o.z + v.z;
            *reinterpret_cast<int4*>(a + i) = o;
            run = o.w + v.w;
        } else {
            for (int u = 0; u < 4; ++u)
                if (i + u < m) {
                    const int v = a[i + u];
                    a[i + u] = run;
                    run += v;
                }
        }
    }
    if (threadIdx.x == 0) {
        if (d_total) *d_total = total;
    }
}
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_single(int* a, const int* d_n, int n_host, int n_cap, int tile, int mult,
                                                            int* d_total) {
    scan_single_body(a, d_n, n_host, n_cap, tile, mult, d_total);
}

__device__ inline void scan_tiles_body(int blk, const int* in, int* out, const int* d_n, int n_host, int n_cap,
                                       const int* sums) {
    const int n = dev_n(d_n, n_host, n_cap);
    const int base = blk * SCAN_TILE;
    if (base >= n) return;
    __shared__ int lds[SCAN_BLOCK / 64 + 1];
    int v[SCAN_ITEMS];
    int local = 0;
    const int i0 = base + threadIdx.x * SCAN_ITEMS;
#pragma unroll
    for (int r = 0; r < SCAN_ITEMS; ++r) {
        v[r] = (i0 + r < n) ? in[i0 + r] : 0;
        local += v[r];
    }
    int total;
    int run = block_exclusive_scan<SCAN_BLOCK / 64>(local, lds, &total) + sums[blk];
#pragma unroll
    for (int r = 0; r < SCAN_ITEMS; ++r) {
        if (i0 + r < n) out[i0 + r] = run;
        run += v[r];
    }
}
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_tiles(const int* in, int* out, const int* d_n, int n_host,
                                                           int n_cap, const int* sums) {
    scan_tiles_body(blockIdx.x, in, out, d_n, n_host, n_cap, sums);
}

void exclusive_scan(const int* in, int* out, const int* d_n, int n_host, int n_cap, int* d_total,
                    const SortScratch& s, hipStream_t st) {
    const int cap = d_n ? n_cap : n_host;
    if (cap <= 0) {
        if (d_total) fill_int(d_total, 0, 1, st);
        return;
    }
    const int tiles = (cap + SCAN_TILE - 1) / SCAN_TILE;
    hipLaunchKernelGGL(k_scan_tile_sums, dim3(tiles), dim3(SCAN_BLOCK), 0, st, in, d_n, n_host, cap, s.scan_tmp);
    hipLaunchKernelGGL(k_scan_single, dim3(1), dim3(SCAN_BLOCK), 0, st, s.scan_tmp, d_n, n_host, cap, SCAN_TILE, 1,
                       d_total);
    hipLaunchKernelGGL(k_scan_tiles, dim3(tiles), dim3(SCAN_BLOCK), 0, st, in, out, d_n, n_host, cap, s.scan_tmp);
}

// ---------------------------------------------------------------------------------------------
// radix sort
// ---------------------------------------------------------------------------------------------
template <int RS_ITEMS>
__device__ inline void rs_upsweep_body(int b, const uint32_t* keys, const int* d_n, int n_host, int n_cap, int shift,
                                       int nbits, int* hist) {
    constexpr int RS_TILE = RS_BLOCK * RS_ITEMS;
    const int n = dev_n(d_n, n_host, n_cap);
    const int ntiles = (n + RS_TILE - 1) / RS_TILE;
    if (b >= ntiles) return;
    const int radix = 1 << nbits;
    const uint32_t mask = (uint32_t)radix - 1u;
    __shared__ int h[256];
    for (int d = threadIdx.x; d < radix; d += RS_BLOCK) h[d] = 0;
    __syncthreads();
    const int base = b * RS_TILE;
#pragma unroll
    for (int r = 0; r < RS_ITEMS; ++r) {
        int i = base + r * RS_BLOCK + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & mask], 1);
    }
    __syncthreads();
    for (int d = threadIdx.x; d < radix; d += RS_BLOCK) hist[d * ntiles + b] = h[d];
}
template <int RS_ITEMS>
__global__ __launch_bounds__(RS_BLOCK) void k_rs_upsweep(const uint32_t* keys, const int* d_n, int n_host, int n_cap, int shift,
                                                          int nbits, int* hist) {
    rs_upsweep_body<RS_ITEMS>(blockIdx.x, keys, d_n, n_host, n_cap, shift, nbits, hist);
}

template <int RS_ITEMS>
__device__ __attribute__((always_inline)) inline void rs_downsweep_body(int b, const uint32_t* __restrict__ kin,
                                                                         const int* __restrict__ vin,
                                         uint32_t* __restrict__ kout, int* __restrict__ vout, const int* d_n,
                                         int n_host, int n_cap, int shift, int nbits, const int* __restrict__ hist) {
    constexpr int RS_TILE = RS_BLOCK * RS_ITEMS;
    const int n = dev_n(d_n, n_host, n_cap);
    const int ntiles = (n + RS_TILE - 1) / RS_TILE;
    if (b >= ntiles) return;
    const int radix = 1 << nbits;
    const uint32_t mask = (uint32_t)radix - 1u;
    const int lane = lane_id(), w = (int)(threadIdx.x >> 6);

    __shared__ int wh[RS_WAVES][256];      // per-wave running digit counts, later wave prefixes
    __shared__ int tile_off[256];          // digit offset inside the sorted tile
    __shared__ int gbase[256];             // global offset of this tile's first key of a digit
    __shared__ uint32_t sk[RS_TILE];
    __shared__ int sv[RS_TILE];
    __shared__ int lds_scan[RS_WAVES + 1];

    for (int d = threadIdx.x; d < 256; d += RS_BLOCK) {
#pragma unroll
        for (int q = 0; q < RS_WAVES; ++q) wh[q][d] = 0;
    }
    __syncthreads();

    const int base = b * RS_TILE + w * (RS_ITEMS * 64);
    uint32_t key[RS_ITEMS];
    int val[RS_ITEMS], dig[RS_ITEMS], rank[RS_ITEMS];
#pragma unroll
    for (int r = 0; r < RS_ITEMS; ++r) {
        const int i = base + r * 64 + lane;
        const bool valid = i < n;
        key[r] = valid ? kin[i] : 0xffffffffu;
        val[r] = valid ? vin[i] : 0;
        const int d = (int)((key[r] >> shift) & mask);
        dig[r] = valid ? d : -1;
        // lanes of this round holding the same digit
        unsigned long long peers = __ballot(valid);
        for (int bit = 0; bit < nbits; ++bit) {
            const unsigned long long bb = __ballot((d >> bit) & 1);
            peers &= ((d >> bit) & 1) ? bb : ~bb;
        }
        const unsigned long long lower = peers & ((1ull << lane) - 1ull);
        int cnt = 0;
        if (valid) cnt = wh[w][d];
        rank[r] = cnt + __popcll(lower);
        if (valid && lower == 0ull) wh[w][d] = cnt + __popcll(peers);
    }
    __syncthreads();
    // per digit: prefix over waves, tile total, global base
    int tot = 0;
    const int d = threadIdx.x;
    if (d < radix) {
        int run = 0;
#pragma unroll
        for (int q = 0; q < RS_WAVES; ++q) {
            int t = wh[q][d];
            wh[q][d] = run;
            run += t;
        }
        tot = run;
        gbase[d] = hist[d * ntiles + b];
    }
    int total;
    int off = block_exclusive_scan<RS_WAVES>(d < radix ? tot : 0, lds_scan, &total);
    if (d < radix) tile_off[d] = off;
    __syncthreads();
    // stage the tile in digit order
#pragma unroll
    for (int r = 0; r < RS_ITEMS; ++r) {
        if (dig[r] >= 0) {
            const int pos = tile_off[dig[r]] + wh[w][dig[r]] + rank[r];
            sk[pos] = key[r];
            sv[pos] = val[r];
        }
    }
    __syncthreads();
    const int m = min(RS_TILE, n - b * RS_TILE);
    for (int i = threadIdx.x; i < m; i += RS_BLOCK) {
        const uint32_t k = sk[i];
        const int dd = (int)((k >> shift) & mask);
        const int g = gbase[dd] + (i - tile_off[dd]);
        kout[g] = k;
        vout[g] = sv[i];
    }
}
template <int RS_ITEMS>
__global__ __launch_bounds__(RS_BLOCK) void k_rs_downsweep(const uint32_t* __restrict__ kin, const int* __restrict__ vin,
                                                            uint32_t* __restrict__ kout, int* __restrict__ vout,
                                                            const int* d_n, int n_host, int n_cap, int shift, int nbits,
                                                            const int* __restrict__ hist) {
    rs_downsweep_body<RS_ITEMS>(blockIdx.x, kin, vin, kout, vout, d_n, n_host, n_cap, shift, nbits, hist);
}

template <int I>
static int radix_sort_tiles(uint32_t* k0, int* v0, const int* d_n, int n_host, int cap, int bits, const SortScratch& s,
                            hipStream_t st) {
    constexpr int TILE = RS_BLOCK * I;
    const int tiles = (cap + TILE - 1) / TILE;
    const int passes = (bits + 7) / 8;
    const int width = (bits + passes - 1) / passes;
    uint32_t* ka = k0;
    int* va = v0;
    uint32_t* kb = s.k1;
    int* vb = s.v1;
    int which = 0;
    for (int p = 0; p < passes; ++p) {
        const int shift = p * width;
        const int nb = min(width, bits - shift);
        hipLaunchKernelGGL(k_rs_upsweep<I>, dim3(tiles), dim3(RS_BLOCK), 0, st, ka, d_n, n_host, cap, shift, nb, s.hist);
        // exclusive scan of the digit-major histogram: (ntiles * radix) entries
        hipLaunchKernelGGL(k_scan_single, dim3(1), dim3(SCAN_BLOCK), 0, st, s.hist, d_n, n_host, cap, TILE, 1 << nb,
                           (int*)nullptr);
        hipLaunchKernelGGL(k_rs_downsweep<I>, dim3(tiles), dim3(RS_BLOCK), 0, st, ka, va, kb, vb, d_n, n_host, cap, shift,
                           nb, s.hist);
        uint32_t* tk = ka; ka = kb; kb = tk;
        int* tv = va; va = vb; vb = tv;
        which ^= 1;
    }
    return which;
}

int radix_sort_pairs(uint32_t* k0, int* v0, const int* d_n, int n_host, int n_cap, int bits, const SortScratch& s,
                     hipStream_t st, int items) {
    const int cap = d_n ? n_cap : n_host;
    if (cap <= 0 || bits <= 0) return 0;
    switch (items) {
        case 4: return radix_sort_tiles<4>(k0, v0, d_n, n_host, cap, bits, s, st);
        case 16: return radix_sort_tiles<16>(k0, v0, d_n, n_host, cap, bits, s, st);
        default: return radix_sort_tiles<8>(k0, v0, d_n, n_host, cap, bits, s, st);
    }
}

// ---------------------------------------------------------------------------------------------
// batched forms: the same bodies, one work item per (job, block)
// ---------------------------------------------------------------------------------------------
int radix_tile(int items) { return RS_BLOCK * (radix_items_valid(items) ? items : 8); }
int radix_passes(int bits) { return bits > 0 ? (bits + 7) / 8 : 0; }
int radix_pass_bits(int bits, int pass, int* shift) {
    const int passes = radix_passes(bits);
    const int width = (bits + passes - 1) / passes;
    *shift = pass * width;
    return min(width, bits - *shift);
}
int scan_tile() { return SCAN_TILE; }

struct OpFill {
    const FillJob* jobs;
    __device__ void operator()(int j, int b) const { fill_int_body(b, jobs[j].p, jobs[j].value, jobs[j].n); }
};
struct OpScanSums {
    const ScanJob* jobs;
    __device__ void operator()(int j, int b) const {
        const ScanJob& q = jobs[j];
        scan_tile_sums_body(b, q.in, q.d_n, q.n_host, q.d_n ? q.n_cap : q.n_host, q.tmp);
    }
};
struct OpScanSingle {
    const ScanJob* jobs;
    __device__ void operator()(int j, int) const {
        const ScanJob& q = jobs[j];
        scan_single_body(q.tmp, q.d_n, q.n_host, q.d_n ? q.n_cap : q.n_host, SCAN_TILE, 1, q.d_total);
    }
};
struct OpScanTiles {
    const ScanJob* jobs;
    __device__ void operator()(int j, int b) const {
        const ScanJob& q = jobs[j];
        scan_tiles_body(b, q.in, q.out, q.d_n, q.n_host, q.d_n ? q.n_cap : q.n_host, q.tmp);
    }
};
template <int I> struct OpUpsweep {
    const SortJob* jobs;
    __device__ void operator()(int j, int b) const {
        const SortJob& q = jobs[j];
        rs_upsweep_body<I>(b, q.kin, q.d_n, q.n_host, q.n_cap, q.shift, q.nbits, q.hist);
    }
};
template <int I> struct OpHistScan {
    const SortJob* jobs;
    __device__ void operator()(int j, int) const {
        const SortJob& q = jobs[j];
        scan_single_body(q.hist, q.d_n, q.n_host, q.n_cap, RS_BLOCK * I, 1 << q.nbits, nullptr);
    }
};
template <int I> struct OpDownsweep {
    const SortJob* jobs;
    __device__ void operator()(int j, int b) const {
        const SortJob& q = jobs[j];
        rs_downsweep_body<I>(b, q.kin, q.vin, q.kout, q.vout, q.d_n, q.n_host, q.n_cap, q.shift, q.nbits, q.hist);
    }
};

void fill_int_batched(const FillJob* jobs, WorkList blocks, hipStream_t st) {
    launch_batched<256>(OpFill{jobs}, blocks, 256, st);
}
void exclusive_scan_batched(const ScanJob* jobs, WorkList tiles, WorkList singles, hipStream_t st) {
    launch_batched<SCAN_BLOCK>(OpScanSums{jobs}, tiles, SCAN_BLOCK, st);
    launch_batched<SCAN_BLOCK>(OpScanSingle{jobs}, singles, SCAN_BLOCK, st);
    launch_batched<SCAN_BLOCK>(OpScanTiles{jobs}, tiles, SCAN_BLOCK, st);
}
template <int I> static void radix_pass_tiles(const SortJob* jobs, WorkList tiles, WorkList singles, hipStream_t st) {
    launch_batched<RS_BLOCK>(OpUpsweep<I>{jobs}, tiles, RS_BLOCK, st);
    launch_batched<SCAN_BLOCK>(OpHistScan<I>{jobs}, singles, SCAN_BLOCK, st);
    launch_batched<RS_BLOCK>(OpDownsweep<I>{jobs}, tiles, RS_BLOCK, st);
}
void radix_pass_batched(int items, const SortJob* jobs, WorkList tiles, WorkList singles, hipStream_t st) {
    switch (items) {
        case 4: radix_pass_tiles<4>(jobs, tiles, singles, st); break;
        case 16: radix_pass_tiles<16>(jobs, tiles, singles, st); break;
        default: radix_pass_tiles<8>(jobs, tiles, singles, st); break;
    }
}

}  // namespace mpsr
