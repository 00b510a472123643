// mpsr_internal.h -- types shared by the HIP translation units of libmpsr.
//
// Everything here is MI355X (gfx950, wave64) code.  Kernels are templated on the build
// precision R (float: the reference's Scalar, with its sqrt14 distances; double: the FP64 build
// with exact square roots).  Particle state lives in structure-of-arrays form in HBM; the
// neighbour kernels read 16 B (float4) or 32 B (double4) position/velocity records gathered
// into cell-sorted order so that one load brings a whole neighbour.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>
#include "../../include/mpsr.h"

namespace mpsr {

template <typename R> struct Vec4;
template <> struct Vec4<float> { using T = float4; };
template <> struct Vec4<double> { using T = double4; };
template <typename R> using R4 = typename Vec4<R>::T;

// ---- the reference's scalar helpers, restated for the device --------------------------------

// solver/Common/inc/Utility_Inline.h:13-26 (sqrt14): one Newton step from the 0x5f3759df guess
__host__ __device__ inline float sqrt14(float x) {
    const float xhalf = 0.5f * x;
    int32_t i = __builtin_bit_cast(int32_t, x);
    i = 0x5f3759df - (i >> 1);
    const float y = __builtin_bit_cast(float, i);
    return x * y * (1.5f - xhalf * y * y);
}
__host__ __device__ inline float build_sqrt(float x) { return sqrt14(x); }
__host__ __device__ inline double build_sqrt(double x) { return sqrt(x); }

// Utility_Inline.h:171-183 GetWeight
template <typename R> __host__ __device__ inline R weight(R re, R r) {
    if (r > re) return R(0);
    if (r > 0.0) return re / r - 1;
    return R(1e6);
}

// Utility_Inline.h:199-336 (gradient) and :513-686 (Laplacian) wall weight tables.  The
// reference compares a float ratio against double literals, so the knots are compared in double.
struct ZTable { int n; double knot[25]; double a[25]; double c[25]; };

template <typename R> __host__ __device__ inline R ztable_eval(const ZTable& t, R r, R l0) {
    const R x = r / l0;
    const double xd = (double)x;
    for (int s = 0; s < t.n; ++s) {
        if (xd <= t.knot[s]) {
            const R a = (R)t.a[s], b = (R)t.knot[s], c = (R)t.c[s];
            return a * (x - b) + c;
        }
    }
    return R(0);
}

// Per-launch constants, computed on the host from mpsr_params with the reference's own
// expressions (Kern = Source/DeviceProcess_Kernel.h).
template <typename R> struct Consts {
    // grid (CGridBox)
    R bminx, bminy, bminz;
    R cminx, cminy, cminz, cmaxx, cmaxy, cmaxz;
    R cell;
    int gx, gy, gz, ncell;
    R tol, l0, dt;
    // explicit predictor, Kern :283-441
    R re_lap, mind_lap, lapc, rc, art_p, art_coeff, mu_wall;
    R gx_acc, gy_acc, gz_acc;
    int expl, has_drag;
    int has_sph;       // spheres present: the predictor adds their reaction
    // PND / EOS, Kern :702-801
    R re_grad, mind_grad, eos_coef, n0, beta_n0;
    // gradient corrector, Kern :804-912
    R c1, c2, c3, vmax;
    // drag, Kern :106-281
    R density, cp, thermal_step;
    int ntri;
    // heat conduction, Kern :485-632
    R re_temp, mind_temp, enth, kf, inv_kf, rf, tmax, tmin, visc_a, visc_b;
    // the radix variant's terms: Smagorinsky viscosity (radix DeviceProcess_Kernel.h:1273-1476)
    // and the negative-pressure gradient (:1214-1241); both off unless set by the caller
    int turb, negp;
    R turb_g;                 // Dimension / ConstantParticleDensity
    R turb_cs, turb_fw;       // SmagorinskyConstant, FilterWidth
    R turb_numax, turb_kmax;  // TurbulanViscosityRatio, TurbulanceIntensity
    R turb_tk;                // 2 Cnue l0 l0 / Cepsilon
};

// per-model negative-pressure fields (radix CTriangleParameters, DataType.h:160-168)
template <typename R> struct NegModel {
    int active;   // RotationInRPM != 0 || NegativePressureCheck
    R infl, pgas, dx, dy, dz;
};

// position -> cell coordinates, Utility_Inline.h:1636-1674 (C truncation toward zero)
template <typename R>
__device__ inline bool cell_coords(const Consts<R>& c, R x, R y, R z, int& i, int& j, int& k) {
    R rx = x - c.bminx, ry = y - c.bminy, rz = z - c.bminz;
    rx /= c.cell; ry /= c.cell; rz /= c.cell;
    if (!(fabs((double)rx) < 1.0e9 && fabs((double)ry) < 1.0e9 && fabs((double)rz) < 1.0e9)) return false;
    i = (int)rx; j = (int)ry; k = (int)rz;
    return true;
}
template <typename R> __device__ inline long long cell_of(const Consts<R>& c, int i, int j, int k) {
    return (long long)c.gy * c.gz * i + (long long)c.gz * j + k;
}
// RegisterParticleTopology's binning key: the cell of a live particle (id >= 0) whose cell lies
// inside the grid, else ncell ("not binned", sorted after every cell)
template <typename R> __device__ inline uint32_t bin_key(const Consts<R>& c, int id, R x, R y, R z) {
    uint32_t key = (uint32_t)c.ncell;
    if (id < 0) return key;
    int i, j, k;
    if (cell_coords(c, x, y, z, i, j, k)) {
        const long long m = cell_of(c, i, j, k);
        if (m >= 0 && m < c.ncell) key = (uint32_t)m;
    }
    return key;
}
// Utility_Inline.h:1936-1975 IsInclude
template <typename R> __device__ inline bool include(const Consts<R>& c, R x, R y, R z) {
    if (!isfinite(x) || !isfinite(y) || !isfinite(z)) return false;
    if (x < c.cminx - c.tol) return false;
    if (y < c.cminy - c.tol) return false;
    if (z < c.cminz - c.tol) return false;
    if (x >= c.cmaxx + c.tol) return false;
    if (y >= c.cmaxy + c.tol) return false;
    if (z >= c.cmaxz + c.tol) return false;
    return true;
}

// Slot-order particle state (the reference's c_daParticle* / c_daOutputParticle* arrays).
template <typename R> struct PState {
    int* id;
    int* type;
    R *px, *py, *pz, *vx, *vy, *vz;
    R *p, *n, *temp, *nu, *solid;
};

// most triangles of a per-cell triangle walk (nearest_triangle_listed); a walk entry's key keeps
// the walk position in its low 10 bits
static constexpr int kStencilTris = 1024;
static constexpr unsigned kWalkPosMask = 1023u;

// Triangle record on the device: three vertices, the stored normal, the model, the model's
// drag flag.  64 B for float, 112 B for double.

template <typename R> struct alignas(16) TriD {
    R v[9];
    R nrm[3];
    int model;
    int is_drag;
    int pad[2];
};

// Rotation of one model for one step (Kern :930-998), cos/sin evaluated on the host.
template <typename R> struct RotCoef {
    R a, b, c, u, v, w, cs, sn, omc;
    int active;
};

// ---- wave64 / block primitives --------------------------------------------------------------
__device__ inline int lane_id() { return (int)(threadIdx.x & 63); }

__device__ inline int wave_inclusive_scan(int v) {
    const int lane = lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}
__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ inline float wave_sum(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ inline int wave_sum(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Exclusive scan over a block of NW waves; `lds` holds NW + 1 ints.  Returns the exclusive
// prefix of v and stores the block total in *total.
template <int NW> __device__ inline int block_exclusive_scan(int v, int* lds, int* total) {
    const int lane = lane_id(), w = (int)(threadIdx.x >> 6);
    const int inc = wave_inclusive_scan(v);
    if (lane == 63) lds[w] = inc;
    __syncthreads();
    if (w == 0) {
        int s = lane < NW ? lds[lane] : 0;
        int si = wave_inclusive_scan(s);
        if (lane < NW) lds[lane] = si - s;
        if (lane == NW - 1) lds[NW] = si;
    }
    __syncthreads();
    const int r = inc - v + lds[w];
    *total = lds[NW];
    __syncthreads();
    return r;
}

// ---- launch helpers shared by the translation units ------------------------------------------
inline unsigned grid_for(long long n, int block) { return (unsigned)((n + block - 1) / block); }

// XCD-contiguous workgroup order for the neighbour kernels.  The GPU deals the workgroups of a
// launch to its 8 XCDs in turn (workgroup b runs on XCD b % 8) and every XCD has its own L2, so in
// launch order each L2 sees every eighth block of sorted particles and fetches nearly the whole
// candidate set.  Launched with xcd_grid blocks, the workgroups of one XCD take one contiguous run of
// the work blocks instead, and the cells a block's candidates come from are mostly the ones its
// neighbouring blocks on the same L2 read.
static constexpr int kXcds = 8;
inline unsigned xcd_grid(long long n, int block) {
    return (grid_for(n, block) + kXcds - 1) / kXcds * kXcds;
}
// the work block of this workgroup in a launch of xcd_grid(n, blockDim.x) workgroups, -1 for padding
__device__ inline int xcd_block(int n) {
    const unsigned nb = ((unsigned)n + blockDim.x - 1) / blockDim.x;
    const unsigned per = (nb + kXcds - 1) / kXcds;
    const unsigned b = (blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
    return b < nb ? (int)b : -1;
}

// PPE matrix layout (SELL-64/4): rows in slices of 64, every slice as wide as its longest row
// rounded up to a multiple of 4, entries kept in groups of 4 per row: entry k of the slice's row
// L sits at slice_base + 256 (k / 4) + 4 L + k % 4.  A lane's 4 consecutive entries are 16
// contiguous bytes (int32 / FP32) and a wave's group of 4 entries is one contiguous 1 KB run.
__host__ __device__ inline int sell_at(int slice_base, int lane, int k) {
    return slice_base + ((k >> 2) << 8) + (lane << 2) + (k & 3);
}
typedef int v4i __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef double v2d __attribute__((ext_vector_type(2)));
// four consecutive matrix entries (16-byte aligned), loaded non-temporally or stored at once
__device__ inline void quad_load(const int* p, int* o) {
    const v4i v = __builtin_nontemporal_load(reinterpret_cast<const v4i*>(p));
    o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w;
}
__device__ inline void quad_load(const float* p, float* o) {
    const v4f v = __builtin_nontemporal_load(reinterpret_cast<const v4f*>(p));
    o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w;
}
__device__ inline void quad_load(const double* p, double* o) {
    const v2d a = __builtin_nontemporal_load(reinterpret_cast<const v2d*>(p));
    const v2d b = __builtin_nontemporal_load(reinterpret_cast<const v2d*>(p + 2));
    o[0] = a.x; o[1] = a.y; o[2] = b.x; o[3] = b.y;
}
// four consecutive 16-bit matrix column codes (8-byte aligned), loaded non-temporally
__device__ inline void quad_load(const uint16_t* p, int* o) {
    const unsigned long long v = __builtin_nontemporal_load(reinterpret_cast<const unsigned long long*>(p));
    o[0] = (int)(v & 0xffffu);
    o[1] = (int)((v >> 16) & 0xffffu);
    o[2] = (int)((v >> 32) & 0xffffu);
    o[3] = (int)(v >> 48);
}
__device__ inline void quad_store(uint16_t* p, int a, int b, int c, int d) {
    const unsigned long long v = (unsigned long long)(uint16_t)a | ((unsigned long long)(uint16_t)b << 16) |
                                 ((unsigned long long)(uint16_t)c << 32) | ((unsigned long long)(uint16_t)d << 48);
    *reinterpret_cast<unsigned long long*>(p) = v;
}
__device__ inline void quad_store(int* p, int a, int b, int c, int d) {
    v4i v; v.x = a; v.y = b; v.z = c; v.w = d;
    *reinterpret_cast<v4i*>(p) = v;
}
__device__ inline void quad_store(float* p, float a, float b, float c, float d) {
    v4f v; v.x = a; v.y = b; v.z = c; v.w = d;
    *reinterpret_cast<v4f*>(p) = v;
}
__device__ inline void quad_store(double* p, double a, double b, double c, double d) {
    v2d u; u.x = a; u.y = b;
    v2d v; v.x = c; v.y = d;
    reinterpret_cast<v2d*>(p)[0] = u;
    reinterpret_cast<v2d*>(p)[1] = v;
}

// sort.hip
struct SortScratch {
    uint32_t* k1;
    int* v1;
    int* hist;
    int hist_cap;      // ints
    int* scan_tmp;     // block sums for the generic scan
    int scan_tmp_cap;
};
size_t radix_hist_ints(int n_cap);
size_t scan_tmp_ints(int n_cap);
// Stable LSD radix sort of (key, value) pairs on `bits` low key bits.  n is *d_n when d_n is
// non-null (device-resident count, n_cap bounds it), else n_host.  The result ends in (k0, v0)
// or in (s.k1, s.v1); the return value says which (0 or 1).
// `items` (4, 8 or 16) sets the tile of 256 * items keys; every tile gives the same result.
int radix_sort_pairs(uint32_t* k0, int* v0, const int* d_n, int n_host, int n_cap, int bits,
                     const SortScratch& s, hipStream_t st, int items = 8);
bool radix_items_valid(int items);
// Exclusive scan of n ints (n = *d_n or n_host); *d_total (if non-null) receives the sum.
void exclusive_scan(const int* in, int* out, const int* d_n, int n_host, int n_cap, int* d_total,
                    const SortScratch& s, hipStream_t st);
void fill_int(int* p, int value, long long n, hipStream_t st);

// ---- batched launches ----------------------------------------------------------------------
// One launch runs a kernel body for several domains.  Work item i of a launch is (domain or job,
// block of it); every item is one workgroup, which runs that block of the domain's body.
struct WorkList {
    const int2* items = nullptr;
    int n = 0;
};
template <int MAXT, typename Op>
__global__ __launch_bounds__(MAXT) void k_batched(Op op, const int2* __restrict__ items) {
    const int2 w = items[blockIdx.x];
    op(w.x, w.y);
}
template <int MAXT, typename Op> inline void launch_batched(const Op& op, WorkList w, int block, hipStream_t st) {
    if (w.n > 0) hipLaunchKernelGGL((k_batched<MAXT, Op>), dim3(w.n), dim3(block), 0, st, op, w.items);
}
// one pass of one domain's radix sort (the arguments of a pass of radix_sort_pairs)
struct SortJob {
    const uint32_t* kin;
    const int* vin;
    uint32_t* kout;
    int* vout;
    const int* d_n;
    int n_host, n_cap, shift, nbits;
    int* hist;
};
// one domain's exclusive scan (the arguments of exclusive_scan)
struct ScanJob {
    const int* in;
    int* out;
    const int* d_n;
    int n_host, n_cap;
    int* d_total;
    int* tmp;
};
struct FillJob {
    int* p;
    int value;
    int n;
};
int radix_tile(int items);                  // keys per sort tile
// passes of radix_sort_pairs over `bits` key bits (an odd count leaves the result in the scratch)
__host__ __device__ inline int radix_passes_dev(int bits) { return bits > 0 ? (bits + 7) / 8 : 0; }
int radix_passes(int bits);
int radix_pass_bits(int bits, int pass, int* shift);
int scan_tile();                            // ints per tile of the exclusive scan
constexpr int kScanBlock = 1024;            // threads of the scan kernels
// fills: one work item per (job, block of 256 ints)
void fill_int_batched(const FillJob* jobs, WorkList blocks, hipStream_t st);
// exclusive scans: tiles = (job, tile of scan_tile() ints), singles = (job, 0)
void exclusive_scan_batched(const ScanJob* jobs, WorkList tiles, WorkList singles, hipStream_t st);
// one radix pass of several sorts: tiles = (job, tile of radix_tile(items) keys), singles = (job, 0)
void radix_pass_batched(int items, const SortJob* jobs, WorkList tiles, WorkList singles, hipStream_t st);

}  // namespace mpsr

namespace mpsr {

// All device pointers of one handle.  Slot-order state (in/out) is the reference's layout;
// everything prefixed s* is in cell-sorted order (index t), valid between two binnings.
// Largest block of the per-particle kernels; the launch block (64-512) is a runtime choice.
static constexpr int kMaxBlock = 512;

// A per-row neighbour list in the SELL-64/4 layout (sorted index t): row t's entries at
// sell_at(64 * width * (t / 64), t % 64, k), n[t] of them (-1: the row was skipped; above width:
// only the first width are listed).
struct NbList {
    const int* q = nullptr;
    const int* n = nullptr;
    int width = 0;
};
// gradient-radius neighbours listed per row by the PND pass of the EOS pressure (about 32 on the
// initial lattice at 2.1 l0; rows with more walk the stencil)
static constexpr int kGradWidth = 64;

template <typename R> struct Dev {
    int N = 0, cap = 0;
    int block = 256;                  // threads per block of the per-particle kernels (64..512)
    int sort_items = 8;               // radix-sort tile = 256 * sort_items keys (4, 8, 16)
    int ntri = 0, nmodel = 0, ndrag = 0;
    Consts<R> c{};
    PState<R> in{}, out{};
    // particle binning
    uint32_t* keys = nullptr;
    int* vals = nullptr;
    const uint32_t* skeys = nullptr;  // sorted keys (points into keys or sort.k1)
    const int* perm = nullptr;        // sorted index -> slot
    int* cell_start = nullptr;
    int* cell_end = nullptr;
    R4<R>* spos = nullptr;
    R4<R>* svel = nullptr;
    R* spres = nullptr;
    int key_bits = 0;
    // walls
    TriD<R>* tris = nullptr;
    mpsr_tri_params* models = nullptr;
    mpsr_drag_params* drag = nullptr;
    RotCoef<R>* rot = nullptr;        // [nmodel] for the current step, or the table base when
    const int* step_ctr = nullptr;    // ... this device step counter picks the row (graph replays)
    int* tri_emit = nullptr;          // [ntri + 1] emitted cell count per triangle
    int* tri_off = nullptr;           // [ntri + 1]
    int* pair_n = nullptr;            // device count of emitted (cell, triangle) pairs
    int* uniq_n = nullptr;            // device count after de-duplication
    uint32_t* pair_cell = nullptr;    // [pair_cap]
    int* pair_tri = nullptr;          // [pair_cap]
    uint32_t* pair_cell_tmp = nullptr;
    int* pair_tri_tmp = nullptr;
    int* pair_flag = nullptr;         // [pair_cap]
    int* pair_pos = nullptr;          // [pair_cap]
    uint32_t* uniq_cell = nullptr;    // [pair_cap]
    int* uniq_tri = nullptr;          // [pair_cap] -> triangle ids listed per cell
    int pair_cap = 0;
    int* overflow = nullptr;          // set when a march emits more than pair_cap pairs
    int* cell_tri_start = nullptr;    // [ncell]
    int* cell_tri_count = nullptr;    // [ncell], at most 100
    int2* cell_utri = nullptr;        // [utri_cap] pool of the cells' triangle walks: (id, bound key)
    int2* cell_uhead = nullptr;       // [ncell] (pool offset, count) of each cell's walk, count -1: none
    int* cell_work = nullptr;         // [ncell] cells whose walk is being written
    int* cell_work_n = nullptr;       // [2] work-list length, pool entries reserved
    int utri_cap = 0;
    const int* step_base = nullptr;   // batched step: device index of the first replayed step
    int thermal_step = 1;             // batched step: heat conduction every this many steps
    // nearest wall and drag, sorted order
    R* stl_mag = nullptr;
    R4<R>* stl_dir = nullptr;
    int* stl_id = nullptr;
    R4<R>* drag_acc = nullptr;
    R4<R>* sph_dv = nullptr;          // [cap] sorted: fluid velocity change from the spheres
    R* turb_nu = nullptr;             // [cap] slot order: turbulent viscosity (zeroed at allocation)
    R* turb_k = nullptr;              // [cap] slot order: turbulent kinetic energy
    NegModel<R>* negm = nullptr;      // [nmodel] negative-pressure fields
    R* drag_temp = nullptr;
    // EOS pressure: the PND pass lists each row's gradient-radius neighbours in walk order for the
    // corrector (SELL-64/4, kGradWidth per row)
    int* gnb_q = nullptr;             // [cap rounded up to 64, * kGradWidth]
    int* gnb_n = nullptr;             // [cap]
    NbList nbl{};                     // batched step: the list the corrector walks (none: the stencil)
    // radix-sort ping-pong buffers: one set for the particle keys (the sorted keys and the
    // permutation may end up in them and must live until the next binning), one for the
    // triangle pairs
    SortScratch sort{};
    SortScratch tsort{};
};

// particles.hip
template <typename R> void launch_bin_hash(const Dev<R>& d, hipStream_t st);
template <typename R> void launch_bin_ranges(Dev<R>& d, hipStream_t st);
template <typename R> void launch_explicit(const Dev<R>& d, hipStream_t st);
template <typename R> void launch_relocate(const Dev<R>& d, hipStream_t st);
template <typename R> void launch_pnd(const Dev<R>& d, int eos, hipStream_t st);
// PND on the PPE path, with each Poisson row's active flag, neighbour bound, diagonal and
// Dirichlet flag from the same neighbour pass, and its first nb_width Laplacian-radius neighbours
// and weights (SELL-64/4 by sorted index, row width nb_width)
template <typename R>
void launch_pnd_ppe(const Dev<R>& d, int* active, int* cnt, double* diag, int* dir, int* nbq, int* nbn,
                    int nb_width, hipStream_t st);
// The Laplacian-radius neighbour list the PPE path's PND pass writes after the second binning:
// per sorted row its first `width` neighbours in the canonical order (SELL-64/4 by sorted index)
// and n[t] = the number of neighbours (all of them, listed or not), -1 for a row the pass skipped
// The corrector; with a neighbour list of the same binning, rows whose neighbours are all listed
// walk the list instead of the 27-cell stencil (the same neighbours in the same order)
template <typename R> void launch_gradient(const Dev<R>& d, hipStream_t st, NbList nb = NbList{});
// CalcTemperatureFactor; with step_ctr the launch returns at once unless (*step_base + *step_ctr)
// is a multiple of thermal_step (graph replays), else the caller decides
template <typename R> void launch_thermal(const Dev<R>& d, const int* step_base, const int* step_ctr, int thermal_step,
                                          hipStream_t st);
template <typename R> void launch_turbulence(const Dev<R>& d, hipStream_t st);
template <typename R> void launch_neighbor_count(const Dev<R>& d, R re, int* counts, hipStream_t st);
template <typename R> void launch_neighbor_fill(const Dev<R>& d, R re, const int* offsets, int* nbr, hipStream_t st);
// walls.hip
template <typename R> void launch_tri_bin(Dev<R>& d, hipStream_t st);
// key bits of the (cell, triangle) pair sort of a grid of ncell cells
__host__ __device__ inline int tri_pair_bits(long long ncell) {
    int bits = 1;
    while ((1ll << bits) < ncell) ++bits;
    return bits;
}
template <typename R> void launch_tri_count_total(Dev<R>& d, hipStream_t st);
template <typename R> void launch_stl_distance(const Dev<R>& d, hipStream_t st);
template <typename R> void launch_drag(const Dev<R>& d, hipStream_t st);
template <typename R> void launch_translate(const Dev<R>& d, hipStream_t st);
template <typename R> void launch_rotate(const Dev<R>& d, hipStream_t st);
void launch_step_advance(int* ctr, hipStream_t st);
// spheres.hip: rigid spheres with Schiller-Naumann drag
template <typename R> struct SphDev {
    int S = 0;
    R4<R>* pos = nullptr;              // [S]
    R4<R>* vel = nullptr;              // [S]
    R4<R>* force = nullptr;            // [S] drag of the last phase
    R* wsum = nullptr;                 // [S] sum of the fluid weights at the centre
    uint32_t* keys = nullptr;          // [S]
    int* vals = nullptr;               // [S]
    SortScratch sort{};
    const int* perm = nullptr;         // sorted -> sphere index
    int* cs = nullptr;                 // [ncell]
    int* ce = nullptr;                 // [ncell]
    R d = 0, rho_s = 0, rho_f = 0, nu_f = 0, mf = 0;
};
template <typename R> void launch_spheres(Dev<R>& d, SphDev<R>& s, hipStream_t st);
// inflow.hip
template <typename R>
void launch_bucket_reuse(const Dev<R>& d, const PState<R>& b, int B, int* flag, int* rank, int* dead, hipStream_t st);
template <typename R> void launch_bucket_append(const Dev<R>& d, const PState<R>& b, int first, int count, hipStream_t st);
template <typename R> void launch_reset_walls(const Dev<R>& d, const TriD<R>* tris0, int step, hipStream_t st);

// ---- batched forms of the step kernels ------------------------------------------------------
// devs: device table of the domains; a work list of (domain, block).  P: blocks of `block`
// particles; T64 / T256: blocks of 64 / 256 triangles; pairs: blocks of 256 of the pair capacity;
// one: one item per domain.  Each domain's Dev holds the buffers its last sort left the keys and
// permutation in (skeys, perm), and for replays its rotation table and step counter.
template <typename R> struct BatchLaunch {
    const Dev<R>* devs = nullptr;
    int block = 256;
};
template <typename R> void launch_bin_hash_b(const BatchLaunch<R>& b, WorkList p, hipStream_t st);
template <typename R> void launch_bin_ranges_b(const BatchLaunch<R>& b, WorkList p, hipStream_t st);
template <typename R> void launch_explicit_b(const BatchLaunch<R>& b, WorkList p, hipStream_t st);
template <typename R> void launch_relocate_b(const BatchLaunch<R>& b, WorkList p, hipStream_t st);
template <typename R> void launch_pnd_b(const BatchLaunch<R>& b, WorkList p, hipStream_t st);   // EOS pressure
template <typename R> void launch_gradient_b(const BatchLaunch<R>& b, WorkList p, hipStream_t st);
template <typename R> void launch_thermal_b(const BatchLaunch<R>& b, WorkList p, hipStream_t st);
template <typename R> void launch_turbulence_b(const BatchLaunch<R>& b, WorkList p, hipStream_t st);
// the PND pass of the PPE path for the PPE buckets of a batch: work item (row set j, block), with
// row set j the Poisson rows of the bucket devs[rows[j].dev]
template <typename R> struct PndRows {
    int dev;
    int* active;
    int* cnt;
    double* diag;
    int* dir;
    int* nbq;
    int* nbn;
    int nb_width;
};
template <typename R> void launch_pnd_ppe_b(const BatchLaunch<R>& b, const PndRows<R>* rows, WorkList p, hipStream_t st);
template <typename R> void launch_tri_count_b(const BatchLaunch<R>& b, WorkList t64, hipStream_t st);
template <typename R> void launch_tri_fill_b(const BatchLaunch<R>& b, WorkList t64, hipStream_t st);
// the pair kernels of launch_tri_bin after the pair sort; `pass` = 0 flags, 1 compacts, 2 ranges
template <typename R> void launch_pairs_b(const BatchLaunch<R>& b, int pass, WorkList pairs, hipStream_t st);
// each bucket's cell walks (c256: (bucket, block of 256 cells) of the buckets with particles);
// work holds one int2 per such cell, nwork one int
template <typename R>
void launch_cell_walks_b(const BatchLaunch<R>& b, WorkList c256, int2* work, int* nwork, hipStream_t st);
template <typename R> void launch_stl_distance_b(const BatchLaunch<R>& b, WorkList p, hipStream_t st);
template <typename R> void launch_drag_b(const BatchLaunch<R>& b, WorkList p, hipStream_t st);
template <typename R> void launch_translate_b(const BatchLaunch<R>& b, WorkList t256, hipStream_t st);
template <typename R> void launch_rotate_b(const BatchLaunch<R>& b, WorkList t256, hipStream_t st);
template <typename R> void launch_step_advance_b(const BatchLaunch<R>& b, WorkList one, hipStream_t st);

}  // namespace mpsr
