// noalloc.hip -- the cell binning of mpsr_register_particle_topology for callers that cannot
// allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc).
//
// The caller passes device arrays for the particle ids and positions (FP32, slot order) and for
// the results: sorted cell keys, the sorted -> slot permutation and the per-cell ranges.  The
// radix sort's second key/value buffers, its digit histograms and the scan partials live in one
// file-scope device array of fixed size (`scratch`, MPSR_NOALLOC_MAX_PARTICLES particles), which
// every GPU the library runs on holds until the process ends.  Host code never passes the array's
// name to a kernel: its device address is looked up once per device (hipGetSymbolAddress, with
// that device current) and that address is what the kernels receive.
//
// Every call on one GPU shares that GPU's scratch and leaves the last call's contents there;
// nothing in it has to start at zero (every histogram and scan entry is written before it is
// read).  Calls that overlap on one GPU race.
#include <mutex>
#include "mpsr_internal.h"

namespace mpsr {

static constexpr size_t NA_MAX = MPSR_NOALLOC_MAX_PARTICLES;
// radix_hist_ints(NA_MAX) and scan_tmp_ints(NA_MAX), checked against sort.hip at set-up
static constexpr size_t NA_HIST = (size_t)256 * ((NA_MAX + 1023) / 1024 + 1);
static constexpr size_t NA_SCAN = (NA_MAX + 4095) / 4096 + 2 + 64;
static constexpr size_t NA_INTS = 2 * NA_MAX + NA_HIST + NA_SCAN;
static constexpr int NA_DEVICES = 64;

// k1 [NA_MAX] | v1 [NA_MAX] | hist [NA_HIST] | scan_tmp [NA_SCAN]
static __device__ __align__(256) int scratch[NA_INTS];

static std::mutex na_mu;
static int* na_ptr[NA_DEVICES] = {};

// the scratch's device address on `device` (current), looked up on the first call there
static hipError_t scratch_address(int device, int** out) {
    std::lock_guard<std::mutex> lock(na_mu);
    if (!na_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        na_ptr[device] = (int*)p;
    }
    *out = na_ptr[device];
    return hipSuccess;
}

// RegisterParticleTopology's key per slot (the same bin_key as k_bin_hash), value = slot
__global__ __launch_bounds__(256) void k_na_keys(int N, Consts<float> c, const int* __restrict__ id,
                                                 const float* __restrict__ px, const float* __restrict__ py,
                                                 const float* __restrict__ pz, uint32_t* keys, int* vals) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= N) return;
    keys[s] = bin_key(c, id[s], px[s], py[s], pz[s]);
    vals[s] = s;
}

__global__ __launch_bounds__(256) void k_na_copy(int N, const uint32_t* __restrict__ ka, const int* __restrict__ va,
                                                 uint32_t* kb, int* vb) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N) return;
    kb[t] = ka[t];
    vb[t] = va[t];
}

// cellStart / cellEnd from the change points of the sorted keys (k_bin_ranges without its
// gather of the neighbour records)
__global__ __launch_bounds__(256) void k_na_ranges(int N, int ncell, const uint32_t* __restrict__ sk, int* cs, int* ce) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N) return;
    const uint32_t k = sk[t];
    const uint32_t kp = t > 0 ? sk[t - 1] : 0xffffffffu;
    const uint32_t sent = (uint32_t)ncell;
    if (k != kp) {
        if (k != sent) cs[k] = t;
        if (t > 0 && kp != sent) ce[kp] = t;
    }
    if (t == N - 1 && k != sent) ce[k] = N;
}

}  // namespace mpsr

using namespace mpsr;

extern "C" mpsr_status mpsr_register_particle_topology_noalloc(const mpsr_gridbox* gb, int n, const int* id,
                                                               const float* px, const float* py, const float* pz,
                                                               unsigned int* sorted_keys, int* perm, int* cell_start,
                                                               int* cell_end, int cell_capacity, int device,
                                                               void* stream) {
    if (!gb || n < 0 || (n > 0 && (!id || !px || !py || !pz || !sorted_keys || !perm)) || !cell_start || !cell_end)
        return MPSR_PARAMERR;
    if ((size_t)n > NA_MAX) return MPSR_PARAMERR;   // more particles than the fixed scratch holds
    if (device < 0 || device >= NA_DEVICES) return MPSR_PARAMERR;
    const long long ncell_ll = (long long)gb->m_GridSize.x * gb->m_GridSize.y * gb->m_GridSize.z;
    if (gb->m_GridSize.x <= 0 || gb->m_GridSize.y <= 0 || gb->m_GridSize.z <= 0 || ncell_ll >= (1ll << 31) - 1 ||
        ncell_ll > (long long)cell_capacity || !(gb->m_CellSize > 0.0f))
        return MPSR_PARAMERR;
    if (radix_hist_ints((int)NA_MAX) > NA_HIST || scan_tmp_ints((int)NA_MAX) > NA_SCAN) return MPSR_ETCERR;
    if (hipSetDevice(device) != hipSuccess) return MPSR_GPUERR;
    int* base = nullptr;
    if (scratch_address(device, &base) != hipSuccess) return MPSR_GPUERR;
    const hipStream_t st = (hipStream_t)stream;
    const int ncell = (int)ncell_ll;
    Consts<float> c{};
    c.bminx = gb->m_BufferedZone.m_MinBound.x;
    c.bminy = gb->m_BufferedZone.m_MinBound.y;
    c.bminz = gb->m_BufferedZone.m_MinBound.z;
    c.cell = gb->m_CellSize;
    c.gx = gb->m_GridSize.x;
    c.gy = gb->m_GridSize.y;
    c.gz = gb->m_GridSize.z;
    c.ncell = ncell;
    int bits = 1;
    while ((1ll << bits) <= (long long)ncell) ++bits;   // keys 0..ncell (sentinel = ncell)
    SortScratch ss{};
    ss.k1 = (uint32_t*)base;
    ss.v1 = base + NA_MAX;
    ss.hist = base + 2 * NA_MAX;
    ss.hist_cap = (int)NA_HIST;
    ss.scan_tmp = base + 2 * NA_MAX + NA_HIST;
    ss.scan_tmp_cap = (int)NA_SCAN;
    fill_int(cell_start, -1, ncell, st);
    fill_int(cell_end, -1, ncell, st);
    if (n > 0) {
        hipLaunchKernelGGL(k_na_keys, dim3(grid_for(n, 256)), dim3(256), 0, st, n, c, id, px, py, pz,
                           (uint32_t*)sorted_keys, perm);
        const int which = radix_sort_pairs((uint32_t*)sorted_keys, perm, nullptr, n, n, bits, ss, st, 8);
        if (which)
            hipLaunchKernelGGL(k_na_copy, dim3(grid_for(n, 256)), dim3(256), 0, st, n, ss.k1, ss.v1,
                               (uint32_t*)sorted_keys, perm);
        hipLaunchKernelGGL(k_na_ranges, dim3(grid_for(n, 256)), dim3(256), 0, st, n, ncell,
                           (const uint32_t*)sorted_keys, cell_start, cell_end);
    }
    return hipGetLastError() == hipSuccess ? MPSR_NOERR : MPSR_GPUERR;
}
