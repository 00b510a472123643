// inflow.hip -- bucket inflow and linear-wall reset, on the device.
//
//   k_dead_flags / k_bucket_reuse / k_bucket_append
//                      CMultiGPUHandler::InsertBucket (Src MultiGPUHandler.cpp:949-981): the
//                      bucket's particles go, in order, into the first dead slots (output ID < 0)
//                      in ascending slot order; the rest is appended after the last slot when the
//                      capacity allows, otherwise dropped.  A reused or appended slot takes
//                      ID = slot index, position, velocity, temperature, density, viscosity,
//                      solid rate and type from the bucket.  A reused slot keeps its pressure
//                      (InsertBucket does not write it); an appended slot starts at 0.
//   k_reset_walls      ResetWallPosition_Kernel (Kern :913-925): triangles of a linearly moving
//                      model return to their initial position when (step + 1) is a multiple of
//                      the model's reset interval.
//
// The reference does the insertion on the host (device-to-host copy of every particle, a serial
// loop, host-to-device copy back).  Here it stays on the device: a flag + exclusive scan ranks
// the dead slots, one thread per slot fills its bucket row, and only the dead-slot count comes
// back to the host (the new particle count sizes the following launches).
#include "mpsr_internal.h"

namespace mpsr {

__global__ __launch_bounds__(256) void k_dead_flags(int N, const int* __restrict__ id, int* flag) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= N) return;
    flag[s] = id[s] < 0 ? 1 : 0;
}

template <typename R>
__device__ inline void put_bucket_row(PState<R>& o, int s, const PState<R>& b, int j) {
    o.id[s] = s;
    o.px[s] = b.px[j]; o.py[s] = b.py[j]; o.pz[s] = b.pz[j];
    o.vx[s] = b.vx[j]; o.vy[s] = b.vy[j]; o.vz[s] = b.vz[j];
    o.temp[s] = b.temp[j];
    o.n[s] = b.n[j];
    o.nu[s] = b.nu[j];
    o.solid[s] = b.solid[j];
    o.type[s] = b.type[j];
}

template <typename R>
__global__ __launch_bounds__(256) void k_bucket_reuse(int N, int B, const int* __restrict__ rank, PState<R> out,
                                                       PState<R> b) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= N) return;
    if (out.id[s] >= 0) return;
    const int j = rank[s];
    if (j < B) put_bucket_row(out, s, b, j);
}

// An appended slot holds no particle in the input arrays until the next relocation copies it
// there: its input ID is -1, the other input fields 0.
template <typename R>
__global__ __launch_bounds__(256) void k_bucket_append(int N, int first, int count, PState<R> in, PState<R> out,
                                                        PState<R> b) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const int s = N + k;
    put_bucket_row(out, s, b, first + k);
    out.p[s] = R(0);
    in.id[s] = -1;
    in.type[s] = 0;
    in.px[s] = R(0); in.py[s] = R(0); in.pz[s] = R(0);
    in.vx[s] = R(0); in.vy[s] = R(0); in.vz[s] = R(0);
    in.p[s] = R(0); in.n[s] = R(0); in.temp[s] = R(0); in.nu[s] = R(0); in.solid[s] = R(0);
}

template <typename R>
__global__ __launch_bounds__(256) void k_reset_walls(int T, int step, TriD<R>* tris, const TriD<R>* __restrict__ tris0,
                                                      const mpsr_tri_params* __restrict__ models) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T) return;
    const mpsr_tri_params& m = models[tris[t].model];
    if (!m.LinearMovableOrNot || m.LinearResetInterval <= 0) return;
    if ((step + 1) % m.LinearResetInterval == 0) tris[t] = tris0[t];
}

template <typename R>
void launch_bucket_reuse(const Dev<R>& d, const PState<R>& b, int B, int* flag, int* rank, int* dead, hipStream_t st) {
    if (d.N <= 0) {
        fill_int(dead, 0, 1, st);
        return;
    }
    hipLaunchKernelGGL(k_dead_flags, dim3(grid_for(d.N, 256)), dim3(256), 0, st, d.N, d.out.id, flag);
    exclusive_scan(flag, rank, nullptr, d.N, d.N, dead, d.sort, st);
    if (B > 0)
        hipLaunchKernelGGL(k_bucket_reuse<R>, dim3(grid_for(d.N, 256)), dim3(256), 0, st, d.N, B, rank, d.out, b);
}

template <typename R> void launch_bucket_append(const Dev<R>& d, const PState<R>& b, int first, int count, hipStream_t st) {
    if (count <= 0) return;
    hipLaunchKernelGGL(k_bucket_append<R>, dim3(grid_for(count, 256)), dim3(256), 0, st, d.N, first, count, d.in, d.out,
                       b);
}

template <typename R> void launch_reset_walls(const Dev<R>& d, const TriD<R>* tris0, int step, hipStream_t st) {
    if (d.ntri <= 0) return;
    hipLaunchKernelGGL(k_reset_walls<R>, dim3(grid_for(d.ntri, 256)), dim3(256), 0, st, d.ntri, step, d.tris, tris0,
                       d.models);
}

#define MPSR_INSTANTIATE(R)                                                                                    \
    template void launch_bucket_reuse<R>(const Dev<R>&, const PState<R>&, int, int*, int*, int*, hipStream_t); \
    template void launch_bucket_append<R>(const Dev<R>&, const PState<R>&, int, int, hipStream_t);             \
    template void launch_reset_walls<R>(const Dev<R>&, const TriD<R>*, int, hipStream_t);
MPSR_INSTANTIATE(float)
MPSR_INSTANTIATE(double)

}  // namespace mpsr
