// neighbors.h -- the canonical 27-cell neighbour walk shared by the particle and PPE kernels.
#pragma once
#include "mpsr_internal.h"

namespace mpsr {

template <typename R> __device__ inline R4<R> mk4(R x, R y, R z, R w);
template <> __device__ inline float4 mk4<float>(float x, float y, float z, float w) { return make_float4(x, y, z, w); }
template <> __device__ inline double4 mk4<double>(double x, double y, double z, double w) { return make_double4(x, y, z, w); }

// every candidate t != self in the 27-cell stencil around (x, y, z), canonical order
template <typename R, typename F>
__device__ inline void for_candidates(const Consts<R>& c, const int* __restrict__ cs, const int* __restrict__ ce,
                                      R x, R y, R z, int self, F&& f) {
    int ci, cj, ck;
    if (!cell_coords(c, x, y, z, ci, cj, ck)) return;
    const long long cid = cell_of(c, ci, cj, ck);
    if (cid < 0 || cid >= c.ncell) return;
    for (int k = -1; k <= 1; ++k)
        for (int j = -1; j <= 1; ++j)
            for (int i = -1; i <= 1; ++i) {
                const long long m = cell_of(c, ci + i, cj + j, ck + k);
                if (m < 0 || m >= c.ncell) continue;
                const int a = cs[m];
                if (a < 0) continue;
                const int b = ce[m];
                for (int t = a; t < b; ++t)
                    if (t != self) f(t);
            }
}

template <typename R>
__device__ inline R dist3(R x, R y, R z, R qx, R qy, R qz) {
    return build_sqrt((x - qx) * (x - qx) + (y - qy) * (y - qy) + (z - qz) * (z - qz));
}

}  // namespace mpsr
