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

// for_candidates restricted to the stencil cells whose box lies within `reach` of (x, y, z), with
// a 1 % margin over the build's distance (sqrt14 errs by at most 0.18 %, the cell boxes by a few
// ulp): a skipped cell holds no particle closer than `reach`, and the cells that remain are
// visited in the same order, so every sum over the neighbours within `reach` is unchanged
template <typename R, typename F>
__device__ inline void for_candidates_within(const Consts<R>& c, const int* __restrict__ cs,
                                             const int* __restrict__ ce, R x, R y, R z, int self, R reach,
                                             F&& f) {
    int ci, cj, ck;
    if (!cell_coords(c, x, y, z, ci, cj, ck)) return;
    const long long cid = cell_of(c, ci, cj, ck);
    if (cid < 0 || cid >= c.ncell) return;
    const R m = reach * (R)1.01;
    const R m2 = m * m;
    // squared distance to the cell slabs at offsets -1 and +1 per axis (0 for the own slab)
    const R lx = c.bminx + (R)ci * c.cell, ly = c.bminy + (R)cj * c.cell, lz = c.bminz + (R)ck * c.cell;
    const R xm = x - lx, xp = lx + c.cell - x, ym = y - ly, yp = ly + c.cell - y, zm = z - lz, zp = lz + c.cell - z;
    const R xm2 = xm * xm, xp2 = xp * xp, ym2 = ym * ym, yp2 = yp * yp, zm2 = zm * zm, zp2 = zp * zp;
    for (int k = -1; k <= 1; ++k) {
        const R dz2 = k < 0 ? zm2 : (k > 0 ? zp2 : (R)0);
        for (int j = -1; j <= 1; ++j) {
            const R dyz2 = dz2 + (j < 0 ? ym2 : (j > 0 ? yp2 : (R)0));
            if (dyz2 > m2) continue;
            for (int i = -1; i <= 1; ++i) {
                if (dyz2 + (i < 0 ? xm2 : (i > 0 ? xp2 : (R)0)) > m2) continue;
                const long long mm = cell_of(c, ci + i, cj + j, ck + k);
                if (mm < 0 || mm >= c.ncell) continue;
                const int a = cs[mm];
                if (a < 0) continue;
                const int b = ce[mm];
                for (int t = a; t < b; ++t)
                    if (t != self) f(t);
            }
        }
    }
}

template <typename R>
__device__ inline R dist3(R x, R y, R z, R qx, R qy, R qz) {
    return build_sqrt((x - qx) * (x - qx) + (y - qy) * (y - qy) + (z - qz) * (z - qz));
}

// one stencil cell's candidates t in [a, b), t != self, as f(t, spos[t]) in order, the records
// loaded in groups of G (then 4) issued together, so G gathers are in flight per lane instead of one
template <int G, typename R, typename F>
__device__ inline void cell_records(const R4<R>* __restrict__ spos, int a, int b, int self, F& f) {
    int t = a;
    if constexpr (G == 8) {
        for (; t + 8 <= b; t += 8) {
            R4<R> Q[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) Q[u] = spos[t + u];
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (t + u != self) f(t + u, Q[u]);
        }
    }
    for (; t + 4 <= b; t += 4) {
        const R4<R> Q0 = spos[t], Q1 = spos[t + 1], Q2 = spos[t + 2], Q3 = spos[t + 3];
        if (t != self) f(t, Q0);
        if (t + 1 != self) f(t + 1, Q1);
        if (t + 2 != self) f(t + 2, Q2);
        if (t + 3 != self) f(t + 3, Q3);
    }
    for (; t < b; ++t)
        if (t != self) f(t, spos[t]);
}

// for_candidates with the records loaded ahead: every candidate t != self of the 27-cell stencil in
// the canonical order, as f(t, spos[t]) (cell_records per cell).  G = 8 suits a kernel whose
// registers allow it (the PND pass: 6 % faster than 4; the predictor loses occupancy with it and
// keeps 4).
template <int G = 4, typename R, typename F>
__device__ inline void for_candidates_rec(const Consts<R>& c, const int* __restrict__ cs, const int* __restrict__ ce,
                                          const R4<R>* __restrict__ spos, R x, R y, R z, int self, F&& f) {
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
                cell_records<G, R>(spos, a, ce[m], self, f);
            }
}

// for_candidates_within with the records loaded ahead as in for_candidates_rec: the stencil cells
// within `reach` (the same test, so the same cells in the same order), f(t, spos[t])
template <int G = 4, typename R, typename F>
__device__ inline void for_candidates_rec_within(const Consts<R>& c, const int* __restrict__ cs,
                                                 const int* __restrict__ ce, const R4<R>* __restrict__ spos, R x,
                                                 R y, R z, int self, R reach, F&& f) {
    int ci, cj, ck;
    if (!cell_coords(c, x, y, z, ci, cj, ck)) return;
    const long long cid = cell_of(c, ci, cj, ck);
    if (cid < 0 || cid >= c.ncell) return;
    const R m = reach * (R)1.01;
    const R m2 = m * m;
    const R lx = c.bminx + (R)ci * c.cell, ly = c.bminy + (R)cj * c.cell, lz = c.bminz + (R)ck * c.cell;
    const R xm = x - lx, xp = lx + c.cell - x, ym = y - ly, yp = ly + c.cell - y, zm = z - lz, zp = lz + c.cell - z;
    const R xm2 = xm * xm, xp2 = xp * xp, ym2 = ym * ym, yp2 = yp * yp, zm2 = zm * zm, zp2 = zp * zp;
    for (int k = -1; k <= 1; ++k) {
        const R dz2 = k < 0 ? zm2 : (k > 0 ? zp2 : (R)0);
        for (int j = -1; j <= 1; ++j) {
            const R dyz2 = dz2 + (j < 0 ? ym2 : (j > 0 ? yp2 : (R)0));
            if (dyz2 > m2) continue;
            for (int i = -1; i <= 1; ++i) {
                if (dyz2 + (i < 0 ? xm2 : (i > 0 ? xp2 : (R)0)) > m2) continue;
                const long long mm = cell_of(c, ci + i, cj + j, ck + k);
                if (mm < 0 || mm >= c.ncell) continue;
                const int a = cs[mm];
                if (a < 0) continue;
                cell_records<G, R>(spos, a, ce[mm], self, f);
            }
        }
    }
}

// for_candidates_rec in groups: g(t0, Q, m) for runs of up to 4 consecutive candidates t0 ..
// t0 + m - 1 of one stencil cell, in the canonical order, with Q[u] = spos[t0 + u] (Q[u] for u >= m
// repeats the last record and is to be ignored; self is not skipped: the caller tests it).  A
// caller that reads a second record of its neighbours loads them for the whole group before it
// uses them, so one wait covers the group instead of one per neighbour.
template <typename R, typename F>
__device__ inline void for_candidate_groups(const Consts<R>& c, const int* __restrict__ cs,
                                            const int* __restrict__ ce, const R4<R>* __restrict__ spos, R x, R y,
                                            R z, F&& g) {
    int ci, cj, ck;
    if (!cell_coords(c, x, y, z, ci, cj, ck)) return;
    const long long cid = cell_of(c, ci, cj, ck);
    if (cid < 0 || cid >= c.ncell) return;
    for (int k = -1; k <= 1; ++k)
        for (int j = -1; j <= 1; ++j)
            for (int i = -1; i <= 1; ++i) {
                const long long mcell = cell_of(c, ci + i, cj + j, ck + k);
                if (mcell < 0 || mcell >= c.ncell) continue;
                const int a = cs[mcell];
                if (a < 0) continue;
                const int b = ce[mcell];
                int t = a;
                for (; t + 4 <= b; t += 4) {
                    R4<R> Q[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) Q[u] = spos[t + u];
                    g(t, Q, 4);
                }
                if (t < b) {
                    const int m = b - t;
                    R4<R> Q[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) Q[u] = spos[t + (u < m ? u : m - 1)];
                    g(t, Q, m);
                }
            }
}

// The canonical walk in two phases per run of up to 32 candidates of one stencil cell: first one
// bit per candidate whose squared distance is below (1.01 reach)^2, then f(q) for the set bits in
// increasing order.  The bound holds every candidate within reach (the build's root errs by at
// most 0.18 %), the cells skipped are those of for_candidates_within, and self is dropped, so f
// sees the candidates of the canonical walk that pass the bound, in the same order: it tests the
// distance itself and every sum over the neighbours within reach is unchanged.  A lane's neighbour
// work runs for its own neighbours only (the second phase takes as many turns as the lane of the
// wave with the most bits).  It pays where that work is light against the stencil and the radius
// small against the cell: the EOS pressure's PND pass (gradient radius 2.1 l0 in 3.1 l0 cells).
// The predictor and the PPE path's PND pass (Laplacian radius) ran 8-10 % slower with it, also
// with the next neighbour's records loaded ahead.
template <typename R, typename F>
__device__ inline void for_near_candidates(const Consts<R>& c, const int* __restrict__ cs,
                                           const int* __restrict__ ce, const R4<R>* __restrict__ spos, R x, R y,
                                           R z, int self, R reach, F&& f) {
    int ci, cj, ck;
    if (!cell_coords(c, x, y, z, ci, cj, ck)) return;
    const long long cid = cell_of(c, ci, cj, ck);
    if (cid < 0 || cid >= c.ncell) return;
    const R m = reach * (R)1.01;
    const R m2 = m * m;
    const R lx = c.bminx + (R)ci * c.cell, ly = c.bminy + (R)cj * c.cell, lz = c.bminz + (R)ck * c.cell;
    const R xm = x - lx, xp = lx + c.cell - x, ym = y - ly, yp = ly + c.cell - y, zm = z - lz, zp = lz + c.cell - z;
    const R xm2 = xm * xm, xp2 = xp * xp, ym2 = ym * ym, yp2 = yp * yp, zm2 = zm * zm, zp2 = zp * zp;
    auto near = [&](const R4<R>& Q) {
        return (x - Q.x) * (x - Q.x) + (y - Q.y) * (y - Q.y) + (z - Q.z) * (z - Q.z) < m2;
    };
    for (int k = -1; k <= 1; ++k) {
        const R dz2 = k < 0 ? zm2 : (k > 0 ? zp2 : (R)0);
        for (int j = -1; j <= 1; ++j) {
            const R dyz2 = dz2 + (j < 0 ? ym2 : (j > 0 ? yp2 : (R)0));
            if (dyz2 > m2) continue;
            for (int i = -1; i <= 1; ++i) {
                if (dyz2 + (i < 0 ? xm2 : (i > 0 ? xp2 : (R)0)) > m2) continue;
                const long long mm = cell_of(c, ci + i, cj + j, ck + k);
                if (mm < 0 || mm >= c.ncell) continue;
                const int a = cs[mm];
                if (a < 0) continue;
                const int b = ce[mm];
                for (int base = a; base < b; base += 32) {
                    const int n = b - base < 32 ? b - base : 32;
                    unsigned bits = 0;
                    int u = 0;
                    for (; u + 4 <= n; u += 4) {
                        R4<R> Q[4];
#pragma unroll
                        for (int v = 0; v < 4; ++v) Q[v] = spos[base + u + v];
#pragma unroll
                        for (int v = 0; v < 4; ++v)
                            if (near(Q[v])) bits |= 1u << (u + v);
                    }
                    for (; u < n; ++u)
                        if (near(spos[base + u])) bits |= 1u << u;
                    if (self >= base && self < base + n) bits &= ~(1u << (self - base));
                    while (bits) {
                        const int v = __builtin_ctz(bits);
                        bits &= bits - 1u;
                        f(base + v);
                    }
                }
            }
        }
    }
}

}  // namespace mpsr
