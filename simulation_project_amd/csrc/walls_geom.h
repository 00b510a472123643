// walls_geom.h -- distance from a point to an STL triangle and the nearest-wall search, shared
// by the wall kernels and the sphere kernels.
#pragma once
#include "mpsr_internal.h"

namespace mpsr {

template <typename R> struct P3 { R x, y, z; };
template <typename R> __device__ inline P3<R> p3(R x, R y, R z) { return P3<R>{x, y, z}; }
template <typename R> __device__ inline P3<R> psub(P3<R> a, P3<R> b) { return p3(a.x - b.x, a.y - b.y, a.z - b.z); }
template <typename R> __device__ inline P3<R> padd(P3<R> a, P3<R> b) { return p3(a.x + b.x, a.y + b.y, a.z + b.z); }
template <typename R> __device__ inline R pdot(P3<R> a, P3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename R> __device__ inline R pmag(P3<R> a) { return build_sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }

// ---------------------------------------------------------------------------------------------
// distance to a triangle, Utility_Inline.h:46-170
// ---------------------------------------------------------------------------------------------
template <typename R> __device__ inline R segment_ratio(P3<R> p1, P3<R> p2, P3<R> p3_) {
    const P3<R> dv = psub(p2, p1);
    R den = dv.x * dv.x + dv.y * dv.y + dv.z * dv.z;
    const R num = (p3_.x - p1.x) * (p2.x - p1.x) + (p3_.y - p1.y) * (p2.y - p1.y) + (p3_.z - p1.z) * (p2.z - p1.z);
    if (den == 0.0) den = (R)1e-5;
    return num / den;
}
// CalcDistanceVectorLine: beyond the ends it returns L1 - P (u < 0) and L0 - P (u > 1)
template <typename R> __device__ inline P3<R> line_vec(P3<R> l0, P3<R> l1, P3<R> p) {
    const R u = segment_ratio(l0, l1, p);
    P3<R> d = p3<R>(0, 0, 0);
    if (u >= 0 && u <= 1) {
        const P3<R> m = psub(l1, l0);
        const P3<R> q = p3(l0.x + u * m.x, l0.y + u * m.y, l0.z + u * m.z);
        d = psub(p, q);
    }
    if (u < 0) d = psub(l1, p);
    if (u > 1) d = psub(l0, p);
    return d;
}
template <typename R> __device__ inline R tri_distance(const TriD<R>& t, P3<R> p, P3<R>* out) {
    const P3<R> v0 = p3(t.v[0], t.v[1], t.v[2]), v1 = p3(t.v[3], t.v[4], t.v[5]), v2 = p3(t.v[6], t.v[7], t.v[8]);
    // CalcNormalDistance with the stored normal and the centroid divided by the double 3.0
    const P3<R> n = p3(t.nrm[0], t.nrm[1], t.nrm[2]);
    P3<R> cc = p3<R>(0, 0, 0);
    cc.x += v0.x; cc.y += v0.y; cc.z += v0.z;
    cc.x += v1.x; cc.y += v1.y; cc.z += v1.z;
    cc.x += v2.x; cc.y += v2.y; cc.z += v2.z;
    cc.x = (R)((double)cc.x / 3.0);
    cc.y = (R)((double)cc.y / 3.0);
    cc.z = (R)((double)cc.z / 3.0);
    const R ln = n.x * cc.x + n.y * cc.y + n.z * cc.z;
    const R rn = n.x * p.x + n.y * p.y + n.z * p.z;
    const R den = n.x * n.x + n.y * n.y + n.z * n.z;
    const R sden = build_sqrt(den);
    P3<R> dv = p3<R>(0, 0, 0);
    if (!(sden == 0.0)) {
        const R coef = (rn - ln) / sden;
        dv.x = coef / sden * n.x;
        dv.y = coef / sden * n.y;
        dv.z = coef / sden * n.z;
    }
    const P3<R> P = p3(p.x - dv.x, p.y - dv.y, p.z - dv.z);
    const P3<R> e0 = psub(v2, v0), e1 = psub(v1, v0), e2 = psub(P, v0);
    const R d00 = pdot(e0, e0), d01 = pdot(e0, e1), d02 = pdot(e0, e2), d11 = pdot(e1, e1), d12 = pdot(e1, e2);
    const R inv = 1 / (d00 * d11 - d01 * d01);
    const R u = (d11 * d02 - d01 * d12) * inv;
    const R v = (d00 * d12 - d01 * d02) * inv;
    if (u >= 0 && v >= 0 && u + v > 1) dv = line_vec(v1, v2, p);
    if (u < 0) dv = line_vec(v0, v1, p);
    if (v < 0) dv = line_vec(v0, v2, p);
    *out = dv;
    return pmag(dv);
}

// nearest triangle of the wanted kind (drag or not) over the 27-cell stencil
template <typename R>
__device__ inline int nearest_triangle(const Consts<R>& c, R x, R y, R z, const int* __restrict__ tstart,
                                       const int* __restrict__ tcount, const int* __restrict__ tids,
                                       const TriD<R>* __restrict__ tris, int want_drag, R* best, P3<R>* bdir) {
    *best = (R)1e8;
    *bdir = p3<R>(0, 0, 0);
    int bid = -1;
    int ci, cj, ck;
    if (!cell_coords(c, x, y, z, ci, cj, ck)) return bid;
    const long long cid = cell_of(c, ci, cj, ck);
    if (cid < 0 || cid >= c.ncell) return bid;
    const P3<R> p = p3(x, y, z);
    for (int k = -1; k <= 1; ++k)
        for (int j = -1; j <= 1; ++j)
            for (int i = -1; i <= 1; ++i) {
                const long long m = cell_of(c, ci + i, cj + j, ck + k);
                if (m < 0 || m >= c.ncell) continue;
                const int n = tcount[m];
                const int a = tstart[m];
                for (int e = 0; e < n; ++e) {
                    const int tid = tids[a + e];
                    const TriD<R>& tr = tris[tid];
                    if (tr.is_drag != want_drag) continue;
                    if (tid != bid) {
                        P3<R> dv;
                        const R dm = tri_distance(tr, p, &dv);
                        if (dm < *best) {
                            *best = dm;
                            *bdir = dv;
                            bid = tid;
                        }
                    }
                }
            }
    return bid;
}

// A lower bound of tri_distance(t, p) over the points p of the box [lo, lo + size]^3, less a margin
// of size / 100 (0 where no bound holds), as a float no larger than the bound.  tri_distance ends
// on a point of an edge, a vertex, or the point P = p - (distance along the stored normal) when the
// barycentric coordinates of P put it inside the triangle.  The first two lie in the triangle's
// bounding box.  P lies in the triangle too when the stored normal is the triangle's own (to within
// an angle whose sine times the longest edge is at most size / 1000) and the triangle is not
// degenerate (twice its area at least 1 % of the longest edge squared); rounding moves it by far
// less than the margin.  The bound is then the distance between the box and the bounding box;
// for other triangles it is 0.
template <typename R>
__device__ inline float tri_box_bound(const TriD<R>& t, R lox, R loy, R loz, R size) {
    const double ax = t.v[0], ay = t.v[1], az = t.v[2], bx = t.v[3], by = t.v[4], bz = t.v[5], cx = t.v[6],
                 cy = t.v[7], cz = t.v[8];
    const double e1x = bx - ax, e1y = by - ay, e1z = bz - az, e2x = cx - ax, e2y = cy - ay, e2z = cz - az;
    const double e3x = cx - bx, e3y = cy - by, e3z = cz - bz;
    const double gx = e1y * e2z - e1z * e2y, gy = e1z * e2x - e1x * e2z, gz = e1x * e2y - e1y * e2x;
    const double nx = t.nrm[0], ny = t.nrm[1], nz = t.nrm[2];
    const double hx = ny * gz - nz * gy, hy = nz * gx - nx * gz, hz = nx * gy - ny * gx;   // n x g
    const double gg = gx * gx + gy * gy + gz * gz, nn = nx * nx + ny * ny + nz * nz, hh = hx * hx + hy * hy + hz * hz;
    const double L2 = fmax(e1x * e1x + e1y * e1y + e1z * e1z,
                           fmax(e2x * e2x + e2y * e2y + e2z * e2z, e3x * e3x + e3y * e3y + e3z * e3z));
    const double sz = (double)size;
    // sin^2(n, g) L2 <= (size / 1000)^2, and |g|^2 >= (L2 / 100)^2
    const bool normal_ok = hh * L2 <= 1e-6 * sz * sz * nn * gg && nn > 0.0;
    const bool shape_ok = gg >= 1e-4 * L2 * L2 && L2 > 0.0;
    if (!(normal_ok && shape_ok)) return 0.0f;
    const double lx = lox, ly = loy, lz = loz;
    auto gap = [&](double a, double b, double c, double lo) {
        const double tmin = fmin(a, fmin(b, c)), tmax = fmax(a, fmax(b, c));
        return fmax(0.0, fmax(tmin - (lo + sz), lo - tmax));
    };
    const double dx = gap(ax, bx, cx, lx), dy = gap(ay, by, cy, ly), dz = gap(az, bz, cz, lz);
    const double lb = sqrt(dx * dx + dy * dy + dz * dz) - 0.01 * sz;
    if (!(lb > 0.0)) return 0.0f;
    return (float)(lb * 0.999999);
}

// A triangle spans several cells, so the 27-cell walk of nearest_triangle meets it in several
// lists.  Met again, its distance is the same and only a strictly nearer triangle replaces the
// running nearest, so the result is the triangle of least distance, the first met of those of
// equal distance: the least (distance, position of first meeting).  The cell walk kernels write
// each cell's triangles once (head[cell] = (pool offset, count)), with their positions of first
// meeting, ordered by tri_box_bound over the cell's box; a particle inside that box (to within half the bound's margin) stops at the
// first triangle whose bound exceeds its nearest distance so far, as no later one can be nearer or
// as near.  A particle outside it (the grid's truncating cell index puts particles up to one cell
// below the grid into the first cell) takes every listed triangle.  A cell without a list (count
// -1: no particles when the lists were written, or more than kStencilTris triangles) takes the
// full walk.
template <typename R>
__device__ inline int nearest_triangle_listed(const Consts<R>& c, R x, R y, R z, const int2* __restrict__ ulist,
                                              const int2* __restrict__ uhead, const int* __restrict__ tstart,
                                              const int* __restrict__ tcount, const int* __restrict__ tids,
                                              const TriD<R>* __restrict__ tris, int want_drag, R* best, P3<R>* bdir) {
    int ci = 0, cj = 0, ck = 0;
    long long cid = -1;
    if (uhead && cell_coords(c, x, y, z, ci, cj, ck)) cid = cell_of(c, ci, cj, ck);
    const int2 h = (cid >= 0 && cid < c.ncell) ? uhead[cid] : make_int2(0, -1);
    const int n = h.y;
    if (n < 0) return nearest_triangle(c, x, y, z, tstart, tcount, tids, tris, want_drag, best, bdir);
    *best = (R)1e8;
    *bdir = p3<R>(0, 0, 0);
    int bid = -1, bpos = 0;
    const P3<R> p = p3(x, y, z);
    const R slack = c.cell * (R)0.005;
    const R lox = c.bminx + (R)ci * c.cell, loy = c.bminy + (R)cj * c.cell, loz = c.bminz + (R)ck * c.cell;
    const bool cut = x >= lox - slack && x <= lox + c.cell + slack && y >= loy - slack &&
                     y <= loy + c.cell + slack && z >= loz - slack && z <= loz + c.cell + slack;
    const int2* L = ulist + h.x;
    for (int e = 0; e < n; ++e) {
        const int2 en = L[e];
        if (cut && (R)__uint_as_float((unsigned)en.y & ~kWalkPosMask) > *best) break;
        const TriD<R>& tr = tris[en.x];
        if (tr.is_drag != want_drag) continue;
        P3<R> dv;
        const R dm = tri_distance(tr, p, &dv);
        const int pos = (int)((unsigned)en.y & kWalkPosMask);
        if (dm < *best || (dm == *best && bid >= 0 && pos < bpos)) {
            *best = dm;
            *bdir = dv;
            bid = en.x;
            bpos = pos;
        }
    }
    return bid;
}

// The wall of the negative-pressure term (radix DeviceProcess_Kernel.h:1110-1134): every triangle
// of the 27-cell stencil's lists, drag or not; the first listed triangle of each cell replaces the
// running nearest (the reference's l == 0 branch) unless it is that triangle already, the others
// replace it when strictly nearer.  Returns the id (-1 for none) and the distance in *best.
template <typename R>
__device__ inline int negp_triangle(const Consts<R>& c, R x, R y, R z, const int* __restrict__ tstart,
                                    const int* __restrict__ tcount, const int* __restrict__ tids,
                                    const TriD<R>* __restrict__ tris, R* best) {
    *best = (R)1e8;
    int nid = -1;
    int ci, cj, ck;
    if (!cell_coords(c, x, y, z, ci, cj, ck)) return nid;
    const long long cid = cell_of(c, ci, cj, ck);
    if (cid < 0 || cid >= c.ncell) return nid;
    const P3<R> p = p3(x, y, z);
    for (int k = -1; k <= 1; ++k)
        for (int j = -1; j <= 1; ++j)
            for (int i = -1; i <= 1; ++i) {
                const long long m = cell_of(c, ci + i, cj + j, ck + k);
                if (m < 0 || m >= c.ncell) continue;
                const int n = tcount[m];
                const int a = tstart[m];
                for (int l = 0; l < n; ++l) {
                    const int tid = tids[a + l];
                    if (tid >= 0 && tid < c.ntri && tid != nid) {
                        P3<R> dv;
                        const R dm = tri_distance(tris[tid], p, &dv);
                        if (l == 0) { *best = dm; nid = tid; }
                        if (dm < *best) { *best = dm; nid = tid; }
                    }
                }
            }
    return nid;
}

// Deterministic pow, exp(y log x) in double rounded once to R, from IEEE-exact operations only
// (+ - * /, frexp, ldexp, floor) so that host and device give the same bits whatever their libm:
//   log x = e ln2 + 2 atanh(s), s = (m - 1)/(m + 1), m in [sqrt(1/2), sqrt(2)), series to s^23
//   exp z = 2^k exp(r), z = k ln2 + r (Cody-Waite split of ln2), |r| <= ln2/2, Taylor to r^15
__device__ inline double det_log(double x) {
    int e = 0;
    double m = frexp(x, &e);
    if (m < 0.70710678118654752440) {
        m = m * 2.0;
        e = e - 1;
    }
    const double s = (m - 1.0) / (m + 1.0), s2 = s * s;
    double p = 1.0 / 23.0;
    for (int k = 10; k >= 0; --k) p = p * s2 + 1.0 / (double)(2 * k + 1);
    const double ln2_hi = 6.93147180369123816490e-01, ln2_lo = 1.90821492927058770002e-10;
    return (double)e * ln2_hi + ((double)e * ln2_lo + 2.0 * s * p);
}
__device__ inline double det_exp(double z) {
    if (z > 709.0) return INFINITY;
    if (z < -745.0) return 0.0;
    const double ln2_hi = 6.93147180369123816490e-01, ln2_lo = 1.90821492927058770002e-10;
    const double k = floor(z / 0.69314718055994530942 + 0.5);
    const double r = (z - k * ln2_hi) - k * ln2_lo;
    double p = 1.0;
    for (int j = 15; j >= 1; --j) p = 1.0 + p * r / (double)j;
    return ldexp(p, (int)k);
}
template <typename R> __device__ inline R powd(R x, R y) {
    if (y == R(0)) return R(1);
    if (x == R(0)) return y > R(0) ? R(0) : (R)INFINITY;
    return (R)det_exp((double)y * det_log((double)x));
}

}  // namespace mpsr
