// spheres.hip -- rigid spheres carried by the fluid through Schiller-Naumann drag (no counterpart
// in the reference; the drum configuration adds 2,000 spheres of diameter 5 l0, density 2500).
//
// One phase per step, after the first binning (sorted fluid positions and velocities) and before
// the predictor:
//   k_sph_keys / sort / k_sph_ranges   spheres by cell (stable by sphere index)
//   k_sph_sample   per sphere (one wave): fluid velocity at the centre, u_f = sum w v / sum w over the fluid
//                  particles within re_lap (MPS weight, canonical neighbour order), and the drag
//                  F = 3 pi mu d (u_f - u_s)(1 + 0.15 Re^0.687) for Re < 1000, else Cd = 0.44
//   k_sph_react    per fluid particle: dv = -sum_s F_s (w_js / W_s) dt / m_f over the spheres in
//                  reach (a gather over the sphere cell lists: deterministic, and the reaction
//                  sums to -F_s dt over the fluid, momentum is conserved)
//   k_sph_move     per sphere: v += dt (F / m_s + (1 - rho_f / rho_s) g), x += dt v; a centre
//                  closer than d/2 to the nearest non-drag wall is pushed out along the wall
//                  normal and loses its inward normal velocity
// The predictor adds dv to the fluid velocity.  One wave per sphere (sampling and moving) or one
// thread per particle (the reaction).
#include "mpsr_internal.h"
#include "neighbors.h"
#include "walls_geom.h"

namespace mpsr {

template <typename R>
__global__ __launch_bounds__(256) void k_sph_keys(int S, Consts<R> c, const R4<R>* __restrict__ pos, uint32_t* keys,
                                                   int* vals) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= S) return;
    const R4<R> P = pos[k];
    uint32_t key = (uint32_t)c.ncell;
    int i, j, l;
    if (cell_coords(c, P.x, P.y, P.z, i, j, l)) {
        const long long m = cell_of(c, i, j, l);
        if (m >= 0 && m < c.ncell) key = (uint32_t)m;
    }
    keys[k] = key;
    vals[k] = k;
}

__global__ __launch_bounds__(256) void k_sph_ranges(int S, int ncell, const uint32_t* __restrict__ sk, int* cs, int* ce) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= S) return;
    const uint32_t k = sk[t];
    if (k == (uint32_t)ncell) return;
    if (t == 0 || sk[t - 1] != k) cs[k] = t;
    if (t == S - 1 || sk[t + 1] != k) ce[k] = t + 1;
}

// Schiller-Naumann drag on a sphere of diameter d at relative velocity du, evaluated in double
// (correctly rounded square root on host and device alike) and rounded once to R
template <typename R> __device__ inline void sn_force(R dux, R duy, R duz, R d, R rho, R nu, R* f) {
    const double ux = dux, uy = duy, uz = duz, dd = d, rr = rho, nn = nu;
    const double mag = sqrt(ux * ux + uy * uy + uz * uz);
    const double re = mag * dd / nn;
    double k;
    if (re < 1000.0) k = 3.0 * M_PI * rr * nn * dd * (1.0 + 0.15 * det_exp(0.687 * det_log(re > 0 ? re : 1.0)) * (re > 0 ? 1.0 : 0.0));
    else k = 0.5 * rr * 0.44 * (M_PI * dd * dd / 4.0) * mag;
    f[0] = (R)(k * ux);
    f[1] = (R)(k * uy);
    f[2] = (R)(k * uz);
}

// One wave per sphere: the lanes test the candidates of a stencil cell in parallel (lane L takes
// the cell's L-th, (L+64)-th, ... particle), then every lane adds the hits in ascending candidate
// order, read out of the hit lanes by shuffle, so the sums are the sequential canonical-order sums.
template <typename R>
__global__ __launch_bounds__(256) void k_sph_sample(int S, Consts<R> c, const R4<R>* __restrict__ pos,
                                                     const R4<R>* __restrict__ vel, const R4<R>* __restrict__ spos,
                                                     const R4<R>* __restrict__ svel, const int* __restrict__ cs,
                                                     const int* __restrict__ ce, R d, R rho, R nu, R4<R>* force,
                                                     R* wsum) {
    const int k = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const int lane = lane_id();
    if (k >= S) return;
    const R4<R> X = pos[k];
    const R re = c.re_lap, mind = c.mind_lap;
    R w = 0, ux = 0, uy = 0, uz = 0;
    int ci, cj, ck;
    if (cell_coords(c, X.x, X.y, X.z, ci, cj, ck)) {
        const long long cid = cell_of(c, ci, cj, ck);
        if (cid >= 0 && cid < c.ncell) {
            for (int kk = -1; kk <= 1; ++kk)
                for (int jj = -1; jj <= 1; ++jj)
                    for (int ii = -1; ii <= 1; ++ii) {
                        const long long m = cell_of(c, ci + ii, cj + jj, ck + kk);
                        if (m < 0 || m >= c.ncell) continue;
                        const int a = cs[m];
                        if (a < 0) continue;
                        const int b = ce[m];
                        for (int base = a; base < b; base += 64) {
                            const int q = base + lane;
                            R wq = 0, vx = 0, vy = 0, vz = 0;
                            bool hit = false;
                            if (q < b) {
                                const R4<R> Q = spos[q];
                                R dd = dist3(X.x, X.y, X.z, Q.x, Q.y, Q.z);
                                if (dd < re) {
                                    if (dd < mind) dd = mind;
                                    wq = weight(re, dd);
                                    const R4<R> V = svel[q];
                                    vx = V.x * wq;
                                    vy = V.y * wq;
                                    vz = V.z * wq;
                                    hit = true;
                                }
                            }
                            unsigned long long m64 = __ballot(hit);
                            while (m64) {
                                const int h = __ffsll((long long)m64) - 1;
                                m64 &= m64 - 1ull;
                                w = w + __shfl(wq, h, 64);
                                ux = ux + __shfl(vx, h, 64);
                                uy = uy + __shfl(vy, h, 64);
                                uz = uz + __shfl(vz, h, 64);
                            }
                        }
                    }
        }
    }
    if (lane != 0) return;
    wsum[k] = w;
    R f[3] = {0, 0, 0};
    if (w > 0) {
        const R4<R> Vs = vel[k];
        sn_force<R>(ux / w - Vs.x, uy / w - Vs.y, uz / w - Vs.z, d, rho, nu, f);
    }
    force[k] = mk4<R>(f[0], f[1], f[2], R(0));
}

template <typename R>
__global__ __launch_bounds__(256) void k_sph_react(int N, Consts<R> c, const uint32_t* __restrict__ skeys,
                                                    const R4<R>* __restrict__ spos, const R4<R>* __restrict__ pos,
                                                    const int* __restrict__ sperm, const int* __restrict__ scs,
                                                    const int* __restrict__ sce, const R4<R>* __restrict__ force,
                                                    const R* __restrict__ wsum, R mf, R4<R>* dv) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N) return;
    R ax = 0, ay = 0, az = 0;
    if (skeys[t] != (uint32_t)c.ncell) {
        const R4<R> P = spos[t];
        const R re = c.re_lap, mind = c.mind_lap;
        // spheres are listed per cell like the particles: the same 27-cell walk over their ranges
        for_candidates(c, scs, sce, P.x, P.y, P.z, -1, [&](int u) {
            const int k = sperm[u];
            const R4<R> X = pos[k];
            R dd = dist3(P.x, P.y, P.z, X.x, X.y, X.z);
            const R w = wsum[k];
            if (dd < re && w > 0) {
                if (dd < mind) dd = mind;
                const R cf = -(weight(re, dd) / w) * c.dt / mf;
                const R4<R> F = force[k];
                ax = ax + F.x * cf;
                ay = ay + F.y * cf;
                az = az + F.z * cf;
            }
        });
    }
    dv[t] = mk4<R>(ax, ay, az, R(0));
}

// One wave per sphere.  The nearest non-drag wall of the moved centre is nearest_triangle's: the
// lanes take the triangles of each stencil cell's list in turn (lane L the L-th, (L+64)-th, ...),
// each keeps its first strictly nearest, and the wave keeps the smallest distance at the smallest
// position in the traversal: the triangle the sequential walk keeps (it replaces its best only when
// strictly nearer, and a listing repeated in another cell is never strictly nearer).
template <typename R>
__global__ __launch_bounds__(256) void k_sph_move(int S, Consts<R> c, R4<R>* pos, R4<R>* vel,
                                                   const R4<R>* __restrict__ force, const int* __restrict__ tstart,
                                                   const int* __restrict__ tcount, const int* __restrict__ tids,
                                                   const TriD<R>* __restrict__ tris, R d, R rho_s, R rho_f) {
    const int k = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const int lane = lane_id();
    if (k >= S) return;   // the same for the whole wave
    const R ms = rho_s * (R)M_PI * d * d * d / (R)6;
    const R inv = (R)1 / ms;
    const R buoy = (R)1 - rho_f / rho_s;
    const R4<R> F = force[k];
    const R ax = F.x * inv + c.gx_acc * buoy, ay = F.y * inv + c.gy_acc * buoy, az = F.z * inv + c.gz_acc * buoy;
    const R4<R> V0 = vel[k], X0 = pos[k];
    R vx = V0.x + ax * c.dt, vy = V0.y + ay * c.dt, vz = V0.z + az * c.dt;
    R x = X0.x + vx * c.dt, y = X0.y + vy * c.dt, z = X0.z + vz * c.dt;
    // this lane's first strictly nearest triangle, and its position in the traversal
    R best = (R)1e8;
    P3<R> dir = p3<R>(0, 0, 0);
    int at = 0x7fffffff;
    int ci, cj, ck;
    if (cell_coords(c, x, y, z, ci, cj, ck)) {
        const long long cid = cell_of(c, ci, cj, ck);
        if (cid >= 0 && cid < c.ncell) {
            const P3<R> p = p3(x, y, z);
            int seen = 0;   // traversal positions before this cell
            for (int kk = -1; kk <= 1; ++kk)
                for (int jj = -1; jj <= 1; ++jj)
                    for (int ii = -1; ii <= 1; ++ii) {
                        const long long m = cell_of(c, ci + ii, cj + jj, ck + kk);
                        if (m < 0 || m >= c.ncell) continue;
                        const int n = tcount[m], a0 = tstart[m];
                        for (int e = lane; e < n; e += 64) {
                            const TriD<R>& tr = tris[tids[a0 + e]];
                            if (tr.is_drag != 0) continue;
                            P3<R> dv;
                            const R dm = tri_distance(tr, p, &dv);
                            if (dm < best) {
                                best = dm;
                                dir = dv;
                                at = seen + e;
                            }
                        }
                        seen += n;
                    }
        }
    }
    // the wave's smallest distance, ties to the earliest position
    for (int o = 32; o > 0; o >>= 1) {
        const R ob = __shfl_xor(best, o, 64);
        const int oa = __shfl_xor(at, o, 64);
        const R odx = __shfl_xor(dir.x, o, 64), ody = __shfl_xor(dir.y, o, 64), odz = __shfl_xor(dir.z, o, 64);
        if (ob < best || (ob == best && oa < at)) {
            best = ob;
            at = oa;
            dir = p3<R>(odx, ody, odz);
        }
    }
    const R half = d / (R)2;
    if (best < half && best > 0) {
        const R nx = -dir.x / best, ny = -dir.y / best, nz = -dir.z / best;
        const R push = half - best;
        x = x + nx * push;
        y = y + ny * push;
        z = z + nz * push;
        const R vn = vx * nx + vy * ny + vz * nz;
        if (vn < 0) {
            vx = vx - nx * vn;
            vy = vy - ny * vn;
            vz = vz - nz * vn;
        }
    }
    if (lane == 0) {
        pos[k] = mk4<R>(x, y, z, R(0));
        vel[k] = mk4<R>(vx, vy, vz, R(0));
    }
}

template <typename R> void launch_spheres(Dev<R>& d, SphDev<R>& s, hipStream_t st) {
    if (s.S <= 0) return;
    const dim3 gs(grid_for(s.S, 256)), b(256);
    hipLaunchKernelGGL(k_sph_keys<R>, gs, b, 0, st, s.S, d.c, s.pos, s.keys, s.vals);
    const int which = radix_sort_pairs(s.keys, s.vals, nullptr, s.S, s.S, d.key_bits, s.sort, st);
    const uint32_t* sk = which ? s.sort.k1 : s.keys;
    s.perm = which ? s.sort.v1 : s.vals;
    fill_int(s.cs, -1, d.c.ncell, st);
    fill_int(s.ce, -1, d.c.ncell, st);
    hipLaunchKernelGGL(k_sph_ranges, gs, b, 0, st, s.S, (int)d.c.ncell, sk, s.cs, s.ce);
    hipLaunchKernelGGL(k_sph_sample<R>, dim3(grid_for((long long)s.S * 64, 256)), b, 0, st, s.S, d.c, s.pos, s.vel,
                       d.spos, d.svel, d.cell_start,
                       d.cell_end, s.d, s.rho_f, s.nu_f, s.force, s.wsum);
    if (d.N > 0)
        hipLaunchKernelGGL(k_sph_react<R>, dim3(grid_for(d.N, 256)), b, 0, st, d.N, d.c, d.skeys, d.spos, s.pos, s.perm,
                           s.cs, s.ce, s.force, s.wsum, s.mf, d.sph_dv);
    hipLaunchKernelGGL(k_sph_move<R>, dim3(grid_for((long long)s.S * 64, 256)), b, 0, st, s.S, d.c, s.pos, s.vel,
                       s.force, d.cell_tri_start,
                       d.cell_tri_count, d.uniq_tri, d.tris, s.d, s.rho_s, s.rho_f);
}

template void launch_spheres<float>(Dev<float>&, SphDev<float>&, hipStream_t);
template void launch_spheres<double>(Dev<double>&, SphDev<double>&, hipStream_t);

}  // namespace mpsr
