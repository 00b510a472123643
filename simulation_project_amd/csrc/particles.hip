// particles.hip -- particle binning and the three MPS neighbour operators.
//
//   k_bin_hash      RegisterParticleTopology_Kernel (Source/DeviceProcess_Kernel.h:6-58):
//                   zone filter, in -> out copy, cell key per slot
//   (radix sort)    thrust::sort_by_key (radix/.../DeviceProcess.cu:207-220)
//   k_bin_ranges    reorderDataAndFindCellStartD (radix/.../DeviceProcess_Kernel.h:54-124):
//                   cellStart/cellEnd, gather of position and velocity into sorted order
//   k_explicit      CalcExplicitly_Kernel (Kern :283-441)
//   k_relocate      RelocateParticleData_Kernel (Kern :61-92)
//   k_pnd           CalcPressureExplicit_Kernel (Kern :702-801)
//   k_gradient      CalcImplicitExplicitly_Kernel (Kern :804-912), optionally with the radix
//                   variant's negative-pressure term (radix/.../DeviceProcess_Kernel.h:1214-1245)
//   k_turbulence    CalcTurbulenceViscosity_Kernel (radix/.../DeviceProcess_Kernel.h:1273-1476)
//
// The operator kernels run one thread per sorted index t.  Neighbours are visited in the
// canonical order: stencil z outer, x inner (Kern :317-330), ascending sorted index inside a
// cell.  Consecutive threads share cells, so a wave walks the same 27 ranges and its neighbour
// records come from L2 in whole 16 B (float4) or 32 B (double4) loads.  Results go to the
// slot-order arrays through the sorted->slot permutation.
#include <cfloat>
#include "mpsr_internal.h"
#include "neighbors.h"
#include "walls_geom.h"

namespace mpsr {

__constant__ ZTable kZgrad = {
    21,
    {0.1, 0.2, 0.3, 0.4, 0.5, 0.6, 0.7, 0.8, 0.9, 1.0, 1.1, 1.2, 1.3, 1.4, 1.5, 1.6, 1.7, 1.8, 1.9, 2.0, 2.1},
    {-5.0474610E+03, -1.1272200E+02, -4.3342000E+01, -2.6328000E+01, -1.9648400E+01, -1.5636200E+01,
     -1.2591500E+01, -1.0988800E+01, -9.0692000E+00, -7.5385000E+00, -6.9148000E+00, -5.9144000E+00,
     -5.4670000E+00, -5.0626000E+00, -4.6948000E+00, -3.6959000E+00, -2.7022800E+00, -2.4819500E+00,
     -1.4080400E+00, -5.5263200E-01, -4.9999800E-01},
    {2.9725900E+01, 1.8453700E+01, 1.4119500E+01, 1.1486700E+01, 9.5218600E+00, 7.9582400E+00,
     6.6990900E+00, 5.6002100E+00, 4.6932900E+00, 3.9394400E+00, 3.2479600E+00, 2.6565200E+00,
     2.1098200E+00, 1.6035600E+00, 1.1340800E+00, 7.6449000E-01, 4.9426200E-01, 2.4606700E-01,
     1.0526300E-01, 4.9999800E-02, 0.0000000E+00}};

__constant__ ZTable kZlap = {
    25,
    {0.1, 0.2, 0.3, 0.4, 0.5, 0.6, 0.7, 0.8, 0.9, 1.0, 1.1, 1.2, 1.3, 1.4, 1.5, 1.6, 1.7, 1.8, 1.9, 2.0,
     2.4, 2.8, 3.2, 3.6, 4.0},
    {-9.6379E+03, -2.3979E+02, -1.0777E+02, -7.4088E+01, -6.0259E+01, -5.2803E+01, -4.8681E+01,
     -4.4630E+01, -4.1171E+01, -3.8583E+01, -3.6846E+01, -3.5476E+01, -3.4183E+01, -3.1589E+01,
     -2.9258E+01, -2.7047E+01, -2.5690E+01, -2.3305E+01, -2.0982E+01, -1.9268E+01, -1.6887E+01,
     -1.1091E+01, -6.5068E+00, -3.7565E+00, -1.3253E+00},
    {1.1497E+02, 9.0990E+01, 8.0213E+01, 7.2804E+01, 6.6778E+01, 6.1498E+01, 5.6630E+01, 5.2167E+01,
     4.8050E+01, 4.4191E+01, 4.0507E+01, 3.6959E+01, 3.3541E+01, 3.0382E+01, 2.7456E+01, 2.4751E+01,
     2.2182E+01, 1.9852E+01, 1.7754E+01, 1.5827E+01, 9.0719E+00, 4.6355E+00, 2.0327E+00, 5.3011E-01,
     0.0000E+00}};

// ---------------------------------------------------------------------------------------------
template <typename R>
__device__ inline void bin_hash_body(int blk, int N, Consts<R> c, PState<R> in, PState<R> out, uint32_t* keys,
                                     int* vals) {
    const int s = blk * blockDim.x + threadIdx.x;
    if (s >= N) return;
    const int id = in.id[s];
    const R x = in.px[s], y = in.py[s], z = in.pz[s];
    if (id < 0 || !include(c, x, y, z)) {
        out.id[s] = -1;
    } else {
        out.id[s] = id;
        out.px[s] = x; out.py[s] = y; out.pz[s] = z;
        out.vx[s] = in.vx[s]; out.vy[s] = in.vy[s]; out.vz[s] = in.vz[s];
        out.p[s] = in.p[s]; out.n[s] = in.n[s]; out.temp[s] = in.temp[s];
        out.nu[s] = in.nu[s]; out.solid[s] = in.solid[s]; out.type[s] = in.type[s];
    }
    keys[s] = bin_key(c, id, x, y, z);
    vals[s] = s;
}
template <typename R>
__global__ __launch_bounds__(kMaxBlock) void k_bin_hash(int N, Consts<R> c, PState<R> in, PState<R> out, uint32_t* keys,
                                                   int* vals) {
    bin_hash_body<R>(blockIdx.x, N, c, in, out, keys, vals);
}

template <typename R>
__device__ inline void bin_ranges_body(int blk, int N, int ncell, const uint32_t* __restrict__ sk,
                                       const int* __restrict__ perm, int* cs, int* ce, PState<R> in,
                                       R4<R>* __restrict__ spos, R4<R>* __restrict__ svel) {
    const int t = blk * blockDim.x + threadIdx.x;
    if (t >= N) return;
    const uint32_t k = sk[t];
    const uint32_t kp = t > 0 ? sk[t - 1] : 0xffffffffu;
    const uint32_t sent = (uint32_t)ncell;
    if (k != kp) {
        if (k != sent) cs[k] = t;
        if (t > 0 && kp != sent) ce[kp] = t;
    }
    if (t == N - 1 && k != sent) ce[k] = N;
    const int s = perm[t];
    spos[t] = mk4<R>(in.px[s], in.py[s], in.pz[s], R(0));
    svel[t] = mk4<R>(in.vx[s], in.vy[s], in.vz[s], R(0));
}
template <typename R>
__global__ __launch_bounds__(kMaxBlock) void k_bin_ranges(int N, int ncell, const uint32_t* __restrict__ sk,
                                                     const int* __restrict__ perm, int* cs, int* ce, PState<R> in,
                                                     R4<R>* __restrict__ spos, R4<R>* __restrict__ svel) {
    bin_ranges_body<R>(blockIdx.x, N, ncell, sk, perm, cs, ce, in, spos, svel);
}

// ---------------------------------------------------------------------------------------------
template <typename R>
__device__ inline void explicit_body(int blk, int N, Consts<R> c, const int* __restrict__ perm, PState<R> in,
                                     PState<R> out, const R4<R>* __restrict__ spos,
                                     const R4<R>* __restrict__ svel, const int* __restrict__ cs,
                                     const int* __restrict__ ce, const R* __restrict__ stl_mag,
                                     const R4<R>* __restrict__ stl_dir, const int* __restrict__ stl_id,
                                     const R4<R>* __restrict__ drag_acc, const R4<R>* __restrict__ sph_dv,
                                     const TriD<R>* __restrict__ tris,
                                     const mpsr_tri_params* __restrict__ models,
                                     const R* __restrict__ turb_nu) {
    const int t = blk * blockDim.x + threadIdx.x;
    if (t >= N) return;
    const int s = perm[t];
    if (out.id[s] < 0) return;
    const R4<R> P = spos[t];
    const R x = P.x, y = P.y, z = P.z;
    if (!include(c, x, y, z)) return;
    const R4<R> V0 = svel[t];
    R lx = 0, ly = 0, lz = 0, ax = 0, ay = 0, az = 0, qx = 0, qy = 0, qz = 0;
    const R re = c.re_lap, mind = c.mind_lap;
    // per group of 4 candidates: the distances first, then the velocities of the hits, loaded
    // together, then the sums in the canonical order
    for_candidate_groups(c, cs, ce, spos, x, y, z, [&](int t0, const R4<R> (&Q)[4], int m) {
        R dd[4];
        bool hit[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            dd[u] = dist3(x, y, z, Q[u].x, Q[u].y, Q[u].z);
            hit[u] = u < m && t0 + u != t && dd[u] < re;
        }
        R4<R> V[4];   // read only where hit[u]
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (hit[u]) V[u] = svel[t0 + u];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (!hit[u]) continue;
            R d = dd[u];
            if (d < mind) d = mind;
            const R w = weight(re, d);
            const R4<R>& V1 = V[u];
            lx += (V1.x - V0.x) * w;
            ly += (V1.y - V0.y) * w;
            lz += (V1.z - V0.z) * w;
            if (c.expl) {
                const R a = c.art_p * w / (d * d);
                ax = ax + (Q[u].x - x) * a;
                ay = ay + (Q[u].y - y) * a;
                az = az + (Q[u].z - z) * a;
                if (d < c.rc) {
                    const R cc = (c.rc - d) / d;
                    qx = qx + (Q[u].x - x) * cc;
                    qy = qy + (Q[u].y - y) * cc;
                    qz = qz + (Q[u].z - z) * cc;
                }
            }
        }
    });
    const R dw = stl_mag[t];
    const int tid = stl_id[t];
    if (tid >= 0 && tid < c.ntri) {
        if (dw > 0 && dw < re) {
            const R ww = ztable_eval(kZlap, dw, c.l0);
            const mpsr_float3 mv = models[tris[tid].model].Velocity;
            const R v1x = (R)mv.x, v1y = (R)mv.y, v1z = (R)mv.z;
            lx += c.mu_wall * (v1x - V0.x) * ww;
            ly += c.mu_wall * (v1y - V0.y) * ww;
            lz += c.mu_wall * (v1z - V0.z) * ww;
            if (c.expl) {
                const R4<R> D = stl_dir[t];
                const R a = c.art_p * ww / (dw * dw);
                ax = ax + D.x * a;
                ay = ay + D.y * a;
                az = az + D.z * a;
                if (dw < c.rc) {
                    const R cc = (c.rc - dw) / dw;
                    qx = qx + D.x * cc;
                    qy = qy + D.y * cc;
                    qz = qz + D.z * cc;
                }
            }
        }
    }
    lx *= c.lapc; ly *= c.lapc; lz *= c.lapc;
    R dvx = 0, dvy = 0, dvz = 0;
    R accx = R(0) + c.gx_acc, accy = R(0) + c.gy_acc, accz = R(0) + c.gz_acc;
    if (c.has_drag) {
        const R4<R> Dg = drag_acc[t];
        dvx = Dg.x; dvy = Dg.y; dvz = Dg.z;
    }
    // with the radix variant's turbulence the viscosity is nu + nu_t (radix
    // DeviceProcess_Kernel.h:545-552); its k-gradient term (:558-562) is scaled by the integer
    // quotient 2/3 = 0 there, so it changes no finite velocity and is not formed
    const R nu = c.turb ? in.nu[s] + turb_nu[s] : in.nu[s];
    lx = lx * nu; ly = ly * nu; lz = lz * nu;
    accx = accx + lx; accy = accy + ly; accz = accz + lz;
    R vx = (V0.x + accx * c.dt) - dvx;
    R vy = (V0.y + accy * c.dt) - dvy;
    R vz = (V0.z + accz * c.dt) - dvz;
    if (c.has_sph) {
        const R4<R> Sd = sph_dv[t];
        vx = vx + Sd.x;
        vy = vy + Sd.y;
        vz = vz + Sd.z;
    }
    if (c.expl) {
        accx = ax * c.art_coeff + qx * (R)0.01;
        accy = ay * c.art_coeff + qy * (R)0.01;
        accz = az * c.art_coeff + qz * (R)0.01;
        vx = vx + accx * c.dt;
        vy = vy + accy * c.dt;
        vz = vz + accz * c.dt;
    }
    if (isfinite(vx) && isfinite(vy) && isfinite(vz)) {
        out.px[s] = x + c.dt * vx;
        out.py[s] = y + c.dt * vy;
        out.pz[s] = z + c.dt * vz;
        out.vx[s] = vx; out.vy[s] = vy; out.vz[s] = vz;
    }
}
template <typename R>
__global__ __launch_bounds__(kMaxBlock) void k_explicit(int N, Consts<R> c, const int* __restrict__ perm, PState<R> in,
                                                   PState<R> out, const R4<R>* __restrict__ spos,
                                                   const R4<R>* __restrict__ svel, const int* __restrict__ cs,
                                                   const int* __restrict__ ce, const R* __restrict__ stl_mag,
                                                   const R4<R>* __restrict__ stl_dir, const int* __restrict__ stl_id,
                                                   const R4<R>* __restrict__ drag_acc, const R4<R>* __restrict__ sph_dv,
                                                   const TriD<R>* __restrict__ tris,
                                                   const mpsr_tri_params* __restrict__ models,
                                                   const R* __restrict__ turb_nu) {
    const int blk = xcd_block(N);
    if (blk < 0) return;
    explicit_body<R>(blk, N, c, perm, in, out, spos, svel, cs, ce, stl_mag, stl_dir, stl_id, drag_acc, sph_dv, tris,
                     models, turb_nu);
}

// ---------------------------------------------------------------------------------------------
template <typename R>
__device__ inline void relocate_body(int blk, int N, Consts<R> c, PState<R> in, PState<R> out) {
    const int s = blk * blockDim.x + threadIdx.x;
    if (s >= N) return;
    if (out.id[s] < 0) return;
    const R x = out.px[s], y = out.py[s], z = out.pz[s];
    if (!include(c, x, y, z)) {
        in.id[s] = -1;
        in.px[s] = c.cmaxx + c.tol;
        in.py[s] = c.cmaxy + c.tol;
        in.pz[s] = c.cmaxz + c.tol;
        return;
    }
    in.id[s] = out.id[s];
    in.px[s] = x; in.py[s] = y; in.pz[s] = z;
    in.vx[s] = out.vx[s]; in.vy[s] = out.vy[s]; in.vz[s] = out.vz[s];
    in.p[s] = out.p[s]; in.n[s] = out.n[s]; in.temp[s] = out.temp[s];
    in.nu[s] = out.nu[s]; in.solid[s] = out.solid[s]; in.type[s] = out.type[s];
}
template <typename R>
__global__ __launch_bounds__(kMaxBlock) void k_relocate(int N, Consts<R> c, PState<R> in, PState<R> out) {
    relocate_body<R>(blockIdx.x, N, c, in, out);
}

// ---------------------------------------------------------------------------------------------
// the last, partial group of a row's neighbour list, padded with its last index
__device__ inline void list_tail(int* nbq, int lbase, int lane, int kept, int q1, int q2, int q3) {
    if (kept & 3) {
        const int used = kept & 3;
        int qq[4] = {q3, q3, q3, q3};
        if (used == 2) qq[0] = q2;
        if (used == 3) { qq[0] = q1; qq[1] = q2; }
        quad_store(nbq + sell_at(lbase, lane, kept - used), qq[0], qq[1], qq[2], qq[3]);
    }
}
// PND (and the EOS pressure when eos != 0).  spres[t] receives the slot's out-pressure after the
// kernel, which is what the gradient kernel reads for neighbours (Kern :860).
//
// With PPE the same pass also gives the pressure Poisson rows their start: the row is active
// (unknown pressure) when the particle is alive, binned, inside the zone and its PND exceeds
// beta n0; over the Laplacian radius it counts the neighbours (an upper bound of the row's
// entries: the Dirichlet ones are dropped when the row is filled) and sums their weights in FP64
// for the diagonal.  A row that is not active or has no neighbour is a Dirichlet row (p = 0,
// diagonal 1).  The PND sum sees the gradient-radius neighbours in the same order as without PPE,
// so both give the same bits.  The sorted indices of the first nb_width Laplacian-radius neighbours
// go to a per-row list (SELL-64/4 by sorted index) that the fill and the corrector read instead of
// scanning again; the fill forms their weights again from the positions, with the same operations.
// Without PPE and with nbq given, the gradient-radius neighbours go to the list instead (nb_width
// of them, nbn[t] their count, -1 for a skipped row), in walk order, for the corrector.
template <typename R, bool PPE>
__device__ inline void pnd_body(int blk, int N, Consts<R> c, int eos, const int* __restrict__ perm, PState<R> in,
                                PState<R> out, const R4<R>* __restrict__ spos, R* __restrict__ spres,
                                const int* __restrict__ cs, const int* __restrict__ ce,
                                const R* __restrict__ stl_mag, const int* __restrict__ stl_id,
                                const uint32_t* __restrict__ skeys, int* __restrict__ active,
                                int* __restrict__ rcnt, double* __restrict__ rdiag, int* __restrict__ rdir,
                                int* __restrict__ nbq, int* __restrict__ nbn,
                                int nb_width) {
    const int t = blk * blockDim.x + threadIdx.x;
    if (t >= N) return;
    const int s = perm[t];
    const R4<R> P = spos[t];
    const R x = P.x, y = P.y, z = P.z;
    auto dirichlet = [&]() {
        if constexpr (PPE) {
            active[t] = 0;
            rcnt[t] = 0;
            rdiag[t] = 1.0;
            rdir[t] = 1;
            nbn[t] = -1;
        } else if (nbn) {
            nbn[t] = -1;
        }
    };
    if (!include(c, x, y, z)) {
        in.p[s] = 0;
        spres[t] = out.p[s];
        dirichlet();
        return;
    }
    int ci, cj, ck;
    long long cid = -1;
    if (cell_coords(c, x, y, z, ci, cj, ck)) cid = cell_of(c, ci, cj, ck);
    if (cid < 0 || cid >= c.ncell) {
        in.p[s] = 0;
        spres[t] = out.p[s];
        dirichlet();
        return;
    }
    const R re = c.re_grad, mind = c.mind_grad;
    R dens = 0;
    int nl = 0;
    double wl = 0.0;
    if constexpr (PPE) {
        const R rl = c.re_lap, mindl = c.mind_lap;
        const int lbase = (t >> 6) * 64 * nb_width, lane = t & 63;
        int q0 = 0, q1 = 0, q2 = 0, q3 = 0;
        // the root only for candidates inside 1.01 times the larger radius (sqrt14 errs by at most
        // 0.18 %): a wave whose lanes all miss skips it
        const R reach = (rl > re ? rl : re) * (R)1.01;
        const R reach2 = reach * reach;
        for_candidates_rec<8>(c, cs, ce, spos, x, y, z, t, [&](int q, const R4<R>& Q) {
            const R d2 = (x - Q.x) * (x - Q.x) + (y - Q.y) * (y - Q.y) + (z - Q.z) * (z - Q.z);
            if (!(d2 < reach2)) return;
            const R d0 = build_sqrt(d2);
            if (d0 < re) {
                R d = d0;
                if (d < mind) d = mind;
                dens = dens + weight(re, d);
            }
            if (d0 < rl) {
                R d = d0;
                if (d < mindl) d = mindl;
                const R w = weight(rl, d);
                wl += (double)w;
                if (nl < nb_width) {
                    q0 = q1; q1 = q2; q2 = q3; q3 = q;
                    if ((nl & 3) == 3) {
                        const int e = sell_at(lbase, lane, nl - 3);
                        quad_store(nbq + e, q0, q1, q2, q3);
                    }
                }
                ++nl;
            }
        });
        list_tail(nbq, lbase, lane, nl < nb_width ? nl : nb_width, q1, q2, q3);
    } else if (nbq) {
        const int lbase = (t >> 6) * 64 * nb_width, lane = t & 63;
        int q0 = 0, q1 = 0, q2 = 0, q3 = 0;
        for_near_candidates(c, cs, ce, spos, x, y, z, t, re, [&](int q) {
            const R4<R> Q = spos[q];
            R d = dist3(x, y, z, Q.x, Q.y, Q.z);
            if (d < re) {
                if (d < mind) d = mind;
                dens = dens + weight(re, d);
                if (nl < nb_width) {
                    q0 = q1; q1 = q2; q2 = q3; q3 = q;
                    if ((nl & 3) == 3) quad_store(nbq + sell_at(lbase, lane, nl - 3), q0, q1, q2, q3);
                }
                ++nl;
            }
        });
        list_tail(nbq, lbase, lane, nl < nb_width ? nl : nb_width, q1, q2, q3);
        nbn[t] = nl;
    } else {
        for_candidates_within(c, cs, ce, x, y, z, t, re, [&](int q) {
            const R4<R> Q = spos[q];
            R d = dist3(x, y, z, Q.x, Q.y, Q.z);
            if (d < re) {
                if (d < mind) d = mind;
                dens = dens + weight(re, d);
            }
        });
    }
    const R riw = stl_mag[t];
    R ww = 0;
    const int tid = stl_id[t];
    if (tid >= 0 && tid < c.ntri)
        if (riw >= 0 && riw <= re) ww = ztable_eval(kZgrad, riw, c.l0);
    dens += ww;
    out.n[s] = dens;
    if constexpr (PPE) {
        const bool act = out.id[s] >= 0 && skeys[t] != (uint32_t)c.ncell && (double)dens > (double)c.beta_n0;
        const bool dr = !(act && wl > 0.0);
        active[t] = act ? 1 : 0;
        nbn[t] = nl;
        rcnt[t] = dr ? 0 : nl;
        rdiag[t] = dr ? 1.0 : wl;
        rdir[t] = dr ? 1 : 0;
    }
    R pout = out.p[s];
    if (eos) {
        R pr = 0;
        if (dens > c.beta_n0) pr = c.eos_coef * (dens - c.n0);
        if (isfinite(pr)) {
            out.p[s] = pr;
            in.p[s] = pr;
            pout = pr;
        }
    }
    spres[t] = pout;
}
template <typename R, bool PPE>
__global__ __launch_bounds__(kMaxBlock) void k_pnd(int N, Consts<R> c, int eos, const int* __restrict__ perm, PState<R> in,
                                              PState<R> out, const R4<R>* __restrict__ spos, R* __restrict__ spres,
                                              const int* __restrict__ cs, const int* __restrict__ ce,
                                              const R* __restrict__ stl_mag, const int* __restrict__ stl_id,
                                              const uint32_t* __restrict__ skeys, int* __restrict__ active,
                                              int* __restrict__ rcnt, double* __restrict__ rdiag, int* __restrict__ rdir,
                                              int* __restrict__ nbq, int* __restrict__ nbn,
                                              int nb_width) {
    pnd_body<R, PPE>(blockIdx.x, N, c, eos, perm, in, out, spos, spres, cs, ce, stl_mag, stl_id, skeys, active, rcnt,
                     rdiag, rdir, nbq, nbn, nb_width);
}

// ---------------------------------------------------------------------------------------------
template <typename R>
__device__ inline void gradient_body(int blk, int N, Consts<R> c, const int* __restrict__ perm, PState<R> in,
                                     PState<R> out, const R4<R>* __restrict__ spos,
                                     const R4<R>* __restrict__ svel, const R* __restrict__ spres,
                                     const int* __restrict__ cs, const int* __restrict__ ce,
                                     const R* __restrict__ stl_mag, const R4<R>* __restrict__ stl_dir,
                                     const int* __restrict__ stl_id, const int* __restrict__ tstart,
                                     const int* __restrict__ tcount, const int* __restrict__ tids,
                                     const TriD<R>* __restrict__ tris,
                                     const NegModel<R>* __restrict__ negm, NbList nb) {
    const int t = blk * blockDim.x + threadIdx.x;
    if (t >= N) return;
    const int s = perm[t];
    if (out.id[s] < 0) return;
    const R4<R> P = spos[t];
    const R x = P.x, y = P.y, z = P.z;
    if (!include(c, x, y, z)) return;
    R gx = 0, gy = 0, gz = 0, wx = 0, wy = 0, wz = 0;
    const R pi = in.p[s];
    const R re = c.re_grad, mind = c.mind_grad;
    auto term = [&](const R4<R>& Q, R pq) {
        R d = dist3(x, y, z, Q.x, Q.y, Q.z);
        if (d < re) {
            if (d < mind) d = mind;
            const R w = weight(re, d);
            const R pc = (R)(0.5 * (double)(pq + pi) * (double)w / (double)(d * d));
            gx = gx + (Q.x - x) * pc;
            gy = gy + (Q.y - y) * pc;
            gz = gz + (Q.z - z) * pc;
        }
    };
    const int nn = nb.n ? nb.n[t] : -1;
    if (nn >= 0 && nn <= nb.width) {
        // every neighbour within the gradient radius is listed (by the PPE path's PND pass within
        // the Laplacian radius, which is not smaller), in the stencil walk's order, so the same
        // distances pick the same neighbours in the same order.
        // A group's records (position and pressure of its 4 entries; the last group is padded with
        // a valid index) are loaded together with the next group's indices and waited for once.
        const int lbase = (t >> 6) * 64 * nb.width, lane = t & 63;
        int q4[4] = {t, t, t, t};
        if (nn > 0) quad_load(nb.q + sell_at(lbase, lane, 0), q4);
        for (int j = 0; j < nn; j += 4) {
            R4<R> Q[4];
            R pq[4];
            int n4[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) Q[u] = spos[q4[u]];
#pragma unroll
            for (int u = 0; u < 4; ++u) pq[u] = spres[q4[u]];
            quad_load(nb.q + sell_at(lbase, lane, j + 4 < nn ? j + 4 : j), n4);
#pragma unroll
            for (int u = 0; u < 4; ++u) asm volatile("" ::"v"(Q[u].x), "v"(Q[u].y), "v"(Q[u].z), "v"(pq[u]));
            const int m = nn - j < 4 ? nn - j : 4;
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (u < m) term(Q[u], pq[u]);
#pragma unroll
            for (int u = 0; u < 4; ++u) q4[u] = n4[u];
        }
    } else {
        for_candidates_within(c, cs, ce, x, y, z, t, re, [&](int q) { term(spos[q], spres[q]); });
    }
    gx *= c.c1; gy *= c.c1; gz *= c.c1;
    const int tid = stl_id[t];
    if (tid >= 0 && tid < c.ntri) {
        const R r = stl_mag[t];
        if (r <= c.l0 && r > 0) {
            const R dr = c.l0 - r;
            const R mc = c.c2 * dr / r;
            const R4<R> D = stl_dir[t];
            wx = mc * (-D.x);
            wy = mc * (-D.y);
            wz = mc * (-D.z);
        }
    }
    R grx = gx + wx, gry = gy + wy, grz = gz + wz;
    if (c.negp) {
        // radix DeviceProcess_Kernel.h:1214-1245: gas pressure PressureGas acting from the nearest
        // wall of a rotating or flagged model within InfluenceRegion l0, along its DirectionVector
        R nx = 0, ny = 0, nz = 0, best;
        const int nid = negp_triangle(c, x, y, z, tstart, tcount, tids, tris, &best);
        if (nid != -1 && best != (R)1e8) {
            const NegModel<R> m = negm[tris[nid].model];
            if (m.active && best <= m.infl * c.l0) {
                const R iw = ztable_eval(kZgrad, c.l0, c.l0);
                const R pc = (R)(0.5 * (double)(m.pgas + pi) * (double)iw / (double)(c.l0 * c.l0));
                nx = pc * m.dx;
                ny = pc * m.dy;
                nz = pc * m.dz;
            }
        }
        nx *= c.c1; ny *= c.c1; nz *= c.c1;
        grx = grx + nx; gry = gry + ny; grz = grz + nz;
    }
    const R ix = c.c3 * grx, iy = c.c3 * gry, iz = c.c3 * grz;
    if (isfinite(ix) && isfinite(iy) && isfinite(iz)) {
        out.px[s] = x + ix * c.dt;
        out.py[s] = y + iy * c.dt;
        out.pz[s] = z + iz * c.dt;
        const R4<R> V = svel[t];
        R vx = V.x + ix, vy = V.y + iy, vz = V.z + iz;
        // Utility_Inline.h:2008-2027 CorrectVelocity, per component
        if (fabs(vx) > c.vmax) vx = c.vmax * vx / fabs(vx);
        if (fabs(vy) > c.vmax) vy = c.vmax * vy / fabs(vy);
        if (fabs(vz) > c.vmax) vz = c.vmax * vz / fabs(vz);
        out.vx[s] = vx; out.vy[s] = vy; out.vz[s] = vz;
    }
}
template <typename R>
__global__ __launch_bounds__(kMaxBlock) void k_gradient(int N, Consts<R> c, const int* __restrict__ perm, PState<R> in,
                                                   PState<R> out, const R4<R>* __restrict__ spos,
                                                   const R4<R>* __restrict__ svel, const R* __restrict__ spres,
                                                   const int* __restrict__ cs, const int* __restrict__ ce,
                                                   const R* __restrict__ stl_mag, const R4<R>* __restrict__ stl_dir,
                                                   const int* __restrict__ stl_id, const int* __restrict__ tstart,
                                                   const int* __restrict__ tcount, const int* __restrict__ tids,
                                                   const TriD<R>* __restrict__ tris,
                                                   const NegModel<R>* __restrict__ negm, NbList nb) {
    const int blk = xcd_block(N);
    if (blk < 0) return;
    gradient_body<R>(blk, N, c, perm, in, out, spos, svel, spres, cs, ce, stl_mag, stl_dir, stl_id, tstart, tcount,
                     tids, tris, negm, nb);
}

// ---------------------------------------------------------------------------------------------
// neighbour lists (inspection and PPE structure): count, then fill with slot ids
template <typename R>
__global__ __launch_bounds__(256) void k_nbr_count(int N, Consts<R> c, R re, const R4<R>* __restrict__ spos,
                                                    const int* __restrict__ cs, const int* __restrict__ ce,
                                                    int* counts) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N) return;
    const R4<R> P = spos[t];
    int n = 0;
    for_candidates_within(c, cs, ce, P.x, P.y, P.z, t, re, [&](int q) {
        const R4<R> Q = spos[q];
        if (dist3(P.x, P.y, P.z, Q.x, Q.y, Q.z) < re) ++n;
    });
    counts[t] = n;
}
template <typename R>
__global__ __launch_bounds__(256) void k_nbr_fill(int N, Consts<R> c, R re, const int* __restrict__ perm,
                                                   const R4<R>* __restrict__ spos, const int* __restrict__ cs,
                                                   const int* __restrict__ ce, const int* __restrict__ off, int* nbr) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N) return;
    const R4<R> P = spos[t];
    int e = off[t];
    for_candidates_within(c, cs, ce, P.x, P.y, P.z, t, re, [&](int q) {
        const R4<R> Q = spos[q];
        if (dist3(P.x, P.y, P.z, Q.x, Q.y, Q.z) < re) nbr[e++] = perm[q];
    });
}

// ---------------------------------------------------------------------------------------------
// CalcTemperatureFactor_Kernel (Kern :485-632), one thread per sorted particle, slot-order state:
// heat from the neighbours within re_T, (T_j - T_i) w / (r^2 (1/k_f + R_f/r)) summed in the
// canonical order, and from the nearest non-drag wall within re_T, (T_wall - T_i) Z(r) /
// (r^2 (1/k + R_w/r)) with k the harmonic mean of the fluid and wall conductivities; the drag
// wall's temperature increment is added, each heat term counts only above FLT_EPSILON (the
// reference's SCALAR_EPSILON in both precisions), T += EnthalpyCoefficient (Hf + Hw), limited to
// [MinWallTemperature, MaxWallTemperature]; the out arrays get T, A T + B and the enthalpy term.
template <typename R>
__device__ inline void thermal_body(int blk, int N, Consts<R> c, const int* __restrict__ perm, PState<R> in,
                                    PState<R> out, const R4<R>* __restrict__ spos,
                                    const int* __restrict__ cs, const int* __restrict__ ce,
                                    const R* __restrict__ stl_mag, const int* __restrict__ stl_id,
                                    const TriD<R>* __restrict__ tris,
                                    const mpsr_tri_params* __restrict__ models,
                                    const R* __restrict__ drag_temp, const int* __restrict__ step_base,
                                    const int* __restrict__ step_ctr, int thermal_step) {
    if (step_ctr && ((*step_base + *step_ctr) % thermal_step) != 0) return;
    const int t = blk * blockDim.x + threadIdx.x;
    if (t >= N) return;
    const int s = perm[t];
    if (out.id[s] < 0) return;
    const R4<R> P = spos[t];
    if (!include(c, P.x, P.y, P.z)) return;
    int ci, cj, ck;
    if (!cell_coords(c, P.x, P.y, P.z, ci, cj, ck)) return;
    const long long cid = cell_of(c, ci, cj, ck);
    if (cid < 0 || cid >= c.ncell) return;
    const R t0 = in.temp[s];
    const R re = c.re_temp, mind = c.mind_temp;
    R hf = 0, hw = 0;
    for_candidates_rec_within<4>(c, cs, ce, spos, P.x, P.y, P.z, t, re, [&](int q, const R4<R>& Q) {
        R d = dist3(P.x, P.y, P.z, Q.x, Q.y, Q.z);
        if (d < re) {
            if (d < mind) d = mind;
            const R num = (in.temp[perm[q]] - t0) * weight(re, d);
            const R den = d * d * (c.inv_kf + c.rf / d);
            hf = hf + num / den;
        }
    });
    const R dist = stl_mag[t];
    const int tid = stl_id[t];
    if (tid >= 0 && tid < c.ntri && dist > 0 && dist < re) {
        const mpsr_tri_params& m = models[tris[tid].model];
        const R rw = (R)m.WallThermalResistivity, kw = (R)m.WallThermalConductivity;
        const R k = (R)2 * c.kf * kw / (c.kf + kw);
        const R num = ((R)m.Temperature - t0) * ztable_eval(kZgrad, dist, c.l0);
        const R den = dist * dist * ((R)1 / k + rw / dist);
        hw = num / den;
    }
    R temp = t0;
    if (c.has_drag) temp = temp + drag_temp[t];
    R comb = 0;
    if (fabs(hf) > (R)FLT_EPSILON) comb = comb + hf;
    if (fabs(hw) > (R)FLT_EPSILON) comb = comb + hw;
    temp = temp + c.enth * comb;
    if (temp > c.tmax) temp = c.tmax;
    if (temp < c.tmin) temp = c.tmin;
    out.nu[s] = c.visc_a * temp + c.visc_b;
    out.temp[s] = temp;
    out.solid[s] = c.enth * comb;
}
template <typename R>
__global__ __launch_bounds__(kMaxBlock) void k_thermal(int N, Consts<R> c, const int* __restrict__ perm, PState<R> in,
                                                    PState<R> out, const R4<R>* __restrict__ spos,
                                                    const int* __restrict__ cs, const int* __restrict__ ce,
                                                    const R* __restrict__ stl_mag, const int* __restrict__ stl_id,
                                                    const TriD<R>* __restrict__ tris,
                                                    const mpsr_tri_params* __restrict__ models,
                                                    const R* __restrict__ drag_temp, const int* __restrict__ step_base,
                                                    const int* __restrict__ step_ctr, int thermal_step) {
    const int blk = xcd_block(N);
    if (blk < 0) return;
    thermal_body<R>(blk, N, c, perm, in, out, spos, cs, ce, stl_mag, stl_id, tris, models, drag_temp, step_base,
                    step_ctr, thermal_step);
}

template <typename R> void launch_bin_hash(const Dev<R>& d, hipStream_t st) {
    if (d.N <= 0) return;
    hipLaunchKernelGGL(k_bin_hash<R>, dim3(grid_for(d.N, d.block)), dim3(d.block), 0, st, d.N, d.c, d.in, d.out, d.keys,
                       d.vals);
}
template <typename R> void launch_bin_ranges(Dev<R>& d, hipStream_t st) {
    if (d.N <= 0) return;
    const int which = radix_sort_pairs(d.keys, d.vals, nullptr, d.N, d.N, d.key_bits, d.sort, st, d.sort_items);
    d.skeys = which ? d.sort.k1 : d.keys;
    d.perm = which ? d.sort.v1 : d.vals;
    fill_int(d.cell_start, -1, d.c.ncell, st);
    fill_int(d.cell_end, -1, d.c.ncell, st);
    hipLaunchKernelGGL(k_bin_ranges<R>, dim3(grid_for(d.N, d.block)), dim3(d.block), 0, st, d.N, d.c.ncell, d.skeys,
                       d.perm, d.cell_start, d.cell_end, d.in, d.spos, d.svel);
}
// ---------------------------------------------------------------------------------------------
// CalcTurbulenceViscosity_Kernel (radix DeviceProcess_Kernel.h:1273-1476): the velocity gradient
// tensor over the gradient radius (:1324-1397), S = A^2 + B^2 + C^2 + 2 D^2 + 2 E^2 + 2 F^2 of its
// symmetric part (:1405-1429), wall damping 1 - exp(-r_wall / re) with exp at single precision
// as the reference's expf (:1437-1440), nu_t = (Cs damp f)^2 sqrt(2 S) clamped by
// TurbulanViscosityRatio, k = S 2 Cnue l0^2 / Cepsilon clamped by TurbulanceIntensity
// (:1443-1474).  The vorticity and density sums there feed nothing and are not formed.  Results go
// to slot order; a particle whose cell lies outside the grid gets 0 (the reference leaves the
// array as it was).
template <typename R>
__device__ inline void turbulence_body(int blk, int N, Consts<R> c, const int* __restrict__ perm,
                                       PState<R> out, const R4<R>* __restrict__ spos,
                                       const R4<R>* __restrict__ svel, const int* __restrict__ cs,
                                       const int* __restrict__ ce, const R* __restrict__ stl_mag,
                                       R* __restrict__ turb_nu, R* __restrict__ turb_k) {
    const int t = blk * blockDim.x + threadIdx.x;
    if (t >= N) return;
    const int s = perm[t];
    if (out.id[s] < 0) return;
    const R4<R> P = spos[t];
    const R x = P.x, y = P.y, z = P.z;
    if (!include(c, x, y, z)) {
        turb_nu[s] = 0;
        return;
    }
    int ci, cj, ck;
    long long cid = -1;
    if (cell_coords(c, x, y, z, ci, cj, ck)) cid = cell_of(c, ci, cj, ck);
    if (cid < 0 || cid >= c.ncell) {
        turb_nu[s] = 0;
        turb_k[s] = 0;
        return;
    }
    const R4<R> V0 = svel[t];
    R txx = 0, txy = 0, txz = 0, tyx = 0, tyy = 0, tyz = 0, tzx = 0, tzy = 0, tzz = 0;
    const R re = c.re_grad, mind = c.mind_grad;
    for_candidates_rec_within<4>(c, cs, ce, spos, x, y, z, t, re, [&](int q, const R4<R>& Q) {
        R d = dist3(x, y, z, Q.x, Q.y, Q.z);
        if (d < re) {
            if (d < mind) d = mind;
            const R w = weight(re, d);
            const R4<R> V1 = svel[q];
            const R ux = V1.x - V0.x, uy = V1.y - V0.y, uz = V1.z - V0.z;
            const R rx = Q.x - x, ry = Q.y - y, rz = Q.z - z;
            const R dd = d * d;
            if constexpr (sizeof(R) == 4) {
                // nine quotients by one divisor: n / dd == (float)((double)n * RN(1 / (double)dd)) bit
                // for bit.  The double product is within 2^-51 of n / dd, and a quotient of two
                // 24-bit floats that is not a float midpoint (none is) lies more than 2^-48 of its
                // size from every midpoint, so both round to the same float.
                const double yd = 1.0 / (double)dd;
                auto qt = [&](R n) { return (R)((double)n * yd); };
                txx = txx + qt(ux * rx * w); txy = txy + qt(ux * ry * w); txz = txz + qt(ux * rz * w);
                tyx = tyx + qt(uy * rx * w); tyy = tyy + qt(uy * ry * w); tyz = tyz + qt(uy * rz * w);
                tzx = tzx + qt(uz * rx * w); tzy = tzy + qt(uz * ry * w); tzz = tzz + qt(uz * rz * w);
            } else {
                txx = txx + ux * rx * w / dd; txy = txy + ux * ry * w / dd; txz = txz + ux * rz * w / dd;
                tyx = tyx + uy * rx * w / dd; tyy = tyy + uy * ry * w / dd; tyz = tyz + uy * rz * w / dd;
                tzx = tzx + uz * rx * w / dd; tzy = tzy + uz * ry * w / dd; tzz = tzz + uz * rz * w / dd;
            }
        }
    });
    const R g = c.turb_g;
    txx = g * txx; txy = g * txy; txz = g * txz;
    tyx = g * tyx; tyy = g * tyy; tyz = g * tyz;
    tzx = g * tzx; tzy = g * tzy; tzz = g * tzz;
    const R D = (R)0.5 * (txy + tyx), E = (R)0.5 * (txz + tzx), F = (R)0.5 * (tyz + tzy);
    const R S = txx * txx + tyy * tyy + tzz * tzz + (R)2 * D * D + (R)2 * E * E + (R)2 * F * F;
    const R ratio = -(stl_mag[t] / re);
    const R damp = (R)1 - (R)(float)det_exp((double)(float)ratio);
    const R cd = c.turb_cs * damp;
    const R coef = (cd * c.turb_fw) * (cd * c.turb_fw);
    R nut = 0;
    if (S > 0) nut = coef * sqrt((R)2 * S);
    if (nut > c.turb_numax) nut = c.turb_numax;
    R k = S * c.turb_tk;
    if (k > c.turb_kmax) k = c.turb_kmax;
    turb_nu[s] = nut;
    turb_k[s] = k;
}
template <typename R>
__global__ __launch_bounds__(kMaxBlock) void k_turbulence(int N, Consts<R> c, const int* __restrict__ perm,
                                                     PState<R> out, const R4<R>* __restrict__ spos,
                                                     const R4<R>* __restrict__ svel, const int* __restrict__ cs,
                                                     const int* __restrict__ ce, const R* __restrict__ stl_mag,
                                                     R* __restrict__ turb_nu, R* __restrict__ turb_k) {
    const int blk = xcd_block(N);
    if (blk < 0) return;
    turbulence_body<R>(blk, N, c, perm, out, spos, svel, cs, ce, stl_mag, turb_nu, turb_k);
}

template <typename R> void launch_explicit(const Dev<R>& d, hipStream_t st) {
    if (d.N <= 0) return;
    hipLaunchKernelGGL(k_explicit<R>, dim3(xcd_grid(d.N, d.block)), dim3(d.block), 0, st, d.N, d.c, d.perm, d.in, d.out,
                       d.spos, d.svel, d.cell_start, d.cell_end, d.stl_mag, d.stl_dir, d.stl_id, d.drag_acc, d.sph_dv,
                       d.tris, d.models, d.turb_nu);
}
template <typename R> void launch_turbulence(const Dev<R>& d, hipStream_t st) {
    if (d.N <= 0 || !d.turb_nu) return;
    hipLaunchKernelGGL(k_turbulence<R>, dim3(xcd_grid(d.N, d.block)), dim3(d.block), 0, st, d.N, d.c, d.perm, d.out,
                       d.spos, d.svel, d.cell_start, d.cell_end, d.stl_mag, d.turb_nu, d.turb_k);
}
template <typename R> void launch_relocate(const Dev<R>& d, hipStream_t st) {
    if (d.N <= 0) return;
    hipLaunchKernelGGL(k_relocate<R>, dim3(grid_for(d.N, d.block)), dim3(d.block), 0, st, d.N, d.c, d.in, d.out);
}
template <typename R>
void launch_pnd_ppe(const Dev<R>& d, int* active, int* cnt, double* diag, int* dir, int* nbq, int* nbn,
                    int nb_width, hipStream_t st) {
    if (d.N <= 0) return;
    hipLaunchKernelGGL((k_pnd<R, true>), dim3(grid_for(d.N, d.block)), dim3(d.block), 0, st, d.N, d.c, 0, d.perm, d.in,
                       d.out, d.spos, d.spres, d.cell_start, d.cell_end, d.stl_mag, d.stl_id, d.skeys, active, cnt, diag,
                       dir, nbq, nbn, nb_width);
}
template <typename R> void launch_pnd(const Dev<R>& d, int eos, hipStream_t st) {
    if (d.N <= 0) return;
    hipLaunchKernelGGL((k_pnd<R, false>), dim3(grid_for(d.N, d.block)), dim3(d.block), 0, st, d.N, d.c, eos, d.perm, d.in, d.out,
                       d.spos, d.spres, d.cell_start, d.cell_end, d.stl_mag, d.stl_id, nullptr, nullptr, nullptr,
                       nullptr, nullptr, d.gnb_q, d.gnb_n, kGradWidth);
}
template <typename R> void launch_gradient(const Dev<R>& d, hipStream_t st, NbList nb) {
    if (d.N <= 0) return;
    hipLaunchKernelGGL(k_gradient<R>, dim3(xcd_grid(d.N, d.block)), dim3(d.block), 0, st, d.N, d.c, d.perm, d.in, d.out,
                       d.spos, d.svel, d.spres, d.cell_start, d.cell_end, d.stl_mag, d.stl_dir, d.stl_id,
                       d.cell_tri_start, d.cell_tri_count, d.uniq_tri, d.tris, d.negm, nb);
}
template <typename R>
void launch_thermal(const Dev<R>& d, const int* step_base, const int* step_ctr, int thermal_step, hipStream_t st) {
    if (d.N <= 0) return;
    hipLaunchKernelGGL(k_thermal<R>, dim3(xcd_grid(d.N, d.block)), dim3(d.block), 0, st, d.N, d.c, d.perm, d.in, d.out,
                       d.spos, d.cell_start, d.cell_end, d.stl_mag, d.stl_id, d.tris, d.models, d.drag_temp, step_base,
                       step_ctr, thermal_step > 0 ? thermal_step : 1);
}
template <typename R> void launch_neighbor_count(const Dev<R>& d, R re, int* counts, hipStream_t st) {
    if (d.N <= 0) return;
    hipLaunchKernelGGL(k_nbr_count<R>, dim3(grid_for(d.N, 256)), dim3(256), 0, st, d.N, d.c, re, d.spos,
                       d.cell_start, d.cell_end, counts);
}
template <typename R> void launch_neighbor_fill(const Dev<R>& d, R re, const int* offsets, int* nbr, hipStream_t st) {
    if (d.N <= 0) return;
    hipLaunchKernelGGL(k_nbr_fill<R>, dim3(grid_for(d.N, 256)), dim3(256), 0, st, d.N, d.c, re, d.perm, d.spos,
                       d.cell_start, d.cell_end, offsets, nbr);
}

// ---- batched forms: one work item = one block of one domain's particles ----------------------
template <typename R> struct OpBinHash {
    const Dev<R>* dv;
    __device__ void operator()(int k, int b) const {
        const Dev<R>& d = dv[k];
        bin_hash_body<R>(b, d.N, d.c, d.in, d.out, d.keys, d.vals);
    }
};
template <typename R> struct OpBinRanges {
    const Dev<R>* dv;
    __device__ void operator()(int k, int b) const {
        const Dev<R>& d = dv[k];
        bin_ranges_body<R>(b, d.N, d.c.ncell, d.skeys, d.perm, d.cell_start, d.cell_end, d.in, d.spos, d.svel);
    }
};
template <typename R> struct OpExplicit {
    const Dev<R>* dv;
    __device__ void operator()(int k, int b) const {
        const Dev<R>& d = dv[k];
        explicit_body<R>(b, d.N, d.c, d.perm, d.in, d.out, d.spos, d.svel, d.cell_start, d.cell_end, d.stl_mag,
                         d.stl_dir, d.stl_id, d.drag_acc, d.sph_dv, d.tris, d.models, d.turb_nu);
    }
};
template <typename R> struct OpRelocate {
    const Dev<R>* dv;
    __device__ void operator()(int k, int b) const {
        const Dev<R>& d = dv[k];
        relocate_body<R>(b, d.N, d.c, d.in, d.out);
    }
};
template <typename R> struct OpPndEos {
    const Dev<R>* dv;
    __device__ void operator()(int k, int b) const {
        const Dev<R>& d = dv[k];
        pnd_body<R, false>(b, d.N, d.c, 1, d.perm, d.in, d.out, d.spos, d.spres, d.cell_start, d.cell_end, d.stl_mag,
                           d.stl_id, nullptr, nullptr, nullptr, nullptr, nullptr, d.gnb_q, d.gnb_n, kGradWidth);
    }
};
template <typename R> struct OpGradient {
    const Dev<R>* dv;
    __device__ void operator()(int k, int b) const {
        const Dev<R>& d = dv[k];
        gradient_body<R>(b, d.N, d.c, d.perm, d.in, d.out, d.spos, d.svel, d.spres, d.cell_start, d.cell_end,
                         d.stl_mag, d.stl_dir, d.stl_id, d.cell_tri_start, d.cell_tri_count, d.uniq_tri, d.tris,
                         d.negm, d.nbl);
    }
};
template <typename R> struct OpThermal {
    const Dev<R>* dv;
    __device__ void operator()(int k, int b) const {
        const Dev<R>& d = dv[k];
        thermal_body<R>(b, d.N, d.c, d.perm, d.in, d.out, d.spos, d.cell_start, d.cell_end, d.stl_mag, d.stl_id,
                        d.tris, d.models, d.drag_temp, d.step_base, d.step_ctr, d.thermal_step);
    }
};
template <typename R> struct OpTurbulence {
    const Dev<R>* dv;
    __device__ void operator()(int k, int b) const {
        const Dev<R>& d = dv[k];
        turbulence_body<R>(b, d.N, d.c, d.perm, d.out, d.spos, d.svel, d.cell_start, d.cell_end, d.stl_mag,
                           d.turb_nu, d.turb_k);
    }
};
template <typename R> struct OpPndPpe {
    const Dev<R>* dv;
    const PndRows<R>* rows;
    __device__ void operator()(int j, int b) const {
        const PndRows<R>& q = rows[j];
        const Dev<R>& d = dv[q.dev];
        pnd_body<R, true>(b, d.N, d.c, 0, d.perm, d.in, d.out, d.spos, d.spres, d.cell_start, d.cell_end, d.stl_mag,
                          d.stl_id, d.skeys, q.active, q.cnt, q.diag, q.dir, q.nbq, q.nbn, q.nb_width);
    }
};
template <typename R> void launch_pnd_ppe_b(const BatchLaunch<R>& b, const PndRows<R>* rows, WorkList p, hipStream_t st) {
    launch_batched<kMaxBlock>(OpPndPpe<R>{b.devs, rows}, p, b.block, st);
}
template <typename R> void launch_bin_hash_b(const BatchLaunch<R>& b, WorkList p, hipStream_t st) {
    launch_batched<kMaxBlock>(OpBinHash<R>{b.devs}, p, b.block, st);
}
template <typename R> void launch_bin_ranges_b(const BatchLaunch<R>& b, WorkList p, hipStream_t st) {
    launch_batched<kMaxBlock>(OpBinRanges<R>{b.devs}, p, b.block, st);
}
template <typename R> void launch_explicit_b(const BatchLaunch<R>& b, WorkList p, hipStream_t st) {
    launch_batched<kMaxBlock>(OpExplicit<R>{b.devs}, p, b.block, st);
}
template <typename R> void launch_relocate_b(const BatchLaunch<R>& b, WorkList p, hipStream_t st) {
    launch_batched<kMaxBlock>(OpRelocate<R>{b.devs}, p, b.block, st);
}
template <typename R> void launch_pnd_b(const BatchLaunch<R>& b, WorkList p, hipStream_t st) {
    launch_batched<kMaxBlock>(OpPndEos<R>{b.devs}, p, b.block, st);
}
template <typename R> void launch_gradient_b(const BatchLaunch<R>& b, WorkList p, hipStream_t st) {
    launch_batched<kMaxBlock>(OpGradient<R>{b.devs}, p, b.block, st);
}
template <typename R> void launch_thermal_b(const BatchLaunch<R>& b, WorkList p, hipStream_t st) {
    launch_batched<kMaxBlock>(OpThermal<R>{b.devs}, p, b.block, st);
}
template <typename R> void launch_turbulence_b(const BatchLaunch<R>& b, WorkList p, hipStream_t st) {
    launch_batched<kMaxBlock>(OpTurbulence<R>{b.devs}, p, b.block, st);
}

#define MPSR_INSTANTIATE(R)                                                                   \
    template void launch_bin_hash<R>(const Dev<R>&, hipStream_t);                             \
    template void launch_bin_ranges<R>(Dev<R>&, hipStream_t);                                 \
    template void launch_explicit<R>(const Dev<R>&, hipStream_t);                             \
    template void launch_relocate<R>(const Dev<R>&, hipStream_t);                             \
    template void launch_pnd<R>(const Dev<R>&, int, hipStream_t);                             \
    template void launch_pnd_ppe<R>(const Dev<R>&, int*, int*, double*, int*, int*, int*, int, hipStream_t); \
    template void launch_gradient<R>(const Dev<R>&, hipStream_t, NbList);                     \
    template void launch_thermal<R>(const Dev<R>&, const int*, const int*, int, hipStream_t);   \
    template void launch_turbulence<R>(const Dev<R>&, hipStream_t);                           \
    template void launch_neighbor_count<R>(const Dev<R>&, R, int*, hipStream_t);              \
    template void launch_neighbor_fill<R>(const Dev<R>&, R, const int*, int*, hipStream_t);   \
    template void launch_bin_hash_b<R>(const BatchLaunch<R>&, WorkList, hipStream_t);         \
    template void launch_bin_ranges_b<R>(const BatchLaunch<R>&, WorkList, hipStream_t);       \
    template void launch_explicit_b<R>(const BatchLaunch<R>&, WorkList, hipStream_t);         \
    template void launch_relocate_b<R>(const BatchLaunch<R>&, WorkList, hipStream_t);         \
    template void launch_pnd_b<R>(const BatchLaunch<R>&, WorkList, hipStream_t);              \
    template void launch_gradient_b<R>(const BatchLaunch<R>&, WorkList, hipStream_t);         \
    template void launch_thermal_b<R>(const BatchLaunch<R>&, WorkList, hipStream_t);          \
    template void launch_turbulence_b<R>(const BatchLaunch<R>&, WorkList, hipStream_t);       \
    template void launch_pnd_ppe_b<R>(const BatchLaunch<R>&, const PndRows<R>*, WorkList, hipStream_t);
MPSR_INSTANTIATE(float)
MPSR_INSTANTIATE(double)

}  // namespace mpsr
