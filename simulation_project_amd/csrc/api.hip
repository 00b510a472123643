// api.hip -- the handle, the step driver and the C ABI of include/mpsr.h.
//
// mpsr_step is one body of CMultiGPUHandler::CalculateNew's loop
// (Source/MultiGPUHandler.cpp:362-492) with the output and bucket branches left to the caller:
//   RegisterParticleTopology  -> hash, radix sort, cell ranges, gather
//   RegisterTriangleTopology  -> triangle binning
//   CalcSTLDistance
//   CalculateDragEffect       (when drag STLs exist)
//   CalcExplicitly
//   RelocateAllParticleTopology
//   RegisterParticleTopology
//   CalcSTLDistance
//   CalcPressureExplicit      (EOS) or PND + PPE solve
//   CalcPressureExplicitGradient
//   RelocateAllParticleTopology
//   UpdateTrianglePosition, RotateTrianglePosition(i)
// All launches go to one HIP stream; there is no host synchronisation inside a step (the
// reference synchronises every step, MultiGPUHandler.cpp:1009-1018), so mpsr_step_many queues
// `count` steps back to back and waits once.
#include <cstring>
#include <string>
#include <vector>
#include <algorithm>
#include "mpsr_internal.h"
#include "ppe.h"
#include "mpnn.h"

namespace mpsr {

static thread_local std::string g_err;
static mpsr_status fail(mpsr_status s, const std::string& m) {
    g_err = m;
    return s;
}
#define HIPC(x)                                                                                      \
    do {                                                                                             \
        hipError_t e_ = (x);                                                                         \
        if (e_ != hipSuccess) return fail(MPSR_GPUERR, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

enum Phase { PH_BIN, PH_TRI, PH_STL, PH_DRAG, PH_EXPL, PH_RELOC, PH_PND, PH_PPE, PH_GRAD, PH_WALL, PH_SPH, PH_THERM, PH_TURB,
             PH_COUNT };

struct EngineBase {
    virtual ~EngineBase() {}
    virtual mpsr_status step_many(int first, int count) = 0;
    virtual mpsr_status phase(int which, int arg) = 0;
    virtual mpsr_status get_particles(int which, mpsr_particles* o) = 0;
    virtual mpsr_status set_particles(const mpsr_particles* p) = 0;
    virtual mpsr_status get_triangles(mpsr_triangle* o, int n) = 0;
    virtual mpsr_status get_binning(int* k, int* p, int* cs, int* ce) = 0;
    virtual mpsr_status get_triangle_cells(int* cs, int* cc, int* ids, int cap, int* total) = 0;
    virtual mpsr_status get_stl(double* m, double* dir, int* id) = 0;
    virtual mpsr_status get_neighbors(double r, int* off, int* nbr, int cap, int* total) = 0;
    virtual mpsr_status get_ppe(double* rhs, double* x, int* dir) = 0;
    virtual mpsr_status get_ppe_matrix(int* ptr, int* col, double* val, double* diag, int cap, int* nnz) = 0;
    virtual mpsr_status set_dt(float dt, float courant) = 0;
    virtual mpsr_status set_ppe(const mpsr_ppe_options* o) = 0;
    virtual mpsr_status sync() = 0;
    virtual mpsr_status set_buckets(const mpsr_particles* b, const int* interval, const int* total, int count) = 0;
    virtual mpsr_status bucket_phase(int step, int* added) = 0;
    virtual int count() const = 0;
    virtual int ppe_max_iter() const = 0;   // the PCG iteration cap of a PPE domain, 0 without the PPE
    virtual mpsr_status mpnn_set(const float* w) = 0;
    virtual mpsr_status mpnn_get(float* w) = 0;
    virtual mpsr_status mpnn_train(double lr, double beta, int update, float* grads, double* parts) = 0;
    virtual mpsr_status mpnn_rows(float* X, float* Y, double* scale) = 0;
    virtual mpsr_status set_spheres(const mpsr_spheres* s) = 0;
    virtual mpsr_status get_spheres(mpsr_spheres* s) = 0;
    // launch shape of the per-particle kernels and the particle sort (results do not depend on it)
    virtual mpsr_status set_launch(int block, int sort_items) = 0;
    // the radix variant's turbulence and negative-pressure terms
    virtual mpsr_status set_radix(const mpsr_radix_terms* t, const mpsr_negative_pressure_model* m, int nm) = 0;
    virtual mpsr_status get_turbulence(double* nu, double* k, int n) = 0;
    // batched form: this domain as one bucket of a batched step (BatchGroup)
    virtual mpsr_status batch_check() = 0;
    virtual mpsr_status batch_prepare(int first, int count) = 0;
    virtual mpsr_status batch_finish() = 0;
    virtual mpsr_status snapshot(int restore) = 0;
    virtual hipStream_t stream() const = 0;
    mpsr_gridbox gb{};
    mpsr_timings tm{};
    int profiling = 0;
    int device = 0;
};

template <typename R> static Consts<R> make_consts(const mpsr_params& P, const mpsr_gridbox& gb, int ntri, int ndrag) {
    Consts<R> c{};
    c.bminx = (R)gb.m_BufferedZone.m_MinBound.x;
    c.bminy = (R)gb.m_BufferedZone.m_MinBound.y;
    c.bminz = (R)gb.m_BufferedZone.m_MinBound.z;
    c.cminx = (R)gb.m_ComputeZone.m_MinBound.x;
    c.cminy = (R)gb.m_ComputeZone.m_MinBound.y;
    c.cminz = (R)gb.m_ComputeZone.m_MinBound.z;
    c.cmaxx = (R)gb.m_ComputeZone.m_MaxBound.x;
    c.cmaxy = (R)gb.m_ComputeZone.m_MaxBound.y;
    c.cmaxz = (R)gb.m_ComputeZone.m_MaxBound.z;
    c.cell = (R)gb.m_CellSize;
    c.gx = gb.m_GridSize.x;
    c.gy = gb.m_GridSize.y;
    c.gz = gb.m_GridSize.z;
    c.ncell = c.gx * c.gy * c.gz;
    c.tol = (R)P.Tolerance;
    c.l0 = (R)P.InitialParticleDistance;
    c.dt = (R)P.Dt;
    c.re_lap = (R)P.LaplacianInfluenceRadiusCoefficient * (R)P.InitialParticleDistance;
    c.mind_lap = (R)((double)c.re_lap * 1e-3);
    c.lapc = (R)((2.0 * P.Dimension) / (double)((R)P.LambdaValue * (R)P.ConstantParticleDensity));
    c.rc = (R)P.r_coeff * c.l0;
    c.art_p = (R)P.ArtificialPressure;
    c.art_coeff = -(R)P.Dimension / ((R)P.ConstantParticleDensity * (R)P.Density);
    c.mu_wall = (R)P.WallFrictionCoeffieicnt;
    c.gx_acc = (R)P.GravityConstant.x;
    c.gy_acc = (R)P.GravityConstant.y;
    c.gz_acc = (R)P.GravityConstant.z;
    c.expl = P.bExplicitPressure ? 1 : 0;
    c.has_drag = ndrag > 0 ? 1 : 0;
    c.re_grad = (R)P.GradientInfluenceRadiusCoefficient * (R)P.InitialParticleDistance;
    c.mind_grad = (R)((double)c.re_grad * 1e-3);
    c.eos_coef = (R)P.SonicVelocity * (R)P.SonicVelocity * (R)P.Density / (R)P.ConstantParticleDensity;
    c.n0 = (R)P.ConstantParticleDensity;
    c.beta_n0 = (R)P.ConstantParticleDensity * (R)P.FreeSurfaceThreshold;
    c.c1 = (R)(2 * P.Dimension) / (R)P.ConstantParticleDensity;
    c.c2 = (R)P.Density / ((R)P.Dt * (R)P.Dt);
    c.c3 = -(R)P.Dt / (R)P.Density;
    c.vmax = ((R)P.InitialParticleDistance * (R)P.CourantNumber / (R)P.Dt) * build_sqrt((R)(1.0 / 3.0));
    c.density = (R)P.Density;
    c.cp = (R)P.FruidSpecificHeat;
    c.thermal_step = (R)P.ThermalStep;
    c.ntri = ntri;
    // CalcTemperatureFactor_Kernel (Kern :497-505): EnthalpyCoefficient = Dt 2 d / (n0 rho Cp),
    // ReTemperature = coefficient * l0, MinDist = ReTemperature * 1e-3 (a double product)
    c.re_temp = (R)P.TemperatureInfluenceRadiusCoefficient * (R)P.InitialParticleDistance;
    c.mind_temp = (R)((double)c.re_temp * 1e-3);
    c.enth = (R)P.Dt * (R)2 * (R)P.Dimension /
             ((R)P.ConstantParticleDensity * (R)P.Density * (R)P.FruidSpecificHeat);
    c.kf = (R)P.FruidThermalConductivity;
    c.inv_kf = (R)1 / (R)P.FruidThermalConductivity;
    c.rf = (R)P.FruidThermalResistance;
    c.tmax = (R)P.MaxWallTemperature;
    c.tmin = (R)P.MinWallTemperature;
    c.visc_a = (R)P.ViscosityUpdateParameterA;
    c.visc_b = (R)P.ViscosityUpdateParameterB;
    return c;
}

// RotateTrianglePosition_Kernel's per-model angle logic (Kern :945-971) and the axis of
// CalcNormal(Center, Second) (Utility_Inline.h:58-67), evaluated once per model and step.
template <typename R> static RotCoef<R> rot_coef(const mpsr_tri_params& m, const mpsr_params& P, int step) {
    RotCoef<R> r{};
    r.active = 0;
    const R dt = (R)P.Dt;
    if (!m.RotatingOrNot) return r;
    if ((R)step * dt < (R)m.RotationStartTime) return r;
    const R start = (R)m.RotationStartTime, reach = (R)m.TimeToReachRPM, deg = (R)m.AngleOfRotationInDegree;
    const int req = (int)((reach - start) / dt);
    const R passed = dt * (R)step;
    R ang = 0;
    if (passed < start || deg == 0) ang = 0;
    else if (passed >= start && passed < reach) ang = (deg / (R)(req + 1)) * ((R)step - (start / dt));
    else ang = deg;
    if (ang == 0) return r;
    const R cx = (R)m.CenterOfRotation.x, cy = (R)m.CenterOfRotation.y, cz = (R)m.CenterOfRotation.z;
    const R qx = (R)m.SecondPointOfRotation.x, qy = (R)m.SecondPointOfRotation.y, qz = (R)m.SecondPointOfRotation.z;
    const R mg = build_sqrt((qx - cx) * (qx - cx) + (qy - cy) * (qy - cy) + (qz - cz) * (qz - cz));
    R ux = 0, uy = 0, uz = 0;
    if (mg > 0) { ux = (qx - cx) / mg; uy = (qy - cy) / mg; uz = (qz - cz) / mg; }
    const R th = (R)(((double)ang * 3.141592653589793238) / 180.0);
    const R cs = std::cos(th), sn = std::sin(th);
    r.a = cx; r.b = cy; r.c = cz; r.u = ux; r.v = uy; r.w = uz;
    r.cs = cs; r.sn = sn; r.omc = 1 - cs;
    r.active = 1;
    return r;
}

// PND and lambda of the ideal cubic lattice at radius re (in l0), exact square roots.
static void lattice_n0_lambda(double re, double* n0, double* lam) {
    const int m = (int)std::ceil(re) + 1;
    double sw = 0.0, swr2 = 0.0;
    for (int i = -m; i <= m; ++i)
        for (int j = -m; j <= m; ++j)
            for (int k = -m; k <= m; ++k) {
                if (!i && !j && !k) continue;
                const double r = std::sqrt((double)(i * i + j * j + k * k));
                if (r < re) {
                    const double w = re / r - 1.0;
                    sw += w;
                    swr2 += w * r * r;
                }
            }
    *n0 = sw;
    *lam = swr2 / sw;
}

// PND of the ideal cubic lattice seen from a site, with the build's distance and weight at re:
// the reference density of the PPE source term.
template <typename R> static double lattice_pnd(R l0, R re) {
    const int m = (int)std::ceil((double)re / (double)l0) + 1;
    const R mind = (R)((double)re * 1e-3);
    double sum = 0.0;
    for (int i = -m; i <= m; ++i)
        for (int j = -m; j <= m; ++j)
            for (int k = -m; k <= m; ++k) {
                if (!i && !j && !k) continue;
                const R x = (R)i * l0, y = (R)j * l0, z = (R)k * l0;
                R d = build_sqrt((R(0) - x) * (R(0) - x) + (R(0) - y) * (R(0) - y) + (R(0) - z) * (R(0) - z));
                if (d < re) {
                    if (d < mind) d = mind;
                    sum += (double)weight(re, d);
                }
            }
    return sum;
}

template <typename T> static hipError_t dalloc(T** p, size_t n) {
    return hipMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(T));
}

template <typename R> struct Engine : EngineBase {
    Dev<R> d;
    mpsr_params P{};
    hipStream_t st = nullptr;
    std::vector<void*> allocs;
    RotCoef<R>* rot_table = nullptr;   // device [rot_steps * nmodel]
    int rot_steps = 0;
    std::vector<mpsr_tri_params> models_h;
    int consts_ver = 0;                // bumped when the per-launch constants change
    std::vector<RotCoef<R>> rot_host;
    int pressure_mode = 0;
    mpsr_ppe_options ppe_opt{};
    PpeSolver<R>* ppe = nullptr;
    double n0_lap = 0, lambda_lap = 0, n0_src = 0;
    hipEvent_t ev_a = nullptr, ev_b = nullptr;
    // bucket inflow (multi-bucket build of the reference)
    struct BucketD {
        PState<R> s{};
        int count = 0, interval = 1, total = 0, added = 0;
    };
    std::vector<BucketD> buckets;
    int* bk_flag = nullptr;            // [cap] dead-slot flags
    int* bk_rank = nullptr;            // [cap] exclusive scan of the flags
    int* bk_dead = nullptr;            // [1] dead-slot count
    int* bk_dead_h = nullptr;          // pinned copy of it
    TriD<R>* tris0 = nullptr;          // initial triangles (ResetWallPosition)
    SphDev<R> sph{};                   // rigid spheres (Schiller-Naumann drag)
    int sph_cap = 0;
    std::vector<hipEvent_t> pev;      // profiling events, two per phase launch
    std::vector<int> pev_phase;
    size_t pev_used = 0;

    ~Engine() override {
        if (ppe) ppe_destroy(ppe);
        if (bk_dead_h) (void)hipHostFree(bk_dead_h);
        for (void* p : allocs) (void)hipFree(p);
        for (hipEvent_t e : pev) (void)hipEventDestroy(e);
        if (ev_a) (void)hipEventDestroy(ev_a);
        if (ev_b) (void)hipEventDestroy(ev_b);
        if (st) (void)hipStreamDestroy(st);
    }
    template <typename T> mpsr_status alloc(T** p, size_t n) {
        HIPC(dalloc(p, n));
        allocs.push_back((void*)*p);
        return MPSR_NOERR;
    }
#define ALLOC(p, n)                                  \
    do {                                             \
        mpsr_status s_ = alloc(&(p), (size_t)(n));   \
        if (s_ != MPSR_NOERR) return s_;             \
    } while (0)

    mpsr_status alloc_state(PState<R>& s, int cap) {
        ALLOC(s.id, cap); ALLOC(s.type, cap);
        ALLOC(s.px, cap); ALLOC(s.py, cap); ALLOC(s.pz, cap);
        ALLOC(s.vx, cap); ALLOC(s.vy, cap); ALLOC(s.vz, cap);
        ALLOC(s.p, cap); ALLOC(s.n, cap); ALLOC(s.temp, cap); ALLOC(s.nu, cap); ALLOC(s.solid, cap);
        return MPSR_NOERR;
    }
    mpsr_status alloc_sort(SortScratch& ss, int cap) {
        ALLOC(ss.k1, cap);
        ALLOC(ss.v1, cap);
        ss.hist_cap = (int)radix_hist_ints(cap);
        ALLOC(ss.hist, ss.hist_cap);
        ss.scan_tmp_cap = (int)scan_tmp_ints(cap);
        ALLOC(ss.scan_tmp, ss.scan_tmp_cap);
        return MPSR_NOERR;
    }

    mpsr_status upload_fields(const mpsr_particles* pp, std::initializer_list<PState<R>*> dst_states) {
        const int n = pp->n;
        std::vector<int> ids(n), type(n);
        std::vector<R> f[11];
        for (auto& v : f) v.assign(n, R(0));
        for (int s = 0; s < n; ++s) {
            ids[s] = pp->id ? pp->id[s] : s;
            type[s] = pp->type ? pp->type[s] : 0;
            for (int a = 0; a < 3; ++a) {
                f[a][s] = (R)pp->position[3 * s + a];
                f[3 + a][s] = (R)pp->velocity[3 * s + a];
            }
            f[6][s] = pp->pressure ? (R)pp->pressure[s] : R(0);
            f[7][s] = pp->density ? (R)pp->density[s] : R(0);
            f[8][s] = pp->temperature ? (R)pp->temperature[s] : R(0);
            f[9][s] = pp->kinetic_viscosity ? (R)pp->kinetic_viscosity[s] : R(0);
            f[10][s] = pp->solid_phase_rate ? (R)pp->solid_phase_rate[s] : R(0);
        }
        if (n == 0) return MPSR_NOERR;
        for (PState<R>* S : dst_states) {
            HIPC(hipMemcpy(S->id, ids.data(), n * sizeof(int), hipMemcpyHostToDevice));
            HIPC(hipMemcpy(S->type, type.data(), n * sizeof(int), hipMemcpyHostToDevice));
            R* dst[11] = {S->px, S->py, S->pz, S->vx, S->vy, S->vz, S->p, S->n, S->temp, S->nu, S->solid};
            for (int a = 0; a < 11; ++a) HIPC(hipMemcpy(dst[a], f[a].data(), n * sizeof(R), hipMemcpyHostToDevice));
        }
        return MPSR_NOERR;
    }
    mpsr_status upload_state(const mpsr_particles* pp) {
        mpsr_status s = upload_fields(pp, {&d.in, &d.out});
        if (s != MPSR_NOERR) return s;
        d.N = pp->n;
        return MPSR_NOERR;
    }

    // Bucket inflow: InsertBucket + SetAddedBucket + ResetWallPosition of BucketAddition
    // (MultiGPUHandler.cpp:943-981, 1085-1123), for every bucket due at this step.
    mpsr_status set_buckets(const mpsr_particles* b, const int* interval, const int* total, int count) override {
        HIPC(hipSetDevice(device));
        if (count < 0 || (count > 0 && (!b || !interval || !total))) return fail(MPSR_PARAMERR, "invalid bucket arguments");
        for (int j = 0; j < count; ++j) {
            if (interval[j] <= 0) return fail(MPSR_PARAMERR, "bucket addition interval must be positive");
            if (b[j].n < 0 || (b[j].n > 0 && !b[j].position)) return fail(MPSR_PARAMERR, "bucket particles missing");
        }
        HIPC(hipStreamSynchronize(st));
        buckets.clear();
        for (int j = 0; j < count; ++j) {
            BucketD B;
            if (alloc_state(B.s, std::max(1, b[j].n)) != MPSR_NOERR) return MPSR_GPUERR;
            mpsr_status s = upload_fields(&b[j], {&B.s});
            if (s != MPSR_NOERR) return s;
            B.count = b[j].n;
            B.interval = interval[j];
            B.total = total[j];
            buckets.push_back(B);
        }
        if (count > 0 && !bk_flag) {
            ALLOC(bk_flag, d.cap);
            ALLOC(bk_rank, d.cap);
            ALLOC(bk_dead, 1);
            HIPC(hipHostMalloc((void**)&bk_dead_h, sizeof(int)));
        }
        return MPSR_NOERR;
    }
    int count() const override { return d.N; }
    int ppe_max_iter() const override {
        return pressure_mode == MPSR_PRESSURE_PPE ? (ppe_opt.max_iter > 0 ? ppe_opt.max_iter : 20000) : 0;
    }
    mpsr_status mpnn_set(const float* w) override {
        HIPC(hipSetDevice(device));
        if (!ppe) return fail(MPSR_PARAMERR, "the network needs a handle in PPE mode");
        HIPC(hipStreamSynchronize(st));
        mpsr_status s = ppe_mpnn_set(ppe, w, st);
        return s == MPSR_NOERR ? s : fail(s, "network weights could not be set");
    }
    mpsr_status mpnn_get(float* w) override {
        HIPC(hipSetDevice(device));
        if (!ppe) return fail(MPSR_PARAMERR, "the network needs a handle in PPE mode");
        mpsr_status s = ppe_mpnn_get(ppe, w, st);
        return s == MPSR_NOERR ? s : fail(s, "no network weights set");
    }
    mpsr_status mpnn_train(double lr, double beta, int update, float* grads, double* parts) override {
        HIPC(hipSetDevice(device));
        if (!ppe) return fail(MPSR_PARAMERR, "the network needs a handle in PPE mode");
        mpsr_status s = ppe_mpnn_train(ppe, d, lr, beta, update, grads, parts, st);
        return s == MPSR_NOERR ? s : fail(s, "network training step needs weights and a solved system");
    }
    mpsr_status mpnn_rows(float* X, float* Y, double* scale) override {
        HIPC(hipSetDevice(device));
        if (!ppe) return fail(MPSR_PARAMERR, "the network needs a handle in PPE mode");
        mpsr_status s = ppe_mpnn_rows(ppe, d, X, Y, scale, st);
        return s == MPSR_NOERR ? s : fail(s, "network rows need weights and a solved system");
    }
    mpsr_status bucket_phase(int step, int* added) override {
        HIPC(hipSetDevice(device));
        bool a = false;
        mpsr_status s = bucket_addition(step, a);
        if (s != MPSR_NOERR) return s;
        HIPC(hipGetLastError());
        HIPC(hipStreamSynchronize(st));
        if (added) *added = a ? 1 : 0;
        return MPSR_NOERR;
    }
    mpsr_status set_spheres(const mpsr_spheres* in) override {
        HIPC(hipSetDevice(device));
        if (!in || in->n < 0 || (in->n > 0 && !in->position)) return fail(MPSR_PARAMERR, "spheres missing");
        if (in->n > 0 && (!(in->diameter > 0) || !(in->density > 0) || !(in->fluid_viscosity > 0)))
            return fail(MPSR_PARAMERR, "sphere diameter, density and fluid viscosity must be positive");
        HIPC(hipStreamSynchronize(st));
        const int S = in->n;
        if (S > sph_cap) {
            ALLOC(sph.pos, S); ALLOC(sph.vel, S); ALLOC(sph.force, S); ALLOC(sph.wsum, S);
            ALLOC(sph.keys, S); ALLOC(sph.vals, S);
            if (alloc_sort(sph.sort, S) != MPSR_NOERR) return MPSR_GPUERR;
            if (!sph.cs) { ALLOC(sph.cs, d.c.ncell); ALLOC(sph.ce, d.c.ncell); }
            sph_cap = S;
        }
        if (!d.sph_dv) ALLOC(d.sph_dv, d.cap);
        std::vector<R4<R>> p(std::max(1, S)), v(std::max(1, S));
        for (int k = 0; k < S; ++k) {
            p[k] = R4<R>{(R)in->position[3 * k], (R)in->position[3 * k + 1], (R)in->position[3 * k + 2], R(0)};
            v[k] = in->velocity ? R4<R>{(R)in->velocity[3 * k], (R)in->velocity[3 * k + 1], (R)in->velocity[3 * k + 2], R(0)}
                                : R4<R>{R(0), R(0), R(0), R(0)};
        }
        if (S > 0) {
            HIPC(hipMemcpy(sph.pos, p.data(), sizeof(R4<R>) * S, hipMemcpyHostToDevice));
            HIPC(hipMemcpy(sph.vel, v.data(), sizeof(R4<R>) * S, hipMemcpyHostToDevice));
            HIPC(hipMemset(sph.force, 0, sizeof(R4<R>) * S));
        }
        sph.S = S;
        sph.d = (R)in->diameter;
        sph.rho_s = (R)in->density;
        sph.rho_f = (R)P.Density;
        sph.nu_f = (R)in->fluid_viscosity;
        const R l0 = (R)P.InitialParticleDistance;
        sph.mf = (R)P.Density * l0 * l0 * l0;
        d.c.has_sph = S > 0 ? 1 : 0;
        return MPSR_NOERR;
    }
    mpsr_status get_spheres(mpsr_spheres* o) override {
        HIPC(hipSetDevice(device));
        if (!o) return MPSR_PARAMERR;
        HIPC(hipStreamSynchronize(st));
        const int n = std::min(o->n, sph.S);
        std::vector<R4<R>> p(std::max(1, n)), v(std::max(1, n)), f(std::max(1, n));
        if (n > 0) {
            HIPC(hipMemcpy(p.data(), sph.pos, sizeof(R4<R>) * n, hipMemcpyDeviceToHost));
            HIPC(hipMemcpy(v.data(), sph.vel, sizeof(R4<R>) * n, hipMemcpyDeviceToHost));
            HIPC(hipMemcpy(f.data(), sph.force, sizeof(R4<R>) * n, hipMemcpyDeviceToHost));
        }
        for (int k = 0; k < n; ++k) {
            const R4<R>* src[3] = {&p[k], &v[k], &f[k]};
            double* dst[3] = {o->position, o->velocity, o->force};
            for (int a = 0; a < 3; ++a)
                if (dst[a]) {
                    dst[a][3 * k] = (double)src[a]->x;
                    dst[a][3 * k + 1] = (double)src[a]->y;
                    dst[a][3 * k + 2] = (double)src[a]->z;
                }
        }
        o->n = sph.S;
        o->diameter = (double)sph.d;
        o->density = (double)sph.rho_s;
        o->fluid_viscosity = (double)sph.nu_f;
        return MPSR_NOERR;
    }
    // Returns whether a bucket was added at this step (the wall translation is then skipped).
    mpsr_status bucket_addition(int step, bool& added) {
        added = false;
        for (BucketD& B : buckets) {
            if (!(B.added < B.total && (step + 1) % B.interval == 0)) continue;
            added = true;
            launch_bucket_reuse(d, B.s, B.count, bk_flag, bk_rank, bk_dead, st);
            HIPC(hipMemcpyAsync(bk_dead_h, bk_dead, sizeof(int), hipMemcpyDeviceToHost, st));
            HIPC(hipStreamSynchronize(st));
            const int used = std::min(*bk_dead_h, B.count);
            if (used < B.count && (long long)d.N + B.count - used <= (long long)d.cap) {
                launch_bucket_append(d, B.s, used, B.count - used, st);
                d.N += B.count - used;
            }
            launch_reset_walls(d, tris0, step, st);
            ++B.added;
        }
        return MPSR_NOERR;
    }

    mpsr_status init(const mpsr_create_info* ci) {
        P = ci->params;
        if (P.Dimension != 3) return fail(MPSR_PARAMERR, "only the 3D operators are built (Dimension must be 3)");
        if (P.bUpdateTemperature && P.ThermalStep < 1)
            return fail(MPSR_PARAMERR, "ThermalStep must be at least 1 when the temperature is updated");
        if (!ci->particles || ci->particles->n < 0) return fail(MPSR_PARAMERR, "particles missing");
        if (ci->num_triangles > 0 && (!ci->triangles || ci->num_models <= 0 || !ci->models))
            return fail(MPSR_PARAMERR, "triangles need their model parameters");
        for (int t = 0; t < ci->num_triangles; ++t) {
            const int m = ci->triangles[t].ModelIDNumber;
            if (m < 0 || m >= ci->num_models) return fail(MPSR_PARAMERR, "triangle model id out of range");
            const mpsr_tri_params& mp = ci->models[m];
            if (mp.isDrag && (mp.DragTriangleID < 0 || mp.DragTriangleID >= ci->num_drag))
                return fail(MPSR_PARAMERR, "drag model without drag parameters");
        }
        mpsr_status s = mpsr_make_gridbox(&P, &ci->compute_zone, &gb);
        if (s != MPSR_NOERR) return fail(s, "invalid compute zone");
        device = ci->device;
        HIPC(hipSetDevice(device));
        HIPC(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        HIPC(hipEventCreate(&ev_a));
        HIPC(hipEventCreate(&ev_b));
        const int N = ci->particles->n;
        const int cap = std::max(ci->max_particles, N);
        d.cap = cap;
        d.ntri = ci->num_triangles;
        d.nmodel = ci->num_models;
        d.ndrag = ci->num_drag;
        d.c = make_consts<R>(P, gb, d.ntri, d.ndrag);
        pressure_mode = ci->pressure_mode;
        if (ci->ppe.coarse_cells < 0 || ci->ppe.coarse_cells > 8)
            return fail(MPSR_PARAMERR, "ppe.coarse_cells must be 0 (default 2) or 1..8");
        ppe_opt = ci->ppe;
        int bits = 1;
        while ((1ll << bits) <= (long long)d.c.ncell) ++bits;  // keys 0..ncell (sentinel = ncell)
        d.key_bits = bits;
        if (alloc_state(d.in, cap) != MPSR_NOERR || alloc_state(d.out, cap) != MPSR_NOERR) return MPSR_GPUERR;
        ALLOC(d.keys, cap); ALLOC(d.vals, cap);
        ALLOC(d.cell_start, d.c.ncell); ALLOC(d.cell_end, d.c.ncell);
        ALLOC(d.spos, cap); ALLOC(d.svel, cap); ALLOC(d.spres, cap);
        ALLOC(d.stl_mag, cap); ALLOC(d.stl_dir, cap); ALLOC(d.stl_id, cap);
        ALLOC(d.drag_acc, cap); ALLOC(d.drag_temp, cap);
        HIPC(hipMemset(d.drag_acc, 0, sizeof(R4<R>) * cap));
        HIPC(hipMemset(d.drag_temp, 0, sizeof(R) * cap));
        s = upload_state(ci->particles);
        if (s != MPSR_NOERR) return s;
        // walls
        models_h.assign(ci->models, ci->models + ci->num_models);
        ALLOC(d.models, std::max(1, d.nmodel));
        if (d.nmodel) HIPC(hipMemcpy(d.models, ci->models, sizeof(mpsr_tri_params) * d.nmodel, hipMemcpyHostToDevice));
        ALLOC(d.drag, std::max(1, d.ndrag));
        if (d.ndrag) HIPC(hipMemcpy(d.drag, ci->drag, sizeof(mpsr_drag_params) * d.ndrag, hipMemcpyHostToDevice));
        std::vector<TriD<R>> th(std::max(1, d.ntri));
        for (int t = 0; t < d.ntri; ++t) {
            const mpsr_triangle& T = ci->triangles[t];
            TriD<R> q{};
            for (int k = 0; k < 3; ++k) {
                q.v[3 * k] = (R)T.Vertex[k].x;
                q.v[3 * k + 1] = (R)T.Vertex[k].y;
                q.v[3 * k + 2] = (R)T.Vertex[k].z;
            }
            q.nrm[0] = (R)T.Normal.x; q.nrm[1] = (R)T.Normal.y; q.nrm[2] = (R)T.Normal.z;
            q.model = T.ModelIDNumber;
            q.is_drag = ci->models[T.ModelIDNumber].isDrag ? 1 : 0;
            th[t] = q;
        }
        ALLOC(d.tris, std::max(1, d.ntri));
        HIPC(hipMemcpy(d.tris, th.data(), sizeof(TriD<R>) * std::max(1, d.ntri), hipMemcpyHostToDevice));
        ALLOC(tris0, std::max(1, d.ntri));
        HIPC(hipMemcpy(tris0, th.data(), sizeof(TriD<R>) * std::max(1, d.ntri), hipMemcpyHostToDevice));
        ALLOC(d.tri_emit, d.ntri + 1); ALLOC(d.tri_off, d.ntri + 1);
        ALLOC(d.pair_n, 1); ALLOC(d.uniq_n, 1); ALLOC(d.overflow, 1);
        HIPC(hipMemset(d.overflow, 0, sizeof(int)));
        HIPC(hipMemset(d.uniq_n, 0, sizeof(int)));
        ALLOC(d.cell_tri_start, d.c.ncell); ALLOC(d.cell_tri_count, d.c.ncell);
        HIPC(hipMemset(d.cell_tri_count, 0, sizeof(int) * d.c.ncell));
        ALLOC(d.cell_uhead, d.c.ncell);
        ALLOC(d.cell_work, d.c.ncell); ALLOC(d.cell_work_n, 2);
        HIPC(hipMemset(d.cell_uhead, 0xff, sizeof(int2) * d.c.ncell));   // count -1: no walk yet
        // capacity of the (cell, triangle) pair buffer from the walls as given
        int pair_total = 0;
        if (d.ntri > 0) {
            // count pass alone, with a scan scratch sized for the triangles
            if (alloc_sort(d.tsort, d.ntri + 1) != MPSR_NOERR) return MPSR_GPUERR;
            launch_tri_count_total(d, st);
            HIPC(hipStreamSynchronize(st));
            HIPC(hipMemcpy(&pair_total, d.pair_n, sizeof(int), hipMemcpyDeviceToHost));
        }
        d.pair_cap = std::max(64, pair_total * 2 + 65536);
        // the cell walks' pool: a cell reserves at most the length of its 27 stencil lists, so 16
        // pair_cap entries hold every reservation while the pairs stay below 16/27 of pair_cap (more
        // than the initial count); a cell that finds the pool full takes the full walk
        d.utri_cap = (int)std::min<long long>(16LL * d.pair_cap, 1LL << 30);
        ALLOC(d.cell_utri, d.utri_cap);
        ALLOC(d.pair_cell, d.pair_cap); ALLOC(d.pair_tri, d.pair_cap);
        ALLOC(d.pair_flag, d.pair_cap); ALLOC(d.pair_pos, d.pair_cap);
        ALLOC(d.uniq_cell, d.pair_cap); ALLOC(d.uniq_tri, d.pair_cap);
        if (alloc_sort(d.sort, cap) != MPSR_NOERR) return MPSR_GPUERR;
        if (alloc_sort(d.tsort, d.pair_cap) != MPSR_NOERR) return MPSR_GPUERR;
        // rotation coefficients table
        rot_steps = 64;
        ALLOC(rot_table, (size_t)rot_steps * std::max(1, d.nmodel));
        // Laplacian-radius lattice constants for the PPE
        double n0l, laml;
        lattice_n0_lambda((double)(R)P.LaplacianInfluenceRadiusCoefficient, &n0l, &laml);
        const double l0 = (double)(R)P.InitialParticleDistance;
        n0_lap = ppe_opt.n0_lap > 0 ? ppe_opt.n0_lap : n0l;
        lambda_lap = ppe_opt.lambda_lap > 0 ? ppe_opt.lambda_lap : laml * l0 * l0;
        n0_src = lattice_pnd<R>(d.c.l0, d.c.re_grad);
        if (pressure_mode == MPSR_PRESSURE_PPE) {
            ppe = ppe_create<R>(d, P, ppe_opt, n0_lap, lambda_lap, n0_src);
            if (!ppe) return fail(MPSR_MEMERR, "PPE solver allocation failed");
        } else {
            ALLOC(d.gnb_q, ((size_t)std::max(cap, 1) + 63) / 64 * 64 * kGradWidth);
            ALLOC(d.gnb_n, std::max(cap, 1));
        }
        HIPC(hipDeviceSynchronize());
        return MPSR_NOERR;
    }

    // ---- timing -----------------------------------------------------------------------------
    // profiling 1 brackets every phase with events, 2 only the PPE phase (each event costs the
    // stream a few microseconds)
    bool pev_open = false;
    void mark_begin(int ph) {
        pev_open = false;
        if (!profiling || (profiling == 2 && ph != PH_PPE)) return;
        if (pev_used + 2 > pev.size()) {
            for (int k = 0; k < 64; ++k) {
                hipEvent_t e;
                if (hipEventCreate(&e) != hipSuccess) return;
                pev.push_back(e);
                pev_phase.push_back(0);
            }
        }
        pev_phase[pev_used] = ph;
        (void)hipEventRecord(pev[pev_used], st);
        pev_open = true;
    }
    void mark_end() {
        if (!pev_open) return;
        (void)hipEventRecord(pev[pev_used + 1], st);
        pev_used += 2;
        pev_open = false;
    }

    // ---- phases -------------------------------------------------------------------------------
    long long call_iters = 0;          // PCG iterations of the current step / step_many call
    int call_rebuilds = 0;             // ... its coarse-operator rebuilds
    double call_relres_max = 0.0;      // ... and its worst final relative residual
    int bin_epoch = 0;                 // binnings so far
    int nb_epoch = -1;                 // the binning the PPE neighbour list was written for
    int gnb_epoch = -1;                // ... and the EOS pressure's gradient list
    void do_bin() {
        ++bin_epoch;
        mark_begin(PH_BIN);
        launch_bin_hash(d, st);
        launch_bin_ranges(d, st);
        mark_end();
    }
    void do_tri() {
        mark_begin(PH_TRI);
        launch_tri_bin(d, st);
        mark_end();
    }
    void do_stl() {
        mark_begin(PH_STL);
        launch_stl_distance(d, st);
        mark_end();
    }
    void do_drag() {
        mark_begin(PH_DRAG);
        launch_drag(d, st);
        mark_end();
    }
    // CalcTemperatureFactor every ThermalStep steps (MultiGPUHandler.cpp:385-390)
    void do_thermal(int step) {
        if (!P.bUpdateTemperature || step % P.ThermalStep != 0) return;
        mark_begin(PH_THERM);
        launch_thermal(d, nullptr, nullptr, P.ThermalStep, st);
        mark_end();
    }
    // CalcTurbulenceViscosity (radix MultiGPUHandler.cpp:1421-1427): after the temperature
    void do_turbulence() {
        if (!d.c.turb) return;
        mark_begin(PH_TURB);
        launch_turbulence(d, st);
        mark_end();
    }
    void do_spheres() {
        if (sph.S <= 0) return;
        mark_begin(PH_SPH);
        launch_spheres(d, sph, st);
        mark_end();
    }
    void do_explicit() {
        mark_begin(PH_EXPL);
        launch_explicit(d, st);
        mark_end();
    }
    void do_relocate() {
        mark_begin(PH_RELOC);
        launch_relocate(d, st);
        mark_end();
    }
    mpsr_status do_pressure() {
        if (pressure_mode == MPSR_PRESSURE_PPE) {
            mark_begin(PH_PND);
            ppe_start(ppe, d, st);
            nb_epoch = bin_epoch;
            mark_end();
            mark_begin(PH_PPE);
            mpsr_status s = ppe_solve(ppe, d, ppe_opt, profiling, st);
            mark_end();
            if (s != MPSR_NOERR) return fail(s, std::string("PPE solve failed: ") + (ppe_why(ppe)[0] ? ppe_why(ppe) : "device error"));
            int it = 0, rb = 0;
            double rr = 0.0;
            ppe_last(ppe, &it, &rr, &rb);
            call_iters += it;
            call_rebuilds += rb;
            call_relres_max = std::max(call_relres_max, rr);
            return s;
        }
        mark_begin(PH_PND);
        launch_pnd(d, 1, st);
        gnb_epoch = bin_epoch;
        mark_end();
        return MPSR_NOERR;
    }
    void do_gradient() {
        mark_begin(PH_GRAD);
        // the PND pass listed every row's neighbours (the PPE path's over the Laplacian radius, the
        // EOS path's over the gradient radius); a list serves while the binning it was written for
        // is the current one
        NbList nb{};
        if (ppe && nb_epoch == bin_epoch) nb = ppe_neighbors(ppe);
        else if (d.gnb_q && gnb_epoch == bin_epoch) nb = NbList{d.gnb_q, d.gnb_n, kGradWidth};
        launch_gradient(d, st, nb);
        mark_end();
    }
    mpsr_status upload_rot(int first, int count) {
        if (count > rot_steps) {
            rot_steps = count;
            RotCoef<R>* p = nullptr;
            ALLOC(p, (size_t)rot_steps * std::max(1, d.nmodel));
            rot_table = p;
        }
        rot_host.assign((size_t)count * std::max(1, d.nmodel), RotCoef<R>{});
        for (int k = 0; k < count; ++k)
            for (int m = 0; m < d.nmodel; ++m) rot_host[(size_t)k * d.nmodel + m] = rot_coef<R>(models_h[m], P, first + k);
        HIPC(hipMemcpyAsync(rot_table, rot_host.data(), sizeof(RotCoef<R>) * rot_host.size(), hipMemcpyHostToDevice, st));
        return MPSR_NOERR;
    }
    void do_walls(int k, bool translate) {
        mark_begin(PH_WALL);
        if (translate) launch_translate(d, st);
        d.rot = rot_table + (size_t)k * std::max(1, d.nmodel);
        launch_rotate(d, st);
        mark_end();
    }

    mpsr_status one_step(int k, int step) {
        do_bin();
        do_tri();
        do_stl();
        if (d.ndrag > 0) do_drag();
        do_thermal(step);
        do_turbulence();
        do_spheres();
        do_explicit();
        do_relocate();
        do_bin();
        do_stl();
        mpsr_status s = do_pressure();
        if (s != MPSR_NOERR) return s;
        do_gradient();
        do_relocate();
        bool added = false;
        if (!buckets.empty()) {
            s = bucket_addition(step, added);
            if (s != MPSR_NOERR) return s;
            if (added) do_relocate();
        }
        do_walls(k, !added);
        return MPSR_NOERR;
    }

    mpsr_status step_many(int first, int count) override {
        HIPC(hipSetDevice(device));
        if (count <= 0) return MPSR_NOERR;
        pev_used = 0;
        call_iters = 0;
        call_rebuilds = 0;
        call_relres_max = 0.0;
        if (ppe) ppe_begin_call(ppe);
        mpsr_status s = upload_rot(first, count);
        if (s != MPSR_NOERR) return s;
        // the host copy above must finish before rot_host is refilled by a later call
        HIPC(hipEventRecord(ev_a, st));
        for (int k = 0; k < count; ++k) {
            s = one_step(k, first + k);
            if (s != MPSR_NOERR) return s;
            HIPC(hipGetLastError());
        }
        HIPC(hipEventRecord(ev_b, st));
        HIPC(hipEventSynchronize(ev_b));
        float ms = 0.f;
        HIPC(hipEventElapsedTime(&ms, ev_a, ev_b));
        collect_timings(ms, count);
        int ovf = 0;
        HIPC(hipMemcpy(&ovf, d.overflow, sizeof(int), hipMemcpyDeviceToHost));
        if (ovf) return fail(MPSR_MEMERR, "triangle/cell pair buffer overflow: walls grew beyond the set-up capacity");
        if (ppe) {
            mpsr_status ps = ppe_check(ppe, &tm);
            if (ps != MPSR_NOERR) return fail(ps, std::string("PPE solve failed: ") + ppe_why(ppe));
        }
        return MPSR_NOERR;
    }

    // ---- launch shape and the batched form --------------------------------------------------
    int* step_ctr_buf = nullptr;       // device step counter of graph replays
    int* step_base_buf = nullptr;      // device index of the first replayed step
    std::vector<char*> snap;           // device copy of the in/out state and the triangles
    int snap_n = 0;

    mpsr_status set_radix(const mpsr_radix_terms* t, const mpsr_negative_pressure_model* m, int nm) override {
        HIPC(hipSetDevice(device));
        if (!t) return fail(MPSR_PARAMERR, "radix terms: NULL");
        if (t->bTurbulance && !(t->Cepsilon != 0.0f))
            return fail(MPSR_PARAMERR, "radix terms: Cepsilon must be nonzero with bTurbulance");
        if (t->bNegativePressure && (!m || nm != d.nmodel))
            return fail(MPSR_PARAMERR, "radix terms: bNegativePressure needs one negative-pressure entry per model");
        HIPC(hipStreamSynchronize(st));
        if (t->bTurbulance && !d.turb_nu) {
            ALLOC(d.turb_nu, d.cap);
            ALLOC(d.turb_k, d.cap);
            HIPC(hipMemset(d.turb_nu, 0, sizeof(R) * d.cap));
            HIPC(hipMemset(d.turb_k, 0, sizeof(R) * d.cap));
        }
        if (t->bNegativePressure) {
            std::vector<NegModel<R>> h(std::max(1, d.nmodel));
            for (int i = 0; i < d.nmodel; ++i) {
                NegModel<R> q{};
                q.active = ((R)m[i].RotationInRPM != 0 || m[i].NegativePressureCheck) ? 1 : 0;
                q.infl = (R)m[i].InfluenceRegion;
                q.pgas = (R)m[i].PressureGas;
                q.dx = (R)m[i].DirectionVector.x;
                q.dy = (R)m[i].DirectionVector.y;
                q.dz = (R)m[i].DirectionVector.z;
                h[i] = q;
            }
            if (!d.negm) ALLOC(d.negm, std::max(1, d.nmodel));
            HIPC(hipMemcpy(d.negm, h.data(), sizeof(NegModel<R>) * h.size(), hipMemcpyHostToDevice));
        }
        // radix DeviceProcess_Kernel.h:1304, 1405-1466: the expressions of the kernel's constants
        Consts<R>& c = d.c;
        c.turb = t->bTurbulance ? 1 : 0;
        c.negp = t->bNegativePressure ? 1 : 0;
        c.turb_g = (R)P.Dimension / (R)P.ConstantParticleDensity;
        c.turb_cs = (R)t->SmagorinskyConstant;
        c.turb_fw = (R)t->FilterWidth;
        c.turb_numax = (R)t->TurbulanViscosityRatio;
        c.turb_kmax = (R)t->TurbulanceIntensity;
        c.turb_tk = (R)2 * (R)t->Cnue * (R)P.InitialParticleDistance * (R)P.InitialParticleDistance / (R)t->Cepsilon;
        ++consts_ver;
        return MPSR_NOERR;
    }
    mpsr_status get_turbulence(double* nu, double* k, int n) override {
        HIPC(hipStreamSynchronize(st));
        if (n < 0 || n > d.cap) return fail(MPSR_PARAMERR, "row count exceeds the allocated slots");
        if (!d.turb_nu) return fail(MPSR_PARAMERR, "turbulence is off (mpsr_set_radix_terms)");
        std::vector<R> a, b;
        mpsr_status ds = d2h(a, d.turb_nu, (size_t)n);
        if (ds == MPSR_NOERR) ds = d2h(b, d.turb_k, (size_t)n);
        if (ds != MPSR_NOERR) return ds;
        for (int s = 0; s < n; ++s) {
            if (nu) nu[s] = (double)a[s];
            if (k) k[s] = (double)b[s];
        }
        return MPSR_NOERR;
    }
    mpsr_status set_launch(int block, int items) override {
        if (block < 64 || block > kMaxBlock || block % 64)
            return fail(MPSR_PARAMERR, "launch block must be a multiple of 64 in [64, 512]");
        if (!radix_items_valid(items)) return fail(MPSR_PARAMERR, "sort items per thread must be 4, 8 or 16");
        d.block = block;
        d.sort_items = items;
        return MPSR_NOERR;
    }
    mpsr_status batch_check() override {
        if (pressure_mode == MPSR_PRESSURE_PPE &&
            (ppe_opt.solver != MPSR_PCG_FP64 || ppe_opt.precond != MPSR_PRECOND_JACOBI ||
             ppe_opt.warm_start == MPSR_WARM_LEARNED))
            return fail(MPSR_PARAMERR, "batched step: a PPE bucket solves with the FP64 Jacobi-PCG (solver "
                                       "MPSR_PCG_FP64, preconditioner MPSR_PRECOND_JACOBI, no learned start); "
                                       "other buckets use the explicit (EOS) pressure");
        if (sph.S > 0) return fail(MPSR_PARAMERR, "batched step: rigid spheres are not part of the batched form");
        if (!buckets.empty()) return fail(MPSR_PARAMERR, "batched step: bucket inflow changes the particle count");
        return MPSR_NOERR;
    }
    // rotation table for steps first .. first+count-1 and the step counter at 0
    mpsr_status batch_prepare(int first, int count) override {
        HIPC(hipSetDevice(device));
        mpsr_status c = batch_check();
        if (c != MPSR_NOERR) return c;
        if (!step_ctr_buf) ALLOC(step_ctr_buf, 1);
        if (!step_base_buf) ALLOC(step_base_buf, 1);
        mpsr_status s = upload_rot(first, count);
        if (s != MPSR_NOERR) return s;
        fill_int(step_ctr_buf, 0, 1, st);
        fill_int(step_base_buf, first, 1, st);
        HIPC(hipStreamSynchronize(st));   // rot_host is refilled by the next call
        return MPSR_NOERR;
    }
    mpsr_status batch_finish() override {
        int ovf = 0;
        HIPC(hipMemcpy(&ovf, d.overflow, sizeof(int), hipMemcpyDeviceToHost));
        if (ovf) return fail(MPSR_MEMERR, "triangle/cell pair buffer overflow: walls grew beyond the set-up capacity");
        if (ppe) (void)ppe_check(ppe, &tm);   // the last batched solve's iterations and residual
        return MPSR_NOERR;
    }
    // restore = 0: copy the state (in and out arrays, triangles) aside; 1: copy it back
    mpsr_status snapshot(int restore) override {
        HIPC(hipSetDevice(device));
        std::vector<std::pair<void*, size_t>> f;
        const size_t n = (size_t)d.cap;
        for (PState<R>* ps : {&d.in, &d.out}) {
            f.push_back({ps->id, n * sizeof(int)});
            f.push_back({ps->type, n * sizeof(int)});
            for (R* q : {ps->px, ps->py, ps->pz, ps->vx, ps->vy, ps->vz, ps->p, ps->n, ps->temp, ps->nu, ps->solid})
                f.push_back({q, n * sizeof(R)});
        }
        f.push_back({d.tris, sizeof(TriD<R>) * std::max(1, d.ntri)});
        if (restore && snap.empty()) return fail(MPSR_PARAMERR, "no snapshot to restore");
        if (snap.empty()) {
            for (auto& x : f) {
                char* b = nullptr;
                ALLOC(b, x.second);
                snap.push_back(b);
            }
        }
        for (size_t i = 0; i < f.size(); ++i) {
            if (restore) HIPC(hipMemcpyAsync(f[i].first, snap[i], f[i].second, hipMemcpyDeviceToDevice, st));
            else HIPC(hipMemcpyAsync(snap[i], f[i].first, f[i].second, hipMemcpyDeviceToDevice, st));
        }
        if (restore) d.N = snap_n;
        else snap_n = d.N;
        HIPC(hipStreamSynchronize(st));
        return MPSR_NOERR;
    }
    hipStream_t stream() const override { return st; }

    void collect_timings(float total_ms, int count) {
        mpsr_timings t{};
        t.step_ms = total_ms / count;
        if (profiling) {
            double acc[PH_COUNT] = {0};
            for (size_t e = 0; e + 1 < pev_used; e += 2) {
                float ms = 0.f;
                if (hipEventElapsedTime(&ms, pev[e], pev[e + 1]) == hipSuccess) acc[pev_phase[e]] += ms;
            }
            t.bin_ms = acc[PH_BIN] / count;
            t.tri_ms = acc[PH_TRI] / count;
            t.stl_ms = acc[PH_STL] / count;
            t.drag_ms = acc[PH_DRAG] / count;
            t.sphere_ms = acc[PH_SPH] / count;
            t.thermal_ms = acc[PH_THERM] / count;
            t.turbulence_ms = acc[PH_TURB] / count;
            t.explicit_ms = acc[PH_EXPL] / count;
            t.relocate_ms = acc[PH_RELOC] / count;
            t.pnd_ms = acc[PH_PND] / count;
            t.ppe_ms = acc[PH_PPE] / count;
            t.gradient_ms = acc[PH_GRAD] / count;
            t.wall_ms = acc[PH_WALL] / count;
        }
        t.ppe_iterations = tm.ppe_iterations;
        t.ppe_refinements = tm.ppe_refinements;
        t.ppe_rel_residual = tm.ppe_rel_residual;
        t.ppe_initial_rel_residual = tm.ppe_initial_rel_residual;
        t.spmv_ms = tm.spmv_ms;
        t.spmv_launches = tm.spmv_launches;
        t.spmv_bytes = tm.spmv_bytes;
        t.ppe_nnz = tm.ppe_nnz;
        t.ppe_stored = tm.ppe_stored;
        t.ppe_rows = tm.ppe_rows;
        t.ppe_fallbacks = tm.ppe_fallbacks;
        t.ppe_coarse_rows = tm.ppe_coarse_rows;
        t.ppe_coarse_rebuilt = tm.ppe_coarse_rebuilt;
        t.ppe_setup_ms = tm.ppe_setup_ms;
        t.ppe_learned = tm.ppe_learned;
        t.ppe_learned_fallback = tm.ppe_learned_fallback;
        t.steps = count;
        t.ppe_iterations_total = call_iters;
        t.ppe_rebuilds = call_rebuilds;
        t.ppe_rel_residual_max = call_relres_max;
        tm = t;
    }

    mpsr_status phase(int which, int arg) override {
        HIPC(hipSetDevice(device));
        pev_used = 0;
        const int saved = profiling;
        profiling = 0;
        mpsr_status s = MPSR_NOERR;
        switch (which) {
            case 0: do_bin(); break;
            case 1: do_tri(); break;
            case 2: do_stl(); break;
            case 3: if (d.ndrag > 0) do_drag(); break;
            case 4: do_explicit(); break;
            case 5: do_relocate(); break;
            case 6:
                launch_pnd(d, 1, st);
                gnb_epoch = bin_epoch;
                break;
            case 7:
                if (!ppe) {
                    ppe = ppe_create<R>(d, P, ppe_opt, n0_lap, lambda_lap, n0_src);
                    if (!ppe) { s = fail(MPSR_MEMERR, "PPE solver allocation failed"); break; }
                }
                ppe_start(ppe, d, st);
                nb_epoch = bin_epoch;
                s = ppe_solve(ppe, d, ppe_opt, 0, st);
                if (s != MPSR_NOERR)
                    s = fail(s, std::string("PPE solve failed: ") + (ppe_why(ppe)[0] ? ppe_why(ppe) : "device error"));
                break;
            case 8: do_gradient(); break;
            case 9: launch_translate(d, st); break;
            case 10:
                s = upload_rot(arg, 1);
                if (s == MPSR_NOERR) { d.rot = rot_table; launch_rotate(d, st); }
                break;
            case 11: do_spheres(); break;
            case 12: launch_thermal(d, nullptr, nullptr, P.ThermalStep, st); break;
            case 13:
                if (!d.turb_nu) { s = fail(MPSR_PARAMERR, "turbulence is off (mpsr_set_radix_terms)"); break; }
                launch_turbulence(d, st);
                break;
            default: s = fail(MPSR_PARAMERR, "unknown phase");
        }
        profiling = saved;
        if (s != MPSR_NOERR) return s;
        HIPC(hipGetLastError());
        HIPC(hipStreamSynchronize(st));
        if (which == 7 && ppe) {
            mpsr_status ps = ppe_check(ppe, &tm);
            if (ps != MPSR_NOERR) return fail(ps, std::string("PPE solve failed: ") + ppe_why(ppe));
        }
        return MPSR_NOERR;
    }

    template <typename T> mpsr_status d2h(std::vector<T>& v, const T* p, size_t n) {
        v.resize(n);
        if (n) HIPC(hipMemcpy(v.data(), p, n * sizeof(T), hipMemcpyDeviceToHost));
        return MPSR_NOERR;
    }
#define D2H(v, p, n)                                   \
    do {                                               \
        mpsr_status s_ = d2h(v, p, (size_t)(n));       \
        if (s_ != MPSR_NOERR) return s_;               \
    } while (0)

    mpsr_status sync() override {
        HIPC(hipSetDevice(device));
        HIPC(hipStreamSynchronize(st));
        return MPSR_NOERR;
    }

    mpsr_status get_particles(int which, mpsr_particles* o) override {
        HIPC(hipStreamSynchronize(st));
        const PState<R>& S = which ? d.out : d.in;
        const int n = std::min(o->n, d.N);
        std::vector<int> ids, type;
        std::vector<R> f[11];
        D2H(ids, S.id, n); D2H(type, S.type, n);
        const R* src[11] = {S.px, S.py, S.pz, S.vx, S.vy, S.vz, S.p, S.n, S.temp, S.nu, S.solid};
        for (int a = 0; a < 11; ++a) D2H(f[a], src[a], n);
        for (int s = 0; s < n; ++s) {
            if (o->id) o->id[s] = ids[s];
            if (o->type) o->type[s] = type[s];
            for (int a = 0; a < 3; ++a) {
                if (o->position) o->position[3 * s + a] = (double)f[a][s];
                if (o->velocity) o->velocity[3 * s + a] = (double)f[3 + a][s];
            }
            if (o->pressure) o->pressure[s] = f[6][s];
            if (o->density) o->density[s] = f[7][s];
            if (o->temperature) o->temperature[s] = f[8][s];
            if (o->kinetic_viscosity) o->kinetic_viscosity[s] = f[9][s];
            if (o->solid_phase_rate) o->solid_phase_rate[s] = f[10][s];
        }
        return MPSR_NOERR;
    }
    mpsr_status set_particles(const mpsr_particles* p) override {
        if (!p || p->n < 0 || p->n > d.cap) return fail(MPSR_PARAMERR, "particle count exceeds the allocated slots");
        HIPC(hipStreamSynchronize(st));
        return upload_state(p);
    }
    mpsr_status get_triangles(mpsr_triangle* o, int n) override {
        HIPC(hipStreamSynchronize(st));
        std::vector<TriD<R>> th;
        D2H(th, d.tris, std::min(n, d.ntri));
        for (size_t t = 0; t < th.size(); ++t) {
            o[t].ModelIDNumber = th[t].model;
            o[t].Normal = {(float)th[t].nrm[0], (float)th[t].nrm[1], (float)th[t].nrm[2]};
            for (int k = 0; k < 3; ++k)
                o[t].Vertex[k] = {(float)th[t].v[3 * k], (float)th[t].v[3 * k + 1], (float)th[t].v[3 * k + 2]};
        }
        return MPSR_NOERR;
    }
    mpsr_status get_binning(int* k, int* p, int* cs, int* ce) override {
        HIPC(hipStreamSynchronize(st));
        if (!d.perm) return fail(MPSR_CALCERR, "no binning yet");
        std::vector<uint32_t> kk;
        std::vector<int> pp, a, b;
        D2H(kk, d.skeys, d.N); D2H(pp, d.perm, d.N);
        D2H(a, d.cell_start, d.c.ncell); D2H(b, d.cell_end, d.c.ncell);
        for (int t = 0; t < d.N; ++t) {
            if (k) k[t] = (int)kk[t];
            if (p) p[t] = pp[t];
        }
        if (cs) std::memcpy(cs, a.data(), sizeof(int) * a.size());
        if (ce) std::memcpy(ce, b.data(), sizeof(int) * b.size());
        return MPSR_NOERR;
    }
    mpsr_status get_triangle_cells(int* cs, int* cc, int* ids, int cap, int* total) override {
        HIPC(hipStreamSynchronize(st));
        std::vector<int> a, b, u;
        D2H(a, d.cell_tri_start, d.c.ncell); D2H(b, d.cell_tri_count, d.c.ncell);
        int un = 0;
        HIPC(hipMemcpy(&un, d.uniq_n, sizeof(int), hipMemcpyDeviceToHost));
        if (d.ntri > 0 && un > 0) D2H(u, d.uniq_tri, std::min(un, d.pair_cap));
        // compact to the capped per-cell lists in cell order
        int e = 0;
        for (int c = 0; c < d.c.ncell; ++c) {
            if (cs) cs[c] = e;
            const int n = b[c];
            if (cc) cc[c] = n;
            for (int q = 0; q < n; ++q, ++e)
                if (ids && e < cap) ids[e] = u[a[c] + q];
        }
        if (total) *total = e;
        return MPSR_NOERR;
    }
    mpsr_status get_stl(double* m, double* dir, int* id) override {
        HIPC(hipStreamSynchronize(st));
        std::vector<R> mm;
        std::vector<R4<R>> dd;
        std::vector<int> ii, pp;
        D2H(mm, d.stl_mag, d.N); D2H(dd, d.stl_dir, d.N); D2H(ii, d.stl_id, d.N); D2H(pp, d.perm, d.N);
        for (int t = 0; t < d.N; ++t) {
            const int s = pp[t];
            if (m) m[s] = mm[t];
            if (dir) { dir[3 * s] = dd[t].x; dir[3 * s + 1] = dd[t].y; dir[3 * s + 2] = dd[t].z; }
            if (id) id[s] = ii[t];
        }
        return MPSR_NOERR;
    }
    mpsr_status get_neighbors(double r, int* off, int* nbr, int cap, int* total) override {
        HIPC(hipStreamSynchronize(st));
        if (!d.perm) return fail(MPSR_CALCERR, "no binning yet");
        const R re = (R)r * d.c.l0;
        int *cnt = nullptr, *doff = nullptr, *dn = nullptr;
        HIPC(hipMalloc(&cnt, sizeof(int) * (d.N + 1)));
        HIPC(hipMalloc(&doff, sizeof(int) * (d.N + 1)));
        launch_neighbor_count(d, re, cnt, st);
        std::vector<int> c(d.N);
        HIPC(hipStreamSynchronize(st));
        HIPC(hipMemcpy(c.data(), cnt, sizeof(int) * d.N, hipMemcpyDeviceToHost));
        std::vector<int> o(d.N + 1, 0);
        for (int t = 0; t < d.N; ++t) o[t + 1] = o[t] + c[t];
        const int tot = o[d.N];
        HIPC(hipMemcpy(doff, o.data(), sizeof(int) * (d.N + 1), hipMemcpyHostToDevice));
        HIPC(hipMalloc(&dn, sizeof(int) * std::max(1, tot)));
        launch_neighbor_fill(d, re, doff, dn, st);
        std::vector<int> nb(tot), pp;
        HIPC(hipStreamSynchronize(st));
        if (tot) HIPC(hipMemcpy(nb.data(), dn, sizeof(int) * tot, hipMemcpyDeviceToHost));
        D2H(pp, d.perm, d.N);
        (void)hipFree(cnt); (void)hipFree(doff); (void)hipFree(dn);
        // rows in slot order
        std::vector<int> row_of_slot(d.N);
        for (int t = 0; t < d.N; ++t) row_of_slot[pp[t]] = t;
        int e = 0;
        off[0] = 0;
        for (int s = 0; s < d.N; ++s) {
            const int t = row_of_slot[s];
            for (int q = o[t]; q < o[t + 1]; ++q, ++e)
                if (e < cap) nbr[e] = nb[q];
            off[s + 1] = e;
        }
        if (total) *total = tot;
        return MPSR_NOERR;
    }
    mpsr_status get_ppe(double* rhs, double* x, int* dir) override {
        HIPC(hipStreamSynchronize(st));
        if (!ppe) return fail(MPSR_CALCERR, "no PPE solve yet");
        return ppe_get(ppe, d, rhs, x, dir);
    }
    mpsr_status get_ppe_matrix(int* ptr, int* col, double* val, double* dg, int cap, int* nnz) override {
        HIPC(hipStreamSynchronize(st));
        if (!ppe) return fail(MPSR_CALCERR, "no PPE solve yet");
        return ppe_get_matrix(ppe, d, ptr, col, val, dg, cap, nnz);
    }
    mpsr_status set_dt(float dt, float courant) override {
        HIPC(hipStreamSynchronize(st));
        P.Dt = dt;
        P.CourantNumber = courant;
        // the constants set by mpsr_set_spheres and mpsr_set_radix_terms do not depend on dt
        const Consts<R> old = d.c;
        d.c = make_consts<R>(P, gb, d.ntri, d.ndrag);
        d.c.has_sph = old.has_sph;
        d.c.turb = old.turb;
        d.c.negp = old.negp;
        d.c.turb_g = old.turb_g;
        d.c.turb_cs = old.turb_cs;
        d.c.turb_fw = old.turb_fw;
        d.c.turb_numax = old.turb_numax;
        d.c.turb_kmax = old.turb_kmax;
        d.c.turb_tk = old.turb_tk;
        ++consts_ver;
        if (ppe) ppe_update_params(ppe, P);
        return MPSR_NOERR;
    }
    mpsr_status set_ppe(const mpsr_ppe_options* o) override {
        if (o->coarse_cells < 0 || o->coarse_cells > 8)
            return fail(MPSR_PARAMERR, "ppe.coarse_cells must be 0 (default 2) or 1..8");
        ppe_opt = *o;
        return MPSR_NOERR;
    }
};

// ---------------------------------------------------------------------------------------------
// The batched step of the buckets of one precision: every kernel of the EOS step is one launch
// over a work list of (bucket, block) items of all buckets, so a step of all buckets is the same
// few dozen launches whatever their number, and small buckets share the GPU with large ones
// inside each launch.  The kernels read each bucket's buffers, grid and constants from a device
// table of the buckets' Dev structs; the work lists and the per-bucket sort, scan and fill jobs
// sit in one device block built once per launch shape.
// ---------------------------------------------------------------------------------------------
template <typename R> struct BatchGroup {
    std::vector<Engine<R>*> eng;       // largest bucket first
    std::vector<Dev<R>> host;
    Dev<R>* devs = nullptr;
    char* tab = nullptr;
    size_t tab_bytes = 0;
    int block = 0, items = 0;
    long long sig = -1;                // launch shape and optional phases the tables were built for
    struct Pass {
        const SortJob* jobs = nullptr;
        WorkList tiles, singles;
    };
    WorkList wp, wt64, wt256, wpairs, wone, wdrag, wtherm, wturb, wc256;
    int2* walk_items = nullptr;        // (bucket, cell) of the cell walks, one per cell of the buckets
    int* walk_n = nullptr;
    size_t walk_cap = 0;
    std::vector<Pass> psort, tsort;
    const FillJob* fill_bin = nullptr;
    const FillJob* fill_tri = nullptr;
    WorkList fill_bin_w, fill_tri_w;
    const ScanJob* scan_emit = nullptr;
    const ScanJob* scan_flag = nullptr;
    WorkList scan_emit_t, scan_emit_1, scan_flag_t, scan_flag_1;
    WorkList wp_eos;                   // particle blocks of the buckets with the EOS pressure
    PpeBatch<R>* pb = nullptr;         // the PPE buckets' solves (null: none)

    ~BatchGroup() {
        if (devs) (void)hipFree(devs);
        if (tab) (void)hipFree(tab);
        if (walk_items) (void)hipFree(walk_items);
        if (walk_n) (void)hipFree(walk_n);
        ppe_batch_destroy(pb);
    }
    bool has_ppe() const {
        for (Engine<R>* e : eng)
            if (e->pressure_mode == MPSR_PRESSURE_PPE) return true;
        return false;
    }
    BatchLaunch<R> launch() const { return BatchLaunch<R>{devs, block}; }

    // the shape of the tables: launch shape and which buckets run the optional phases
    long long signature(int blk, int itm) const {
        long long h = (long long)blk * 131 + itm;
        for (Engine<R>* e : eng) {
            h = h * 8 + (e->d.ndrag > 0 ? 1 : 0) + (e->P.bUpdateTemperature ? 2 : 0) +
                (e->d.c.turb && e->d.turb_nu ? 4 : 0);
            if (e->pressure_mode == MPSR_PRESSURE_PPE) {
                long long t = 0;
                std::memcpy(&t, &e->ppe_opt.rel_tol, sizeof(t));
                h = h * 1000003 + (t ^ ((long long)e->ppe_opt.max_iter << 3) ^ e->ppe_opt.warm_start);
            }
        }
        return h;
    }

    // every bucket's Dev for the batched kernels: rotation table and step counter of replays, and
    // the buffers the particle sort leaves the keys and permutation in
    mpsr_status refresh() {
        host.resize(eng.size());
        for (size_t k = 0; k < eng.size(); ++k) {
            Engine<R>* e = eng[k];
            Dev<R>& d = e->d;
            const bool odd = d.N > 0 && (radix_passes(d.key_bits) & 1);
            d.skeys = odd ? d.sort.k1 : d.keys;
            d.perm = odd ? d.sort.v1 : d.vals;
            Dev<R> x = d;
            x.rot = e->rot_table;
            x.step_ctr = e->step_ctr_buf;
            x.step_base = e->step_base_buf;
            x.thermal_step = e->P.ThermalStep > 0 ? e->P.ThermalStep : 1;
            // the corrector walks the list of the step's own PND pass
            if (e->pressure_mode == MPSR_PRESSURE_PPE) x.nbl = e->ppe ? ppe_neighbors(e->ppe) : NbList{};
            else x.nbl = d.gnb_q ? NbList{d.gnb_q, d.gnb_n, kGradWidth} : NbList{};
            host[k] = x;
        }
        if (!devs) HIPC(hipMalloc((void**)&devs, sizeof(Dev<R>) * std::max<size_t>(1, eng.size())));
        HIPC(hipMemcpy(devs, host.data(), sizeof(Dev<R>) * host.size(), hipMemcpyHostToDevice));
        return MPSR_NOERR;
    }

    mpsr_status build(int blk, int itm) {
        const long long sg = signature(blk, itm);
        if (tab && sg == sig) return MPSR_NOERR;
        block = blk;
        items = itm;
        // host images of the lists and jobs, packed into one device block
        std::vector<char> img;
        auto put = [&](const void* p, size_t n) {
            const size_t off = (img.size() + 15) & ~(size_t)15;
            img.resize(off + n);
            if (n) std::memcpy(img.data() + off, p, n);
            return off;
        };
        struct Ref {
            size_t off;
            int n;
        };
        auto items_of = [&](const std::vector<int2>& v) { return Ref{put(v.data(), v.size() * sizeof(int2)), (int)v.size()}; };
        std::vector<int2> p, t64, t256, pairs, one, drag, therm, turb, peos, c256;
        const int n = (int)eng.size();
        size_t ncells = 0;
        for (int k = 0; k < n; ++k) {
            const Dev<R>& d = host[k];
            const int nb = (int)grid_for(d.N, block);
            for (int j = 0; j < nb; ++j) {
                p.push_back(make_int2(k, j));
                if (eng[k]->pressure_mode != MPSR_PRESSURE_PPE) peos.push_back(make_int2(k, j));
                if (d.ndrag > 0) drag.push_back(make_int2(k, j));
                if (eng[k]->P.bUpdateTemperature) therm.push_back(make_int2(k, j));
                if (d.c.turb && d.turb_nu) turb.push_back(make_int2(k, j));
            }
            if (d.ntri > 0) {
                for (int j = 0; j < (int)grid_for(d.ntri, 64); ++j) t64.push_back(make_int2(k, j));
                for (int j = 0; j < (int)grid_for(d.ntri, 256); ++j) t256.push_back(make_int2(k, j));
                for (int j = 0; j < (int)grid_for(d.pair_cap, 256); ++j) pairs.push_back(make_int2(k, j));
            }
            if (d.N > 0) {
                for (int j = 0; j < (int)grid_for(d.c.ncell, 256); ++j) c256.push_back(make_int2(k, j));
                ncells += (size_t)d.c.ncell;
            }
            one.push_back(make_int2(k, 0));
        }
        const Ref rp = items_of(p), r64 = items_of(t64), r256 = items_of(t256), rpairs = items_of(pairs),
                  rone = items_of(one), rdrag = items_of(drag), rtherm = items_of(therm), rturb = items_of(turb),
                  rpeos = items_of(peos), rc256 = items_of(c256);
        // fills: cell ranges of the particle binning; triangle counts per cell (and the pair count of
        // a bucket without triangles)
        std::vector<FillJob> fb, ft;
        std::vector<int2> fbw, ftw;
        for (int k = 0; k < n; ++k) {
            const Dev<R>& d = host[k];
            auto add = [&](std::vector<FillJob>& jobs, std::vector<int2>& w, int* ptr, int value, int cnt) {
                const int j = (int)jobs.size();
                jobs.push_back(FillJob{ptr, value, cnt});
                for (int b = 0; b < (int)grid_for(cnt, 256); ++b) w.push_back(make_int2(j, b));
            };
            if (d.N > 0) {
                add(fb, fbw, d.cell_start, -1, (int)d.c.ncell);
                add(fb, fbw, d.cell_end, -1, (int)d.c.ncell);
            }
            add(ft, ftw, d.cell_tri_count, 0, (int)d.c.ncell);
            add(ft, ftw, d.cell_work_n, 0, 2);
            if (d.ntri <= 0) add(ft, ftw, d.uniq_n, 0, 1);
        }
        const size_t ofb = put(fb.data(), fb.size() * sizeof(FillJob)), oft = put(ft.data(), ft.size() * sizeof(FillJob));
        const Ref rfbw = items_of(fbw), rftw = items_of(ftw);
        // scans of the triangle pass: emitted pairs per triangle, then the unique-pair flags
        std::vector<ScanJob> se, sf;
        std::vector<int2> se_t, se_1, sf_t, sf_1;
        for (int k = 0; k < n; ++k) {
            const Dev<R>& d = host[k];
            if (d.ntri <= 0) continue;
            const int je = (int)se.size();
            se.push_back(ScanJob{d.tri_emit, d.tri_off, nullptr, d.ntri, d.ntri, d.pair_n, d.tsort.scan_tmp});
            for (int b = 0; b < (int)grid_for(d.ntri, scan_tile()); ++b) se_t.push_back(make_int2(je, b));
            se_1.push_back(make_int2(je, 0));
            const int jf = (int)sf.size();
            sf.push_back(ScanJob{d.pair_flag, d.pair_pos, d.pair_n, 0, d.pair_cap, d.uniq_n, d.tsort.scan_tmp});
            for (int b = 0; b < (int)grid_for(d.pair_cap, scan_tile()); ++b) sf_t.push_back(make_int2(jf, b));
            sf_1.push_back(make_int2(jf, 0));
        }
        const size_t ose = put(se.data(), se.size() * sizeof(ScanJob)), osf = put(sf.data(), sf.size() * sizeof(ScanJob));
        const Ref rse_t = items_of(se_t), rse_1 = items_of(se_1), rsf_t = items_of(sf_t), rsf_1 = items_of(sf_1);
        // radix passes: the particle sort (the buckets' key bits, `items` keys per thread) and the
        // pair sort (8 per thread, device-resident pair count), pass by pass over the buckets that
        // have that many passes
        struct PassImg {
            size_t jobs;
            Ref tiles, singles;
        };
        auto passes = [&](bool pairs_sort, std::vector<PassImg>& out) {
            const int tile = radix_tile(pairs_sort ? 8 : items);
            for (int pass = 0;; ++pass) {
                std::vector<SortJob> jobs;
                std::vector<int2> tiles, singles;
                for (int k = 0; k < n; ++k) {
                    const Dev<R>& d = host[k];
                    const int bits = pairs_sort ? tri_pair_bits(d.c.ncell) : d.key_bits;
                    const int cap = pairs_sort ? d.pair_cap : d.N;
                    if ((pairs_sort ? d.ntri <= 0 : d.N <= 0) || bits <= 0 || cap <= 0) continue;
                    if (pass >= radix_passes(bits)) continue;
                    uint32_t* k0 = pairs_sort ? d.pair_cell : d.keys;
                    int* v0 = pairs_sort ? d.pair_tri : d.vals;
                    const SortScratch& sc = pairs_sort ? d.tsort : d.sort;
                    const bool even = (pass & 1) == 0;
                    SortJob q{};
                    q.kin = even ? k0 : sc.k1;
                    q.vin = even ? v0 : sc.v1;
                    q.kout = even ? sc.k1 : k0;
                    q.vout = even ? sc.v1 : v0;
                    q.d_n = pairs_sort ? d.pair_n : nullptr;
                    q.n_host = pairs_sort ? 0 : d.N;
                    q.n_cap = cap;
                    q.nbits = radix_pass_bits(bits, pass, &q.shift);
                    q.hist = sc.hist;
                    const int j = (int)jobs.size();
                    jobs.push_back(q);
                    for (int b = 0; b < (int)grid_for(cap, tile); ++b) tiles.push_back(make_int2(j, b));
                    singles.push_back(make_int2(j, 0));
                }
                if (jobs.empty()) break;
                out.push_back(PassImg{put(jobs.data(), jobs.size() * sizeof(SortJob)), items_of(tiles), items_of(singles)});
            }
        };
        std::vector<PassImg> pp, tp;
        passes(false, pp);
        passes(true, tp);
        // upload and resolve the offsets
        if (tab && img.size() > tab_bytes) {
            (void)hipFree(tab);
            tab = nullptr;
        }
        if (!tab) {
            tab_bytes = std::max<size_t>(16, img.size());
            HIPC(hipMalloc((void**)&tab, tab_bytes));
        }
        HIPC(hipMemcpy(tab, img.data(), img.size(), hipMemcpyHostToDevice));
        auto wl = [&](Ref r) { return WorkList{reinterpret_cast<const int2*>(tab + r.off), r.n}; };
        wp = wl(rp); wt64 = wl(r64); wt256 = wl(r256); wpairs = wl(rpairs); wone = wl(rone);
        wdrag = wl(rdrag); wtherm = wl(rtherm); wturb = wl(rturb); wp_eos = wl(rpeos); wc256 = wl(rc256);
        if (ncells > walk_cap) {
            if (walk_items) (void)hipFree(walk_items);
            walk_items = nullptr;
            walk_cap = 0;
            HIPC(hipMalloc((void**)&walk_items, sizeof(int2) * ncells));
            walk_cap = ncells;
        }
        if (!walk_n) HIPC(hipMalloc((void**)&walk_n, sizeof(int)));
        fill_bin = reinterpret_cast<const FillJob*>(tab + ofb);
        fill_tri = reinterpret_cast<const FillJob*>(tab + oft);
        fill_bin_w = wl(rfbw);
        fill_tri_w = wl(rftw);
        scan_emit = reinterpret_cast<const ScanJob*>(tab + ose);
        scan_flag = reinterpret_cast<const ScanJob*>(tab + osf);
        scan_emit_t = wl(rse_t); scan_emit_1 = wl(rse_1); scan_flag_t = wl(rsf_t); scan_flag_1 = wl(rsf_1);
        psort.clear();
        tsort.clear();
        for (const PassImg& q : pp) psort.push_back(Pass{reinterpret_cast<const SortJob*>(tab + q.jobs), wl(q.tiles), wl(q.singles)});
        for (const PassImg& q : tp) tsort.push_back(Pass{reinterpret_cast<const SortJob*>(tab + q.jobs), wl(q.tiles), wl(q.singles)});
        // the PPE buckets' solver tables
        std::vector<PpeSolver<R>*> solvers;
        std::vector<int> pdev;
        std::vector<mpsr_ppe_options> popt;
        for (int k = 0; k < n; ++k)
            if (eng[k]->pressure_mode == MPSR_PRESSURE_PPE) {
                solvers.push_back(eng[k]->ppe);
                pdev.push_back(k);
                popt.push_back(eng[k]->ppe_opt);
            }
        if (!solvers.empty()) {
            if (!pb) pb = ppe_batch_create<R>();
            mpsr_status ps = ppe_batch_build(pb, solvers.data(), pdev.data(), host.data(), popt.data(),
                                             (int)solvers.size(), block);
            if (ps != MPSR_NOERR) return fail(ps, "batched step: PPE solver tables");
        }
        sig = sg;
        return MPSR_NOERR;
    }

    void bin(hipStream_t st) {
        const BatchLaunch<R> b = launch();
        fill_int_batched(fill_bin, fill_bin_w, st);
        launch_bin_hash_b(b, wp, st);
        for (const Pass& q : psort) radix_pass_batched(items, q.jobs, q.tiles, q.singles, st);
        launch_bin_ranges_b(b, wp, st);
    }
    void tri(hipStream_t st) {
        const BatchLaunch<R> b = launch();
        fill_int_batched(fill_tri, fill_tri_w, st);
        launch_tri_count_b(b, wt64, st);
        exclusive_scan_batched(scan_emit, scan_emit_t, scan_emit_1, st);
        launch_tri_fill_b(b, wt64, st);
        for (const Pass& q : tsort) radix_pass_batched(8, q.jobs, q.tiles, q.singles, st);
        launch_pairs_b(b, 0, wpairs, st);
        exclusive_scan_batched(scan_flag, scan_flag_t, scan_flag_1, st);
        launch_pairs_b(b, 1, wpairs, st);
        launch_pairs_b(b, 2, wpairs, st);
        launch_cell_walks_b(b, wc256, walk_items, walk_n, st);
    }
    // one step of every bucket in the order of Engine::one_step, the rotation row of each bucket
    // from its device step counter, which the step then advances.  head: up to the pressure (with
    // the PPE buckets' assembly, start and `iters` PCG iterations); tail: from the PPE write-back on.
    void head(hipStream_t st, int iters) {
        const BatchLaunch<R> b = launch();
        bin(st);
        tri(st);
        launch_stl_distance_b(b, wp, st);
        launch_drag_b(b, wdrag, st);
        launch_thermal_b(b, wtherm, st);
        launch_turbulence_b(b, wturb, st);
        launch_explicit_b(b, wp, st);
        launch_relocate_b(b, wp, st);
        bin(st);
        launch_stl_distance_b(b, wp, st);
        launch_pnd_b(b, wp_eos, st);
        if (pb) {
            ppe_batch_pnd(pb, b, st);
            ppe_batch_head(pb, b, iters, st);
        }
    }
    void iterate(hipStream_t st, int iters) {
        if (pb) ppe_batch_iterate(pb, iters, st);
    }
    void tail(hipStream_t st) {
        const BatchLaunch<R> b = launch();
        if (pb) ppe_batch_tail(pb, b, st);
        launch_gradient_b(b, wp, st);
        launch_relocate_b(b, wp, st);
        launch_translate_b(b, wt256, st);
        launch_rotate_b(b, wt256, st);
        launch_step_advance_b(b, wone, st);
    }
};

}  // namespace mpsr

using namespace mpsr;

struct mpsr_handle {
    EngineBase* e;
    bool borrowed = false;   // owned by a batch: mpsr_destroy leaves it alone
};

// Many independent domains ("buckets") advanced together: each keeps its own engine (grid, walls,
// precision, buffers); a step of all of them is the batched step of each precision group (every
// kernel one launch over all buckets of that precision), recorded once into a HIP graph, and each
// step is then one graph launch.  The rotation row of a replay comes from each bucket's device step
// counter.
struct mpsr_batch {
    static constexpr int kHeadIters = 32;   // PCG iterations of the PPE buckets in the step's first graph
    static constexpr int kChunkIters = 16;  // ... and in each further chunk until all have stopped
    std::vector<mpsr_handle*> h;
    std::vector<int> order;            // largest bucket first
    int device = 0;
    hipStream_t main = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    // without PPE buckets the whole step is graph[0]; with them: head, PCG chunk, tail
    hipGraph_t graph[3] = {};
    hipGraphExec_t exec[3] = {};
    bool dirty = true;
    int block = 256, items = 8;
    size_t nodes = 0;
    float ms_per_step = 0.f;
    BatchGroup<float> g32;
    BatchGroup<double> g64;
    bool ppe() const { return g32.has_ppe() || g64.has_ppe(); }
    void drop_graph() {
        for (int i = 0; i < 3; ++i) {
            if (exec[i]) (void)hipGraphExecDestroy(exec[i]);
            if (graph[i]) (void)hipGraphDestroy(graph[i]);
            exec[i] = nullptr;
            graph[i] = nullptr;
        }
        dirty = true;
    }
    // part 0: head (the whole step without PPE buckets), 1: a PCG chunk, 2: tail
    void enqueue(int part) {
        const bool p = ppe();
        if (part == 0) {
            if (!g32.eng.empty()) g32.head(main, kHeadIters);
            if (!g64.eng.empty()) g64.head(main, kHeadIters);
            if (!p) {
                if (!g32.eng.empty()) g32.tail(main);
                if (!g64.eng.empty()) g64.tail(main);
            }
        } else if (part == 1) {
            g32.iterate(main, kChunkIters);
            g64.iterate(main, kChunkIters);
        } else {
            if (!g32.eng.empty()) g32.tail(main);
            if (!g64.eng.empty()) g64.tail(main);
        }
    }
    ~mpsr_batch() {
        drop_graph();
        for (mpsr_handle* x : h) {
            delete x->e;
            delete x;
        }
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (main) (void)hipStreamDestroy(main);
    }
};

namespace {
// records the step's graphs (one, or head / PCG chunk / tail with PPE buckets); on any failure the
// capture is ended and discarded so that no stream is left capturing
mpsr_status batch_capture(mpsr_batch* b) {
    b->drop_graph();
    HIPC(hipSetDevice(b->device));
    b->nodes = 0;
    const int parts = b->ppe() ? 3 : 1;
    for (int i = 0; i < parts; ++i) {
        HIPC(hipStreamBeginCapture(b->main, hipStreamCaptureModeThreadLocal));
        b->enqueue(i);
        const hipError_t he = hipGetLastError();
        hipGraph_t g = nullptr;
        const hipError_t ee = hipStreamEndCapture(b->main, &g);
        if (he != hipSuccess || ee != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            (void)hipGetLastError();
            b->drop_graph();
            return fail(MPSR_GPUERR, std::string("batched step capture: ") + hipGetErrorString(he != hipSuccess ? he : ee));
        }
        b->graph[i] = g;
        HIPC(hipGraphInstantiate(&b->exec[i], g, nullptr, nullptr, 0));
        size_t n = 0;
        HIPC(hipGraphGetNodes(g, nullptr, &n));
        b->nodes += n;
    }
    b->dirty = false;
    return MPSR_NOERR;
}

// rotation tables and step counters of every bucket, the device tables of both groups, and the
// work lists when the launch shape or an optional phase changed (the graph then records anew)
mpsr_status batch_prepare_all(mpsr_batch* b, int first, int count) {
    for (mpsr_handle* x : b->h) {
        mpsr_status s = x->e->batch_prepare(first, count);
        if (s != MPSR_NOERR) return s;
    }
    const long long s32 = b->g32.sig, s64 = b->g64.sig;
    mpsr_status s = MPSR_NOERR;
    if (!b->g32.eng.empty()) {
        s = b->g32.refresh();
        if (s == MPSR_NOERR) s = b->g32.build(b->block, b->items);
    }
    if (s == MPSR_NOERR && !b->g64.eng.empty()) {
        s = b->g64.refresh();
        if (s == MPSR_NOERR) s = b->g64.build(b->block, b->items);
    }
    if (s != MPSR_NOERR) return s;
    if (b->g32.sig != s32 || b->g64.sig != s64) b->dirty = true;
    return MPSR_NOERR;
}

// the PPE buckets' last solves: a failed one (breakdown, not converged, matrix too wide) is the
// call's error, naming the bucket
template <typename R> mpsr_status group_ppe_finish(mpsr_batch* b, BatchGroup<R>& g) {
    if (!g.pb) return MPSR_NOERR;
    int failed = -1;
    std::string why;
    mpsr_status s = ppe_batch_finish(g.pb, &failed, &why);
    if (s != MPSR_NOERR) return fail(s, "batched step: reading the PPE state");
    if (failed < 0) return MPSR_NOERR;
    int k = -1;
    for (int i = 0, seen = -1; i < (int)g.eng.size(); ++i)
        if (g.eng[i]->pressure_mode == MPSR_PRESSURE_PPE && ++seen == failed) k = i;
    int bucket = -1;
    for (size_t i = 0; i < b->h.size(); ++i)
        if (k >= 0 && b->h[i]->e == g.eng[k]) bucket = (int)i;
    return fail(MPSR_CALCERR, "batched step: bucket " + std::to_string(bucket) + ": PPE solve failed: " + why);
}
mpsr_status batch_ppe_finish(mpsr_batch* b) {
    mpsr_status s = group_ppe_finish(b, b->g32);
    if (s == MPSR_NOERR) s = group_ppe_finish(b, b->g64);
    return s;
}

mpsr_status batch_finish_all(mpsr_batch* b, int count) {
    HIPC(hipEventRecord(b->e1, b->main));
    HIPC(hipEventSynchronize(b->e1));
    float ms = 0.f;
    HIPC(hipEventElapsedTime(&ms, b->e0, b->e1));
    b->ms_per_step = ms / count;
    const mpsr_status ps = batch_ppe_finish(b);
    const std::string why = g_err;
    for (mpsr_handle* x : b->h) {
        mpsr_status s = x->e->batch_finish();
        if (s != MPSR_NOERR) return s;
    }
    return ps == MPSR_NOERR ? MPSR_NOERR : fail(ps, why);
}

mpsr_status batch_run(mpsr_batch* b, int first, int count, bool graph) {
    HIPC(hipSetDevice(b->device));
    if (count <= 0) return MPSR_NOERR;
    mpsr_status s = batch_prepare_all(b, first, count);
    if (s != MPSR_NOERR) return s;
    if (graph && (b->dirty || !b->exec)) {
        s = batch_capture(b);
        if (s != MPSR_NOERR) return s;
    }
    HIPC(hipEventRecord(b->e0, b->main));
    auto part = [&](int i) -> mpsr_status {
        if (graph) HIPC(hipGraphLaunch(b->exec[i], b->main));
        else b->enqueue(i);
        return MPSR_NOERR;
    };
    // the PCG chunks a step may need: every bucket stops at its max_iter at the latest
    int max_chunks = 2;
    for (mpsr_handle* x : b->h) max_chunks = std::max(max_chunks, x->e->ppe_max_iter() / mpsr_batch::kChunkIters + 2);
    for (int k = 0; k < count; ++k) {
        s = part(0);
        if (s != MPSR_NOERR) return s;
        if (!b->ppe()) continue;
        for (int c = 0; c <= max_chunks; ++c) {
            bool a32 = true, a64 = true;
            if (b->g32.pb) s = ppe_batch_converged(b->g32.pb, b->main, &a32);
            if (s == MPSR_NOERR && b->g64.pb) s = ppe_batch_converged(b->g64.pb, b->main, &a64);
            if (s != MPSR_NOERR) return fail(s, "batched step: reading the PCG flags");
            if (a32 && a64) break;
            s = part(1);
            if (s != MPSR_NOERR) return s;
        }
        s = part(2);
        if (s != MPSR_NOERR) return s;
    }
    HIPC(hipGetLastError());
    return batch_finish_all(b, count);
}

mpsr_status batch_set_launch(mpsr_batch* b, int block, int items) {
    for (mpsr_handle* x : b->h) {
        mpsr_status s = x->e->set_launch(block, items);
        if (s != MPSR_NOERR) return s;
    }
    b->block = block;
    b->items = items;
    b->drop_graph();
    return MPSR_NOERR;
}
}  // namespace

#define GUARD(h)                                                       \
    do {                                                               \
        if (!(h) || !(h)->e) return fail(MPSR_PARAMERR, "null handle"); \
    } while (0)

extern "C" {

const char* mpsr_last_error(void) { return g_err.c_str(); }

mpsr_status mpsr_create(const mpsr_create_info* info, mpsr_handle** out) {
    if (!info || !out) return fail(MPSR_PARAMERR, "null argument");
    *out = nullptr;
    EngineBase* e = nullptr;
    mpsr_status s;
    if (info->precision == MPSR_FP64) {
        auto* en = new Engine<double>();
        s = en->init(info);
        e = en;
    } else if (info->precision == MPSR_FP32) {
        auto* en = new Engine<float>();
        s = en->init(info);
        e = en;
    } else {
        return fail(MPSR_PARAMERR, "unknown precision");
    }
    if (s != MPSR_NOERR) {
        delete e;
        return s;
    }
    *out = new mpsr_handle{e};
    return MPSR_NOERR;
}
void mpsr_destroy(mpsr_handle* h) {
    if (!h || h->borrowed) return;
    delete h->e;
    delete h;
}
mpsr_status mpsr_step(mpsr_handle* h, int i) { GUARD(h); return h->e->step_many(i, 1); }
mpsr_status mpsr_set_launch(mpsr_handle* h, int block, int sort_items) { GUARD(h); return h->e->set_launch(block, sort_items); }
mpsr_status mpsr_set_radix_terms(mpsr_handle* h, const mpsr_radix_terms* t, const mpsr_negative_pressure_model* m, int nm) {
    GUARD(h);
    return h->e->set_radix(t, m, nm);
}
mpsr_status mpsr_calc_turbulence_viscosity(mpsr_handle* h) { GUARD(h); return h->e->phase(13, 0); }
mpsr_status mpsr_get_turbulence(mpsr_handle* h, double* nu, double* k, int n) {
    GUARD(h);
    return h->e->get_turbulence(nu, k, n);
}

mpsr_status mpsr_create_batch(const mpsr_create_info* infos, int nbuckets, mpsr_batch** out) {
    if (!infos || !out || nbuckets <= 0) return fail(MPSR_PARAMERR, "mpsr_create_batch: no buckets");
    *out = nullptr;
    auto* b = new mpsr_batch();
    b->device = infos[0].device;
    for (int i = 0; i < nbuckets; ++i) {
        if (infos[i].device != b->device) {
            delete b;
            return fail(MPSR_PARAMERR, "mpsr_create_batch: every bucket must be on the batch's device");
        }
        mpsr_handle* h = nullptr;
        mpsr_status s = mpsr_create(&infos[i], &h);
        if (s == MPSR_NOERR) {
            h->borrowed = true;
            b->h.push_back(h);
            s = h->e->batch_check();
        }
        if (s != MPSR_NOERR) {
            const std::string why = g_err;
            delete b;
            return fail(s, "bucket " + std::to_string(i) + ": " + why);
        }
    }
    if (hipSetDevice(b->device) != hipSuccess || hipStreamCreateWithFlags(&b->main, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&b->e0) != hipSuccess || hipEventCreate(&b->e1) != hipSuccess) {
        delete b;
        return fail(MPSR_GPUERR, "mpsr_create_batch: stream or event creation failed");
    }
    b->order.resize(b->h.size());
    for (size_t i = 0; i < b->h.size(); ++i) b->order[i] = (int)i;
    std::stable_sort(b->order.begin(), b->order.end(),
                     [&](int x, int y) { return b->h[x]->e->count() > b->h[y]->e->count(); });
    for (const int i : b->order) {
        if (auto* e32 = dynamic_cast<Engine<float>*>(b->h[i]->e)) b->g32.eng.push_back(e32);
        else if (auto* e64 = dynamic_cast<Engine<double>*>(b->h[i]->e)) b->g64.eng.push_back(e64);
    }
    *out = b;
    return MPSR_NOERR;
}
void mpsr_destroy_batch(mpsr_batch* b) { delete b; }
mpsr_status mpsr_batch_handle(mpsr_batch* b, int bucket, mpsr_handle** h) {
    if (!b || !h || bucket < 0 || bucket >= (int)b->h.size()) return fail(MPSR_PARAMERR, "mpsr_batch_handle: bad bucket");
    *h = b->h[bucket];
    return MPSR_NOERR;
}
mpsr_status mpsr_step_batch(mpsr_batch* b, int first, int count) {
    if (!b) return fail(MPSR_PARAMERR, "null batch");
    return batch_run(b, first, count, true);
}
mpsr_status mpsr_step_batch_eager(mpsr_batch* b, int first, int count) {
    if (!b) return fail(MPSR_PARAMERR, "null batch");
    return batch_run(b, first, count, false);
}
mpsr_status mpsr_batch_set_launch(mpsr_batch* b, int block, int sort_items) {
    if (!b) return fail(MPSR_PARAMERR, "null batch");
    return batch_set_launch(b, block, sort_items);
}
mpsr_status mpsr_batch_info(mpsr_batch* b, float* ms_per_step, int* graph_nodes, int* block, int* sort_items) {
    if (!b) return fail(MPSR_PARAMERR, "null batch");
    if (ms_per_step) *ms_per_step = b->ms_per_step;
    if (graph_nodes) *graph_nodes = (int)b->nodes;
    if (block) *block = b->block;
    if (sort_items) *sort_items = b->items;
    return MPSR_NOERR;
}
mpsr_status mpsr_batch_autotune(mpsr_batch* b, int first, int steps, int* best_block, int* best_items, float* ms_table) {
    if (!b || steps <= 0) return fail(MPSR_PARAMERR, "mpsr_batch_autotune: null batch or no steps");
    static const int blocks[MPSR_TUNE_BLOCKS] = {64, 128, 256, 512};
    static const int items[MPSR_TUNE_ITEMS] = {4, 8, 16};
    for (mpsr_handle* x : b->h) {
        mpsr_status s = x->e->snapshot(0);
        if (s != MPSR_NOERR) return s;
    }
    float best = -1.f;
    int bb = b->block, bi = b->items;
    for (int i = 0; i < MPSR_TUNE_BLOCKS; ++i)
        for (int j = 0; j < MPSR_TUNE_ITEMS; ++j) {
            mpsr_status s = batch_set_launch(b, blocks[i], items[j]);
            if (s == MPSR_NOERR) s = batch_run(b, first, steps, true);   // records the graph, warms up
            for (mpsr_handle* x : b->h)
                if (s == MPSR_NOERR) s = x->e->snapshot(1);
            if (s == MPSR_NOERR) s = batch_run(b, first, steps, true);
            for (mpsr_handle* x : b->h)
                if (s == MPSR_NOERR) s = x->e->snapshot(1);
            if (s != MPSR_NOERR) return s;
            if (ms_table) ms_table[i * MPSR_TUNE_ITEMS + j] = b->ms_per_step;
            if (best < 0.f || b->ms_per_step < best) {
                best = b->ms_per_step;
                bb = blocks[i];
                bi = items[j];
            }
        }
    if (best_block) *best_block = bb;
    if (best_items) *best_items = bi;
    return batch_set_launch(b, bb, bi);
}
mpsr_status mpsr_step_many(mpsr_handle* h, int first, int count) { GUARD(h); return h->e->step_many(first, count); }
mpsr_status mpsr_synchronize(mpsr_handle* h) { GUARD(h); return h->e->sync(); }
mpsr_status mpsr_register_particle_topology(mpsr_handle* h) { GUARD(h); return h->e->phase(0, 0); }
mpsr_status mpsr_register_triangle_topology(mpsr_handle* h) { GUARD(h); return h->e->phase(1, 0); }
mpsr_status mpsr_calc_stl_distance(mpsr_handle* h) { GUARD(h); return h->e->phase(2, 0); }
mpsr_status mpsr_calc_drag_effect(mpsr_handle* h) { GUARD(h); return h->e->phase(3, 0); }
mpsr_status mpsr_calc_explicitly(mpsr_handle* h) { GUARD(h); return h->e->phase(4, 0); }
mpsr_status mpsr_relocate_particle_data(mpsr_handle* h) { GUARD(h); return h->e->phase(5, 0); }
mpsr_status mpsr_calc_pressure_explicit(mpsr_handle* h) { GUARD(h); return h->e->phase(6, 0); }
mpsr_status mpsr_calc_pressure_poisson(mpsr_handle* h) { GUARD(h); return h->e->phase(7, 0); }
mpsr_status mpsr_calc_implicit_explicitly(mpsr_handle* h) { GUARD(h); return h->e->phase(8, 0); }
mpsr_status mpsr_update_triangle_position(mpsr_handle* h) { GUARD(h); return h->e->phase(9, 0); }
mpsr_status mpsr_calc_temperature_factor(mpsr_handle* h) { GUARD(h); return h->e->phase(12, 0); }
mpsr_status mpsr_rotate_triangle_position(mpsr_handle* h, int step) { GUARD(h); return h->e->phase(10, step); }
mpsr_status mpsr_get_particles(mpsr_handle* h, int which, mpsr_particles* o) {
    GUARD(h);
    if (!o) return fail(MPSR_PARAMERR, "null output");
    return h->e->get_particles(which, o);
}
mpsr_status mpsr_set_particles(mpsr_handle* h, const mpsr_particles* p) { GUARD(h); return h->e->set_particles(p); }
mpsr_status mpsr_get_particle_count(mpsr_handle* h, int* n) {
    GUARD(h);
    if (!n) return MPSR_PARAMERR;
    *n = h->e->count();
    return MPSR_NOERR;
}
mpsr_status mpsr_set_buckets(mpsr_handle* h, const mpsr_particles* b, const int* interval, const int* total, int count) {
    GUARD(h);
    return h->e->set_buckets(b, interval, total, count);
}
mpsr_status mpsr_bucket_addition(mpsr_handle* h, int step, int* added) { GUARD(h); return h->e->bucket_phase(step, added); }
mpsr_status mpsr_set_spheres(mpsr_handle* h, const mpsr_spheres* s) { GUARD(h); return h->e->set_spheres(s); }
mpsr_status mpsr_get_spheres(mpsr_handle* h, mpsr_spheres* s) { GUARD(h); return h->e->get_spheres(s); }
mpsr_status mpsr_calc_spheres(mpsr_handle* h) { GUARD(h); return h->e->phase(11, 0); }
int mpsr_mpnn_param_count(void) { return mpsr::mpnn_param_count(); }
mpsr_status mpsr_mpnn_set_weights(mpsr_handle* h, const float* w, int n) {
    GUARD(h);
    if (!w || n != mpsr::mpnn_param_count()) return MPSR_PARAMERR;
    return h->e->mpnn_set(w);
}
mpsr_status mpsr_mpnn_get_weights(mpsr_handle* h, float* w, int n) {
    GUARD(h);
    if (!w || n != mpsr::mpnn_param_count()) return MPSR_PARAMERR;
    return h->e->mpnn_get(w);
}
mpsr_status mpsr_mpnn_train_step(mpsr_handle* h, double lr, double beta, int update, float* grads, double* parts) {
    GUARD(h);
    return h->e->mpnn_train(lr, beta, update, grads, parts);
}
mpsr_status mpsr_mpnn_rows(mpsr_handle* h, float* X, float* Y, double* scale) {
    GUARD(h);
    return h->e->mpnn_rows(X, Y, scale);
}
mpsr_status mpsr_get_triangles(mpsr_handle* h, mpsr_triangle* o, int n) { GUARD(h); return h->e->get_triangles(o, n); }
mpsr_status mpsr_get_gridbox(mpsr_handle* h, mpsr_gridbox* o) {
    GUARD(h);
    *o = h->e->gb;
    return MPSR_NOERR;
}
mpsr_status mpsr_set_dt(mpsr_handle* h, float dt, float courant) { GUARD(h); return h->e->set_dt(dt, courant); }
mpsr_status mpsr_get_timings(mpsr_handle* h, mpsr_timings* o) {
    GUARD(h);
    *o = h->e->tm;
    return MPSR_NOERR;
}
mpsr_status mpsr_set_ppe_options(mpsr_handle* h, const mpsr_ppe_options* o) { GUARD(h); return h->e->set_ppe(o); }
mpsr_status mpsr_set_profiling(mpsr_handle* h, int on) {
    GUARD(h);
    if (on < 0 || on > 2) return MPSR_PARAMERR;
    h->e->profiling = on;
    return MPSR_NOERR;
}
mpsr_status mpsr_get_binning(mpsr_handle* h, int* k, int* p, int* cs, int* ce) { GUARD(h); return h->e->get_binning(k, p, cs, ce); }
mpsr_status mpsr_get_triangle_cells(mpsr_handle* h, int* cs, int* cc, int* ids, int cap, int* total) {
    GUARD(h);
    return h->e->get_triangle_cells(cs, cc, ids, cap, total);
}
mpsr_status mpsr_get_stl_distance(mpsr_handle* h, double* m, double* dir, int* id) { GUARD(h); return h->e->get_stl(m, dir, id); }
mpsr_status mpsr_get_neighbors(mpsr_handle* h, double r, int* off, int* nbr, int cap, int* total) {
    GUARD(h);
    return h->e->get_neighbors(r, off, nbr, cap, total);
}
mpsr_status mpsr_get_ppe(mpsr_handle* h, double* rhs, double* x, int* dir) { GUARD(h); return h->e->get_ppe(rhs, x, dir); }
mpsr_status mpsr_get_ppe_matrix(mpsr_handle* h, int* ptr, int* col, double* val, double* diag, int cap, int* nnz) {
    GUARD(h);
    return h->e->get_ppe_matrix(ptr, col, val, diag, cap, nnz);
}

}  // extern "C"
