// deck.cpp -- the reference's input-deck surface and output frames (host only).
//
//   mpsr_deck_load     CConfiguration::ReadConfigurationFileWithDrag (Src Configuration.cpp:21-421),
//                      IO::Load of the parameter file (Src IO.cpp:66-432), of the particle and bucket
//                      files (IO.cpp:7-64), of the STL models (IO.cpp:456-509 with
//                      solver/Common/inc/STLSerializer.cpp:226-351 and CalcTriangleNormal, IO.cpp:435-453),
//                      and the model set-up of CMultiGPUHandler::Run (Src MultiGPUHandler.cpp:640-786)
//   mpsr_write_frame   IO::SaveAscii / IO::SaveBinary (IO.cpp:635-742)
//   mpsr_dt_schedule_load  DtFrame::Load (Src DtFrame.cpp:11-35), the time-step schedule file
//
// Paths: the reference joins the configuration's input directory and file names with '\'.  Here
// the directory is tried first; when it does not exist (decks written on another machine), the
// files are looked up next to the configuration file and in its IN/ subdirectory, and a file name
// that differs only in letter case is accepted (the decks were written on a case-insensitive file
// system).
#include <algorithm>
#include <array>
#include <cctype>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <dirent.h>
#include <fstream>
#include <sstream>
#include <string>
#include <sys/stat.h>
#include <vector>
#include "../../include/mpsr.h"

namespace {

thread_local std::string g_deck_err;

mpsr_status deck_fail(mpsr_status s, const std::string& m) {
    g_deck_err = m;
    return s;
}

float sqrt_new(float x) {  // solver/Common/inc/STLSerializer.h:8-21 (same scheme as sqrt14)
    const float xhalf = 0.5f * x;
    int32_t i;
    std::memcpy(&i, &x, 4);
    i = 0x5f3759df - (i >> 1);
    float y;
    std::memcpy(&y, &i, 4);
    return x * y * (1.5f - xhalf * y * y);
}

std::string upper(std::string s) {
    for (char& c : s) c = (char)std::toupper((unsigned char)c);
    return s;
}
std::string trim(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) ++a;
    while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
    return s.substr(a, b - a);
}
bool is_dir(const std::string& p) {
    struct stat st;
    return stat(p.c_str(), &st) == 0 && S_ISDIR(st.st_mode);
}
bool is_file(const std::string& p) {
    struct stat st;
    return stat(p.c_str(), &st) == 0 && S_ISREG(st.st_mode);
}
std::string dirname_of(const std::string& p) {
    const size_t k = p.find_last_of("/\\");
    return k == std::string::npos ? std::string(".") : p.substr(0, k);
}
// file `name` in `dir`, exact or differing only in letter case
std::string find_in(const std::string& dir, const std::string& name) {
    if (is_file(dir + "/" + name)) return dir + "/" + name;
    DIR* d = opendir(dir.c_str());
    if (!d) return "";
    std::string hit;
    const std::string want = upper(name);
    while (dirent* e = readdir(d)) {
        if (upper(e->d_name) == want) {
            hit = dir + "/" + e->d_name;
            break;
        }
    }
    closedir(d);
    return hit;
}

struct Triangle3 {
    float n[3];
    float v[3][3];
};

// ASCII STL: solid ... facet normal nx ny nz / outer loop / vertex x3 / endloop / endfacet
bool load_stl_ascii(const std::string& path, std::vector<Triangle3>& out) {
    std::ifstream in(path);
    if (!in) return false;
    std::string tok;
    bool solid = false;
    while (in >> tok) {
        const std::string u = upper(tok);
        if (u == "SOLID") {
            solid = true;
            continue;
        }
        if (u != "FACET") continue;
        Triangle3 t{};
        std::string w;
        in >> w;
        if (upper(w) != "NORMAL") return false;
        in >> t.n[0] >> t.n[1] >> t.n[2];
        in >> w;
        if (upper(w) != "OUTER") return false;
        in >> w;
        if (upper(w) != "LOOP") return false;
        for (int k = 0; k < 3; ++k) {
            in >> w;
            if (upper(w) != "VERTEX") return false;
            in >> t.v[k][0] >> t.v[k][1] >> t.v[k][2];
        }
        in >> w;
        if (upper(w) != "ENDLOOP") return false;
        in >> w;
        if (upper(w) != "ENDFACET") return false;
        if (!in) return false;
        out.push_back(t);
    }
    return solid;
}

// binary STL: 80-byte header, uint32 count, per facet 12 floats and a 2-byte attribute
bool load_stl_binary(const std::string& path, std::vector<Triangle3>& out) {
    std::ifstream in(path, std::ios::binary);
    if (!in) return false;
    char header[80];
    in.read(header, 80);
    uint32_t n = 0;
    in.read((char*)&n, 4);
    if (!in) return false;
    for (uint32_t i = 0; i < n; ++i) {
        float f[12];
        char attr[2];
        in.read((char*)f, sizeof(f));
        in.read(attr, 2);
        if (!in) return false;
        Triangle3 t{};
        std::memcpy(t.n, f, 12);
        std::memcpy(t.v, f + 3, 36);
        out.push_back(t);
    }
    return true;
}

// IO::Load particle file: 12 tokens per row; rows that fail to parse are dropped
bool load_particles(const std::string& path, std::vector<std::array<double, 12>>& rows) {
    std::ifstream in(path);
    if (!in) return false;
    while (true) {
        float x[11];
        int type;
        if (!(in >> x[0] >> x[1] >> x[2] >> x[3] >> x[4] >> x[5] >> x[6] >> x[7] >> x[8] >> x[9] >> x[10] >> type))
            break;
        std::array<double, 12> r;
        for (int k = 0; k < 11; ++k) r[k] = x[k];
        r[11] = (type == 1) ? 1 : (type == 2 ? 2 : 0);
        rows.push_back(r);
    }
    return true;
}

// one word of the positional parameter file with the reference's eof guards
struct ParamReader {
    std::ifstream in;
    bool eof() { return !(in >> std::ws) || in.eof(); }
    template <typename T> void get(T& v) { in >> v; }
    int onoff() {
        std::string s;
        in >> s;
        return s == "ON" ? 1 : 0;
    }
};

}  // namespace

struct mpsr_deck {
    mpsr_params params{};
    std::vector<std::array<double, 12>> rows;
    struct Bucket {
        std::vector<std::array<double, 12>> rows;
        int interval = 0, total = 0;
    };
    std::vector<Bucket> buckets;
    std::vector<int> ids, types;
    std::vector<double> pos, vel, pres, dens, temp, nu, solid;
    mpsr_particles particles{};
    std::vector<mpsr_triangle> tris;
    std::vector<mpsr_tri_params> models;
    std::vector<mpsr_drag_params> drag;
    std::vector<mpsr_float3> model_offset;
    mpsr_box zone{};
    int has_box = 0;
    int start_step = 0, output_interval = 1, ascii = 1;
    std::string output_dir, output_name;
};

extern "C" {

const char* mpsr_deck_last_error(void) { return g_deck_err.c_str(); }

mpsr_status mpsr_deck_load(const char* config_path, mpsr_deck** out) {
    if (!config_path || !out) return deck_fail(MPSR_PARAMERR, "null argument");
    *out = nullptr;
    std::ifstream cf(config_path);
    if (!cf) return deck_fail(MPSR_FILEERR, std::string("cannot open configuration ") + config_path);
    auto* D = new mpsr_deck();
    auto bad = [&](mpsr_status s, const std::string& m) {
        delete D;
        return deck_fail(s, m);
    };
    std::string input_dir, param_file, particle_file, line;
    std::getline(cf, input_dir);
    std::getline(cf, param_file);
    std::getline(cf, particle_file);
    input_dir = trim(input_dir);
    param_file = trim(param_file);
    particle_file = trim(particle_file);
    std::vector<std::string> bucket_files;
    std::vector<int> bucket_int, bucket_tot;
    while (std::getline(cf, line) && !trim(line).empty()) {  // bucket lines (Configuration.cpp:45-61)
        std::istringstream ls(line);
        std::string f;
        int a = 0, b = 0;
        ls >> f >> a >> b;
        bucket_files.push_back(f);
        bucket_int.push_back(a);
        bucket_tot.push_back(b);
    }
    // STL lines (Configuration.cpp:65-336)
    std::vector<std::string> stl_files;
    int drag_count = 0;
    while (std::getline(cf, line) && !trim(line).empty()) {
        std::istringstream ls(line);
        std::string f, ifdrag;
        ls >> f >> ifdrag;
        stl_files.push_back(f);
        ifdrag = upper(ifdrag);
        mpsr_tri_params m{};
        mpsr_float3 off{0, 0, 0};
        if (ifdrag == "D" || ifdrag == "DE" || ifdrag == "DC" || ifdrag == "DF") {
            mpsr_drag_params dp{};
            dp.DragID = drag_count;
            std::string tin, tout;
            ls >> tin >> tout;
            tin = upper(tin);
            ls >> dp.CoefficientofInflueneRegion;
            if (ifdrag == "D" || ifdrag == "DE") {
                dp.DragToken = 'E';
                if (tin == "V") {
                    dp.InputVelocityToken1 = dp.OutputVelocityToken2 = 'V';
                    ls >> dp.VelocityIndex >> dp.VelocityExponential >> dp.VelocityDragPermiabilityConstant;
                } else if (tin == "T") {
                    dp.InputVelocityToken1 = dp.OutputVelocityToken2 = 'T';
                    ls >> dp.TempIndex >> dp.TempExponential >> dp.TempDragPermiabilityConstant;
                } else {  // "VT", "TV" and (the reference's always-true else-if) anything else
                    dp.InputVelocityToken1 = dp.OutputVelocityToken2 = 'B';
                    ls >> dp.VelocityIndex >> dp.VelocityExponential >> dp.VelocityDragPermiabilityConstant;
                    ls >> dp.TempIndex >> dp.TempExponential >> dp.TempDragPermiabilityConstant;
                }
            }
            if (ifdrag == "DF") {
                dp.DragToken = 'F';
                if (tin == "V") {
                    dp.InputVelocityToken1 = dp.OutputVelocityToken2 = 'V';
                    ls >> dp.VelocityMagnitueFactor;
                }
            } else if (ifdrag == "DC") {
                dp.InputVelocityToken1 = dp.OutputVelocityToken2 = 'V';
                dp.DragToken = 'C';
                ls >> dp.ConstantVx >> dp.ConstantVy >> dp.ConstantVz;
            }
            D->drag.push_back(dp);
            m.Temperature = 533;
            m.isDrag = 1;
            m.DragTriangleID = drag_count++;
        } else {
            // velocity, temperature, offset, linear reset, rotation, friction, thermal
            m.Velocity.x = (float)std::atof(ifdrag.c_str());
            ls >> m.Velocity.y >> m.Velocity.z >> m.Temperature;
            ls >> off.x >> off.y >> off.z;
            bool b1 = false, b2 = false;
            int rpm = 0;
            if (ls >> b1) m.LinearMovableOrNot = b1;
            ls >> m.LinearResetInterval;
            if (ls >> b2) m.RotatingOrNot = b2;
            if (ls >> rpm) m.AngleOfRotationInDegree = (float)rpm;  // converted to degrees below
            ls >> m.RotationStartTime >> m.TimeToReachRPM;
            ls >> m.CenterOfRotation.x >> m.CenterOfRotation.y >> m.CenterOfRotation.z;
            ls >> m.SecondPointOfRotation.x >> m.SecondPointOfRotation.y >> m.SecondPointOfRotation.z;
            ls >> m.WallFrictionCoefficient >> m.WallThermalResistivity >> m.WallThermalConductivity;
            m.isDrag = 0;
            m.DragTriangleID = -1;
        }
        D->models.push_back(m);
        D->model_offset.push_back(off);
    }
    std::getline(cf, D->output_dir);
    std::getline(cf, D->output_name);
    D->output_dir = trim(D->output_dir);
    D->output_name = trim(D->output_name);
    if (std::getline(cf, line)) std::istringstream(line) >> D->output_interval;
    if (std::getline(cf, line)) {
        std::istringstream ls(line);
        mpsr_box b{};
        ls >> b.m_MinBound.x >> b.m_MinBound.y >> b.m_MinBound.z >> b.m_MaxBound.x >> b.m_MaxBound.y >> b.m_MaxBound.z;
        const bool pass = std::fabs(b.m_MaxBound.x - b.m_MinBound.x) != 0.0 &&
                          std::fabs(b.m_MaxBound.y - b.m_MinBound.y) != 0.0 &&
                          std::fabs(b.m_MaxBound.z - b.m_MinBound.z) != 0.0;
        if (pass) {
            D->zone = b;
            D->has_box = 1;
        }
    }
    int sx = 1, sy = 1, sz = 1;
    if (std::getline(cf, line)) std::istringstream(line) >> sx >> sy >> sz;
    if ((long long)sx * sy * sz == 0) return bad(MPSR_PARAMERR, "section division zero");
    if (sx * sy * sz != 1) return bad(MPSR_PARAMERR, "more than one section: one GPU runs the whole domain");
    if (std::getline(cf, line)) std::istringstream(line) >> D->start_step;
    if (std::getline(cf, line)) {
        std::string a;
        std::istringstream(line) >> a;
        D->ascii = upper(a) == "BINARY" ? 0 : 1;
    }
    // resolve the input directory
    const std::string cdir = dirname_of(config_path);
    std::string indir = input_dir;
    std::replace(indir.begin(), indir.end(), '\\', '/');
    auto resolve = [&](const std::string& name) -> std::string {
        std::string n = name;
        std::replace(n.begin(), n.end(), '\\', '/');
        for (const std::string& d : {indir, cdir + "/IN", cdir}) {
            if (!d.empty() && is_dir(d)) {
                const std::string p = find_in(d, n);
                if (!p.empty()) return p;
            }
        }
        return "";
    };
    // parameter file (IO.cpp:66-432)
    const std::string ppath = resolve(param_file);
    if (ppath.empty()) return bad(MPSR_FILEERR, "parameter file not found: " + param_file);
    ParamReader R;
    R.in.open(ppath);
    if (!R.in) return bad(MPSR_FILEERR, "cannot open " + ppath);
    mpsr_params& P = D->params;
    R.get(P.Dimension);
    R.get(P.GravityConstant.x); R.get(P.GravityConstant.y); R.get(P.GravityConstant.z);
    R.get(P.Density); R.get(P.ConstantParticleDensity); R.get(P.SurfaceTensionCoefficient);
    R.get(P.InitialParticleDistance); R.get(P.FreeSurfaceThreshold);
    R.get(P.GradientInfluenceRadiusCoefficient); R.get(P.LaplacianInfluenceRadiusCoefficient);
    R.get(P.TemperatureInfluenceRadiusCoefficient);
    R.get(P.FruidInitializeialTemperature); R.get(P.WallInitializeialTemperature);
    R.get(P.FruidThermalResistance); R.get(P.WallThermalResistance);
    R.get(P.FruidThermalConductivity); R.get(P.WallThermalConductivity);
    R.get(P.FruidSpecificHeat); R.get(P.FruidDensity);
    R.get(P.Dt); R.get(P.AnalysisStep); R.get(P.ThermalStep); R.get(P.NegativeCalculationStep); R.get(P.OutputStep);
    R.get(P.LambdaID);
    R.get(P.SORJacobianRelaxationCoefficient); R.get(P.SORLoopMax); R.get(P.SORConvergenceConstant);
    P.bUpdateTemperature = R.onoff();
    if (!R.eof()) { R.get(P.ViscosityUpdateParameterA); R.get(P.ViscosityUpdateParameterB); R.get(P.MaxWallTemperature); R.get(P.MinWallTemperature); }
    if (!R.eof()) R.get(P.CourantNumber);
    if (!R.eof()) { R.get(P.InitialTimeForZeroAngle); R.get(P.FinalTimeToReachAngle); R.get(P.CellSize.z); }
    if (!R.eof()) { R.get(P.MvVelocity.x); R.get(P.MvVelocity.y); R.get(P.MvVelocity.z); }
    if (!R.eof()) R.get(P.OffSet);
    if (!R.eof()) { R.get(P.ParticleGenerationStep); R.get(P.TotalGenerationStep); }
    if (!R.eof()) R.get(P.Alpha);
    if (!R.eof()) R.get(P.AlphaLightWeight);
    if (!R.eof()) { P.bExplicitCollisonCheck = R.onoff(); P.bExplicitCollisionOut = R.onoff(); P.bImplicitCollisonCheck = R.onoff(); }
    if (!R.eof()) { R.get(P.CollisionCheckCoefficient); R.get(P.MinDistanceCoefficient); }
    P.Time1 = P.Dt;
    if (!R.eof()) {
        P.bFreeSurfacePressureUpdate = R.onoff();
        R.get(P.BasePressure); R.get(P.MaxPressure); R.get(P.Time1); R.get(P.TriggerStep); R.get(P.EndStep);
        R.get(P.AnalysisArea);
    }
    if (!R.eof()) P.bExplicitOut = R.onoff();
    if (!R.eof()) R.get(P.WallFrictionCoeffieicnt);
    if (!R.eof()) R.get(P.DragCoefficient);
    if (!R.eof()) P.bTwoPhase = R.onoff();
    P.StopWallStep = INT_MAX;
    if (!R.eof()) R.get(P.StopWallStep);
    if (!R.eof()) R.get(P.Tolerance);
    if (!R.eof()) R.get(P.ImplicitSurfaceThreshold);
    if (!R.eof()) P.bCGSolver = R.onoff();
    if (!R.eof()) P.bUpdateInnerPressure = R.onoff();
    if (!R.eof()) R.get(P.SurfaceInfluenceRadiusCoefficient);
    if (!R.eof()) R.get(P.SurfaceThreshold);
    if (!R.eof()) R.get(P.GradientForce);
    if (!R.eof()) R.get(P.StartStep);
    if (!R.eof()) R.get(P.InitialForce);
    if (!R.eof()) R.get(P.VelocityThreshold);
    if (!R.eof()) P.bExplicitPressure = R.onoff();
    if (!R.eof()) R.get(P.SonicVelocity);
    if (!R.eof()) R.get(P.ArtificialPressure);
    if (!R.eof()) R.get(P.r_coeff);
    if (!R.eof()) P.bImplicitVelocity = R.onoff();
    if (!R.eof()) P.bNJPCGSolver = R.onoff();
    P.PressureMax = FLT_MAX;
    if (!R.eof()) R.get(P.PressureMax);
    P.RLimitCoefficient = FLT_MAX;
    if (!R.eof()) R.get(P.RLimitCoefficient);
    P.SmagorinskyConstant = 0.1f;
    P.FilterWidth = P.InitialParticleDistance;
    if (!R.eof()) { P.bTurbulance = R.onoff(); R.get(P.SmagorinskyConstant); R.get(P.FilterWidth); }
    P.DragPermiabilityConstant = 1;
    if (!R.eof()) R.get(P.DragPermiabilityConstant);
    P.MaxNeighborhoodNum = P.Dimension == 2 ? 81 : 729;
    // STL models (IO.cpp:456-509): triangles of model i carry ModelIDNumber i
    for (size_t i = 0; i < stl_files.size(); ++i) {
        const std::string sp = resolve(stl_files[i]);
        if (sp.empty()) return bad(MPSR_FILEERR, "STL file not found: " + stl_files[i]);
        std::ifstream probe(sp);
        std::string first;
        probe >> first;
        std::vector<Triangle3> ts;
        const bool ok = first == "solid" ? load_stl_ascii(sp, ts) : load_stl_binary(sp, ts);
        if (!ok) return bad(MPSR_FILEERR, "STL load failed: " + sp);
        for (const Triangle3& t : ts) {
            mpsr_triangle T{};
            T.ModelIDNumber = (int)i;
            for (int k = 0; k < 3; ++k) T.Vertex[k] = {t.v[k][0], t.v[k][1], t.v[k][2]};
            T.Normal = {t.n[0], t.n[1], t.n[2]};
            // CalcTriangleNormal: (v1 - v0) x (v2 - v0) / sqrt_new(|.|^2)
            const float a0 = T.Vertex[1].x - T.Vertex[0].x, a1 = T.Vertex[1].y - T.Vertex[0].y, a2 = T.Vertex[1].z - T.Vertex[0].z;
            const float b0 = T.Vertex[2].x - T.Vertex[0].x, b1 = T.Vertex[2].y - T.Vertex[0].y, b2 = T.Vertex[2].z - T.Vertex[0].z;
            const float nx = a1 * b2 - b1 * a2, ny = a2 * b0 - b2 * a0, nz = a0 * b1 - b0 * a1;
            const float ab = sqrt_new(nx * nx + ny * ny + nz * nz);
            if (ab != 0.0f) T.Normal = {nx / ab, ny / ab, nz / ab};
            // SetModelPosition (MultiGPUHandler.cpp:871-889): vertices + the model's offset
            for (int k = 0; k < 3; ++k) {
                T.Vertex[k].x = T.Vertex[k].x + D->model_offset[i].x;
                T.Vertex[k].y = T.Vertex[k].y + D->model_offset[i].y;
                T.Vertex[k].z = T.Vertex[k].z + D->model_offset[i].z;
            }
            D->tris.push_back(T);
        }
    }
    // per-model rotation in degrees per step = RPM * Dt * 6 (MultiGPUHandler.cpp:690)
    for (mpsr_tri_params& m : D->models)
        if (!m.isDrag) m.AngleOfRotationInDegree = (int)m.AngleOfRotationInDegree * P.Dt * 6;
    // particles (and the bucket file, kept for inflow)
    const std::string pp = resolve(particle_file);
    if (pp.empty() || !load_particles(pp, D->rows)) return bad(MPSR_FILEERR, "particle file not found: " + particle_file);
    for (size_t j = 0; j < bucket_files.size(); ++j) {  // LoadMultiBucket (MultiGPUHandler.cpp:1068-1083)
        mpsr_deck::Bucket b;
        const std::string bp = resolve(bucket_files[j]);
        if (bp.empty() || !load_particles(bp, b.rows)) return bad(MPSR_FILEERR, "bucket file not found: " + bucket_files[j]);
        b.interval = bucket_int[j];
        b.total = bucket_tot[j];
        D->buckets.push_back(std::move(b));
    }
    const int n = (int)D->rows.size();
    D->ids.resize(n); D->types.resize(n);
    D->pos.resize(3 * n); D->vel.resize(3 * n);
    D->pres.resize(n); D->dens.resize(n); D->temp.resize(n); D->nu.resize(n); D->solid.resize(n);
    for (int s = 0; s < n; ++s) {
        const auto& r = D->rows[s];
        D->ids[s] = s;
        for (int a = 0; a < 3; ++a) { D->pos[3 * s + a] = r[a]; D->vel[3 * s + a] = r[3 + a]; }
        D->pres[s] = r[6]; D->dens[s] = r[7]; D->temp[s] = r[8]; D->nu[s] = r[9]; D->solid[s] = r[10];
        D->types[s] = (int)r[11];
    }
    D->particles = {n, D->ids.data(), D->pos.data(), D->vel.data(), D->pres.data(), D->dens.data(),
                    D->temp.data(), D->nu.data(), D->solid.data(), D->types.data()};
    P.LambdaValue = mpsr_calculate_lambda(P.InitialParticleDistance);
    if (!D->has_box)
        mpsr_auto_compute_zone(&P, &D->particles, D->tris.data(), (int)D->tris.size(), &D->zone);
    *out = D;
    return MPSR_NOERR;
}

void mpsr_deck_free(mpsr_deck* d) { delete d; }

mpsr_status mpsr_deck_create_info(const mpsr_deck* d, mpsr_create_info* info) {
    if (!d || !info) return deck_fail(MPSR_PARAMERR, "null argument");
    std::memset(info, 0, sizeof(*info));
    info->params = d->params;
    info->compute_zone = d->zone;
    // MaxParticleNum of the multi-bucket build (MultiGPUHandler.cpp:758-765)
    long long maxn = d->particles.n;
    for (const auto& b : d->buckets) maxn += (long long)b.total * (long long)b.rows.size();
    info->max_particles = (int)std::min<long long>(maxn, INT_MAX);
    info->particles = &d->particles;
    info->triangles = d->tris.data();
    info->num_triangles = (int)d->tris.size();
    info->models = d->models.data();
    info->num_models = (int)d->models.size();
    info->drag = d->drag.data();
    info->num_drag = (int)d->drag.size();
    info->precision = MPSR_FP32;
    info->pressure_mode = MPSR_PRESSURE_EOS;
    info->ppe.solver = MPSR_PCG_FP64;
    info->ppe.rel_tol = 1e-8;
    info->ppe.max_iter = 20000;
    return MPSR_NOERR;
}

int mpsr_deck_num_buckets(const mpsr_deck* d) { return d ? (int)d->buckets.size() : 0; }

mpsr_status mpsr_deck_bucket(const mpsr_deck* d, int j, int* count, int* interval, int* total) {
    if (!d || j < 0 || j >= (int)d->buckets.size()) return deck_fail(MPSR_PARAMERR, "bucket index out of range");
    if (count) *count = (int)d->buckets[j].rows.size();
    if (interval) *interval = d->buckets[j].interval;
    if (total) *total = d->buckets[j].total;
    return MPSR_NOERR;
}

mpsr_status mpsr_deck_get_bucket(const mpsr_deck* d, int j, mpsr_particles* out) {
    if (!d || !out || j < 0 || j >= (int)d->buckets.size()) return deck_fail(MPSR_PARAMERR, "bucket index out of range");
    const auto& rows = d->buckets[j].rows;
    if (out->n < (int)rows.size()) return deck_fail(MPSR_PARAMERR, "bucket output arrays too small");
    for (int s = 0; s < (int)rows.size(); ++s) {
        const auto& r = rows[s];
        if (out->id) out->id[s] = s;
        for (int a = 0; a < 3; ++a) {
            if (out->position) out->position[3 * s + a] = r[a];
            if (out->velocity) out->velocity[3 * s + a] = r[3 + a];
        }
        if (out->pressure) out->pressure[s] = r[6];
        if (out->density) out->density[s] = r[7];
        if (out->temperature) out->temperature[s] = r[8];
        if (out->kinetic_viscosity) out->kinetic_viscosity[s] = r[9];
        if (out->solid_phase_rate) out->solid_phase_rate[s] = r[10];
        if (out->type) out->type[s] = (int)r[11];
    }
    out->n = (int)rows.size();
    return MPSR_NOERR;
}

int mpsr_deck_start_step(const mpsr_deck* d) { return d ? d->start_step : 0; }
int mpsr_deck_output_interval(const mpsr_deck* d) { return d ? d->output_interval : 1; }
int mpsr_deck_output_ascii(const mpsr_deck* d) { return d ? d->ascii : 1; }
const char* mpsr_deck_output_name(const mpsr_deck* d) { return d ? d->output_name.c_str() : ""; }
int mpsr_deck_model_offsets(const mpsr_deck* d, mpsr_float3* out, int cap) {
    if (!d) return 0;
    for (int i = 0; i < (int)d->model_offset.size() && i < cap; ++i) out[i] = d->model_offset[i];
    return (int)d->model_offset.size();
}

mpsr_status mpsr_dt_schedule_load(const char* path, mpsr_dt_entry* out, int cap, int* count) {
    if (!path || !count || cap < 0 || (cap > 0 && !out)) return deck_fail(MPSR_PARAMERR, "null argument");
    std::ifstream in(path);
    if (!in.is_open()) return deck_fail(MPSR_FILEERR, std::string("cannot open dt schedule ") + path);
    // records of five whitespace-separated values; a record counts only when the stream is still
    // good after its last value (a last record that ends the file without a line break is dropped,
    // and the first unreadable value ends the list)
    int n = 0;
    while (in.good()) {
        mpsr_dt_entry e{};
        in >> e.step >> e.dt >> e.courant >> e.inflow_step >> e.output_step;
        if (!in.good()) break;
        if (n < cap) out[n] = e;
        ++n;
    }
    *count = n;
    return MPSR_NOERR;
}

mpsr_status mpsr_write_frame(const char* path, int ascii, const mpsr_particles* p, float time, int nmodels,
                             const mpsr_float3* model_pos, const float* angles) {
    if (!path || !p) return deck_fail(MPSR_PARAMERR, "null argument");
    if (ascii) {
        std::FILE* f = std::fopen(path, "w");
        if (!f) return deck_fail(MPSR_FILEERR, std::string("cannot write ") + path);
        std::fprintf(f, "%d\n", p->n);
        for (int i = 0; i < p->n; ++i)
            std::fprintf(f, "%g %g %g %g %g %g %g %g %g %g %g %d\n", (float)p->position[3 * i],
                         (float)p->position[3 * i + 1], (float)p->position[3 * i + 2], (float)p->velocity[3 * i],
                         (float)p->velocity[3 * i + 1], (float)p->velocity[3 * i + 2], (float)p->pressure[i],
                         (float)p->density[i], (float)p->temperature[i], (float)p->kinetic_viscosity[i],
                         (float)p->solid_phase_rate[i], p->type[i]);
        std::fprintf(f, "%g\n%d\n", time, nmodels);
        for (int m = 0; m < nmodels; ++m)
            std::fprintf(f, "%g\t%g\t%g\t%g\n", model_pos[m].x, model_pos[m].y, model_pos[m].z, angles[m]);
        std::fclose(f);
        return MPSR_NOERR;
    }
    std::FILE* f = std::fopen(path, "wb");
    if (!f) return deck_fail(MPSR_FILEERR, std::string("cannot write ") + path);
    std::fwrite(&p->n, 4, 1, f);
    for (int i = 0; i < p->n; ++i) {
        float v[11] = {(float)p->position[3 * i], (float)p->position[3 * i + 1], (float)p->position[3 * i + 2],
                       (float)p->velocity[3 * i], (float)p->velocity[3 * i + 1], (float)p->velocity[3 * i + 2],
                       (float)p->pressure[i], (float)p->density[i], (float)p->temperature[i],
                       (float)p->kinetic_viscosity[i], (float)p->solid_phase_rate[i]};
        std::fwrite(v, 4, 11, f);
        std::fwrite(&p->type[i], 4, 1, f);
    }
    std::fwrite(&time, 4, 1, f);
    std::fwrite(&nmodels, 4, 1, f);
    for (int m = 0; m < nmodels; ++m) {
        std::fwrite(&model_pos[m], 4, 3, f);
        std::fwrite(&angles[m], 4, 1, f);
    }
    std::fclose(f);
    return MPSR_NOERR;
}

}  // extern "C"
