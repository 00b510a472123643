// mpsr_run -- command-line driver of one deck on one GPU, through the C ABI only.
//
//   mpsr_run <configuration file> [--steps K] [--out DIR] [--fp64] [--ppe] [--mixed] [--two-level]
//            [--dt FILE] [--device D] [--no-output] [--quiet]
//
// Mirrors CMultiGPUHandler::Run + CalculateNew (Src MultiGPUHandler.cpp:334-492, 574-840) for a
// single section: load the deck, create the domain, register the buckets, then for
// i = StartStep .. AnalysisStep run the step body.  Steps without output run in batches through
// mpsr_step_many (one host synchronisation per batch).  A step with output (i % interval == 0)
// runs phase by phase, because the frame is taken after the second relocation and before the
// bucket addition and the wall motion of that step:
//   register, triangle register, STL distance, drag, explicit, relocate, register,
//   [explicit frame when bExplicitOut], STL distance, pressure, gradient, relocate,
//   frame, bucket addition, relocate (added) or wall translation, wall rotation.
// Frames are named <OutputName>_<Implicit|Explicit>_<interval>_<NN>.dat (Output,
// MultiGPUHandler.cpp:534-556), time = Dt * interval * frame number; particles outside the
// compute zone are written with position, velocity, pressure, density and temperature 0
// (AddToOutputBuffer with Output = true, CGPUCellSet.cpp:1388-1428).  Model positions follow
// COutputFrame::SetModelPosition (OutputHandler.h:150-166) and the rotation angle is the sum of the
// per-step angles of SetRotationAngle (OutputHandler.h:168-193) over steps 1..i.
//
// --dt FILE is the reference's third argument (Src MultiGPUHandler.cpp:812-826): the time-step
// schedule (DtFrame::Load, Src DtFrame.cpp:11-35).  Before every step i the next record whose step
// is <= i is applied, one record per step (CheckAndUpdateDt, MultiGPUHandler.cpp:1048-1066): the
// new Dt and Courant number go to the device through mpsr_set_dt, as CCellSet::CheckAndUpdateDt
// uploads them (CGPUCellSet.cpp:1260-1282), and the host copy used for frame times and the model
// angles and positions changes with them.  (The reference's live loop calls the handler's version,
// which changes only that host copy and leaves the device at the first Dt.)  Batches of
// mpsr_step_many end before a step that applies a record.
//
// Frames are written by a writer thread (the reference's output thread, OutputHandler.cpp:23-36):
// the step loop copies the particles to host memory and queues the frame; at most two frames wait,
// and the loop goes on stepping while the files are written.
#include <sys/stat.h>

#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mpsr.h"

namespace {

struct Host {
    int n = 0;
    std::vector<int> id, type;
    std::vector<double> pos, vel, p, dens, temp, nu, solid;
    void resize(int m) {
        n = m;
        id.resize(m); type.resize(m);
        pos.resize(3 * (size_t)m); vel.resize(3 * (size_t)m);
        p.resize(m); dens.resize(m); temp.resize(m); nu.resize(m); solid.resize(m);
    }
    mpsr_particles view() {
        return {n, id.data(), pos.data(), vel.data(), p.data(), dens.data(), temp.data(), nu.data(), solid.data(),
                type.data()};
    }
};

int usage() {
    std::fprintf(stderr,
                 "usage: mpsr_run <configuration file> [--steps K] [--out DIR] [--fp64] [--ppe] [--mixed]\n"
                 "                [--two-level] [--dt FILE] [--device D] [--no-output] [--quiet]\n");
    return 2;
}

// one frame waiting for the writer thread: its own copy of the particles and the model data
struct FrameJob {
    std::string path;
    int ascii = 1;
    Host data;
    float time = 0.0f;
    int nmodels = 0;
    std::vector<mpsr_float3> model_pos;
    std::vector<float> angles;
};

class FrameWriter {
public:
    FrameWriter() : th_([this] { run(); }) {}
    ~FrameWriter() { finish(); }
    void push(FrameJob&& j) {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return q_.size() < kMaxQueued; });
        q_.push_back(std::move(j));
        cv_.notify_all();
    }
    // waits for the queued frames; false when one could not be written
    bool finish() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            done_ = true;
        }
        cv_.notify_all();
        if (th_.joinable()) th_.join();
        if (failed_) std::fprintf(stderr, "mpsr_run: frame writer: %s\n", err_.c_str());
        return !failed_;
    }

private:
    static constexpr size_t kMaxQueued = 2;
    void run() {
        for (;;) {
            FrameJob j;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return done_ || !q_.empty(); });
                if (q_.empty()) return;
                j = std::move(q_.front());
                q_.pop_front();
                cv_.notify_all();
            }
            mpsr_particles v = j.data.view();
            const mpsr_status st = mpsr_write_frame(j.path.c_str(), j.ascii, &v, j.time, j.nmodels, j.model_pos.data(),
                                                    j.angles.data());
            if (st != MPSR_NOERR) {
                std::lock_guard<std::mutex> lk(mu_);
                failed_ = true;
                err_ = mpsr_deck_last_error();
            }
        }
    }
    std::mutex mu_;
    std::condition_variable cv_;
    std::deque<FrameJob> q_;
    bool done_ = false;
    bool failed_ = false;
    std::string err_;
    std::thread th_;
};

bool check(mpsr_status s, const char* what) {
    if (s == MPSR_NOERR) return true;
    std::fprintf(stderr, "mpsr_run: %s failed (status %d): %s\n", what, (int)s, mpsr_last_error());
    return false;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) return usage();
    const char* config = argv[1];
    long long steps = -1;
    std::string out_dir;
    int fp64 = 0, ppe = 0, mixed = 0, two_level = 0, device = 0, no_output = 0, quiet = 0;
    std::string dt_file;
    for (int a = 2; a < argc; ++a) {
        const std::string k = argv[a];
        if (k == "--steps" && a + 1 < argc) steps = std::atoll(argv[++a]);
        else if (k == "--dt" && a + 1 < argc) dt_file = argv[++a];
        else if (k == "--out" && a + 1 < argc) out_dir = argv[++a];
        else if (k == "--device" && a + 1 < argc) device = std::atoi(argv[++a]);
        else if (k == "--fp64") fp64 = 1;
        else if (k == "--ppe") ppe = 1;
        else if (k == "--mixed") mixed = 1;
        else if (k == "--two-level") two_level = 1;
        else if (k == "--no-output") no_output = 1;
        else if (k == "--quiet") quiet = 1;
        else return usage();
    }

    mpsr_deck* deck = nullptr;
    mpsr_status st = mpsr_deck_load(config, &deck);
    if (st != MPSR_NOERR) {
        std::fprintf(stderr, "mpsr_run: deck %s: status %d: %s\n", config, (int)st, mpsr_deck_last_error());
        return 1;
    }
    mpsr_create_info info;
    mpsr_deck_create_info(deck, &info);
    info.precision = fp64 ? MPSR_FP64 : MPSR_FP32;
    info.device = device;
    if (ppe) {
        info.pressure_mode = MPSR_PRESSURE_PPE;
        info.ppe.solver = mixed ? MPSR_PCG_MIXED : MPSR_PCG_FP64;
        info.ppe.warm_start = MPSR_WARM_PREVIOUS;
        info.ppe.precond = two_level ? MPSR_PRECOND_TWO_LEVEL : MPSR_PRECOND_JACOBI;
        info.ppe.coarse_cells = 4;
    }
    mpsr_params P = info.params;
    std::vector<mpsr_dt_entry> dts;
    if (!dt_file.empty()) {
        int cnt = 0;
        if (!check(mpsr_dt_schedule_load(dt_file.c_str(), nullptr, 0, &cnt), "dt schedule")) return 1;
        dts.resize(cnt);
        if (cnt > 0 && !check(mpsr_dt_schedule_load(dt_file.c_str(), dts.data(), cnt, &cnt), "dt schedule")) return 1;
    }
    size_t next_dt = 0;
    const int interval = std::max(1, mpsr_deck_output_interval(deck));
    const int ascii = mpsr_deck_output_ascii(deck);
    const int start = mpsr_deck_start_step(deck);
    const long long last = steps >= 0 ? (long long)start + steps - 1 : (long long)P.AnalysisStep;
    const std::string name = mpsr_deck_output_name(deck);
    if (out_dir.empty()) {
        std::string c = config;
        const size_t k = c.find_last_of('/');
        out_dir = (k == std::string::npos ? std::string(".") : c.substr(0, k)) + "/OUT";
    }
    if (!no_output) mkdir(out_dir.c_str(), 0755);

    const int nmodels = info.num_models;
    std::vector<mpsr_tri_params> models(info.models, info.models + nmodels);
    std::vector<mpsr_float3> offsets(std::max(1, nmodels));
    mpsr_deck_model_offsets(deck, offsets.data(), nmodels);

    mpsr_handle* h = nullptr;
    if (!check(mpsr_create(&info, &h), "mpsr_create")) return 1;
    // buckets
    const int nb = mpsr_deck_num_buckets(deck);
    std::vector<Host> bh(nb);
    std::vector<mpsr_particles> bv(nb);
    std::vector<int> bint(nb), btot(nb);
    for (int j = 0; j < nb; ++j) {
        int cnt = 0;
        mpsr_deck_bucket(deck, j, &cnt, &bint[j], &btot[j]);
        bh[j].resize(cnt);
        bv[j] = bh[j].view();
        mpsr_deck_get_bucket(deck, j, &bv[j]);
    }
    if (nb > 0 && !check(mpsr_set_buckets(h, bv.data(), bint.data(), btot.data(), nb), "mpsr_set_buckets")) return 1;
    mpsr_gridbox gb;
    mpsr_get_gridbox(h, &gb);
    mpsr_deck_free(deck);

    std::vector<float> angle(std::max(1, nmodels), 0.0f);
    auto add_rotation = [&](int i) {  // SetRotationAngle(i, ...) for i != 0
        for (int m = 0; m < nmodels; ++m) {
            const float dt = P.Dt, deg = models[m].AngleOfRotationInDegree;
            const float t0 = models[m].RotationStartTime, t1 = models[m].TimeToReachRPM;
            const float passed = dt * i;
            const int req = (int)((t1 - t0) / dt);
            float a;
            if (passed < t0 || deg == 0) a = 0.0f;
            else if (passed >= t0 && passed < t1) a = (deg / (req + 1)) * (i - (t0 / dt));
            else a = deg;
            angle[m] += a;
        }
    };
    int file_no = start / interval;
    FrameWriter writer;
    auto frame = [&](int i, const char* type, int number) -> bool {
        if (no_output) return true;
        int n = 0;
        if (!check(mpsr_get_particle_count(h, &n), "mpsr_get_particle_count")) return false;
        FrameJob job;
        Host& fr = job.data;
        fr.resize(n);
        mpsr_particles v = fr.view();
        if (!check(mpsr_get_particles(h, 1, &v), "mpsr_get_particles")) return false;
        const mpsr_box& z = gb.m_ComputeZone;
        const float tol = P.Tolerance;
        for (int s = 0; s < n; ++s) {
            const float x = (float)fr.pos[3 * s], y = (float)fr.pos[3 * s + 1], w = (float)fr.pos[3 * s + 2];
            const bool in = x >= z.m_MinBound.x - tol && y >= z.m_MinBound.y - tol && w >= z.m_MinBound.z - tol &&
                            x < z.m_MaxBound.x + tol && y < z.m_MaxBound.y + tol && w < z.m_MaxBound.z + tol;
            if (!in) {
                for (int a = 0; a < 3; ++a) fr.pos[3 * s + a] = fr.vel[3 * s + a] = 0.0;
                fr.p[s] = fr.dens[s] = fr.temp[s] = 0.0;
            }
        }
        std::vector<mpsr_float3> mp(std::max(1, nmodels), mpsr_float3{0, 0, 0});
        for (int m = 0; m < nmodels; ++m) {
            if (!models[m].LinearMovableOrNot || models[m].LinearResetInterval <= 0) continue;
            const float lapse = P.Dt * (i % models[m].LinearResetInterval);
            mp[m] = {lapse * models[m].Velocity.x + offsets[m].x, lapse * models[m].Velocity.y + offsets[m].y,
                     lapse * models[m].Velocity.z + offsets[m].z};
        }
        char num[32];
        std::snprintf(num, sizeof(num), number < 10 ? "0%d" : "%d", number);
        job.path = out_dir + "/" + name + "_" + type + "_" + std::to_string(interval) + "_" + num + ".dat";
        job.ascii = ascii;
        job.time = P.Dt * interval * number;
        job.nmodels = nmodels;
        job.model_pos = std::move(mp);
        job.angles.assign(angle.begin(), angle.end());
        writer.push(std::move(job));
        return true;
    };
    // CheckAndUpdateDt: at most one due record per step
    auto update_dt = [&](long long i) -> bool {
        if (next_dt >= dts.size() || dts[next_dt].step < 0 || i < dts[next_dt].step) return true;
        const mpsr_dt_entry& e = dts[next_dt++];
        P.Dt = e.dt;
        P.CourantNumber = e.courant;
        P.ParticleGenerationStep = e.inflow_step;
        P.OutputStep = e.output_step;
        return check(mpsr_set_dt(h, e.dt, e.courant), "mpsr_set_dt");
    };
    auto phase_step = [&](int i) -> bool {
        bool ok = check(mpsr_register_particle_topology(h), "register") &&
                  check(mpsr_register_triangle_topology(h), "triangle register") &&
                  check(mpsr_calc_stl_distance(h), "STL distance");
        if (ok && info.num_drag > 0) ok = check(mpsr_calc_drag_effect(h), "drag");
        ok = ok && check(mpsr_calc_explicitly(h), "explicit") && check(mpsr_relocate_particle_data(h), "relocate") &&
             check(mpsr_register_particle_topology(h), "register");
        if (ok && P.bExplicitOut) ok = frame(i, "Explicit", file_no);
        ok = ok && check(mpsr_calc_stl_distance(h), "STL distance") &&
             check(ppe ? mpsr_calc_pressure_poisson(h) : mpsr_calc_pressure_explicit(h), "pressure") &&
             check(mpsr_calc_implicit_explicitly(h), "gradient") && check(mpsr_relocate_particle_data(h), "relocate");
        if (!ok) return false;
        if (i != 0) add_rotation(i);
        if (!frame(i, "Implicit", file_no)) return false;
        ++file_no;
        int added = 0;
        if (!check(mpsr_bucket_addition(h, i, &added), "bucket addition")) return false;
        if (added) ok = check(mpsr_relocate_particle_data(h), "relocate");
        else ok = check(mpsr_update_triangle_position(h), "wall translation");
        return ok && check(mpsr_rotate_triangle_position(h, i), "wall rotation");
    };

    const auto t0 = std::chrono::steady_clock::now();
    long long i = start;
    while (i <= last) {
        if (!update_dt(i)) return 1;
        if (i % interval == 0) {
            if (!phase_step((int)i)) return 1;
            ++i;
            continue;
        }
        const long long next_out = ((i + interval - 1) / interval) * interval;  // next output step
        long long end = std::min(last + 1, next_out);
        // the batch ends before the next step that applies a schedule record
        if (next_dt < dts.size() && dts[next_dt].step > i) end = std::min(end, (long long)dts[next_dt].step);
        if (next_dt < dts.size() && dts[next_dt].step >= 0 && dts[next_dt].step <= i) end = i + 1;
        if (!check(mpsr_step_many(h, (int)i, (int)(end - i)), "mpsr_step_many")) return 1;
        for (long long k = i; k < end; ++k)
            if (k != 0) add_rotation((int)k);
        i = end;
        if (!quiet) {
            int n = 0;
            mpsr_get_particle_count(h, &n);
            const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            std::printf("step %lld particles %d elapsed %.2f s\n", i - 1, n, s);
            std::fflush(stdout);
        }
    }
    mpsr_synchronize(h);
    if (!writer.finish()) return 1;
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    int n = 0;
    mpsr_get_particle_count(h, &n);
    std::printf("done: steps %lld..%lld, %d particles, %.3f s, %d frames in %s\n", (long long)start, last, n, secs,
                no_output ? 0 : file_no - start / interval, no_output ? "-" : out_dir.c_str());
    mpsr_destroy(h);
    return 0;
}
