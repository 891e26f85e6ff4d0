// Serving: concurrent reconstructions under a device-memory budget.
//
// Jobs wait in arrival order; the head job is admitted when the memory it needs
// fits next to the jobs already running and a concurrency slot is free, so a
// job that arrives while others run starts as soon as memory frees up and its
// work overlaps theirs on the device (each job has its own streams).  A job
// whose resident footprint exceeds the whole budget runs alone with slab
// streaming (volume held on the host).  Each job owns its buffers and seeded
// jitter: its result does not depend on what else runs.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../../include/xpr_hip.h"
#include "params.hpp"
#include "recon.hpp"

int xprint_fail(int code, const char *msg);

namespace xpr {
namespace {

double now_ms()
{
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

Parameters job_parameters(const xpr_job &j, bool streamed)
{
    Parameters p;
    p.recon.method = streamed ? Method::method_1a : Method::naive;
    p.recon.it = j.it;
    p.recon.sit = j.sit;
    p.recon.weight = j.weight;
    p.recon.weight_factor = j.weight_factor;
    p.recon.mlog = j.mlog != 0;
    p.recon.normalize = j.normalize != 0;
    const xpr_geometry_params &g = j.geometry;
    p.geometry.so = float(g.so);
    p.geometry.sd = float(g.sd);
    p.geometry.angle = float(g.angle);
    p.geometry.heli_offset = float(g.heli_offset);
    if (g.angle_list)
        p.geometry.angle_list.assign(g.angle_list, g.angle_list + g.projections);
    p.detector.px = float(g.px);
    p.detector.rx = float(g.rx);
    p.detector.sx = float(g.sx);
    p.detector.sy = float(g.sy);
    p.detector.module_number = g.module_number;
    if (g.has_module_angle) p.detector.module_angle_offset = float(g.module_angle_offset);
    p.detector.module_center_offset_z = float(g.module_center_offset_z);
    p.detector.roll = float(g.roll);
    p.detector.pitch = float(g.pitch);
    p.detector.yaw = float(g.yaw);
    p.xpr.storage = j.storage;
    p.xpr.final_fp32 = j.final_fp32 != 0;
    p.xpr.accumulator = j.acc_bits;
    p.xpr.jitter = j.jitter != 0;
    p.xpr.jitter_seed = j.jitter_seed;
    p.xpr.geometry_fp64 = g.fp64 != 0;
    p.xpr.slab_height = j.slab_height > 0 ? j.slab_height : 64;
    p.xpr.layer_store = "ram";
    p.xpr.threads = 4;
    p.xpr.mem_projections = j.projections;
    p.xpr.mem_width = g.dwidth;
    p.xpr.mem_height = g.dheight;
    p.xpr.mem_count = g.projections;
    p.xpr.mem_volume_out = j.volume_out;
    p.xpr.quiet = true;
    return p;
}

// Device bytes of the naive schedule (ReconstructionN::requiredGPUMemory) and
// of the slab schedule for a job, without building anything.
void job_bytes(const xpr_job &j, uint64_t &resident, uint64_t &streamed)
{
    const xpr_geometry_params &g = j.geometry;
    const double vx = g.px * (g.so / g.sd);
    const uint64_t W = uint64_t(g.dwidth);
    const uint64_t H = uint64_t(g.dheight) + uint64_t(std::fabs(g.heli_offset / vx));
    const uint64_t P = uint64_t(g.dwidth * g.dheight), S = uint64_t(j.sit);
    const uint64_t As = uint64_t(g.projections) / S;
    const uint64_t sv = j.storage == XPR_F32 ? 4 : 2, acc = uint64_t(j.acc_bits / 8);
    const uint64_t pd = uint64_t(g.projections * g.module_number) * 80;
    const uint64_t V = W * W * H;
    resident = V * sv + S * As * P * sv + As * P * (acc + sv) + pd + P * 8;
    if (j.storage != XPR_F32 && j.final_fp32) resident += V * 4 + As * P * 8;
    const uint64_t Hs = std::min<uint64_t>(j.slab_height > 0 ? uint64_t(j.slab_height) : 64, H);
    streamed = 2 * Hs * W * W * (sv + 4) + As * P * (sv + acc + sv + 4) + pd + P * 8;
}

struct JobRun {
    xpr_job job;
    std::vector<double> angles;     // owned copy of the angle list
    xpr_job_stats stats{};
    uint64_t need = 0;
    bool streamed = false;
    bool done = false;
    std::string error;              // message of a failed job
};

} // namespace

struct Server {
    uint64_t budget;
    int max_conc, device;
    std::mutex m;
    std::condition_variable cv;
    std::deque<uint64_t> pending;
    std::map<uint64_t, std::shared_ptr<JobRun>> jobs;
    uint64_t next_id = 1, in_use = 0;
    int running = 0;
    bool stopping = false;
    std::thread dispatcher;
    std::vector<std::thread> workers;

    Server(uint64_t b, int mc, int dev) : budget(b), max_conc(mc), device(dev)
    {
        dispatcher = std::thread([this] { dispatch(); });
    }

    ~Server()
    {
        {
            std::unique_lock<std::mutex> lk(m);
            cv.wait(lk, [this] { return pending.empty() && running == 0; });
            stopping = true;
        }
        cv.notify_all();
        dispatcher.join();
        for (auto &t : workers) t.join();
    }

    void dispatch()
    {
        std::unique_lock<std::mutex> lk(m);
        while (true) {
            cv.wait(lk, [this] {
                if (stopping) return true;
                if (pending.empty() || running >= max_conc) return false;
                const auto &j = jobs[pending.front()];
                return running == 0 || in_use + j->need <= budget;
            });
            if (stopping) return;
            const uint64_t id = pending.front();
            pending.pop_front();
            auto run = jobs[id];
            in_use += run->need;
            ++running;
            run->stats.start_ms = now_ms();
            workers.emplace_back([this, run] { execute(run); });
        }
    }

    void execute(std::shared_ptr<JobRun> run)
    {
        int status = XPR_OK;
        double dev_ms = 0.0;
        std::string error;
        try {
            check(xpr_set_device(device), "set device");
            Parameters p = job_parameters(run->job, run->streamed);
            p.xpr.device = device;
            p.xpr.device_budget_gb = double(run->need) / 1e9 + 1e-3;
            auto recon = make_reconstruction(p);
            recon->exec();
            const Json rec = recon->run_record();
            if (const Json *ms = rec.get("subiteration_ms"))
                for (const Json &v : ms->arr) dev_ms += v.num;
        } catch (const std::exception &e) {
            const std::string what = e.what();
            status = what.find("GPU memory") != std::string::npos ? XPR_E_NOMEM : XPR_E_INVALID;
            error = what;
        }
        {
            std::lock_guard<std::mutex> lk(m);
            run->stats.end_ms = now_ms();
            run->stats.device_ms = dev_ms;
            run->stats.status = status;
            run->error = error;
            run->done = true;
            in_use -= run->need;
            --running;
        }
        cv.notify_all();
    }

    uint64_t submit(const xpr_job &job)
    {
        auto run = std::make_shared<JobRun>();
        run->job = job;
        if (job.geometry.angle_list) {
            run->angles.assign(job.geometry.angle_list, job.geometry.angle_list + job.geometry.projections);
            run->job.geometry.angle_list = run->angles.data();
        }
        uint64_t res = 0, str = 0;
        job_bytes(job, res, str);
        run->streamed = res > budget;
        run->need = run->streamed ? str : res;
        const xpr_geometry_params &g = job.geometry;
        const uint64_t As = uint64_t(g.projections / job.sit);
        const double vx = g.px * (g.so / g.sd);
        const uint64_t H = uint64_t(g.dheight) + uint64_t(std::fabs(g.heli_offset / vx));
        const uint64_t V = uint64_t(g.dwidth) * uint64_t(g.dwidth) * H;
        const uint64_t subs = uint64_t(job.it * job.sit);
        run->stats.voxel_views = V * As * uint64_t(g.module_number) * (2 * subs - 1);
        run->stats.device_bytes = run->need;
        run->stats.streamed = run->streamed;
        run->stats.submit_ms = now_ms();
        std::lock_guard<std::mutex> lk(m);
        const uint64_t id = next_id++;
        jobs[id] = run;
        pending.push_back(id);
        cv.notify_all();
        return id;
    }

    bool wait(uint64_t id, xpr_job_stats *stats, std::string &error)
    {
        std::unique_lock<std::mutex> lk(m);
        auto it = jobs.find(id);
        if (it == jobs.end()) return false;
        auto run = it->second;
        cv.wait(lk, [&] { return run->done; });
        if (stats) *stats = run->stats;
        error = run->error;
        jobs.erase(id);
        return true;
    }
};

} // namespace xpr

struct xpr_server {
    std::unique_ptr<xpr::Server> impl;
};

extern "C" {

int xpr_server_create(xpr_server **out, uint64_t budget_bytes, int max_concurrent, int device)
{
    if (!out || max_concurrent < 1 || budget_bytes == 0)
        return xprint_fail(XPR_E_INVALID, "server: bad arguments");
    *out = new xpr_server{std::make_unique<xpr::Server>(budget_bytes, max_concurrent, device)};
    return XPR_OK;
}

int xpr_job_device_bytes(const xpr_job *job, uint64_t *bytes)
{
    if (!job || !bytes) return xprint_fail(XPR_E_INVALID, "null pointer");
    uint64_t res = 0, str = 0;
    xpr::job_bytes(*job, res, str);
    *bytes = res;
    return XPR_OK;
}

int xpr_server_submit(xpr_server *sv, const xpr_job *job, uint64_t *job_id)
{
    if (!sv || !job || !job_id) return xprint_fail(XPR_E_INVALID, "null pointer");
    if (!job->projections || !job->volume_out)
        return xprint_fail(XPR_E_INVALID, "job without projections or output");
    if (job->sit < 1 || job->it < 1 || job->geometry.projections / job->sit < 1)
        return xprint_fail(XPR_E_INVALID, "job: it, sit or projection count");
    if (job->acc_bits != XPR_ACC_U32 && job->acc_bits != XPR_ACC_U64)
        return xprint_fail(XPR_E_INVALID, "job: acc_bits must be 32 or 64");
    if (job->storage < XPR_F32 || job->storage > XPR_BF16)
        return xprint_fail(XPR_E_UNSUPPORTED, "job: storage type");
    *job_id = sv->impl->submit(*job);
    return XPR_OK;
}

int xpr_server_wait(xpr_server *sv, uint64_t job_id, xpr_job_stats *stats)
{
    if (!sv) return xprint_fail(XPR_E_INVALID, "null server");
    xpr_job_stats st{};
    std::string error;
    if (!sv->impl->wait(job_id, &st, error)) return xprint_fail(XPR_E_INVALID, "unknown job id");
    if (stats) *stats = st;
    return st.status != XPR_OK ? xprint_fail(st.status, error.c_str()) : XPR_OK;
}

int xpr_server_destroy(xpr_server *sv)
{
    delete sv;
    return XPR_OK;
}

} // extern "C"
