// Serving: concurrent reconstructions under a device-memory budget.
//
// Jobs wait in arrival order; the head job is admitted when the memory it needs
// fits next to the jobs already running and a concurrency slot is free.  A job
// whose resident footprint exceeds the whole budget runs alone with slab
// streaming (volume held on the host).  Each job owns its buffers and seeded
// jitter: its result does not depend on what else runs.
//
// Two ways to run admitted jobs:
//  * batch (XPR_SERVE_MODE=batch): continuous batching.  Each admitted job prepares its
//    resident buffers on its own stream (raw projections uploaded and
//    preprocessed on the device), then joins the active set of one scheduler
//    thread.  The scheduler runs steps: per step every active job gets a share
//    of work (forward / backward layer ranges, error, fill, the FP32 switch of
//    its last sub-iteration), shortest remaining work first, each at least one
//    quantum, up to a step target; the step's items go to a table and five
//    persistent kernels replayed from a captured HIP graph run them all.  A
//    job's pieces run in its own sub-iteration order with its own arguments,
//    so its volume is the one the naive schedule gives alone.
//    Jobs whose whole work fits in one step (and u32-accumulator jobs) run
//    on their own streams beside the steps.
//  * streams (default): each admitted job runs the naive schedule on its own
//    streams; the device interleaves the workgroups of concurrent jobs.  On
//    the mixed-size serving workload this gives the lower median latency: the
//    persistent step kernels hold every CU until their queue drains, so small
//    jobs beside them wait, while the batch mode's shortest-work-first steps
//    help mid-size jobs and the tail.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "../../../include/xpr_hip.h"
#include "../batch.hpp"
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
    const uint64_t sv = j.storage == XPR_F32 ? 4 : 2, acc = acc_bytes(j.acc_bits);
    const uint64_t pd = uint64_t(g.projections * g.module_number) * 80;
    const uint64_t V = W * W * H;
    resident = V * sv + S * As * P * sv + As * P * (acc + sv) + pd + P * 8;
    if (j.storage != XPR_F32 && j.final_fp32) resident += V * 4 + As * P * 8;
    const uint64_t Hs = std::min<uint64_t>(j.slab_height > 0 ? uint64_t(j.slab_height) : 64, H);
    streamed = 2 * Hs * W * W * (sv + 4) + As * P * (sv + acc + sv + 4) + pd + P * 8;
}

size_t esize(int dt) { return dt == XPR_F32 ? 4 : 2; }

// ---------------------------------------------------------------------------
// one job of the batch scheduler: resident buffers and its sub-iteration cursor
// ---------------------------------------------------------------------------
struct BatchJob {
    enum Phase { kFwd, kErr, kBwd };

    Parameters p;
    std::unique_ptr<Dataset> ds;
    xpr_stream stream = nullptr;
    int st = XPR_F32;
    bool final32 = false;
    int64_t W = 0, H = 0, P = 0, As = 0, S = 0, IT = 0, M = 1, det_w = 0, det_h = 0;
    float vx = 0.0f;
    std::array<float, 3> orig{};
    std::mt19937 rng;
    DevBuf vol, vol32, sums, err, err32, img32;
    std::vector<DevBuf> images, pds;
    std::vector<std::vector<float>> pd_host;

    // cursor
    int64_t mit = 0, s = 0, layer = 0;
    Phase phase = kFwd;
    bool in_sit = false, complete = false;
    std::array<float, 3> origin{};
    float w = 1.0f;
    int vdt = XPR_F32;

    explicit BatchJob(const Parameters &params) : p(params)
    {
        check(xpr_stream_create(&stream), "stream");
    }
    ~BatchJob()
    {
        {
            AllocStreamScope scope(stream);
            vol.reset();
            vol32.reset();
            sums.reset();
            err.reset();
            err32.reset();
            img32.reset();
            images.clear();
            pds.clear();
        }
        if (stream) xpr_stream_destroy(stream);
    }

    // resident buffers of the naive schedule, images preprocessed on the device
    void prepare()
    {
        AllocStreamScope scope(stream);
        ds = std::make_unique<Dataset>(p);
        ds->load(stream);
        st = p.xpr.storage;
        final32 = st != XPR_F32 && p.xpr.final_fp32;
        W = ds->vwidth();
        H = ds->vheight();
        det_w = ds->width();
        det_h = ds->height();
        P = det_w * det_h;
        As = ds->concurrent();
        S = p.recon.sit;
        IT = p.recon.it;
        M = p.detector.module_number;
        vx = ds->vx();
        orig = ds->orig();
        if (p.xpr.jitter_seed >= 0) {
            rng.seed(uint32_t(p.xpr.jitter_seed));
        } else {
            std::random_device rd;
            rng.seed(rd());
        }
        const uint64_t V = uint64_t(W * W * H);
        images.resize(size_t(S));
        pds.resize(size_t(S));
        pd_host.resize(size_t(S));
        for (int64_t k = 0; k < S; ++k) {
            images[size_t(k)].alloc(uint64_t(As * P) * esize(st));
            ds->upload_sit_images(stream, k, 0, det_h, images[size_t(k)].get(), st);
            pd_host[size_t(k)] = ds->sit_projdata(k);
            pds[size_t(k)].alloc(pd_host[size_t(k)].size() * 4);
            check(xpr_upload(stream, pds[size_t(k)].get(), pd_host[size_t(k)].data(),
                             pd_host[size_t(k)].size() * 4), "upload");
        }
        vol.alloc(V * esize(st));
        check(xpr_fill(stream, vol.get(), st, 1.0f, V), "fill");
        sums.alloc(uint64_t(As * P) * 8);
        check(xpr_fill_u64(stream, sums.as<uint64_t>(), 1ull << 22, uint64_t(As * P)), "fill");
        err.alloc(uint64_t(As * P) * esize(st));
        if (final32) {
            vol32.alloc(V * 4);
            img32.alloc(uint64_t(As * P) * 4);
            err32.alloc(uint64_t(As * P) * 4);
            ds->upload_sit_images(stream, S - 1, 0, det_h, img32.get(), XPR_F32);
        }
        check(xpr_stream_sync(stream), "sync");
    }

    double layer_cost() const { return double(W) * double(W) * double(As) * double(M); }

    // voxel-views left, the shortest-remaining-work key
    double remaining() const
    {
        const double subs_left = double(IT * S - (mit * S + s));
        double done_now = 0.0;
        if (in_sit) done_now = phase == kFwd ? double(layer) : phase == kErr ? double(H) : double(H + layer);
        return (subs_left * 2.0 * double(H) - done_now) * layer_cost();
    }

    double quantum_cost() const { return double(std::min<int64_t>(16, H)) * layer_cost(); }

    xpr_proj_args args(int64_t l0, int64_t l1) const
    {
        xpr_proj_args a{};
        a.det_width = uint32_t(det_w);
        a.n_views = uint32_t(As);
        a.modules = uint32_t(M);
        a.vmin[0] = 0;
        a.vmin[1] = int32_t(l0);
        a.vmin[2] = 0;
        a.vmax[0] = int32_t(W);
        a.vmax[1] = int32_t(l1);
        a.vmax[2] = int32_t(W);
        a.line_offset = 0;
        a.line_rows = int32_t(det_h);
        for (int i = 0; i < 3; ++i) {
            a.origin[i] = origin[size_t(i)];
            a.voxel[i] = vx;
        }
        a.view_index = nullptr;
        return a;
    }

    bool in_last() const { return mit == IT - 1 && s == S - 1; }
    // the last sub-iteration of a 16-bit job runs on the float32 copies
    bool on_f32_copies() const { return final32 && in_last(); }

    // Adds this step's items (at most one per phase) within quota; returns the
    // voxel-views used.  Sets complete when the job's last fill is planned.
    double plan(xpr_batch::Table &t, double quota)
    {
        using namespace xpr_batch;
        double used = 0.0;
        auto add = [&](uint32_t ph) -> Item & {
            Item &it = t.items[ph][t.count[ph]++];
            std::memset(&it, 0, sizeof(it));
            return it;
        };
        if (!in_sit) {
            // setOrigin(orig, vx) once per sub-iteration (src/reconstruction.hpp:268-272)
            for (int i = 0; i < 3; ++i) {
                float d = 0.0f;
                if (p.xpr.jitter) d = std::uniform_real_distribution<float>(0.0f, vx)(rng);
                origin[size_t(i)] = orig[size_t(i)] + d;
            }
            w = p.recon.weight;
            for (int64_t k = 0; k < mit; ++k) w *= p.recon.weight_factor;
            vdt = st;
            if (on_f32_copies()) {
                Item &c = add(kConvert);
                c.src = vol.get();
                c.dst = vol32.get();
                c.dt_src = st;
                c.dt_dst = XPR_F32;
                c.n = uint64_t(W * W * H);
                c.units = elementwise_units(c.n);
                vdt = XPR_F32;
            }
            phase = (mit == 0 && s == 0) ? kErr : kFwd;      // first pass: sums = 1.0
            layer = 0;
            in_sit = true;
        }
        const bool f32 = on_f32_copies();
        const size_t vb = esize(vdt);
        const double lc = layer_cost();
        auto layers_for = [&](double q) {
            int64_t n = int64_t(q / lc);
            n = std::max<int64_t>(n - n % 4, 16);
            return std::min<int64_t>(n, H - layer);
        };
        char *vbase = static_cast<char *>(f32 ? vol32.get() : vol.get());
        if (phase == kFwd) {
            const int64_t n = layers_for(quota - used);
            Item &f = add(xpr_batch::forward_phase(vdt));
            f.dst = vbase + size_t(layer) * size_t(W * W) * vb;
            f.sums = sums.get();
            f.pd = pds[size_t(s)].as<float>();
            f.a = args(layer, layer + n);
            f.dt_dst = vdt;
            f.units = forward_units(f.a);
            used += double(n) * lc;
            layer += n;
            if (layer < H) return used;
            phase = kErr;
            layer = 0;
        }
        if (phase == kErr) {
            Item &e = add(kError);
            e.src = f32 ? img32.get() : images[size_t(s)].get();
            e.dt_src = vdt;
            e.dst = f32 ? err32.get() : err.get();
            e.dt_dst = vdt;
            e.sums = sums.get();
            e.n = uint64_t(As * P);
            e.weight = w;
            e.units = elementwise_units(e.n);
            phase = kBwd;
            layer = 0;
        }
        const int64_t n = layers_for(std::max(quota - used, quantum_cost()));
        Item &b = add(xpr_batch::backward_phase(vdt));
        b.dst = vbase + size_t(layer) * size_t(W * W) * vb;
        b.err = f32 ? err32.get() : err.get();
        b.pd = pds[size_t(s)].as<float>();
        b.a = args(layer, layer + n);
        b.dt_dst = vdt;
        b.dt_aux = vdt;
        b.units = backward_units(b.a);
        used += double(n) * lc;
        layer += n;
        if (layer < H) return used;
        Item &z = add(kFill);                                // setBuffer(sums, 0)
        z.dst = sums.get();
        z.n = uint64_t(As * P);
        z.value = 0;
        z.units = elementwise_units(z.n);
        in_sit = false;
        if (++s == S) {
            s = 0;
            ++mit;
        }
        if (mit == IT) complete = true;
        return used;
    }

    // the volume, de-normalised, into the caller's buffer
    void finish()
    {
        AllocStreamScope scope(stream);
        const uint64_t V = uint64_t(W * W * H);
        float *out = ds->volume_sink();
        DevBuf stage;
        if (final32) {
            check(xpr_download(stream, out, vol32.get(), V * 4), "download");
        } else if (st == XPR_F32) {
            check(xpr_download(stream, out, vol.get(), V * 4), "download");
        } else {
            stage.alloc(V * 4);
            check(xpr_convert(stream, stage.get(), XPR_F32, vol.get(), st, V), "convert");
            check(xpr_download(stream, out, stage.get(), V * 4), "download");
        }
        check(xpr_stream_sync(stream), "sync");
        if (ds->normalize()) {                // saveLayer(finalize), data_loader.hpp:135-139
            const float scale = ds->norm_max() - ds->norm_min();
            for (uint64_t k = 0; k < V; ++k) out[k] = out[k] * scale;
        }
    }
};

struct JobRun {
    xpr_job job;
    std::vector<double> angles;     // owned copy of the angle list
    xpr_job_stats stats{};
    uint64_t need = 0;
    bool streamed = false;
    bool batched = false;
    bool done = false;
    std::string error;              // message of a failed job
    // batch mode
    std::shared_ptr<BatchJob> bj;
    bool gpu_done = false;          // its last step has finished on the device
    double step_ms = 0.0;           // device time of the steps it took part in
};

} // namespace

struct Server {
    uint64_t budget;
    int max_conc, device;
    bool batch_mode = false;
    double step_target = 5e9;       // voxel-views per step
    std::mutex m;
    std::condition_variable cv;
    std::deque<uint64_t> pending;
    std::map<uint64_t, std::shared_ptr<JobRun>> jobs;
    uint64_t next_id = 1, in_use = 0;
    int running = 0;
    bool stopping = false;
    std::thread dispatcher, scheduler;
    std::vector<std::thread> workers;
    std::deque<std::shared_ptr<JobRun>> ready;      // prepared, waiting to join a step
    std::string engine_error;

    Server(uint64_t b, int mc, int dev) : budget(b), max_conc(mc), device(dev)
    {
        if (const char *e = std::getenv("XPR_SERVE_MODE")) batch_mode = std::string(e) == "batch";
        if (const char *e = std::getenv("XPR_SERVE_STEP")) step_target = std::max(1e6, std::atof(e));
        dispatcher = std::thread([this] { dispatch(); });
        if (batch_mode) scheduler = std::thread([this] { schedule(); });
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
        if (scheduler.joinable()) scheduler.join();
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
            // a job whose whole work fits in one step gains nothing from
            // batching: its small launches run at once beside the steps
            run->batched = batch_mode && !run->streamed && run->job.acc_bits == XPR_ACC_U64 &&
                           double(run->stats.voxel_views) > step_target;
            if (run->batched)
                workers.emplace_back([this, run] { execute_batched(run); });
            else
                workers.emplace_back([this, run] { execute(run); });
        }
    }

    void finish_run(const std::shared_ptr<JobRun> &run, int status, const std::string &error,
                    double dev_ms)
    {
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

    static int status_of(const std::string &what)
    {
        return what.find("GPU memory") != std::string::npos ? XPR_E_NOMEM : XPR_E_INVALID;
    }

    // one job on its own streams (streams mode, streamed and u32 jobs)
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
            error = e.what();
            status = status_of(error);
        }
        finish_run(run, status, error, dev_ms);
    }

    // prepare on the job's stream, join the scheduler, download when done
    void execute_batched(std::shared_ptr<JobRun> run)
    {
        int status = XPR_OK;
        std::string error;
        std::shared_ptr<BatchJob> bj;
        try {
            check(xpr_set_device(device), "set device");
            Parameters p = job_parameters(run->job, false);
            p.xpr.device = device;
            bj = std::make_shared<BatchJob>(p);
            bj->prepare();
            {
                std::unique_lock<std::mutex> lk(m);
                if (!engine_error.empty()) throw std::runtime_error(engine_error);
                run->bj = bj;
                ready.push_back(run);
                cv.notify_all();
                cv.wait(lk, [&] { return run->gpu_done; });
                run->bj.reset();
                if (!engine_error.empty()) throw std::runtime_error(engine_error);
            }
            bj->finish();
        } catch (const std::exception &e) {
            error = e.what();
            status = status_of(error);
        }
        bj.reset();                 // frees the job's device memory before its budget returns
        finish_run(run, status, error, run->step_ms);
    }

    // the engine failed: every waiting job ends with its message
    void fail_engine(const std::string &what, std::vector<std::shared_ptr<JobRun>> &active)
    {
        std::lock_guard<std::mutex> lk(m);
        engine_error = what;
        for (auto &r : active) r->gpu_done = true;
        for (auto &r : ready) r->gpu_done = true;
        active.clear();
        ready.clear();
        cv.notify_all();
    }

    // the scheduler thread: plans and launches steps until the server stops
    void schedule()
    {
        std::vector<std::shared_ptr<JobRun>> active, members[2], finishing[2];
        xpr_batch::Engine *eng = nullptr;
        if (xpr_set_device(device) != XPR_OK || xpr_batch::engine_create(&eng) != XPR_OK) {
            const std::string what = std::string("batch engine: ") + xpr_last_error();
            std::unique_lock<std::mutex> lk(m);
            engine_error = what;
            while (!stopping) {
                for (auto &r : ready) r->gpu_done = true;
                ready.clear();
                cv.notify_all();
                cv.wait_for(lk, std::chrono::milliseconds(50));
            }
            return;
        }
        int slot = 0;
        bool busy[2] = {false, false};
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(m);
                cv.wait(lk, [&] {
                    return stopping || !ready.empty() || !active.empty() || busy[0] || busy[1];
                });
                if (stopping && ready.empty() && active.empty() && !busy[0] && !busy[1]) break;
                while (!ready.empty()) {
                    active.push_back(ready.front());
                    ready.pop_front();
                }
            }
            // a failed step: wait until neither slot's replay can still touch
            // a job's buffers, then release every job in flight
            auto fail_all = [&](const std::string &what) {
                xpr_batch::engine_sync(eng);
                busy[0] = busy[1] = false;
                std::lock_guard<std::mutex> lk(m);
                engine_error = what;
                for (int k = 0; k < 2; ++k) {
                    for (auto &r : members[k]) r->gpu_done = true;
                    for (auto &r : finishing[k]) r->gpu_done = true;
                    members[k].clear();
                    finishing[k].clear();
                }
                for (auto &r : active) r->gpu_done = true;
                active.clear();
                cv.notify_all();
            };
            // the slot's previous step: account its time, release finished jobs
            if (busy[slot]) {
                float ms = 0.0f;
                const bool ok = xpr_batch::engine_wait(eng, slot, &ms) == XPR_OK;
                if (!ok) {
                    fail_all(std::string("batch step: ") + xpr_last_error());
                    continue;
                }
                busy[slot] = false;
                std::lock_guard<std::mutex> lk(m);
                for (auto &r : members[slot]) r->step_ms += ms;
                for (auto &r : finishing[slot]) r->gpu_done = true;
                members[slot].clear();
                finishing[slot].clear();
                cv.notify_all();
            }
            if (active.empty()) {
                slot ^= 1;
                continue;
            }
            // plan: shortest remaining work first, each job at least one quantum
            std::sort(active.begin(), active.end(), [](const auto &a, const auto &b) {
                return a->bj->remaining() < b->bj->remaining();
            });
            xpr_batch::Table &t = *xpr_batch::engine_table(eng, slot);
            std::memset(t.count, 0, sizeof(t.count));
            double left = step_target;
            std::vector<std::shared_ptr<JobRun>> keep;
            for (auto &r : active) {
                bool room = true;
                for (int ph = 0; ph < xpr_batch::kPhases; ++ph)
                    room = room && t.count[ph] < uint32_t(xpr_batch::kMaxItems);
                if (!room) {
                    keep.push_back(r);
                    continue;
                }
                left -= r->bj->plan(t, std::max(left, r->bj->quantum_cost()));
                members[slot].push_back(r);
                if (r->bj->complete)
                    finishing[slot].push_back(r);
                else
                    keep.push_back(r);
            }
            for (int ph = 0; ph < xpr_batch::kPhases; ++ph) {
                uint32_t first = 0;
                for (uint32_t i = 0; i < t.count[ph]; ++i) {
                    t.items[ph][i].first = first;
                    first += t.items[ph][i].units;
                }
                t.total[ph] = first;
            }
            active.swap(keep);
            if (xpr_batch::engine_launch(eng, slot) != XPR_OK) {
                fail_all(std::string("batch step: ") + xpr_last_error());
                continue;
            }
            busy[slot] = true;
            slot ^= 1;
        }
        xpr_batch::engine_destroy(eng);
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
    if (job->acc_bits != XPR_ACC_U32 && job->acc_bits != XPR_ACC_U64 && job->acc_bits != XPR_ACC_F32)
        return xprint_fail(XPR_E_INVALID, "job: acc_bits must be XPR_ACC_U32, XPR_ACC_U64 or XPR_ACC_F32");
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
