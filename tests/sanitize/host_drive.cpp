// Driver of the sanitizer builds (tests/test_host_sanitizers.py): runs the
// host engine's schedules and the server through the public C-ABI, linked
// against the host stand-in of the device entry points (host_device.cpp).
// The values it computes are placeholders; what matters is that every run
// ends, returns success, and that the sanitizer reports nothing.
//
//   host_drive <dir>          <dir> holds proj_*.tif written by the test
//   host_drive --race         self-check: must be reported by ThreadSanitizer
//   host_drive --overflow     self-check: must be reported by AddressSanitizer
//
// Each step prints its name first, so a report names the step it came from.
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/xpr_hip.h"

namespace {

std::atomic<int> failures{0};          // expect() runs on several threads

void step(const char *name) { std::printf("step: %s\n", name), std::fflush(stdout); }

void expect(bool ok, const char *what)
{
    if (!ok) {
        std::printf("FAILED: %s (%s)\n", what, xpr_last_error());
        ++failures;
    }
}

std::string params(const std::string &dir, const char *method, const char *store, int slab,
                   double budget_gb, const char *storage, int sit)
{
    char buf[2048];
    std::snprintf(buf, sizeof buf,
                  "{\"recon\": {\"input\": \"%s/in\", \"output\": \"%s/out_%s_%s_%s\", "
                  "\"method\": \"%s\", \"it\": 2, \"sit\": %d},"
                  " \"xpr\": {\"layer_store\": \"%s\", \"slab_height\": %d, "
                  "\"device_budget_gb\": %g, \"storage\": \"%s\", \"jitter_seed\": 5, "
                  "\"quiet\": true}}",
                  dir.c_str(), dir.c_str(), method, store, storage, method, sit, store, slab,
                  budget_gb, storage);
    return buf;
}

} // namespace

int main(int argc, char **argv)
{
    if (argc < 2) {
        std::fprintf(stderr, "usage: host_drive <dir>\n");
        return 2;
    }
    // self-checks of the stand-in: each breaks a rule on purpose and must be
    // reported by the sanitizer of its build (the test asserts the report)
    if (std::strcmp(argv[1], "--race") == 0) {
        // the host rewrites a pinned buffer while its upload may still run
        void *h = nullptr, *d = nullptr;
        xpr_stream s = nullptr;
        xpr_host_alloc(&h, 1 << 16);
        xpr_alloc(&d, 1 << 16);
        xpr_stream_create(&s);
        std::memset(h, 1, 1 << 16);
        xpr_upload(s, d, h, 1 << 16);
        std::memset(h, 2, 1 << 16);
        xpr_stream_sync(s);
        xpr_stream_destroy(s);
        return 0;
    }
    if (std::strcmp(argv[1], "--overflow") == 0) {
        // a forward whose slab extent exceeds the volume's allocation
        void *vol = nullptr, *sums = nullptr, *pd = nullptr;
        xpr_alloc(&vol, 8 * 8 * 8 * 4 - 4);
        xpr_alloc(&sums, 8 * 8 * 8);
        xpr_alloc(&pd, 20 * 4);
        xpr_proj_args a;
        std::memset(&a, 0, sizeof a);
        a.det_width = 8;
        a.n_views = 1;
        a.modules = 1;
        a.vmax[0] = a.vmax[1] = a.vmax[2] = 8;
        a.line_rows = 8;
        xpr_forward(nullptr, vol, XPR_F32, sums, XPR_ACC_U64, (const float *)pd, &a);
        xpr_stream_sync(nullptr);
        return 0;
    }
    const std::string dir = argv[1];
    const int W = 24, H = 16, N = 24;             // the TIFFs the test wrote

    step("template");
    {
        std::vector<char> t(1 << 16);
        expect(xpr_template_json(t.data(), t.size()) == 0, "template");
        char tiny[8];
        expect(xpr_template_json(tiny, sizeof tiny) == -1, "template into a short buffer");
    }

    step("bad parameters");
    {
        char rec[256];
        expect(xpr_reconstruct("{\"recon\": {\"it\": \"x\"}}", rec, sizeof rec) == -2, "bad type");
        expect(xpr_reconstruct("{\"recon\": ", rec, sizeof rec) == -2, "truncated JSON");
        expect(xpr_reconstruct("{\"recon\": {\"method\": \"hartmann\"}}", rec, sizeof rec) != 0,
               "unsupported method");
    }

    // every schedule from the TIFF files, with each layer store
    struct Run { const char *method, *store; int slab; double budget; const char *storage; };
    const Run runs[] = {
        {"naive", "auto", 0, 0.0, "f32"},
        {"method_1a", "ram", 4, 0.0, "f32"},
        {"method_1a", "zstd", 4, 0.0, "f16"},
        {"method_1a", "device", 4, 0.0, "f32"},
        {"method_1b", "ram", 4, 0.0, "f32"},
        {"method_1b", "zstd", 8, 0.0, "bf16"},
        {"method_1c", "ram", 4, 0.0, "f32"},
        {"method_2", "auto", 0, 0.00004, "f32"},   // several chunks
    };
    for (const Run &r : runs) {
        char name[128];
        std::snprintf(name, sizeof name, "reconstruct %s store %s slab %d %s", r.method, r.store,
                      r.slab, r.storage);
        step(name);
        std::vector<char> rec(1 << 16);
        const std::string p = params(dir, r.method, r.store, r.slab, r.budget, r.storage, 4);
        expect(xpr_reconstruct(p.c_str(), rec.data(), rec.size()) == 0, name);
    }

    // in-memory entry, two threads at once (the engine's state is per call)
    step("reconstruct_mem from two threads");
    {
        std::vector<float> proj(size_t(N) * H * W, 0.5f);
        std::vector<float> vol_a(size_t(W) * W * H), vol_b(size_t(W) * W * H);
        int rc_a = -9, rc_b = -9;
        const std::string pa = params(dir, "method_1b", "ram", 4, 0.0, "f32", 4);
        const std::string pb = params(dir, "method_2", "auto", 0, 0.00004, "f32", 4);
        std::thread ta([&] { rc_a = xpr_reconstruct_mem(pa.c_str(), proj.data(), W, H, N, vol_a.data(), nullptr, 0); });
        std::thread tb([&] { rc_b = xpr_reconstruct_mem(pb.c_str(), proj.data(), W, H, N, vol_b.data(), nullptr, 0); });
        ta.join();
        tb.join();
        expect(rc_a == 0 && rc_b == 0, "reconstruct_mem");
    }

    step("plan_chunks");
    {
        int64_t chunks[6 * 64], n = 0;
        const std::string p = params(dir, "method_2", "auto", 0, 0.00004, "f32", 4);
        expect(xpr_plan_chunks(p.c_str(), W, H, N, chunks, 64, &n) == XPR_OK && n > 1, "plan");
        expect(xpr_plan_chunks(p.c_str(), W, H, N, chunks, 1, &n) == XPR_E_INVALID, "short plan buffer");
    }

    // the server, both executions: jobs submitted from two threads while
    // others run, a budget that forces waiting, one job streamed
    for (const char *mode : {"streams", "batch"}) {
        std::string name = std::string("server ") + mode;
        step(name.c_str());
        setenv("XPR_SERVE_MODE", mode, 1);
        setenv("XPR_SERVE_STEP", "2e4", 1);
        const int jobs = 8;
        std::vector<std::vector<float>> proj(jobs), vol(jobs);
        std::vector<xpr_job> spec(jobs);
        for (int j = 0; j < jobs; ++j) {
            const int w = j % 2 ? 16 : 24, n = j % 3 ? 12 : 24;
            proj[j].assign(size_t(n) * w * w, 0.25f + 0.01f * j);
            vol[j].assign(size_t(w) * w * w, 0.0f);
            xpr_job &s = spec[j];
            std::memset(&s, 0, sizeof s);
            s.geometry.so = 100.0;
            s.geometry.sd = 1000.0;
            s.geometry.angle = 360.0;
            s.geometry.px = 0.1;
            s.geometry.module_number = 1;
            s.geometry.dwidth = w;
            s.geometry.dheight = w;
            s.geometry.projections = n;
            s.it = 1;
            s.sit = 3;
            s.weight = 1.0f;
            s.weight_factor = 0.9f;
            s.mlog = 1;
            s.storage = j % 3 == 2 ? XPR_F16 : XPR_F32;
            s.final_fp32 = 1;
            s.acc_bits = j == 5 ? XPR_ACC_U32 : XPR_ACC_U64;
            s.jitter = 1;
            s.jitter_seed = 100 + j;
            s.slab_height = 4;
            s.projections = proj[j].data();
            s.volume_out = vol[j].data();
        }
        uint64_t one = 0;
        expect(xpr_job_device_bytes(&spec[0], &one) == XPR_OK && one > 0, "job bytes");
        xpr_server *sv = nullptr;
        expect(xpr_server_create(&sv, 3 * one, 3, 0) == XPR_OK, "server create");
        std::vector<uint64_t> ids(jobs, 0);
        auto submit = [&](int from, int step_) {
            for (int j = from; j < jobs; j += step_)
                expect(xpr_server_submit(sv, &spec[j], &ids[j]) == XPR_OK, "submit");
        };
        std::thread t0(submit, 0, 2), t1(submit, 1, 2);
        t0.join();
        t1.join();
        for (int j = 0; j < jobs; ++j) {
            xpr_job_stats st;
            expect(xpr_server_wait(sv, ids[j], &st) == XPR_OK && st.status == XPR_OK, "job");
        }
        expect(xpr_server_destroy(sv) == XPR_OK, "server destroy");
        // a job larger than the whole budget runs streamed
        uint64_t id = 0;
        xpr_job big = spec[0];
        expect(xpr_server_create(&sv, one / 2, 2, 0) == XPR_OK, "server create (small budget)");
        expect(xpr_server_submit(sv, &big, &id) == XPR_OK, "submit streamed");
        xpr_job_stats st;
        expect(xpr_server_wait(sv, id, &st) == XPR_OK && st.status == XPR_OK && st.streamed == 1,
               "streamed job");
        expect(xpr_server_destroy(sv) == XPR_OK, "server destroy");
    }

    std::printf("host_drive: %d failure(s)\n", failures.load());
    return failures.load() ? 1 : 0;
}
