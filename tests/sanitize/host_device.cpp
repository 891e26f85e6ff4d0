// Host stand-in for the device half of the C-ABI (include/xpr_hip.h), for the
// sanitizer builds of the host engine and server only (tests/sanitize/).  It
// is never part of the library: it computes nothing the reconstruction needs,
// it only reproduces the ORDERING the HIP runtime guarantees, so that
// AddressSanitizer and ThreadSanitizer see the host code's buffer traffic
// the way the device would perform it:
//
//   * a stream is a worker thread draining a FIFO; every enqueued operation
//     runs there, after the stream's earlier ones;
//   * an event is a completion marker recorded on a stream; waiting on it (or
//     making another stream wait) orders everything after that point;
//   * copies from or to pinned host memory (xpr_host_alloc) are asynchronous
//     like hipMemcpyAsync; copies with pageable host memory complete before
//     the call returns, as the HIP runtime stages them;
//   * a "kernel" is a host loop on the stream's thread that reads and writes
//     exactly the buffers, with the extents, that the real kernel reads and
//     writes (values are placeholders; buffers the kernel only adds to with
//     atomics get atomic adds), so a host access that races with it, or an
//     extent past an allocation, is reported.
//
// Device memory is host memory (malloc), so ASan checks the extents too.
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <functional>
#include <mutex>
#include <map>
#include <string>
#include <thread>

#include "../../include/xpr_hip.h"
#include "../../xpulse_recon_amd/csrc/batch.hpp"

namespace {

thread_local std::string g_error;

int fail(int code, const char *msg)
{
    g_error = msg;
    return code;
}

double now_ms()
{
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

struct Stream {
    std::mutex m;
    std::condition_variable cv;
    std::deque<std::function<void()>> q;
    uint64_t queued = 0, done = 0;
    bool stop = false;
    std::thread worker;

    Stream() : worker([this] { run(); }) {}
    ~Stream()
    {
        {
            std::lock_guard<std::mutex> lk(m);
            stop = true;
        }
        cv.notify_all();
        worker.join();
    }
    void run()
    {
        std::unique_lock<std::mutex> lk(m);
        for (;;) {
            cv.wait(lk, [this] { return stop || !q.empty(); });
            if (q.empty()) return;
            auto f = std::move(q.front());
            q.pop_front();
            lk.unlock();
            f();
            lk.lock();
            ++done;
            cv.notify_all();
        }
    }
    void push(std::function<void()> f)
    {
        std::lock_guard<std::mutex> lk(m);
        q.push_back(std::move(f));
        ++queued;
        cv.notify_all();
    }
    void sync()
    {
        std::unique_lock<std::mutex> lk(m);
        const uint64_t target = queued;
        cv.wait(lk, [&] { return done >= target; });
    }
};

Stream &null_stream()
{
    static Stream s;
    return s;
}

Stream &S(xpr_stream s) { return s ? *static_cast<Stream *>(s) : null_stream(); }

struct Event {
    std::mutex m;
    std::condition_variable cv;
    uint64_t recorded = 0, completed = 0;   // generations
    double t = 0.0;
};

std::mutex g_pinned_m;                       // guards both allocation maps
std::map<const char *, size_t> g_pinned;     // pinned host blocks: start -> bytes
std::map<const void *, size_t> g_device;     // "device" allocations: start -> bytes
uint64_t g_alloc_bytes = 0;

// whether p lies in a block of pinned host memory
bool pinned(const void *p)
{
    const char *c = (const char *)p;
    std::lock_guard<std::mutex> lk(g_pinned_m);
    auto it = g_pinned.upper_bound(c);
    if (it == g_pinned.begin()) return false;
    --it;
    return c < it->first + it->second;
}

size_t dtype_bytes(int dt) { return dt == XPR_F32 ? 4 : 2; }
size_t acc_bytes(int acc) { return acc == XPR_ACC_U64 ? 8 : 4; }

// read (and write back when w) one byte in four of [p, p + bytes)
void touch(const void *p, size_t bytes, bool w)
{
    if (!p || !bytes) return;
    volatile unsigned char *b = (volatile unsigned char *)p;
    unsigned acc = 0;
    for (size_t i = 0; i < bytes; i += 4) {
        acc += b[i];
        if (w) b[i] = (unsigned char)(acc & 0x7f);
    }
}

// atomic adds to every 32-bit word of [p, p + bytes): the accesses of a kernel
// that only adds to a buffer with atomics (the forward's detector sums, the
// backward's empty-voxel counter), which concurrent launches may share
void touch_atomic(void *p, size_t bytes)
{
    if (!p || !bytes) return;
    uint32_t *w = (uint32_t *)p;
    for (size_t i = 0; i < bytes / 4; ++i) __atomic_fetch_add(&w[i], 0u, __ATOMIC_RELAXED);
}

uint64_t slab_voxels(const xpr_proj_args *a)
{
    return uint64_t(a->vmax[0] - a->vmin[0]) * uint64_t(a->vmax[1] - a->vmin[1]) *
           uint64_t(a->vmax[2] - a->vmin[2]);
}

uint64_t det_samples(const xpr_proj_args *a)
{
    return uint64_t(a->n_views) * uint64_t(a->line_rows) * a->det_width;
}

} // namespace

// the library's internal error setter (C++ linkage), used by the host code
int xprint_fail(int code, const char *msg) { return fail(code, msg); }

extern "C" {

int xpr_abi_version(void) { return XPR_ABI_VERSION; }
const char *xpr_last_error(void) { return g_error.c_str(); }
int xpr_set_device(int device) { return device == 0 ? XPR_OK : fail(XPR_E_INVALID, "device"); }

int xpr_mem_info(size_t *free_bytes, size_t *total_bytes)
{
    // 4 GiB of "device memory" unless XPR_FAKE_DEVICE_GB says otherwise
    const char *e = std::getenv("XPR_FAKE_DEVICE_GB");
    const size_t total = size_t((e ? std::atof(e) : 4.0) * (1ull << 30));
    size_t used;
    {
        std::lock_guard<std::mutex> lk(g_pinned_m);
        used = g_alloc_bytes;
    }
    if (free_bytes) *free_bytes = used < total ? total - used : 0;
    if (total_bytes) *total_bytes = total;
    return XPR_OK;
}

int xpr_alloc(void **d, size_t bytes)
{
    if (!d) return fail(XPR_E_INVALID, "null");
    *d = std::malloc(bytes ? bytes : 1);
    if (!*d) return fail(XPR_E_NOMEM, "malloc");
    std::lock_guard<std::mutex> lk(g_pinned_m);
    g_device[*d] = bytes;
    g_alloc_bytes += bytes;
    return XPR_OK;
}
int xpr_free(void *d)
{
    if (!d) return XPR_OK;
    {
        std::lock_guard<std::mutex> lk(g_pinned_m);
        auto it = g_device.find(d);
        if (it == g_device.end()) return fail(XPR_E_INVALID, "free of an unknown device pointer");
        g_alloc_bytes -= it->second;
        g_device.erase(it);
    }
    std::free(d);
    return XPR_OK;
}
int xpr_host_alloc(void **h, size_t bytes)
{
    if (!h) return fail(XPR_E_INVALID, "null");
    *h = std::malloc(bytes ? bytes : 1);
    if (!*h) return fail(XPR_E_NOMEM, "malloc");
    std::lock_guard<std::mutex> lk(g_pinned_m);
    g_pinned[(const char *)*h] = bytes;
    return XPR_OK;
}
int xpr_host_free(void *h)
{
    {
        std::lock_guard<std::mutex> lk(g_pinned_m);
        g_pinned.erase((const char *)h);
    }
    std::free(h);
    return XPR_OK;
}
int xpr_alloc_async(void **d, size_t bytes, xpr_stream) { return xpr_alloc(d, bytes); }
int xpr_free_async(void *d, xpr_stream s)
{
    S(s).push([d] { xpr_free(d); });
    return XPR_OK;
}

static int copy(xpr_stream s, void *dst, const void *src, size_t bytes, const void *host)
{
    if (pinned(host)) {
        S(s).push([=] { std::memcpy(dst, src, bytes); });
    } else {
        // pageable host memory: the runtime stages the copy before returning
        Stream &st = S(s);
        st.push([=] { std::memcpy(dst, src, bytes); });
        st.sync();
    }
    return XPR_OK;
}
int xpr_upload(xpr_stream s, void *dst, const void *src, size_t bytes) { return copy(s, dst, src, bytes, src); }
int xpr_download(xpr_stream s, void *dst, const void *src, size_t bytes) { return copy(s, dst, src, bytes, dst); }

int xpr_stream_create(xpr_stream *s)
{
    *s = new Stream();
    return XPR_OK;
}
int xpr_stream_destroy(xpr_stream s)
{
    if (s) {
        S(s).sync();
        delete static_cast<Stream *>(s);
    }
    return XPR_OK;
}
int xpr_stream_sync(xpr_stream s)
{
    S(s).sync();
    return XPR_OK;
}

int xpr_event_create(xpr_event *e)
{
    *e = new Event();
    return XPR_OK;
}
int xpr_event_destroy(xpr_event e)
{
    delete static_cast<Event *>(e);
    return XPR_OK;
}
int xpr_event_record(xpr_event e, xpr_stream s)
{
    Event *ev = static_cast<Event *>(e);
    uint64_t gen;
    {
        std::lock_guard<std::mutex> lk(ev->m);
        gen = ++ev->recorded;
    }
    S(s).push([ev, gen] {
        std::lock_guard<std::mutex> lk(ev->m);
        ev->t = now_ms();
        if (ev->completed < gen) ev->completed = gen;
        ev->cv.notify_all();
    });
    return XPR_OK;
}
int xpr_stream_wait_event(xpr_stream s, xpr_event e)
{
    Event *ev = static_cast<Event *>(e);
    uint64_t gen;
    {
        std::lock_guard<std::mutex> lk(ev->m);
        gen = ev->recorded;
    }
    S(s).push([ev, gen] {
        std::unique_lock<std::mutex> lk(ev->m);
        ev->cv.wait(lk, [&] { return ev->completed >= gen; });
    });
    return XPR_OK;
}
int xpr_event_sync(xpr_event e)
{
    Event *ev = static_cast<Event *>(e);
    std::unique_lock<std::mutex> lk(ev->m);
    const uint64_t gen = ev->recorded;
    ev->cv.wait(lk, [&] { return ev->completed >= gen; });
    return XPR_OK;
}
int xpr_event_elapsed_ms(xpr_event a, xpr_event b, float *ms)
{
    Event *ea = static_cast<Event *>(a), *eb = static_cast<Event *>(b);
    double ta, tb;
    {
        std::lock_guard<std::mutex> lk(ea->m);
        ta = ea->t;
    }
    {
        std::lock_guard<std::mutex> lk(eb->m);
        tb = eb->t;
    }
    *ms = float(tb - ta);
    return XPR_OK;
}

int xpr_fill_u32(xpr_stream s, uint32_t *d, uint32_t v, uint64_t n)
{
    S(s).push([=] { for (uint64_t i = 0; i < n; ++i) d[i] = v; });
    return XPR_OK;
}
int xpr_fill_u64(xpr_stream s, uint64_t *d, uint64_t v, uint64_t n)
{
    S(s).push([=] { for (uint64_t i = 0; i < n; ++i) d[i] = v; });
    return XPR_OK;
}
int xpr_fill(xpr_stream s, void *d, int dt, float, uint64_t n)
{
    S(s).push([=] { std::memset(d, 0, n * dtype_bytes(dt)); });
    return XPR_OK;
}
int xpr_convert(xpr_stream s, void *dst, int ddt, const void *src, int sdt, uint64_t n)
{
    S(s).push([=] {
        touch(src, n * dtype_bytes(sdt), false);
        touch(dst, n * dtype_bytes(ddt), true);
    });
    return XPR_OK;
}
int xpr_forward(xpr_stream s, const void *vol, int vdt, void *sums, int acc, const float *pd,
                const xpr_proj_args *a)
{
    if (!a || !vol || !sums || !pd) return fail(XPR_E_INVALID, "null");
    const xpr_proj_args c = *a;
    S(s).push([=] {
        touch(pd, size_t(c.n_views) * c.modules * 20 * 4, false);
        touch(vol, slab_voxels(&c) * dtype_bytes(vdt), false);
        touch_atomic(sums, det_samples(&c) * acc_bytes(acc));
    });
    return XPR_OK;
}
int xpr_error(xpr_stream s, const void *img, int idt, const void *sums, int acc, void *err,
              int edt, uint64_t n, float)
{
    S(s).push([=] {
        touch(img, n * dtype_bytes(idt), false);
        touch(sums, n * acc_bytes(acc), false);
        touch(err, n * dtype_bytes(edt), true);
    });
    return XPR_OK;
}
int xpr_backward_max(xpr_stream s, void *vol, int vdt, const void *err, int edt, const float *pd,
                     const xpr_proj_args *a, uint64_t *empty, uint32_t *max_bits)
{
    if (!a || !vol || !err || !pd) return fail(XPR_E_INVALID, "null");
    const xpr_proj_args c = *a;
    S(s).push([=] {
        touch(pd, size_t(c.n_views) * c.modules * 20 * 4, false);
        touch(err, det_samples(&c) * dtype_bytes(edt), false);
        touch(vol, slab_voxels(&c) * dtype_bytes(vdt), true);
        if (empty) touch_atomic(empty, 8);
        if (max_bits) touch_atomic(max_bits, 4 * XPR_MAX_SLOTS);
    });
    return XPR_OK;
}

int xpr_backward(xpr_stream s, void *vol, int vdt, const void *err, int edt, const float *pd,
                 const xpr_proj_args *a, uint64_t *empty)
{
    return xpr_backward_max(s, vol, vdt, err, edt, pd, a, empty, nullptr);
}
int xpr_tune_forward(xpr_stream s, const void *vol, int vdt, void *sums, int acc, const float *pd,
                     const xpr_proj_args *a, int *choice)
{
    if (choice) *choice = 1536;
    return xpr_forward(s, vol, vdt, sums, acc, pd, a);
}
int xpr_tune_backward(xpr_stream s, void *vol, int vdt, const void *err, int edt, const float *pd,
                      const xpr_proj_args *a, int *choice)
{
    if (choice) *choice = 8;
    return xpr_backward(s, vol, vdt, err, edt, pd, a, nullptr);
}
int xpr_tune_reset(void) { return XPR_OK; }
int xpr_q22_to_float(xpr_stream s, const void *sums, int acc, float *out, uint64_t n)
{
    S(s).push([=] {
        touch(sums, n * acc_bytes(acc), false);
        touch(out, n * 4, true);
    });
    return XPR_OK;
}
int xpr_scale(xpr_stream s, float *x, uint64_t n, float)
{
    S(s).push([=] { touch(x, n * 4, true); });
    return XPR_OK;
}
int xpr_minmax_init(xpr_stream s, int32_t *mm)
{
    S(s).push([=] {
        mm[0] = 0x7f7fffff;
        mm[1] = 0x00800000;
    });
    return XPR_OK;
}
int xpr_minmax_normal(xpr_stream s, const float *d, uint64_t n, int32_t *mm)
{
    S(s).push([=] {
        touch(d, n * 4, false);
        touch(mm, 8, true);
        mm[0] = 0;                      // order keys of 0.0 and 1.0
        mm[1] = 0x3f800000;
    });
    return XPR_OK;
}
int xpr_minmax_read(const int32_t *mm, float *mn, float *mx)
{
    for (int i = 0; i < 2; ++i) {
        int32_t k = mm[i];
        int32_t b = k < 0 ? (k ^ 0x7fffffff) : k;
        float f;
        std::memcpy(&f, &b, 4);
        (i == 0 ? *mn : *mx) = f;
    }
    return XPR_OK;
}
int xpr_preprocess(xpr_stream s, const float *in, void *out, int odt, uint64_t n, int, int,
                   float, float)
{
    S(s).push([=] {
        touch(in, n * 4, false);
        touch(out, n * dtype_bytes(odt), true);
    });
    return XPR_OK;
}

int xpr_preprocess_views(xpr_stream s, const float *in, uint64_t stride, void *out, int odt,
                         uint64_t n, uint32_t nv, int, int, float, float)
{
    if (stride < n) return fail(XPR_E_INVALID, "views overlap");
    S(s).push([=] {
        for (uint32_t j = 0; j < nv; ++j) {
            touch(in + uint64_t(j) * stride, n * 4, false);
            touch(static_cast<char *>(out) + uint64_t(j) * n * dtype_bytes(odt), n * dtype_bytes(odt), true);
        }
    });
    return XPR_OK;
}

} // extern "C"

// The batch engine: one stream, two host tables; a launch copies its slot's
// table on the stream (as the graph's first node does) and touches every
// item's buffers.
namespace xpr_batch {

struct Engine {
    Stream stream;
    Table *host[2] = {nullptr, nullptr};
    xpr_event stop[2] = {nullptr, nullptr};
    bool pending[2] = {false, false};
};

uint32_t forward_units(const xpr_proj_args &a)
{
    const int64_t nx = a.vmax[0] - a.vmin[0], nl = a.vmax[1] - a.vmin[1], nk = a.vmax[2] - a.vmin[2];
    return (uint32_t)(((nx + 63) / 64) * ((nk + 15) / 16) * ((nl + 3) / 4));
}

uint32_t backward_units(const xpr_proj_args &a)
{
    const int64_t nx = a.vmax[0] - a.vmin[0], nl = a.vmax[1] - a.vmin[1], nk = a.vmax[2] - a.vmin[2];
    const int64_t rows = ((nl + 3) / 4) * nk;
    return (uint32_t)(((rows + 3) / 4) * ((nx + 127) / 128));
}

int engine_create(Engine **e)
{
    Engine *g = new Engine();
    for (int k = 0; k < 2; ++k) {
        void *p = nullptr;
        xpr_host_alloc(&p, sizeof(Table));
        std::memset(p, 0, sizeof(Table));
        g->host[k] = (Table *)p;
        xpr_event_create(&g->stop[k]);
    }
    *e = g;
    return XPR_OK;
}

Table *engine_table(Engine *e, int slot) { return e->host[slot]; }

int engine_launch(Engine *e, int slot)
{
    const Table *src = e->host[slot];
    e->stream.push([src] {
        Table t;
        std::memcpy(&t, src, sizeof(Table));
        for (int ph = 0; ph < kPhases; ++ph)
            for (uint32_t i = 0; i < t.count[ph] && i < (uint32_t)kMaxItems; ++i) {
                const Item &it = t.items[ph][i];
                switch (ph) {
                case kConvert:
                    touch(it.src, it.n * dtype_bytes(it.dt_src), false);
                    touch(it.dst, it.n * dtype_bytes(it.dt_dst), true);
                    break;
                case kForward:
                case kForwardF16:
                case kForwardBF16:
                    touch(it.dst, slab_voxels(&it.a) * dtype_bytes(it.dt_dst), false);
                    touch_atomic(it.sums, det_samples(&it.a) * 8);
                    break;
                case kError:
                    touch(it.src, it.n * dtype_bytes(it.dt_src), false);
                    touch(it.sums, it.n * 8, false);
                    touch(it.dst, it.n * dtype_bytes(it.dt_dst), true);
                    break;
                case kBackward:
                case kBackwardF16:
                case kBackwardBF16:
                    touch(it.err, det_samples(&it.a) * dtype_bytes(it.dt_aux), false);
                    touch(it.dst, slab_voxels(&it.a) * dtype_bytes(it.dt_dst), true);
                    break;
                default:
                    touch(it.dst, it.n * 8, true);
                    break;
                }
            }
    });
    xpr_event_record(e->stop[slot], &e->stream);
    e->pending[slot] = true;
    return XPR_OK;
}

int engine_wait(Engine *e, int slot, float *ms)
{
    if (e->pending[slot]) xpr_event_sync(e->stop[slot]);
    e->pending[slot] = false;
    if (ms) *ms = 0.0f;
    return XPR_OK;
}

void engine_sync(Engine *e) { e->stream.sync(); }

void engine_destroy(Engine *e)
{
    e->stream.sync();
    for (int k = 0; k < 2; ++k) {
        xpr_host_free(e->host[k]);
        xpr_event_destroy(e->stop[k]);
    }
    delete e;
}

} // namespace xpr_batch
