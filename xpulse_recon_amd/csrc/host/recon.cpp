#include "recon.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <functional>
#include <iostream>
#include <mutex>
#include <set>
#include <thread>

#include "io.hpp"

namespace fs = std::filesystem;

namespace xpr {

void check(int rc, const char *what)
{
    if (rc != XPR_OK) throw XprFailure(std::string(what) + ": " + xpr_last_error());
}

// ---------------------------------------------------------------------------
// buffers
// ---------------------------------------------------------------------------
namespace {
thread_local xpr_stream t_alloc_stream = nullptr;
}
AllocStreamScope::AllocStreamScope(xpr_stream s) : prev_(t_alloc_stream) { t_alloc_stream = s; }
AllocStreamScope::~AllocStreamScope() { t_alloc_stream = prev_; }

DevBuf::~DevBuf() { reset(); }
DevBuf &DevBuf::operator=(DevBuf &&o) noexcept
{
    if (this != &o) {
        reset();
        p_ = o.p_;
        n_ = o.n_;
        s_ = o.s_;
        o.p_ = nullptr;
        o.n_ = 0;
    }
    return *this;
}
void DevBuf::alloc(size_t bytes)
{
    reset();
    s_ = t_alloc_stream;
    if (s_) {
        check(xpr_alloc_async(&p_, bytes, s_), "device allocation");
        check(xpr_stream_sync(s_), "sync");
    } else {
        check(xpr_alloc(&p_, bytes), "device allocation");
    }
    n_ = bytes;
}
void DevBuf::reset()
{
    if (p_) {
        if (s_) {
            xpr_stream_sync(s_);            // work using the buffer is done
            xpr_free_async(p_, s_);
        } else {
            xpr_free(p_);
        }
    }
    p_ = nullptr;
    n_ = 0;
}
HostBuf::~HostBuf()
{
    if (p_) xpr_host_free(p_);
}
void HostBuf::alloc(size_t bytes)
{
    if (p_) xpr_host_free(p_);
    p_ = nullptr;
    check(xpr_host_alloc(&p_, bytes), "pinned host allocation");
    n_ = bytes;
}

namespace {

int io_threads(const Parameters &p)
{
    int t = p.xpr.threads > 0 ? p.xpr.threads : int(std::thread::hardware_concurrency());
    return std::max(1, std::min(t, 16));
}

// parallel for over [0, n) on up to `threads` host threads
void parallel_for(int64_t n, int threads, const std::function<void(int64_t)> &f)
{
    if (n <= 0) return;
    threads = int(std::min<int64_t>(threads, n));
    if (threads <= 1) {
        for (int64_t i = 0; i < n; ++i) f(i);
        return;
    }
    std::atomic<int64_t> next{0};
    std::vector<std::thread> pool;
    std::exception_ptr err;
    std::mutex m;
    for (int t = 0; t < threads; ++t)
        pool.emplace_back([&] {
            for (int64_t i; (i = next.fetch_add(1)) < n;) {
                try {
                    f(i);
                } catch (...) {
                    std::lock_guard<std::mutex> g(m);
                    if (!err) err = std::current_exception();
                }
            }
        });
    for (auto &th : pool) th.join();
    if (err) std::rethrow_exception(err);
}

double now_ms()
{
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

struct Event {
    xpr_event e = nullptr;
    Event() { check(xpr_event_create(&e), "event"); }
    ~Event() { if (e) xpr_event_destroy(e); }
};

} // namespace

// ---------------------------------------------------------------------------
// Dataset
// ---------------------------------------------------------------------------
Dataset::Dataset(Parameters &p) : p_(p)
{
    if (p.xpr.mem_projections || p.xpr.shape_only) {   // in-process job or planning
        dwidth_ = p.xpr.mem_width;
        dheight_ = p.xpr.mem_height;
        nproj_ = p.xpr.mem_count;
        if (dwidth_ <= 0 || dheight_ <= 0 || nproj_ <= 0)
            throw std::runtime_error("empty projection stack");
    }
    // collectFromDirectory + checkTiffFileInfos (dataset.hpp:214-269)
    const bool files = !p.xpr.mem_projections && !p.xpr.shape_only;
    if (files && !fs::is_directory(p.recon.input))
        throw std::runtime_error("input directory not found: " + p.recon.input);
    if (files)
    for (const std::string &f : list_tiffs(p.recon.input)) {
        TiffInfo i = tiff_info(f);
        if (i.sample_format != 3 || i.bits != 32)
            throw std::runtime_error("Trying to load a tiff with an incompatble format (should "
                                     "be floating-point 32bits)");
        if (i.width == 0 || i.height == 0)
            throw std::runtime_error("Trying to load a tiff with an incorrect size.");
        if (dwidth_ == 0 && dheight_ == 0) {
            dwidth_ = i.width;
            dheight_ = i.height;
        }
        if (i.width == dwidth_ && i.height == dheight_) files_.push_back(f);
    }
    if (files) {
        if (files_.empty()) throw std::runtime_error("No valid TIFF file found in the directory.");
        nproj_ = int64_t(files_.size());
    }
    if (nproj_ / p.recon.sit < 1)
        throw std::runtime_error("fewer projections than sub-iterations");
    if (!p.geometry.angle_list.empty() && int64_t(p.geometry.angle_list.size()) != nproj_)
        throw std::runtime_error("The size of the angle list is different from the number of "
                                 "projections");
    if (dwidth_ % p.detector.module_number != 0)
        std::cerr << "Warning: detector width is not a multiple of module_number\n";
    // Geometry (geometry.hpp:35-76)
    xpr_geometry_params g{};
    g.so = p.geometry.so;
    g.sd = p.geometry.sd;
    g.angle = p.geometry.angle;
    g.heli_offset = p.geometry.heli_offset;
    g.px = p.detector.px;
    g.rx = p.detector.rx;
    g.sx = p.detector.sx;
    g.sy = p.detector.sy;
    g.module_angle_offset = p.detector.module_angle_offset.value_or(0.0f);
    g.has_module_angle = p.detector.module_angle_offset.has_value();
    g.module_center_offset_z = p.detector.module_center_offset_z;
    g.roll = p.detector.roll;
    g.pitch = p.detector.pitch;
    g.yaw = p.detector.yaw;
    g.module_number = p.detector.module_number;
    g.dwidth = dwidth_;
    g.dheight = dheight_;
    g.projections = nproj_;
    g.fp64 = p.xpr.geometry_fp64 ? 1 : 0;
    std::vector<double> angles(p.geometry.angle_list.begin(), p.geometry.angle_list.end());
    g.angle_list = angles.empty() ? nullptr : angles.data();
    pd_.resize(size_t(nproj_ * p.detector.module_number * 20));
    xpr_grid grid;
    check(xpr_geometry(&g, pd_.data(), &grid), "geometry");
    vx_ = float(grid.vx);
    vwidth_ = grid.vwidth;
    vheight_ = grid.vheight;
    for (int i = 0; i < 3; ++i) orig_[i] = float(grid.orig[i]);
    if (!p.xpr.shape_only && (!p.xpr.mem_volume_out || p.xpr.layer_store == "zstd")) {
        fs::create_directories(p.recon.output);
        proj_dir_ = (fs::path(p.recon.output) / "projs").string();
    }
    if (!p.xpr.shape_only && !p.recon.proj_output.empty())
        fs::create_directories(p.recon.proj_output);
    if (!p.xpr.quiet && !p.xpr.shape_only)
        std::cout << "Volume dimension : " << vwidth_ << "x" << vheight_ << "x" << vwidth_ << "\n"
                  << "Pixel size : " << p.detector.px << ", Voxel size : " << vx_ << std::endl;
}

Dataset::~Dataset()
{
    for (auto &e : slot_done_)
        if (e) xpr_event_destroy(e);
    if (proj_dir_.empty()) return;
    std::error_code ec;
    fs::remove_all(proj_dir_, ec);     // cleanTempImages (data_loader.hpp:107-111)
}

std::vector<float> Dataset::raw_image(int64_t i) const
{
    if (p_.xpr.mem_projections) {
        const float *src = p_.xpr.mem_projections + size_t(i) * size_t(dwidth_ * dheight_);
        return std::vector<float>(src, src + dwidth_ * dheight_);
    }
    return tiff_read_f32(files_[size_t(i)]);
}

std::vector<int64_t> Dataset::sit_views(int64_t sit) const
{
    std::vector<int64_t> v;
    for (int64_t i = sit; i < nproj_ && int64_t(v.size()) < concurrent(); i += p_.recon.sit)
        v.push_back(i);
    return v;
}

std::vector<float> Dataset::sit_projdata(int64_t sit) const
{
    const int64_t m = p_.detector.module_number;
    std::vector<float> out;
    for (int64_t v : sit_views(sit))
        out.insert(out.end(), pd_.begin() + v * m * 20, pd_.begin() + (v + 1) * m * 20);
    return out;
}

const float *Dataset::raw(int64_t i) const
{
    if (p_.xpr.mem_projections)
        return p_.xpr.mem_projections + size_t(i) * size_t(dwidth_ * dheight_);
    return raw_[size_t(i)].data();
}

void Dataset::stream_views(xpr_stream s, const std::vector<int64_t> &views, int64_t row0,
                           int64_t rows, const std::function<void(const float *, size_t, size_t)> &fn)
{
    if (row0 < 0 || rows <= 0 || row0 + rows > dheight_)
        throw std::runtime_error("detector row window outside the projections");
    const size_t n = size_t(rows * dwidth_);                  // floats per view
    const size_t per = std::max<size_t>(1, (size_t(16) << 20) / (n * 4));
    const size_t chunk = std::min(per, views.size());
    for (auto &h : pin_)
        if (h.bytes() < chunk * n * 4) h.alloc(chunk * n * 4);
    for (auto &e : slot_done_)
        if (!e) check(xpr_event_create(&e), "event");
    AllocStreamScope scope(s);
    DevBuf dev[2] = {DevBuf(chunk * n * 4), DevBuf(chunk * n * 4)};
    for (size_t j0 = 0, k = 0; j0 < views.size(); j0 += chunk, k ^= 1) {
        const size_t c = std::min(chunk, views.size() - j0);
        check(xpr_event_sync(slot_done_[k]), "sync");       // slot's previous transfer is done
        for (size_t j = 0; j < c; ++j)
            std::memcpy(pin_[k].as<float>() + j * n, raw(views[j0 + j]) + row0 * dwidth_, n * 4);
        check(xpr_upload(s, dev[k].get(), pin_[k].get(), c * n * 4), "upload");
        check(xpr_event_record(slot_done_[k], s), "event");
        fn(dev[k].as<float>(), j0, c);
    }
}

void Dataset::load(xpr_stream s)
{
    if (loaded_) return;
    loaded_ = true;
    // the RAM layer store is one pinned block, allocated here once (slabs are
    // copied to and from it by DMA, and save_layer runs on several threads)
    if (ram_layers() && !ram_block_.get())
        ram_block_.alloc(size_t(vwidth_ * vwidth_ * vheight_) * 4);
    if (!p_.xpr.mem_projections) {
        raw_.assign(size_t(nproj_), {});
        parallel_for(nproj_, io_threads(p_), [&](int64_t i) { raw_[size_t(i)] = raw_image(i); });
    }
    if (!p_.recon.normalize) return;
    const size_t n = size_t(dwidth_ * dheight_);
    AllocStreamScope scope(s);
    DevBuf d_mm(8);
    check(xpr_minmax_init(s, d_mm.as<int32_t>()), "minmax init");
    if (device_io()) {                             // projections already in HBM
        for (int64_t i = 0; i < nproj_; ++i)
            check(xpr_minmax_normal(s, p_.xpr.mem_projections + size_t(i) * n, uint64_t(n),
                                    d_mm.as<int32_t>()), "minmax");
    } else {
        std::vector<int64_t> all(static_cast<size_t>(nproj_));
        for (int64_t i = 0; i < nproj_; ++i) all[size_t(i)] = i;
        stream_views(s, all, 0, dheight_, [&](const float *d, size_t, size_t c) {
            check(xpr_minmax_normal(s, d, uint64_t(c * n), d_mm.as<int32_t>()), "minmax");
        });
    }
    int32_t keys[2];
    check(xpr_download(s, keys, d_mm.get(), 8), "download");
    check(xpr_stream_sync(s), "sync");
    check(xpr_minmax_read(keys, &mn_, &mx_), "minmax read");
}

void Dataset::upload_sit_images(xpr_stream s, int64_t sit, int64_t row0, int64_t rows, void *dst,
                                int dt)
{
    load(s);
    const size_t n = size_t(rows * dwidth_);
    char *out = static_cast<char *>(dst);
    if (device_io()) {                              // preprocess straight from HBM
        if (row0 < 0 || rows <= 0 || row0 + rows > dheight_)
            throw std::runtime_error("detector row window outside the projections");
        // the sub-iteration's views s, s+S, ... in one launch
        const auto views = sit_views(sit);
        if (views.empty()) return;
        const uint64_t step = views.size() > 1 ? uint64_t(views[1] - views[0]) : 1;
        check(xpr_preprocess_views(s, p_.xpr.mem_projections + size_t(views[0] * dheight_ + row0) *
                                          size_t(dwidth_),
                                   step * uint64_t(dheight_ * dwidth_), out, dt, uint64_t(n),
                                   uint32_t(views.size()), p_.recon.normalize, p_.recon.mlog,
                                   mn_, mx_), "preprocess");
        return;
    }
    stream_views(s, sit_views(sit), row0, rows, [&](const float *d, size_t j0, size_t c) {
        check(xpr_preprocess(s, d, out + j0 * n * dtype_bytes(dt), dt, uint64_t(c * n),
                             p_.recon.normalize, p_.recon.mlog, mn_, mx_), "preprocess");
    });
}

void Dataset::save_layer(const float *data, int64_t layer, bool finalize)
{
    const size_t n = size_t(vwidth_ * vwidth_);
    if (finalize) {
        save_final_layers(data, layer, layer + 1);
        if (!ram_layers()) {
            std::error_code ec;
            fs::remove(numbered_path(p_.recon.output, "layer", layer, 5, "zstd"), ec);
        }
        return;
    }
    if (ram_layers()) {
        std::memcpy(ram_layer(layer), data, n * 4);
    } else {
        spill_write(numbered_path(p_.recon.output, "layer", layer, 5, "zstd"), data, n);
    }
}

void Dataset::load_layer(int64_t layer, float *out) const
{
    const size_t n = size_t(vwidth_ * vwidth_);
    if (ram_layers()) {
        std::memcpy(out, ram_layer(layer), n * 4);
        return;
    }
    const std::string z = numbered_path(p_.recon.output, "layer", layer, 5, "zstd");
    if (fs::exists(z)) {
        spill_read(z, out, n);
        return;
    }
    int64_t w = 0, h = 0;
    auto v = tiff_read_f32(numbered_path(p_.recon.output, "layer", layer, 5, "tif"), &w, &h);
    if (size_t(w * h) != n) throw std::runtime_error("layer file has the wrong size");
    std::memcpy(out, v.data(), n * 4);
}

void Dataset::ensure_output_dir()
{
    if (!proj_dir_.empty()) return;
    fs::create_directories(p_.recon.output);
    proj_dir_ = (fs::path(p_.recon.output) / "projs").string();
}

float *Dataset::ram_layer(int64_t layer) const
{
    if (!ram_block_.get() || layer < 0 || layer >= vheight_)
        throw std::logic_error("RAM layer store used before Dataset::load");
    return ram_block_.as<float>() + size_t(layer) * size_t(vwidth_ * vwidth_);
}

void Dataset::load_final_layer(int64_t layer, float *out) const
{
    const size_t n = size_t(vwidth_ * vwidth_);
    if (const float *sink = p_.xpr.mem_volume_out) {
        if (device_io()) {
            check(xpr_download(nullptr, out, sink + size_t(layer) * n, n * 4), "download");
            check(xpr_stream_sync(nullptr), "sync");
        } else {
            std::memcpy(out, sink + size_t(layer) * n, n * 4);
        }
        return;
    }
    int64_t w = 0, h = 0;
    auto v = tiff_read_f32(numbered_path(p_.recon.output, "layer", layer, 5, "tif"), &w, &h);
    if (size_t(w * h) != n) throw std::runtime_error("layer file has the wrong size");
    std::memcpy(out, v.data(), n * 4);
}

void Dataset::save_final_layers(const float *vol, int64_t l0, int64_t l1)
{
    const size_t n = size_t(vwidth_ * vwidth_);
    if (p_.xpr.mem_volume_out) {                  // in-process job: into the caller's buffer
        float *dst = p_.xpr.mem_volume_out + size_t(l0) * n;
        const float scale = mx_ - mn_;
        for (size_t k = 0; k < size_t(l1 - l0) * n; ++k)
            dst[k] = p_.recon.normalize ? vol[k] * scale : vol[k];
        return;
    }
    parallel_for(l1 - l0, io_threads(p_), [&](int64_t k) {
        const float *src = vol + size_t(k) * n;
        std::vector<float> tmp;
        if (p_.recon.normalize) {               // saveLayer(finalize), data_loader.hpp:135-139
            tmp.assign(src, src + n);
            for (float &v : tmp) v = v * (mx_ - mn_);
            src = tmp.data();
        }
        tiff_write_f32(numbered_path(p_.recon.output, "layer", l0 + k, 5, "tif"), src, vwidth_,
                       vwidth_);
    });
}

void Dataset::save_projection(const float *img, int64_t w, int64_t h, int64_t id) const
{
    // saveImage (data_loader.hpp:166-173)
    if (p_.recon.proj_output.empty())
        tiff_write_f32(numbered_path(p_.recon.output, "projection", id, 5, "tif"), img, w, h);
    else
        tiff_write_f32(numbered_path(p_.recon.proj_output, "p", id, 10, "tif"), img, w, h);
}

// ---------------------------------------------------------------------------
// Operators
// ---------------------------------------------------------------------------
void Operators::setImageParameters(int64_t w, int64_t h, int64_t n)
{
    a_.det_width = uint32_t(w);
    elements_ = uint64_t(w * h * n);                 // setImageParameters, :208-215
}

void Operators::setVolumeParameters(std::array<int, 3> mn, std::array<int, 3> mx, float voxel)
{
    for (int i = 0; i < 3; ++i) {
        a_.vmin[i] = mn[i];
        a_.vmax[i] = mx[i];
        a_.voxel[i] = voxel;
    }
}

void Operators::forward()
{
    check(xpr_forward(stream_, vol_, vol_dt_, sums_, narrow_ ? XPR_ACC_U32 : acc_, pd_, &a_), "forward");
}

void Operators::error()
{
    check(xpr_error(stream_, img_, img_dt_, sums_, narrow_ ? XPR_ACC_U32 : acc_, err_, err_dt_, elements_, w_),
          "error");
}

void Operators::backward()
{
    check(xpr_backward_max(stream_, vol_, vol_dt_, err_, err_dt_, pd_, &a_, nullptr, max_bits_), "backward");
}

// ---------------------------------------------------------------------------
// Reconstruction base
// ---------------------------------------------------------------------------
Reconstruction::Reconstruction(Parameters &p) : p_(p)
{
    check(xpr_set_device(p.xpr.device), "set device");
    ds_ = std::make_unique<Dataset>(p);
    check(xpr_stream_create(&stream_), "stream");
    check(xpr_stream_create(&copy_), "stream");
    check(xpr_stream_create(&copy_back_), "stream");
    size_t fr = 0, tot = 0;
    check(xpr_mem_info(&fr, &tot), "memory info");
    const uint64_t reserve = 512ull << 20;                 // gpu/context.hpp:104-108
    budget_ = p.xpr.device_budget_gb > 0.0 ? uint64_t(p.xpr.device_budget_gb * 1e9)
                                           : (fr > reserve ? fr - reserve : 0);
    // the arithmetic of the run: Q.22 sums and nearest-pixel sampling are the
    // reference's; float sums and bilinear sampling are opt-in
    note("accumulator", p.xpr.accumulator == XPR_ACC_F32 ? Json::string("f32")
                                                        : Json::number(p.xpr.accumulator));
    note("interp", Json::string(p.xpr.interp == XPR_INTERP_BILINEAR ? "bilinear" : "nearest"));
    if (p.xpr.jitter_seed >= 0) {
        rng_.seed(uint32_t(p.xpr.jitter_seed));
    } else {
        std::random_device rd;                              // src/reconstruction.hpp:31,36
        rng_.seed(rd());
    }
    note("method", Json::string(method_name(p.recon.method)));
    note("device_budget_bytes", Json::number(double(budget_)));
    Json g = Json::array();
    for (int64_t v : {ds_->vwidth(), ds_->vheight(), ds_->vwidth()}) g.arr.push_back(Json::number(double(v)));
    note("grid", g);
    note("voxel_size", Json::number(ds_->vx()));
    note("views_per_subiteration", Json::number(double(ds_->concurrent())));
}

Reconstruction::~Reconstruction()
{
    if (stream_) xpr_stream_destroy(stream_);
    if (copy_) xpr_stream_destroy(copy_);
    if (copy_back_) xpr_stream_destroy(copy_back_);
}

std::array<float, 3> Reconstruction::next_origin()
{
    // setOrigin(orig, {vx, vx, vx}) (src/reconstruction.hpp:268-272)
    std::array<float, 3> o = ds_->orig();
    Json j = Json::array();
    for (int i = 0; i < 3; ++i) {
        float d = 0.0f;
        if (p_.xpr.jitter) d = std::uniform_real_distribution<float>(0.0f, ds_->vx())(rng_);
        o[size_t(i)] += d;
        j.arr.push_back(Json::number(double(d)));
    }
    jitter_log_.arr.push_back(j);
    record_.obj.erase(std::remove_if(record_.obj.begin(), record_.obj.end(),
                                     [](const auto &kv) { return kv.first == "jitter"; }),
                      record_.obj.end());
    note("jitter", jitter_log_);
    return o;
}

void Reconstruction::check_memory(uint64_t need, const char *what)
{
    note("required_device_bytes", Json::number(double(need)));
    if (!p_.xpr.quiet)
        std::cout << "Checking GPU memory requierments : " << need / (1024 * 1024) << "Mo" << std::endl;
    if (need > budget_)
        throw std::runtime_error(std::string("Not enough GPU memory (") + what + ": " +
                                 std::to_string(need >> 20) + " MiB needed, budget " +
                                 std::to_string(budget_ >> 20) + " MiB)");
}

void Reconstruction::log_sit(int64_t mit, int64_t sit, double ms, double fwd_ms, double bwd_ms)
{
    (void)mit;
    (void)sit;
    sit_ms_.arr.push_back(Json::number(ms));
    if (fwd_ms >= 0.0) fwd_ms_.arr.push_back(Json::number(fwd_ms));
    if (bwd_ms >= 0.0) bwd_ms_.arr.push_back(Json::number(bwd_ms));
    for (const char *k : {"subiteration_ms", "forward_pass_ms", "backward_pass_ms"})
        record_.obj.erase(std::remove_if(record_.obj.begin(), record_.obj.end(),
                                         [k](const auto &kv) { return kv.first == k; }),
                          record_.obj.end());
    note("subiteration_ms", sit_ms_);
    note("forward_pass_ms", fwd_ms_);
    note("backward_pass_ms", bwd_ms_);
}

void Reconstruction::tune_shapes(void *vol_slab, int vdt, void *sums, void *err, int edt,
                                 const float *pd, const xpr_proj_args &a)
{
    if (!p_.xpr.autotune) return;
    static std::mutex mu;
    static std::set<std::vector<int64_t>> done;
    const std::vector<int64_t> key = {p_.xpr.device, vdt, edt, p_.xpr.accumulator,
                                      a.vmax[0] - a.vmin[0], a.vmax[2] - a.vmin[2], a.n_views,
                                      a.modules, a.det_width, a.line_rows, a.line_offset != 0};
    {
        std::lock_guard<std::mutex> g(mu);
        if (!done.insert(key).second) return;
    }
    const uint64_t n = uint64_t(a.n_views) * uint64_t(a.line_rows) * uint64_t(a.det_width);
    check(xpr_fill(stream_, err, edt, 0.0f, n), "fill");
    int c = 0;
    check(xpr_tune_backward(stream_, vol_slab, vdt, err, edt, pd, &a, &c), "tune backward");
    note("tuned_backward_voxels_per_lane", Json::number(c));
    check(xpr_tune_forward(stream_, vol_slab, vdt, sums, p_.xpr.accumulator, pd, &a, &c),
          "tune forward");
    note("tuned_forward_tile_cap", Json::number(c));
}

// ---------------------------------------------------------------------------
// helpers shared by the schedules
// ---------------------------------------------------------------------------
namespace {

// upload float32 host data into a device buffer of dtype dt (through a float
// staging buffer when dt is 16-bit)
void upload_as(xpr_stream s, DevBuf &dst, int dt, const float *src, size_t n, DevBuf &stage)
{
    if (dt == XPR_F32) {
        check(xpr_upload(s, dst.get(), src, n * 4), "upload");
    } else {
        if (stage.bytes() < n * 4) stage.alloc(n * 4);
        check(xpr_upload(s, stage.get(), src, n * 4), "upload");
        check(xpr_convert(s, dst.get(), dt, stage.get(), XPR_F32, n), "convert");
    }
    check(xpr_stream_sync(s), "sync");
}

void download_f32(xpr_stream s, float *dst, const void *src, int dt, size_t n, DevBuf &stage)
{
    if (dt == XPR_F32) {
        check(xpr_download(s, dst, src, n * 4), "download");
    } else {
        if (stage.bytes() < n * 4) stage.alloc(n * 4);
        check(xpr_convert(s, stage.get(), XPR_F32, src, dt, n), "convert");
        check(xpr_download(s, dst, stage.get(), n * 4), "download");
    }
    check(xpr_stream_sync(s), "sync");
}

// Final layers [l0, l1) of the volume, from device memory of dtype dt, to
// where the reconstruction writes them: the caller's device or host volume
// buffer, or layer_%05d.tif files; de-normalised when normalising
// (saveLayer(finalize), data_loader.hpp:133-146).  Returns with the data
// written (synchronises s).
void write_final(Dataset &ds, xpr_stream s, const void *dev, int dt, int64_t l0, int64_t l1,
                 DevBuf &stage)
{
    const size_t layer = size_t(ds.vwidth() * ds.vwidth());
    const size_t n = size_t(l1 - l0) * layer;
    const float scale = ds.norm_max() - ds.norm_min();
    float *sink = ds.volume_sink();
    if (sink && ds.device_io()) {
        float *dst = sink + size_t(l0) * layer;
        check(xpr_convert(s, dst, XPR_F32, dev, dt, n), "convert");
        if (ds.normalize()) check(xpr_scale(s, dst, n, scale), "scale");
        check(xpr_stream_sync(s), "sync");
    } else if (sink) {
        float *dst = sink + size_t(l0) * layer;
        download_f32(s, dst, dev, dt, n, stage);
        if (ds.normalize())
            for (size_t k = 0; k < n; ++k) dst[k] = dst[k] * scale;
    } else {
        std::vector<float> tmp(n);
        download_f32(s, tmp.data(), dev, dt, n, stage);
        ds.save_final_layers(tmp.data(), l0, l1);
    }
}

// detector sums set to 1.0 (one: the first pass, reconstruction_n.hpp:60)
// or 0, in the accumulator's own representation
void fill_sums(xpr_stream s, DevBuf &sums, int acc, bool one, uint64_t n)
{
    if (acc == XPR_ACC_U64) check(xpr_fill_u64(s, sums.as<uint64_t>(), one ? 1ull << 22 : 0, n), "fill");
    else if (acc == XPR_ACC_F32) check(xpr_fill(s, sums.get(), XPR_F32, one ? 1.0f : 0.0f, n), "fill");
    else check(xpr_fill_u32(s, sums.as<uint32_t>(), one ? 1u << 22 : 0u, n), "fill");
}

// 32-bit detector sums in the forwards over a resident volume whenever every
// sum provably stays below 2^32 (xpr.narrow_sums): the launch's hit bound
// (xpr_hit_bound over the sub-iteration's views, for any jitter) times the
// largest Q.22 voxel value, which the previous backward reports
// (xpr_backward_max).  Such sums are the same integers as 64-bit ones, so no
// result changes; the 32-bit forward is the lighter kernel.
class NarrowSums {
public:
    // hits: the bound of each sub-iteration's forward over the resident
    // volume; called again for each new volume (method_2's chunks: the
    // maximum is then unknown until the chunk's first backward)
    void init(bool enable, int acc, int storage, std::vector<double> hits)
    {
        on_ = enable && acc == XPR_ACC_U64;
        known_ = false;
        if (!on_) return;
        hits_ = std::move(hits);
        // a 16-bit store rounds the reported value up by less than this factor
        margin_ = storage == XPR_F32 ? 1.0 : 1.0 + 1.0 / 128.0;
        if (!top_.get()) {
            top_.alloc(4 * XPR_MAX_SLOTS);
            top_host_.alloc(4 * XPR_MAX_SLOTS);
        }
    }
    // before the forward of sub-iteration s: whether its sums may be 32-bit
    bool choose(int64_t s)
    {
        bool narrow = false;
        if (on_ && known_) {
            const uint32_t *slot = top_host_.as<uint32_t>();
            const uint32_t bits = *std::max_element(slot, slot + XPR_MAX_SLOTS);
            if (bits <= 0x7f7fffffu) {                 // finite and not negative
                float v;
                std::memcpy(&v, &bits, 4);
                const double q = std::floor(double(v) * margin_ * 4194304.0);
                narrow = hits_[size_t(s)] * q <= 4294967295.0;
            }
        }
        bits_.arr.push_back(Json::number(narrow ? 32 : 64));
        return narrow;
    }
    // around each backward (enqueued on its stream): clear the device maximum's
    // slots, then copy them to pinned host memory, read after the caller's
    // next sync
    uint32_t *before_backward(xpr_stream st)
    {
        if (!on_) return nullptr;
        check(xpr_fill_u32(st, top_.as<uint32_t>(), 0u, XPR_MAX_SLOTS), "fill");
        return top_.as<uint32_t>();
    }
    void after_backward(xpr_stream st)
    {
        if (on_) check(xpr_download(st, top_host_.get(), top_.get(), 4 * XPR_MAX_SLOTS), "download");
    }
    void synced() { known_ = on_; }
    bool on() const { return on_; }
    const Json &bits() const { return bits_; }
    double max_hits() const
    {
        double m = 0.0;
        for (double h : hits_) m = std::max(m, h);
        return m;
    }

private:
    bool on_ = false, known_ = false;
    std::vector<double> hits_;
    double margin_ = 1.0;
    DevBuf top_;
    HostBuf top_host_;
    Json bits_ = Json::array();
};

// the bound of a forward of args a (volume range, detector rows) over each
// sub-iteration's views (grid origin: the bound holds for any jitter)
std::vector<double> launch_hits(const std::vector<std::vector<float>> &pd_host, xpr_proj_args a,
                                const std::array<float, 3> &orig)
{
    std::vector<double> h(pd_host.size(), HUGE_VAL);
    for (int i = 0; i < 3; ++i) a.origin[i] = orig[size_t(i)];
    a.view_index = nullptr;
    for (size_t s = 0; s < pd_host.size(); ++s) check(xpr_hit_bound(pd_host[s].data(), &a, &h[s]), "hit bound");
    return h;
}

} // namespace

float Reconstruction::next_weight(int64_t mit) const
{
    float w = p_.recon.weight;
    for (int64_t k = 0; k < mit; ++k) w *= p_.recon.weight_factor;
    return w;
}

// ---------------------------------------------------------------------------
// naive: everything resident (reconstruction_n.hpp)
// ---------------------------------------------------------------------------
class ReconstructionN : public Reconstruction {
public:
    explicit ReconstructionN(Parameters &p) : Reconstruction(p)
    {
        check_memory(requiredGPUMemory(), "naive");
    }
    uint64_t requiredGPUMemory() override
    {
        const uint64_t sv = dtype_bytes(storage()), V = uint64_t(ds_->vwidth() * ds_->vwidth() * ds_->vheight());
        const uint64_t P = uint64_t(ds_->width() * ds_->height());
        const uint64_t As = uint64_t(ds_->concurrent()), S = uint64_t(p_.recon.sit);
        const uint64_t acc = acc_bytes(p_.xpr.accumulator);
        uint64_t need = V * sv + S * As * P * sv + As * P * (acc + sv) + ds_->projdata().size() * 4;
        if (storage() != XPR_F32 && p_.xpr.final_fp32) need += V * 4 + As * P * 8;
        return need;
    }
    void exec() override
    {
        AllocStreamScope scope(stream_);
        auto &ds = *ds_;
        ds.load(stream_);
        const int st = storage();
        const int64_t W = ds.vwidth(), H = ds.vheight(), P = ds.width() * ds.height();
        const int64_t As = ds.concurrent(), S = p_.recon.sit;
        const uint64_t V = uint64_t(W * W * H);
        const bool final32 = st != XPR_F32 && p_.xpr.final_fp32;
        // device-resident images of every sub-iteration, uploaded and
        // preprocessed on the copy stream ahead of the sub-iteration that
        // first reads them (the host staging of sit s+1 overlaps compute of s)
        DevBuf stage, img32, err32, vol32;
        std::vector<DevBuf> images(static_cast<size_t>(S)), pds(static_cast<size_t>(S));
        std::vector<std::vector<float>> pd_host(static_cast<size_t>(S));
        for (int64_t s = 0; s < S; ++s) {
            images[size_t(s)].alloc(uint64_t(As * P) * dtype_bytes(st));
            pd_host[size_t(s)] = ds.sit_projdata(s);
            pds[size_t(s)].alloc(pd_host[size_t(s)].size() * 4);
            check(xpr_upload(stream_, pds[size_t(s)].get(), pd_host[size_t(s)].data(),
                             pd_host[size_t(s)].size() * 4), "upload");
        }
        if (final32) {
            img32.alloc(uint64_t(As * P) * 4);
            err32.alloc(uint64_t(As * P) * 4);
            vol32.alloc(V * 4);
        }
        DevBuf vol(V * dtype_bytes(st)), sums(uint64_t(As * P) * acc_bytes(p_.xpr.accumulator)),
            err(uint64_t(As * P) * dtype_bytes(st));
        Event ready, up32;
        std::vector<Event> up(static_cast<size_t>(S));
        check(xpr_event_record(ready.e, stream_), "event");
        check(xpr_stream_wait_event(copy_, ready.e), "wait");     // buffers allocated
        auto upload = [&](int64_t s) {
            ds.upload_sit_images(copy_, s, 0, ds.height(), images[size_t(s)].get(), st);
            check(xpr_event_record(up[size_t(s)].e, copy_), "event");
            if (final32 && s == S - 1) {
                ds.upload_sit_images(copy_, s, 0, ds.height(), img32.get(), XPR_F32);
                check(xpr_event_record(up32.e, copy_), "event");
            }
        };
        Operators ops(stream_, p_.xpr.accumulator, p_.xpr.interp);
        ops.setSumImagesBuffer(sums.get());
        ops.setErrorBuffer(err.get(), st);
        ops.setVolumeBuffer(vol.get(), st);
        ops.setImageParameters(ds.width(), ds.height(), As);
        ops.setVolumeParameters({0, 0, 0}, {int(W), int(H), int(W)}, ds.vx());
        ops.setImageOffset(0, int(ds.height()));
        ops.setAngleNumber(As, p_.detector.module_number);
        check(xpr_fill(stream_, vol.get(), st, 1.0f, V), "fill");
        {
            // tile shapes timed on 128 central layers of the all-1.0 volume (on
            // 32 the launch tail can favour a shape that is slower at scale)
            const int64_t l0 = std::max<int64_t>(0, H / 2 - 64), l1 = std::min<int64_t>(H, l0 + 128);
            xpr_proj_args ta = ops.args();
            ta.vmin[1] = int32_t(l0);
            ta.vmax[1] = int32_t(l1);
            for (int i = 0; i < 3; ++i) ta.origin[i] = ds.orig()[size_t(i)];
            tune_shapes(static_cast<char *>(vol.get()) + size_t(l0) * size_t(W * W) * dtype_bytes(st), st,
                        sums.get(), err.get(), st, pds[0].as<float>(), ta);
        }
        fill_sums(stream_, sums, p_.xpr.accumulator, true, uint64_t(As * P));
        upload(0);
        NarrowSums narrow;
        narrow.init(p_.xpr.narrow_sums, p_.xpr.accumulator, st,
                    p_.xpr.narrow_sums ? launch_hits(pd_host, ops.args(), ds.orig()) : std::vector<double>());
        int vdt = st;
        Event e0, e1, fb, fe, bb, be;
        note("launches_per_pass", Json::number(1));
        for (int64_t mit = 0; mit < p_.recon.it; ++mit) {
            ops.setWeight(next_weight(mit));
            for (int64_t s = 0; s < S; ++s) {
                const bool last = mit == p_.recon.it - 1 && s == S - 1;
                if (mit == 0) check(xpr_stream_wait_event(stream_, up[size_t(s)].e), "wait");
                check(xpr_event_record(e0.e, stream_), "event");
                ops.setOrigin(next_origin());
                ops.setImagesBuffer(images[size_t(s)].get(), st);
                ops.setProjDataBuffer(pds[size_t(s)].as<float>());
                if (last && final32) {
                    // final sub-iteration with float32 storage
                    check(xpr_stream_wait_event(stream_, up32.e), "wait");
                    check(xpr_convert(stream_, vol32.get(), XPR_F32, vol.get(), st, V), "convert");
                    ops.setVolumeBuffer(vol32.get(), XPR_F32);
                    ops.setImagesBuffer(img32.get(), XPR_F32);
                    ops.setErrorBuffer(err32.get(), XPR_F32);
                    vdt = XPR_F32;
                }
                const bool fwd = s > 0 || mit > 0;
                ops.setNarrowSums(fwd && narrow.choose(s));
                if (fwd) {
                    check(xpr_event_record(fb.e, stream_), "event");
                    ops.forward();
                    check(xpr_event_record(fe.e, stream_), "event");
                }
                ops.error();
                ops.setMaxBuffer(narrow.before_backward(stream_));
                check(xpr_event_record(bb.e, stream_), "event");
                ops.backward();
                check(xpr_event_record(be.e, stream_), "event");
                narrow.after_backward(stream_);
                fill_sums(stream_, sums, p_.xpr.accumulator, false, uint64_t(As * P));
                check(xpr_event_record(e1.e, stream_), "event");
                if (mit == 0 && s + 1 < S) upload(s + 1);
                check(xpr_event_sync(e1.e), "sync");
                narrow.synced();
                float ms = 0, fms = -1.0f, bms = 0;
                check(xpr_event_elapsed_ms(e0.e, e1.e, &ms), "elapsed");
                if (fwd) check(xpr_event_elapsed_ms(fb.e, fe.e, &fms), "elapsed");
                check(xpr_event_elapsed_ms(bb.e, be.e, &bms), "elapsed");
                log_sit(mit, s, ms, fms, bms);
            }
        }
        note("forward_sum_bits", narrow.bits());
        if (narrow.on()) note("hit_bound", Json::number(narrow.max_hits()));
        const DevBuf &result = vdt == XPR_F32 && final32 ? vol32 : vol;
        write_final(ds, stream_, result.get(), vdt, 0, H, stage);
    }
};

// ---------------------------------------------------------------------------
// slab streaming: method_1a (one IO thread), method_1b (IO thread pool),
// method_1c (plus per-slab view culling).  The volume lives in the layer
// store on the host; each pass streams slabs of slab_height layers through
// two device slab buffers, uploads on the copy stream overlapping compute.
// (reconstruction_lpln.hpp:27-96, reconstruction_lpla.hpp:29-157,
//  reconstruction_lplh.hpp:40-219)
// ---------------------------------------------------------------------------
class ReconstructionSlab : public Reconstruction {
public:
    ReconstructionSlab(Parameters &p, int threads, bool cull) : Reconstruction(p), threads_(threads), cull_(cull)
    {
        resolve();
        check_memory(requiredGPUMemory(), method_name(p.recon.method));
    }
    uint64_t requiredGPUMemory() override { return need(store_, hs_, resident_); }

    void exec() override
    {
        AllocStreamScope scope(stream_);
        auto &ds = *ds_;
        ds.load(stream_);
        st_ = storage();
        sb_ = dtype_bytes(st_);
        const int64_t W = ds.vwidth(), H = ds.vheight(), P = ds.width() * ds.height();
        const int64_t As = ds.concurrent(), S = p_.recon.sit, mods = p_.detector.module_number;
        layer_ = size_t(W * W);
        nslab_ = (H + hs_ - 1) / hs_;
        const size_t slab_n = size_t(hs_) * layer_;
        DevBuf sums(uint64_t(As * P) * acc_bytes(p_.xpr.accumulator)), err(uint64_t(As * P) * sb_);
        std::vector<DevBuf> images(static_cast<size_t>(resident_ ? S : 1));
        for (auto &b : images) b.alloc(uint64_t(As * P) * sb_);
        // ProjData of every sub-iteration and, for method_1c, the views of each
        // slab, uploaded once (the culling widens the slab by a voxel, so it
        // does not depend on the jitter)
        pds_.clear();
        pds_.resize(size_t(S));
        cviews_.clear();
        cviews_.resize(size_t(S));
        std::vector<std::vector<float>> pd_whole(whole_ ? size_t(S) : 0);
        for (int64_t s = 0; s < S; ++s) {
            const auto pd = ds.sit_projdata(s);
            if (whole_) pd_whole[size_t(s)] = pd;
            pds_[size_t(s)].alloc(pd.size() * 4);
            check(xpr_upload(stream_, pds_[size_t(s)].get(), pd.data(), pd.size() * 4), "upload");
            if (!cull_) continue;
            const auto lists = culled(pd, nslab_, hs_);
            cviews_[size_t(s)].resize(lists.size());
            for (size_t b = 0; b < lists.size(); ++b) {
                SlabDev &d = cviews_[size_t(s)][b];
                d.n = int64_t(lists[b].index.size());
                d.pd.alloc(std::max<size_t>(lists[b].pd.size(), 20) * 4);
                d.idx.alloc(std::max<size_t>(lists[b].index.size(), 1) * 4);
                if (d.n > 0) {
                    check(xpr_upload(stream_, d.pd.get(), lists[b].pd.data(), lists[b].pd.size() * 4), "upload");
                    check(xpr_upload(stream_, d.idx.get(), lists[b].index.data(), lists[b].index.size() * 4), "upload");
                }
            }
        }
        // 16-bit storage with the volume in HBM: the last sub-iteration in
        // float32 storage (volume converted, images and error in float32)
        const bool final32 = final32_;
        DevBuf vol32, img32, err32;
        Event up32;
        if (final32) {
            vol32.alloc(uint64_t(H) * layer_ * 4);
            img32.alloc(uint64_t(As * P) * 4);
            err32.alloc(uint64_t(As * P) * 4);
        }
        if (store_ == kDevice) {
            vol_.alloc(uint64_t(H) * layer_ * sb_);
            check(xpr_fill(stream_, vol_.get(), st_, 1.0f, uint64_t(H) * layer_), "fill");
        } else {
            for (int k = 0; k < kRing; ++k) {
                dslab_[k].alloc(slab_n * sb_);
                if (st_ != XPR_F32) dup_[k].alloc(slab_n * 4);
                if (store_ == kZstd) {
                    hin_[k].alloc(slab_n * 4);
                    hout_[k].alloc(slab_n * 4);
                }
            }
            check(xpr_fill(stream_, dslab_[0].get(), st_, 1.0f, slab_n), "fill");
        }
        cvol_ = vol_.get();
        cst_ = st_;
        csb_ = sb_;
        Operators ops(stream_, p_.xpr.accumulator, p_.xpr.interp);
        ops.setSumImagesBuffer(sums.get());
        ops.setErrorBuffer(err.get(), st_);
        ops.setImageParameters(ds.width(), ds.height(), As);
        ops.setImageOffset(0, int(ds.height()));
        ops.setAngleNumber(As, mods);
        {
            // tile shapes timed on a central slab holding 1.0 (error 0 keeps
            // it); whole-volume passes on 128 central layers, like the naive
            // schedule's
            const int64_t b = nslab_ / 2;
            const int64_t l0 = whole_ ? std::max<int64_t>(0, H / 2 - 64) : b * hs_;
            const int64_t l1 = std::min(H, l0 + (whole_ ? 128 : hs_));
            void *v = store_ == kDevice ? static_cast<void *>(static_cast<char *>(vol_.get()) + size_t(l0) * layer_ * sb_)
                                        : dslab_[0].get();
            xpr_proj_args ta = ops.args();
            ta.vmin[0] = 0; ta.vmax[0] = int32_t(W);
            ta.vmin[1] = int32_t(l0); ta.vmax[1] = int32_t(l1);
            ta.vmin[2] = 0; ta.vmax[2] = int32_t(W);
            for (int i = 0; i < 3; ++i) { ta.origin[i] = ds.orig()[size_t(i)]; ta.voxel[i] = ds.vx(); }
            tune_shapes(v, st_, sums.get(), err.get(), st_, pds_[0].as<float>(), ta);
        }
        fill_sums(stream_, sums, p_.xpr.accumulator, true, uint64_t(As * P));
        // the slab buffers are free once the tuning launches above are done
        for (int k = 0; k < kRing; ++k) check(xpr_event_record(kdone_[k].e, stream_), "event");
        stored_ = std::vector<Event>(static_cast<size_t>(nslab_));
        // images: every sub-iteration's resident in HBM, preprocessed on the copy
        // stream ahead of the sub-iteration that first reads them, or one
        // sub-iteration's at a time on the compute stream
        Event ready;
        std::vector<Event> up_img(static_cast<size_t>(resident_ ? S : 0));
        check(xpr_event_record(ready.e, stream_), "event");
        check(xpr_stream_wait_event(copy_, ready.e), "wait");
        auto upload = [&](int64_t s) {
            ds.upload_sit_images(copy_, s, 0, ds.height(), images[size_t(s)].get(), st_);
            check(xpr_event_record(up_img[size_t(s)].e, copy_), "event");
            if (final32 && s == S - 1) {
                ds.upload_sit_images(copy_, s, 0, ds.height(), img32.get(), XPR_F32);
                check(xpr_event_record(up32.e, copy_), "event");
            }
        };
        if (resident_) upload(0);
        note("launches_per_pass", Json::number(whole_ ? 1.0 : double(nslab_)));
        // 32-bit sums where provably exact: whole-volume passes only
        NarrowSums narrow;
        {
            xpr_proj_args wa = ops.args();
            wa.vmin[0] = wa.vmin[1] = wa.vmin[2] = 0;
            wa.vmax[0] = int32_t(W);
            wa.vmax[1] = int32_t(H);
            wa.vmax[2] = int32_t(W);
            for (int i = 0; i < 3; ++i) wa.voxel[i] = ds.vx();
            const bool on = whole_ && p_.xpr.narrow_sums;
            narrow.init(on, p_.xpr.accumulator, st_, on ? launch_hits(pd_whole, wa, ds.orig()) : std::vector<double>());
        }
        Event e0, e1, fb, fe, bb, be;
        for (int64_t mit = 0; mit < p_.recon.it; ++mit) {
            ops.setWeight(next_weight(mit));
            for (int64_t s = 0; s < S; ++s) {
                const bool first = s == 0 && mit == 0;
                const bool last = mit == p_.recon.it - 1 && s == S - 1;
                check(xpr_event_record(e0.e, stream_), "event");
                ops.setOrigin(next_origin());
                if (resident_) {
                    if (mit == 0) check(xpr_stream_wait_event(stream_, up_img[size_t(s)].e), "wait");
                    ops.setImagesBuffer(images[size_t(s)].get(), st_);
                } else {
                    ds.upload_sit_images(stream_, s, 0, ds.height(), images[0].get(), st_);
                    ops.setImagesBuffer(images[0].get(), st_);
                }
                if (last && final32) {
                    if (resident_) check(xpr_stream_wait_event(stream_, up32.e), "wait");
                    else ds.upload_sit_images(stream_, s, 0, ds.height(), img32.get(), XPR_F32);
                    check(xpr_convert(stream_, vol32.get(), XPR_F32, vol_.get(), st_, uint64_t(H) * layer_),
                          "convert");
                    cvol_ = vol32.get();
                    cst_ = XPR_F32;
                    csb_ = 4;
                    ops.setImagesBuffer(img32.get(), XPR_F32);
                    ops.setErrorBuffer(err32.get(), XPR_F32);
                }
                ops.setNarrowSums(!first && narrow.choose(s));
                if (!first) {
                    check(xpr_event_record(fb.e, stream_), "event");
                    pass(false, false, false, s, ops);
                    check(xpr_event_record(fe.e, stream_), "event");
                }
                ops.setAngleNumber(As, mods);
                ops.setViewIndex(nullptr);
                ops.setVolumeParameters({0, 0, 0}, {int(W), int(H), int(W)}, ds.vx());
                ops.error();
                ops.setMaxBuffer(narrow.before_backward(stream_));
                check(xpr_event_record(bb.e, stream_), "event");
                pass(true, first, last, s, ops);
                check(xpr_event_record(be.e, stream_), "event");
                narrow.after_backward(stream_);
                fill_sums(stream_, sums, p_.xpr.accumulator, false, uint64_t(As * P));
                check(xpr_event_record(e1.e, stream_), "event");
                if (resident_ && mit == 0 && s + 1 < S) upload(s + 1);
                check(xpr_event_sync(e1.e), "sync");
                narrow.synced();
                float ms = 0, fms = -1.0f, bms = 0;
                check(xpr_event_elapsed_ms(e0.e, e1.e, &ms), "elapsed");
                if (!first) check(xpr_event_elapsed_ms(fb.e, fe.e, &fms), "elapsed");
                check(xpr_event_elapsed_ms(bb.e, be.e, &bms), "elapsed");
                log_sit(mit, s, ms, fms, bms);
            }
        }
        note("forward_sum_bits", narrow.bits());
        if (narrow.on()) note("hit_bound", Json::number(narrow.max_hits()));
        if (store_ == kDevice) write_final(ds, stream_, cvol_, cst_, 0, H, stage_);
        note("slab_h2d_bytes", Json::number(double(h2d_bytes_)));
        note("slab_d2h_bytes", Json::number(double(d2h_bytes_)));
        check(xpr_stream_sync(copy_back_), "sync");
        check(xpr_stream_sync(copy_), "sync");
        check(xpr_stream_sync(stream_), "sync");
    }

private:
    enum Store { kDevice, kRam, kZstd };
    int threads_;
    bool cull_;
    Store store_ = kZstd;
    int64_t hs_ = 1;                 // slab height
    bool resident_ = false;          // every sub-iteration's images in HBM
    int st_ = XPR_F32;
    size_t sb_ = 4, layer_ = 0;
    bool final32_ = false;           // last sub-iteration with float32 storage
    // HBM layer store without per-slab view lists: a projector pass is one
    // launch over the resident volume whatever the slab height (which then
    // only sets the slab streaming granularity of the host stores)
    bool whole_ = false;
    void *cvol_ = nullptr;           // the HBM volume the kernels update, its dtype and size
    int cst_ = XPR_F32;
    size_t csb_ = 4;
    int64_t nslab_ = 0;
    // dup_[i]: float32 staging of slab buffer i for 16-bit storage, used in
    // both directions (the upload's conversion out of it is done before the
    // slab's kernel, the download's conversion into it after)
    // Slab buffers of the host-store pipeline: with three, a buffer's next
    // upload waits for the download of the slab three places before instead
    // of the one just before, so the uploads and downloads of the backward
    // pass keep both copy directions busy.
    static constexpr int kRing = 3;
    DevBuf vol_, stage_, dslab_[kRing], dup_[kRing];
    HostBuf hin_[kRing], hout_[kRing];
    std::vector<DevBuf> pds_;
    // Events of the slab pipelines, per buffer i: up_ (its upload done),
    // kdone_ (its last kernel, and the conversion of 16-bit values, done),
    // dl_ (its last download done).
    Event up_[kRing], kdone_[kRing], dl_[kRing];
    int64_t held_[kRing] = {-1, -1, -1};   // the slab whose current values buffer i holds
    bool reverse_ = false;                 // the next pass runs its slabs from the last
    std::vector<Event> stored_;            // slab b's last download into the store done
    uint64_t h2d_bytes_ = 0, d2h_bytes_ = 0;  // slab bytes the passes moved

    struct SlabViews {
        std::vector<float> pd;
        std::vector<uint32_t> index;
    };
    struct SlabDev {
        DevBuf pd, idx;
        int64_t n = 0;
    };
    std::vector<std::vector<SlabDev>> cviews_;

    // Device bytes of the schedule for a store, slab height and image residency.
    uint64_t need(Store st, int64_t Hs, bool resident) const
    {
        const uint64_t sv = dtype_bytes(storage()), W = uint64_t(ds_->vwidth()), H = uint64_t(ds_->vheight());
        const uint64_t P = uint64_t(ds_->width() * ds_->height()), As = uint64_t(ds_->concurrent());
        const uint64_t S = uint64_t(p_.recon.sit), acc = acc_bytes(p_.xpr.accumulator);
        const uint64_t layer = W * W;
        // the preprocessing's two upload slots of up to 16 MiB (or one
        // sub-iteration's views), when the projections come from the host
        const uint64_t staging = ds_->device_io() ? 0 : 2 * std::min(std::max<uint64_t>(16ull << 20, P * 4), As * P * 4);
        uint64_t m = As * P * (acc + sv) + (resident ? S : 1) * As * P * sv +
                     ds_->projdata().size() * 4 + staging;
        if (st == kDevice) m += H * layer * sv;
        else m += uint64_t(kRing) * uint64_t(Hs) * layer * (sv + (sv != 4 ? 4 : 0));
        if (st == kDevice && sv != 4 && p_.xpr.final_fp32) m += H * layer * 4 + As * P * 8;
        return m;
    }

    // The layer store, slab height and image residency under the device budget.
    void resolve()
    {
        const std::string &ls = p_.xpr.layer_store;
        const int64_t H = ds_->vheight();
        if (ls == "device") store_ = kDevice;
        else if (ls == "ram") store_ = kRam;
        else if (ls == "zstd") store_ = kZstd;
        else store_ = need(kDevice, H, false) <= budget() ? kDevice : kZstd;
        if (p_.xpr.slab_height > 0) {
            hs_ = std::min<int64_t>(p_.xpr.slab_height, H);
        } else if (store_ == kDevice) {
            hs_ = H;                                // one launch per projector pass
        } else {
            // the tallest slab whose double buffers fit beside the rest
            const uint64_t base = need(store_, 0, false), per = need(store_, 1, false) - base;
            const uint64_t b = budget();
            hs_ = b > base ? int64_t(std::min<uint64_t>((b - base) / per, uint64_t(H))) : 1;
            hs_ = std::max<int64_t>(hs_, 1);
        }
        resident_ = need(store_, hs_, true) <= budget();
        final32_ = store_ == kDevice && storage() != XPR_F32 && p_.xpr.final_fp32;
        whole_ = store_ == kDevice && !cull_;
        note("final_fp32", Json::boolean(final32_));
        p_.xpr.layer_store = store_ == kDevice ? "device" : store_ == kRam ? "ram" : "zstd";
        if (store_ == kZstd) {
            ds_->ensure_output_dir();
            // spill files are zstd frames; without libzstd.so.1 they are raw
            // floats, said in the run record and on stderr
            const bool z = zstd_available();
            note("spill_compression", Json::string(z ? "zstd" : "none"));
            if (!z)
                std::cerr << "Warning: libzstd.so.1 not found, the layer spill files are written "
                             "uncompressed" << std::endl;
        }
        note("layer_store", Json::string(p_.xpr.layer_store));
        note("slab_height", Json::number(double(hs_)));
        note("images_resident", Json::boolean(resident_));
    }

    // Per slab: the views whose projection of the slab can meet detector rows
    // [vpb, vpt) (method_1c's per-layer view lists, geometry.hpp:141-171).  A
    // dropped view holds no in-bounds voxel of the slab, so the sums and the
    // updates are unchanged; only the work shrinks.
    std::vector<SlabViews> culled(const std::vector<float> &pd, int64_t nslab, int64_t Hs)
    {
        auto &ds = *ds_;
        const int64_t W = ds.vwidth(), H = ds.vheight(), mods = p_.detector.module_number;
        const auto o = ds.orig();
        const double vx = ds.vx();
        const int64_t nv = int64_t(pd.size() / size_t(20 * mods));
        std::vector<SlabViews> out(static_cast<size_t>(nslab));
        for (int64_t b = 0; b < nslab; ++b) {
            const int64_t l0 = b * Hs, l1 = std::min(H, l0 + Hs);
            for (int64_t v = 0; v < nv; ++v) {
                bool keep = false;
                for (int64_t m = 0; m < mods && !keep; ++m) {
                    const float *r = pd.data() + 20 * (v * mods + m);
                    double vmin = 1e300, vmax = -1e300;
                    bool front = true;
                    // corners of the slab's voxel positions widened by one voxel
                    // (the origin jitter is below one voxel)
                    for (int c = 0; c < 8; ++c) {
                        const double x = o[0] + ((c & 1) ? double(W) : -1.0) * vx;
                        const double y = o[1] + ((c & 2) ? double(l1) : double(l0) - 1.0) * vx;
                        const double z = o[2] + ((c & 4) ? double(W) : -1.0) * vx;
                        const double Y = double(r[4]) * x + double(r[5]) * y + double(r[6]) * z + double(r[7]);
                        const double Wh = double(r[12]) * x + double(r[13]) * y + double(r[14]) * z + double(r[15]);
                        if (!(Wh > 0)) front = false;
                        vmin = std::min(vmin, Y / Wh);
                        vmax = std::max(vmax, Y / Wh);
                    }
                    keep = !front || !(vmax < double(r[18]) - 1.0 || vmin > double(r[19]) + 1.0);
                }
                if (keep) {
                    out[size_t(b)].pd.insert(out[size_t(b)].pd.end(), pd.begin() + 20 * v * mods,
                                             pd.begin() + 20 * (v + 1) * mods);
                    out[size_t(b)].index.push_back(uint32_t(v));
                }
            }
        }
        return out;
    }

    // the kernel arguments of slab b (all views, or the slab's culled views)
    void set_slab(Operators &ops, int64_t s, int64_t b, void *vol)
    {
        const int64_t l0 = b * hs_, l1 = std::min(ds_->vheight(), l0 + hs_);
        ops.setVolumeBuffer(vol, cst_);
        ops.setVolumeParameters({0, int(l0), 0}, {int(ds_->vwidth()), int(l1), int(ds_->vwidth())},
                                ds_->vx());
        if (cull_) {
            // a slab no view reaches runs with zero views: the forward adds
            // nothing and the backward leaves the slab unchanged
            const SlabDev &v = cviews_[size_t(s)][size_t(b)];
            ops.setProjDataBuffer(v.pd.as<float>());
            ops.setViewIndex(v.idx.as<uint32_t>());
            ops.setAngleNumber(v.n, p_.detector.module_number);
        } else {
            ops.setProjDataBuffer(pds_[size_t(s)].as<float>());
            ops.setViewIndex(nullptr);
            ops.setAngleNumber(ds_->concurrent(), p_.detector.module_number);
        }
    }

    void pass(bool backward, bool first, bool last, int64_t s, Operators &ops)
    {
        if (whole_) {
            // the resident volume in one launch: the sums are integers and each
            // voxel's update is its own, so they equal the slab launches' bit for
            // bit, without a launch tail per slab
            const int64_t W = ds_->vwidth(), H = ds_->vheight();
            ops.setVolumeBuffer(cvol_, cst_);
            ops.setVolumeParameters({0, 0, 0}, {int(W), int(H), int(W)}, ds_->vx());
            ops.setProjDataBuffer(pds_[size_t(s)].as<float>());
            ops.setViewIndex(nullptr);
            ops.setAngleNumber(ds_->concurrent(), p_.detector.module_number);
            if (backward) ops.backward();
            else ops.forward();
            return;
        }
        if (store_ == kDevice) {
            // the layer store is the resident volume: slabs are launched in place
            // (method_1c: each with its own view list)
            for (int64_t b = 0; b < nslab_; ++b) {
                set_slab(ops, s, b, static_cast<char *>(cvol_) + size_t(b * hs_) * layer_ * csb_);
                if (backward) ops.backward();
                else ops.forward();
            }
            return;
        }
        host_pass(backward, first, last, s, ops);
    }

    // Slabs through kRing device buffers, three streams: uploads on copy_,
    // kernels on the compute stream, downloads on copy_back_; the host only
    // enqueues ("ram": DMA straight from and to the pinned store) or also
    // decompresses and compresses slabs while the device works ("zstd").
    //
    // Each pass runs its slabs in the reverse order of the previous one, so
    // the last kRing slabs of a pass, still in their buffers with their
    // current values, open the next pass without an upload (a forward
    // leaves the volume as it was; a backward's buffers hold the values
    // that are on their way back to the store).  The forward adds integers
    // and each backward owns its voxels: the order changes no result.
    void host_pass(bool backward, bool first, bool last, int64_t s, Operators &ops)
    {
        auto &ds = *ds_;
        const int64_t H = ds.vheight(), n = nslab_;
        const bool zstd = store_ == kZstd;
        auto save = [&](int64_t b, int i) {              // zstd: slab b (buffer i) to the spill files
            const int64_t l0 = b * hs_, l1 = std::min(H, l0 + hs_);
            check(xpr_event_sync(dl_[i].e), "sync");
            const float *src = hout_[i].as<float>();
            parallel_for(l1 - l0, threads_, [&](int64_t k) {
                ds.save_layer(src + size_t(k) * layer_, l0 + k, false);
            });
        };
        // this pass's order and buffers: the slabs a buffer holds come first
        // and keep their buffer; the others take the free buffers, then the
        // buffer of the slab kRing places before them
        std::vector<int64_t> ord(static_cast<size_t>(n));
        std::vector<int> slot(static_cast<size_t>(n));
        std::vector<char> resident(static_cast<size_t>(n), 0);
        for (int64_t k = 0; k < n; ++k) ord[size_t(k)] = reverse_ ? n - 1 - k : k;
        reverse_ = !reverse_;
        bool taken[kRing] = {};
        for (int64_t k = 0; k < std::min<int64_t>(n, kRing); ++k)
            for (int j = 0; j < kRing; ++j)
                if (!first && held_[j] == ord[size_t(k)]) {
                    slot[size_t(k)] = j;
                    resident[size_t(k)] = 1;
                    taken[j] = true;
                }
        for (int64_t k = 0, j = 0; k < std::min<int64_t>(n, kRing); ++k) {
            if (resident[size_t(k)]) continue;
            while (taken[j]) ++j;
            slot[size_t(k)] = j;
            taken[j] = true;
        }
        for (int64_t k = kRing; k < n; ++k) slot[size_t(k)] = slot[size_t(k - kRing)];
        for (int j = 0; j < kRing; ++j) held_[j] = -1;
        int64_t prev_b = -1;
        int prev_i = 0;
        for (int64_t k = 0; k < n; ++k) {
            const int64_t b = ord[size_t(k)];
            const int i = slot[size_t(k)];
            const int64_t l0 = b * hs_, l1 = std::min(H, l0 + hs_);
            const size_t nv = size_t(l1 - l0) * layer_;
            if (!resident[size_t(k)]) {
                const bool fill = backward && first;
                const float *src = nullptr;
                if (!fill) {
                    if (zstd) {
                        // hin_[i] is free once its last upload is done
                        check(xpr_event_sync(up_[i].e), "sync");
                        float *dst = hin_[i].as<float>();
                        parallel_for(l1 - l0, threads_, [&](int64_t q) {
                            ds.load_layer(l0 + q, dst + size_t(q) * layer_);
                        });
                        src = dst;
                    } else {
                        src = ds.ram_layer(l0);
                        // the store holds the slab once its last download is done
                        check(xpr_stream_wait_event(copy_, stored_[size_t(b)].e), "wait");
                    }
                }
                // buffer i is free once its last kernel and its last download
                // (a slab that opened this pass without an upload may still be
                // on its way back) are done
                check(xpr_stream_wait_event(copy_, kdone_[i].e), "wait");
                check(xpr_stream_wait_event(copy_, dl_[i].e), "wait");
                if (!fill) h2d_bytes_ += nv * 4;
                if (fill) {
                    check(xpr_fill(copy_, dslab_[i].get(), st_, 1.0f, nv), "fill");
                } else if (st_ == XPR_F32) {
                    check(xpr_upload(copy_, dslab_[i].get(), src, nv * 4), "upload");
                } else {
                    check(xpr_upload(copy_, dup_[i].get(), src, nv * 4), "upload");
                    check(xpr_convert(copy_, dslab_[i].get(), st_, dup_[i].get(), XPR_F32, nv), "convert");
                }
                check(xpr_event_record(up_[i].e, copy_), "event");
                check(xpr_stream_wait_event(stream_, up_[i].e), "wait");
            }
            held_[i] = b;
            set_slab(ops, s, b, dslab_[i].get());
            if (!backward) {
                ops.forward();
                check(xpr_event_record(kdone_[i].e, stream_), "event");
                continue;
            }
            ops.backward();
            if (last) {
                write_final(ds, stream_, dslab_[i].get(), st_, l0, l1, stage_);
                check(xpr_event_record(kdone_[i].e, stream_), "event");
                if (zstd)
                    for (int64_t l = l0; l < l1; ++l) {
                        std::error_code ec;
                        fs::remove(numbered_path(p_.recon.output, "layer", l, 5, "zstd"), ec);
                    }
                continue;
            }
            const void *back = dslab_[i].get();
            if (st_ != XPR_F32) {
                // dup_[i] may still feed the buffer's last upload
                check(xpr_stream_wait_event(stream_, up_[i].e), "wait");
                check(xpr_convert(stream_, dup_[i].get(), XPR_F32, dslab_[i].get(), st_, nv), "convert");
                back = dup_[i].get();
            }
            check(xpr_event_record(kdone_[i].e, stream_), "event");
            check(xpr_stream_wait_event(copy_back_, kdone_[i].e), "wait");
            check(xpr_download(copy_back_, zstd ? hout_[i].get() : ds.ram_layer(l0), back, nv * 4), "download");
            d2h_bytes_ += nv * 4;
            check(xpr_event_record(dl_[i].e, copy_back_), "event");
            check(xpr_event_record(stored_[size_t(b)].e, copy_back_), "event");
            if (zstd && prev_b >= 0) save(prev_b, prev_i);
            prev_b = b;
            prev_i = i;
        }
        if (zstd && prev_b >= 0) save(prev_b, prev_i);
    }
};

// ---------------------------------------------------------------------------
// method_2: independent chunks with padding and cropped detector rows
// (reconstruction_chunk.hpp:36-131, dataset/geometry_chunk.hpp:92-197)
// ---------------------------------------------------------------------------
uint64_t chunk_bytes(const Chunk &c, int64_t vwidth, int64_t dwidth, int64_t concurrent, int storage)
{
    const uint64_t sv = dtype_bytes(storage);
    const uint64_t vol = uint64_t(c.vSize + c.vPadTop + c.vPadBot) * uint64_t(vwidth * vwidth) * sv;
    const uint64_t img = uint64_t(c.iSize * dwidth * concurrent);
    return vol + img * (sv + 8 + sv);
}

std::vector<Chunk> plan_chunks(const Parameters &p, const Dataset &ds, uint64_t budget,
                               std::vector<float> *lmin_out, std::vector<float> *lmax_out)
{
    const int64_t H = ds.vheight(), W = ds.vwidth(), dh = ds.height();
    const auto o = ds.orig();
    const float vx = ds.vx();
    const auto &pd = ds.projdata();
    const int64_t nrec = int64_t(pd.size() / 20);
    // computeLayerYBounds (geometry_chunk.hpp:120-145): the four vertical
    // edges of the volume at the top and bottom of each layer, all views
    std::vector<float> lmin(static_cast<size_t>(H)), lmax(static_cast<size_t>(H));
    const float cx[4] = {o[0], -o[0], o[0], -o[0]}, cz[4] = {-o[2], -o[2], o[2], o[2]};
    for (int64_t l = 0; l < H; ++l) {
        float mn = 1e30f, mx = -1e30f;
        for (int64_t r = 0; r < nrec; ++r) {
            const float *m = pd.data() + 20 * r;
            for (int c = 0; c < 4; ++c)
                for (int e = 0; e < 2; ++e) {
                    const float y = o[1] + float(l + e) * vx;
                    const float Y = m[4] * cx[c] + m[5] * y + m[6] * cz[c] + m[7];
                    const float Wh = m[12] * cx[c] + m[13] * y + m[14] * cz[c] + m[15];
                    const float v = Y / Wh;
                    mn = std::min(mn, v);
                    mx = std::max(mx, v);
                }
        }
        lmin[size_t(l)] = mn;
        lmax[size_t(l)] = mx;
    }
    auto yproj = [&](int64_t off, int64_t h) {
        return std::pair<float, float>{std::min(lmin[size_t(off)], lmin[size_t(off + h - 1)]),
                                       std::max(lmax[size_t(off)], lmax[size_t(off + h - 1)])};
    };
    // updateImageViewport (:164-168): rows [floor(min)-1, ceil(max)+2) within the detector
    auto viewport = [&](Chunk &c) {
        auto mm = yproj(c.vOffset, c.vSize);
        c.iOffset = int64_t(std::max(std::floor(mm.first) - 1.0f, 0.0f));
        c.iSize = int64_t(std::max(std::min(std::ceil(mm.second - float(c.iOffset)) + 2.0f,
                                            float(dh - c.iOffset)), 0.0f));
    };
    // updateVolumePadding (:175-188): neighbouring layers that project into
    // the chunk's rows, scaled by overlap_fct
    auto padding = [&](Chunk &c) {
        c.vPadTop = 0;
        c.vPadBot = 0;
        while (c.vOffset - c.vPadTop > 0 && yproj(c.vOffset - c.vPadTop - 1, 1).second > float(c.iOffset - 1))
            ++c.vPadTop;
        while (c.vOffset + c.vSize + c.vPadBot < H &&
               yproj(c.vOffset + c.vSize + c.vPadBot, 1).first < float(c.iOffset + c.iSize + 1))
            ++c.vPadBot;
        c.vPadTop = int64_t(std::ceil(float(c.vPadTop) * p.method2.overlap_fct));
        c.vPadBot = int64_t(std::ceil(float(c.vPadBot) * p.method2.overlap_fct));
    };
    auto fits = [&](const Chunk &c) {
        return chunk_bytes(c, W, ds.width(), ds.concurrent(), p.xpr.storage) + pd.size() * 4 < budget;
    };
    // computeChunks (:92-114): grow each chunk layer by layer while it fits
    std::vector<Chunk> chunks{{0, 0, 0, 0, 0, 0}};
    for (int64_t l = 0; l < H; ++l) {
        Chunk c = chunks.back();
        ++c.vSize;
        viewport(c);
        padding(c);
        if (fits(c)) {
            chunks.back() = c;
        } else {
            if (c.vSize <= 1)
                throw std::runtime_error("Not enough GPU memory for the chunk division (" +
                                         std::to_string(c.vPadTop) + "-" + std::to_string(c.vSize) + "-" +
                                         std::to_string(c.vPadBot) + ", " +
                                         std::to_string(chunk_bytes(c, W, ds.width(), ds.concurrent(),
                                                                    p.xpr.storage) >> 20) + "Mo)");
            Chunk n{chunks.back().vOffset + chunks.back().vSize, 1, 0, 0, 0, 0};
            viewport(n);
            padding(n);
            if (!fits(n)) throw std::runtime_error("Not enough GPU memory for the chunk division");
            chunks.push_back(n);
        }
    }
    if (lmin_out) *lmin_out = lmin;
    if (lmax_out) *lmax_out = lmax;
    return chunks;
}

class ReconstructionChunk : public Reconstruction {
public:
    explicit ReconstructionChunk(Parameters &p) : Reconstruction(p)
    {
        chunks_ = plan_chunks(p, *ds_, budget());
        const bool talk = !p.xpr.quiet;
        if (talk) std::cout << "Chunks : ";
        Json cj = Json::array();
        for (const auto &c : chunks_) {
            if (talk) std::cout << " [" << c.vOffset << "-" << c.vOffset + c.vSize << "],";
            Json e = Json::array();
            for (int64_t v : {c.vOffset, c.vSize, c.iOffset, c.iSize, c.vPadTop, c.vPadBot})
                e.arr.push_back(Json::number(double(v)));
            cj.arr.push_back(e);
        }
        if (talk) std::cout << std::endl;
        note("chunks", cj);
        check_memory(requiredGPUMemory(), "method_2");
    }
    uint64_t requiredGPUMemory() override
    {
        uint64_t m = 0;
        for (const auto &c : chunks_)
            m = std::max(m, chunk_bytes(c, ds_->vwidth(), ds_->width(), ds_->concurrent(), storage()));
        return m + ds_->projdata().size() * 4;
    }
    void exec() override
    {
        AllocStreamScope scope(stream_);
        auto &ds = *ds_;
        ds.load(stream_);
        const int st = storage();
        const int64_t W = ds.vwidth(), As = ds.concurrent(), S = p_.recon.sit, dw = ds.width();
        const size_t layer = size_t(W * W);
        DevBuf stage;
        std::vector<DevBuf> pds(static_cast<size_t>(S));
        std::vector<std::vector<float>> pd_host(static_cast<size_t>(S));
        for (int64_t s = 0; s < S; ++s) {
            pd_host[size_t(s)] = ds.sit_projdata(s);
            const auto &pd = pd_host[size_t(s)];
            pds[size_t(s)].alloc(pd.size() * 4);
            check(xpr_upload(stream_, pds[size_t(s)].get(), pd.data(), pd.size() * 4), "upload");
        }
        Event e0, e1, fb, fe, bb, be;
        note("launches_per_pass", Json::number(1));
        NarrowSums narrow;             // 32-bit sums where provably exact, per chunk
        double max_hits = 0.0;
        for (const Chunk &c : chunks_) {
            if (!p_.xpr.quiet) std::cout << "[" << c.vOffset << "-" << c.vOffset + c.vSize << "]... " << std::flush;
            const int64_t l0 = c.vOffset - c.vPadTop, l1 = c.vOffset + c.vSize + c.vPadBot;
            const uint64_t nv = uint64_t(l1 - l0) * layer, npx = uint64_t(c.iSize * dw * As);
            DevBuf vol(nv * dtype_bytes(st)), img(npx * dtype_bytes(st)),
                sums(npx * acc_bytes(p_.xpr.accumulator)), err(npx * dtype_bytes(st));
            check(xpr_fill(stream_, vol.get(), st, 1.0f, nv), "fill");
            fill_sums(stream_, sums, p_.xpr.accumulator, true, npx);
            Operators ops(stream_, p_.xpr.accumulator, p_.xpr.interp);
            ops.setSumImagesBuffer(sums.get());
            ops.setErrorBuffer(err.get(), st);
            ops.setImagesBuffer(img.get(), st);
            ops.setVolumeBuffer(vol.get(), st);
            ops.setImageParameters(dw, c.iSize, As);            // elements = W*iSize*A_s (Q10)
            ops.setImageOffset(int(c.iOffset), int(c.iSize));
            ops.setAngleNumber(As, p_.detector.module_number);
            ops.setVolumeParameters({0, int(l0), 0}, {int(W), int(l1), int(W)}, ds.vx());
            narrow.init(p_.xpr.narrow_sums, p_.xpr.accumulator, st,
                        p_.xpr.narrow_sums ? launch_hits(pd_host, ops.args(), ds.orig()) : std::vector<double>());
            if (narrow.on()) max_hits = std::max(max_hits, narrow.max_hits());
            for (int64_t mit = 0; mit < p_.recon.it; ++mit) {
                ops.setWeight(next_weight(mit));
                for (int64_t s = 0; s < S; ++s) {
                    check(xpr_event_record(e0.e, stream_), "event");
                    ds.upload_sit_images(stream_, s, c.iOffset, c.iSize, img.get(), st);
                    ops.setOrigin(next_origin());
                    ops.setProjDataBuffer(pds[size_t(s)].as<float>());
                    const bool fwd = s > 0 || mit > 0;
                    ops.setNarrowSums(fwd && narrow.choose(s));
                    if (fwd) {
                        check(xpr_event_record(fb.e, stream_), "event");
                        ops.forward();
                        check(xpr_event_record(fe.e, stream_), "event");
                    }
                    ops.error();
                    ops.setMaxBuffer(narrow.before_backward(stream_));
                    check(xpr_event_record(bb.e, stream_), "event");
                    ops.backward();
                    check(xpr_event_record(be.e, stream_), "event");
                    narrow.after_backward(stream_);
                    fill_sums(stream_, sums, p_.xpr.accumulator, false, npx);
                    check(xpr_event_record(e1.e, stream_), "event");
                    check(xpr_event_sync(e1.e), "sync");
                    narrow.synced();
                    float ms = 0, fms = -1.0f, bms = 0;
                    check(xpr_event_elapsed_ms(e0.e, e1.e, &ms), "elapsed");
                    if (fwd) check(xpr_event_elapsed_ms(fb.e, fe.e, &fms), "elapsed");
                    check(xpr_event_elapsed_ms(bb.e, be.e, &bms), "elapsed");
                    log_sit(mit, s, ms, fms, bms);
                }
            }
            // keep only the chunk's own layers (reconstruction_chunk.hpp:121-127)
            write_final(ds, stream_,
                        static_cast<const char *>(vol.get()) + size_t(c.vPadTop) * layer * dtype_bytes(st),
                        st, c.vOffset, c.vOffset + c.vSize, stage);
            if (!p_.xpr.quiet) std::cout << "Done." << std::endl;
        }
        note("forward_sum_bits", narrow.bits());
        if (p_.xpr.narrow_sums && p_.xpr.accumulator == XPR_ACC_U64) note("hit_bound", Json::number(max_hits));
    }

private:
    std::vector<Chunk> chunks_;
};

// ---------------------------------------------------------------------------
// forward_exec: reprojection of the saved volume (src/reconstruction.hpp:66-158)
// ---------------------------------------------------------------------------
void Reconstruction::forward_exec()
{
    AllocStreamScope scope(stream_);
    auto &ds = *ds_;
    const int64_t W = ds.vwidth(), H = ds.vheight(), As = ds.concurrent(), S = p_.recon.sit;
    const int64_t dw = ds.width(), dh = ds.height();
    const size_t layer = size_t(W * W);
    const int samples = 32;
    std::vector<float> vol(layer * size_t(H));
    parallel_for(H, io_threads(p_), [&](int64_t l) {
        float *dst = vol.data() + size_t(l) * layer;
        ds.load_final_layer(l, dst);
        for (size_t k = 0; k < layer; ++k) dst[k] /= float(samples);
        for (int64_t j = 0; j < W; ++j)
            for (int64_t i = 0; i < W; ++i)
                if (std::sqrt(std::pow(float(i) - float(W) * 0.5f, 2.0f) +
                              std::pow(float(j) - float(W) * 0.5f, 2.0f)) >= float(W) * 0.5f - 1.0f)
                    dst[size_t(j * W + i)] = 0.0f;
    });
    DevBuf dvol(vol.size() * 4), sums(uint64_t(As * dw * dh) * 8), pdb, fsum(uint64_t(As * dw * dh) * 4);
    check(xpr_upload(stream_, dvol.get(), vol.data(), vol.size() * 4), "upload");
    Operators ops(stream_, 64);
    ops.setSumImagesBuffer(sums.get());
    ops.setVolumeBuffer(dvol.get(), XPR_F32);
    ops.setImageParameters(dw, dh, As);
    ops.setImageOffset(0, int(dh));
    ops.setAngleNumber(As, p_.detector.module_number);
    ops.setVolumeParameters({0, 0, 0}, {int(W), int(H), int(W)}, ds.vx());
    std::vector<float> out(size_t(As * dw * dh));
    for (int64_t s = 0; s < S; ++s) {
        auto pd = ds.sit_projdata(s);
        pdb.alloc(pd.size() * 4);
        check(xpr_upload(stream_, pdb.get(), pd.data(), pd.size() * 4), "upload");
        ops.setProjDataBuffer(pdb.as<float>());
        check(xpr_fill_u64(stream_, sums.as<uint64_t>(), 0, uint64_t(As * dw * dh)), "fill");
        for (int k = 0; k < samples; ++k) {
            ops.setOrigin(next_origin());
            ops.forward();
        }
        check(xpr_q22_to_float(stream_, sums.as<uint64_t>(), 64, fsum.as<float>(), uint64_t(As * dw * dh)),
              "convert sums");
        check(xpr_download(stream_, out.data(), fsum.get(), out.size() * 4), "download");
        check(xpr_stream_sync(stream_), "sync");
        const auto views = ds.sit_views(s);
        for (size_t j = 0; j < views.size(); ++j)
            ds.save_projection(out.data() + j * size_t(dw * dh), dw, dh, views[j]);
    }
}

std::unique_ptr<Reconstruction> make_reconstruction(Parameters &p)
{
    switch (p.recon.method) {
    case Method::naive: return std::make_unique<ReconstructionN>(p);
    case Method::method_1a: return std::make_unique<ReconstructionSlab>(p, 1, false);
    case Method::method_1b: {
        int t = p.xpr.threads > 0 ? p.xpr.threads : 12;                // ASYNC_THREAD
        return std::make_unique<ReconstructionSlab>(p, std::min(t, 16), false);
    }
    case Method::method_1c: return std::make_unique<ReconstructionSlab>(p, io_threads(p), true);
    case Method::method_2: return std::make_unique<ReconstructionChunk>(p);
    case Method::hartmann:
    case Method::ipr:
        throw std::runtime_error(std::string("method '") + method_name(p.recon.method) +
                                 "' (a separate phase-imaging modality) is not part of this build");
    }
    throw std::runtime_error("unknown method");
}

} // namespace xpr
