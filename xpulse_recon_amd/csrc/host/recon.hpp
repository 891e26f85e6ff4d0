// Host engine of the reconstruction: the reference's Dataset / DataLoader
// (dataset.hpp, dataset/data_loader.hpp), its operator API (Reconstruction,
// src/reconstruction.hpp:30-316) and the five memory schedules
// (reconstruction_{n,lpln,lpla,lplh,chunk}.hpp), driving the HIP kernels
// through the C-ABI of include/xpr_hip.h.
#pragma once

#include <array>
#include <functional>
#include <cstdint>
#include <memory>
#include <random>
#include <string>
#include <vector>

#include "../../../include/xpr_hip.h"
#include "json.hpp"
#include "params.hpp"

namespace xpr {

struct XprFailure : std::runtime_error {
    using std::runtime_error::runtime_error;
};
void check(int rc, const char *what);

// Stream for the device allocations of the calling thread: while a scope is
// alive, DevBuf allocates and frees stream-ordered on it (no device-wide
// synchronisation, so concurrent jobs on other streams keep running).
class AllocStreamScope {
public:
    explicit AllocStreamScope(xpr_stream s);
    ~AllocStreamScope();

private:
    xpr_stream prev_;
};

// Device allocation owned by the host engine.
class DevBuf {
public:
    DevBuf() = default;
    explicit DevBuf(size_t bytes) { alloc(bytes); }
    ~DevBuf();
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p_(o.p_), n_(o.n_), s_(o.s_) { o.p_ = nullptr; o.n_ = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept;
    void alloc(size_t bytes);
    void reset();
    void *get() const { return p_; }
    template <typename T> T *as() const { return (T *)p_; }
    size_t bytes() const { return n_; }

private:
    void *p_ = nullptr;
    size_t n_ = 0;
    xpr_stream s_ = nullptr;        // stream it was allocated on (stream-ordered)
};

// Pinned host allocation.
class HostBuf {
public:
    HostBuf() = default;
    explicit HostBuf(size_t bytes) { alloc(bytes); }
    ~HostBuf();
    HostBuf(const HostBuf &) = delete;
    HostBuf &operator=(const HostBuf &) = delete;
    void alloc(size_t bytes);
    void *get() const { return p_; }
    template <typename T> T *as() const { return (T *)p_; }
    size_t bytes() const { return n_; }

private:
    void *p_ = nullptr;
    size_t n_ = 0;
};

inline size_t dtype_bytes(int dt) { return dt == XPR_F32 ? 4 : 2; }

// ---------------------------------------------------------------------------
// Dataset: projections on disk, geometry, preprocessed images and the layer
// store (dataset.hpp:18-270, data_loader.hpp:13-309)
// ---------------------------------------------------------------------------
class Dataset {
public:
    explicit Dataset(Parameters &p);
    ~Dataset();

    int64_t width() const { return dwidth_; }
    int64_t height() const { return dheight_; }
    int64_t projections() const { return nproj_; }
    int64_t concurrent() const { return nproj_ / p_.recon.sit; }     // dataset.hpp:225
    int64_t vwidth() const { return vwidth_; }
    int64_t vheight() const { return vheight_; }
    float vx() const { return vx_; }
    std::array<float, 3> orig() const { return orig_; }

    // ProjData of all (angle, module) records and of one sub-iteration.
    const std::vector<float> &projdata() const { return pd_; }
    std::vector<float> sit_projdata(int64_t sit) const;
    std::vector<int64_t> sit_views(int64_t sit) const;

    // Raw projections in host memory (TIFF sources are read once by the IO
    // threads; in-process jobs use the caller's buffer) and, when normalising,
    // the global min/max over isnormal samples, reduced on the device
    // (initializeTempImages, data_loader.hpp:48-101).  Idempotent.
    void load(xpr_stream s);
    bool quiet() const { return p_.xpr.quiet; }
    // Rows [row0, row0+rows) of the views of one sub-iteration, preprocessed
    // (normalise, -log) on the device into dst [view][row][col] of dtype dt.
    // Uploads go through two pinned staging slots so the host copy of one
    // chunk overlaps the transfer of the other; everything is ordered on s.
    void upload_sit_images(xpr_stream s, int64_t sit, int64_t row0, int64_t rows, void *dst, int dt);
    // The caller's volume buffer of an in-process job ([vheight][vwidth][vwidth]),
    // or null; a device address when device_io().
    float *volume_sink() const { return p_.xpr.mem_volume_out; }
    bool device_io() const { return p_.xpr.mem_on_device; }
    bool normalize() const { return p_.recon.normalize; }
    float norm_min() const { return mn_; }
    float norm_max() const { return mx_; }

    // Layer store between passes (getLayer/saveLayer, data_loader.hpp:118-146):
    // "ram" is one pinned host block of the whole volume (slabs move by DMA
    // straight from and to it), "zstd" per-layer spill files.
    void save_layer(const float *data, int64_t layer, bool finalize);
    void load_layer(int64_t layer, float *out) const;
    bool ram_layers() const { return p_.xpr.layer_store == "ram"; }
    float *ram_layer(int64_t layer) const;
    // The output directory for spill files (created once).
    void ensure_output_dir();
    // Final layers [l0, l1) of host float data: into the host sink or TIFF
    // files, de-normalised (saveLayer(finalize), data_loader.hpp:133-146).
    void save_final_layers(const float *vol, int64_t l0, int64_t l1);
    // The final volume as written (sink or TIFF layers), for reprojection.
    void load_final_layer(int64_t layer, float *out) const;
    void save_projection(const float *img, int64_t w, int64_t h, int64_t id) const;

private:
    Parameters &p_;
    std::vector<std::string> files_;
    int64_t dwidth_ = 0, dheight_ = 0, nproj_ = 0;
    int64_t vwidth_ = 0, vheight_ = 0;
    float vx_ = 0.0f;
    std::array<float, 3> orig_{};
    std::vector<float> pd_;
    std::vector<std::vector<float>> raw_;       // raw projections of TIFF sources
    HostBuf pin_[2];                            // pinned staging slots
    xpr_event slot_done_[2] = {nullptr, nullptr};
    HostBuf ram_block_;                         // layer store when "ram" (pinned)
    float mn_ = 0.0f, mx_ = 1.0f;
    std::string proj_dir_;
    bool loaded_ = false;
    std::vector<float> raw_image(int64_t i) const;
    const float *raw(int64_t i) const;
    // stream the rows of `views` through the staging slots into device
    // memory; fn(device rows, first view, views in chunk) runs after each upload
    void stream_views(xpr_stream s, const std::vector<int64_t> &views, int64_t row0, int64_t rows,
                      const std::function<void(const float *, size_t, size_t)> &fn);
};

// ---------------------------------------------------------------------------
// Operator API: the reference's positional kernel state, kept on the host
// and handed by value to every stateless C-ABI call.
// ---------------------------------------------------------------------------
// bytes per detector sum of an accumulator (XPR_ACC_*)
inline uint64_t acc_bytes(int acc) { return acc == XPR_ACC_U64 ? 8u : 4u; }

class Operators {
public:
    Operators(xpr_stream s, int acc_bits, int interp = XPR_INTERP_NEAREST) : stream_(s), acc_(acc_bits)
    {
        a_.interp = interp;
    }
    void setImagesBuffer(const void *img, int dtype) { img_ = img; img_dt_ = dtype; }
    void setSumImagesBuffer(void *sums) { sums_ = sums; }
    void setErrorBuffer(void *err, int dtype) { err_ = err; err_dt_ = dtype; }
    void setVolumeBuffer(void *vol, int dtype) { vol_ = vol; vol_dt_ = dtype; }
    void setProjDataBuffer(const float *pd) { pd_ = pd; }
    void setImageParameters(int64_t w, int64_t h, int64_t n);
    void setVolumeParameters(std::array<int, 3> mn, std::array<int, 3> mx, float voxel);
    void setAngleNumber(int64_t n, int64_t mods) { a_.n_views = uint32_t(n); a_.modules = uint32_t(mods); }
    void setWeight(float w) { w_ = w; }
    void setOrigin(std::array<float, 3> o) { for (int i = 0; i < 3; ++i) a_.origin[i] = o[i]; }
    void setImageOffset(int offset, int rows) { a_.line_offset = offset; a_.line_rows = rows; }
    void setImageElements(int64_t n) { elements_ = uint64_t(n); }
    void setViewIndex(const uint32_t *idx) { a_.view_index = idx; }
    // 32-bit detector sums for the next forward and error (only in place of
    // 64-bit ones, and only when a bound proves every sum below 2^32: the
    // sums are then the same integers)
    void setNarrowSums(bool narrow) { narrow_ = narrow && acc_ == XPR_ACC_U64; }
    // device uint32 the backward raises to its largest voxel value's bits (or NULL)
    void setMaxBuffer(uint32_t *max_bits) { max_bits_ = max_bits; }
    void forward();
    void error();
    void backward();
    xpr_stream stream() const { return stream_; }
    int acc_bits() const { return acc_; }
    const xpr_proj_args &args() const { return a_; }

private:
    xpr_stream stream_;
    int acc_;
    bool narrow_ = false;
    uint32_t *max_bits_ = nullptr;
    const void *img_ = nullptr;
    void *sums_ = nullptr, *err_ = nullptr, *vol_ = nullptr;
    int img_dt_ = XPR_F32, err_dt_ = XPR_F32, vol_dt_ = XPR_F32;
    const float *pd_ = nullptr;
    xpr_proj_args a_{};
    float w_ = 1.0f;
    uint64_t elements_ = 0;
};

// ---------------------------------------------------------------------------
// Schedules
// ---------------------------------------------------------------------------
struct Chunk {                      // specialized_parameters.hpp:28-32
    int64_t vOffset, vSize, iOffset, iSize, vPadTop, vPadBot;
};

class Reconstruction {
public:
    explicit Reconstruction(Parameters &p);
    virtual ~Reconstruction();
    virtual uint64_t requiredGPUMemory() = 0;      // requieredGPUMemory
    virtual void exec() = 0;
    void forward_exec();                            // src/reconstruction.hpp:66-158
    Json run_record() const { return record_; }
    Dataset &dataset() { return *ds_; }

protected:
    std::array<float, 3> next_origin();             // setOrigin(orig, vx) draws
    uint64_t budget() const { return budget_; }
    void check_memory(uint64_t need, const char *what);
    int storage() const { return p_.xpr.storage; }
    float next_weight(int64_t mit) const;
    void note(const std::string &k, Json v) { record_.set(k, std::move(v)); }
    // sub-iteration device time, and the time of its forward and backward
    // projector passes (negative: no such pass)
    void log_sit(int64_t mit, int64_t sit, double ms, double fwd_ms = -1.0, double bwd_ms = -1.0);
    // Times the projectors' tile shapes (xpr_tune_*) for launches of this
    // shape, once per shape in the process, when xpr.autotune is set.  vol_slab
    // must hold 1.0 (it keeps it: the error used is 0); sums is overwritten.
    void tune_shapes(void *vol_slab, int vdt, void *sums, void *err, int edt, const float *pd,
                     const xpr_proj_args &a);

    Parameters &p_;
    std::unique_ptr<Dataset> ds_;
    xpr_stream stream_ = nullptr, copy_ = nullptr, copy_back_ = nullptr;
    uint64_t budget_ = 0;
    std::mt19937 rng_;
    Json record_ = Json::object();
    Json jitter_log_ = Json::array();
    Json sit_ms_ = Json::array();
    Json fwd_ms_ = Json::array();
    Json bwd_ms_ = Json::array();
};

std::unique_ptr<Reconstruction> make_reconstruction(Parameters &p);

// Chunk planner (GeometryChunk, dataset/geometry_chunk.hpp:92-197).
std::vector<Chunk> plan_chunks(const Parameters &p, const Dataset &ds, uint64_t budget_bytes,
                               std::vector<float> *layer_min = nullptr,
                               std::vector<float> *layer_max = nullptr);
uint64_t chunk_bytes(const Chunk &c, int64_t vwidth, int64_t dwidth, int64_t concurrent,
                     int storage);

} // namespace xpr
