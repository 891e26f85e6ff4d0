// Reconstruction parameters: the reference's JSON surface (Parameters,
// dataset/parameters.hpp:19-27; ReconstructionData / GeometryData /
// DetectorData, parameters/global_parameters.hpp:8-54; ReconMethod2Data /
// ReconHartmannData / ReconIPRData, parameters/specialized_parameters.hpp:8-39)
// plus an "xpr" section for the options this build adds.
#pragma once

#include <cstdint>
#include <optional>
#include <string>
#include <vector>

#include "json.hpp"

namespace xpr {

enum class Method { naive, method_1a, method_1b, method_1c, method_2, hartmann, ipr };
const char *method_name(Method m);

struct ReconSection {                       // global_parameters.hpp:8-21
    std::string input = "./";
    std::string output = "./recon/";
    std::string proj_output = "";
    Method method = Method::method_1b;
    int64_t it = 1, sit = 10;
    bool mlog = true;
    float weight = 1.0f, weight_factor = 0.9f;
    bool normalize = false;
    bool simulation = false;
};

struct GeometrySection {                    // global_parameters.hpp:23-36
    float so = 100.0f, sd = 1000.0f, angle = 360.0f;
    float heli_offset = 0.0f, heli_step = 1.0f;
    std::vector<float> angle_list;
};

struct DetectorSection {                    // global_parameters.hpp:46-54
    float px = 0.1f, rx = 0.0f, sx = 0.0f, sy = 0.0f;   // rx in pixels
    int64_t module_number = 1;
    std::optional<float> module_angle_offset;
    float module_center_offset_z = 0.0f;
    // module rotation (README "modules" entries; not applied by the reference)
    float roll = 0.0f, pitch = 0.0f, yaw = 0.0f;
};

struct Method2Section {                     // specialized_parameters.hpp:34-39
    float overlap_fct = 1.0f;
};

struct HartmannSection {                    // parse only (modality not built)
    std::string devx = "./", devy = "./";
    float iterations = 0, iterations_fct = 1;
    int64_t iterations_step = 1;
    std::string method = "gs";
};

struct IprSection {                         // parse only (modality not built)
    float iterations = 0, iterations_fct = 1;
    int64_t iterations_step = 1;
    float energy_kev = 1;
};

// Options of this build (no counterpart in the reference).
struct XprSection {
    int storage = 0;                 // 0 f32, 1 f16, 2 bf16 for volume and projections
    bool final_fp32 = true;          // last sub-iteration with FP32 storage when storage != f32
    int accumulator = 64;            // detector sums: 64 or 32 (the reference's uint32) bit
                                     // Q.22, or XPR_ACC_F32 ("f32": float sums, opt-in)
    int interp = 0;                  // backprojector sampling: XPR_INTERP_NEAREST (the
                                     // reference's) or XPR_INTERP_BILINEAR ("bilinear", opt-in)
    bool jitter = true;              // per-sub-iteration origin jitter U[0, vx)^3
    int64_t jitter_seed = -1;        // -1: std::random_device like the reference
    bool geometry_fp64 = false;      // FP64 ProjData construction (default: float32 as glm)
    int64_t slab_height = 0;         // layers per slab for method_1a/1b/1c; 0: from the budget
    double device_budget_gb = 0.0;   // 0: free device memory
    std::string layer_store = "auto";// method_1a/1b/1c layer store between passes: "device"
                                     // (HBM), "ram" (pinned host memory), "zstd" (host spill
                                     // files) or "auto" (device when it fits the budget, else zstd)
    int threads = 0;                 // host IO threads (0: hardware concurrency, max 16)
    std::string run_log = "";        // JSON run record (jitter, timings, memory)
    bool autotune = true;            // time the projectors' tile shapes once per launch shape
    bool narrow_sums = true;         // 32-bit detector sums in a forward over a resident volume when
                                     // a bound proves every sum below 2^32 (the same integers)
    int device = 0;
    // in-process jobs (not JSON): projections from memory, volume to memory
    const float *mem_projections = nullptr;   // [n][h][w] raw images
    int64_t mem_width = 0, mem_height = 0, mem_count = 0;
    float *mem_volume_out = nullptr;          // [vheight][vwidth][vwidth]
    bool mem_on_device = false;               // mem_projections / mem_volume_out are device
                                              // addresses (inputs resident in HBM)
    bool shape_only = false;                  // geometry of mem_width x mem_height x mem_count
                                              // projections only: no data, no output files
    bool quiet = false;                       // no progress output
};

struct Parameters {
    ReconSection recon;
    GeometrySection geometry;
    DetectorSection detector;
    Method2Section method2;
    HartmannSection hartmann;
    IprSection ipr;
    XprSection xpr;
};

// Reads a file path or inline JSON (a string starting with '{'), applies the
// clamps of initializeParameters (parameters.hpp:34-48).  Throws on error.
Parameters load_parameters(const std::string &file_or_content);
Parameters parameters_from_json(const Json &j);
Json parameters_to_json(const Parameters &p);

} // namespace xpr
