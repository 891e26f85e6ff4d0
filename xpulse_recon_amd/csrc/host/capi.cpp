// In-process entry to the XPRecon drop-in (the CLI's body, src/main.cpp:69-137),
// for callers that load the library instead of spawning the executable.
#include <cstdint>
#include <cstring>
#include <iostream>
#include <string>

#include "json.hpp"
#include "params.hpp"
#include "recon.hpp"

namespace {

int run(const char *params, const float *projections, int64_t width, int64_t height,
        int64_t count, float *volume_out, char *record, size_t record_cap, bool on_device = false)
{
    if (record && record_cap) record[0] = '\0';
    if (!params) return -1;                                  // RETURN_CODE_ERROR_ARGUMENTS
    xpr::Parameters p;
    try {
        p = xpr::load_parameters(params);
    } catch (const std::exception &e) {
        std::cerr << e.what() << std::endl;
        return -2;                                           // RETURN_CODE_ERROR_PARAMETERS
    }
    if (projections) {                                       // in-memory source and sink
        p.xpr.mem_projections = projections;
        p.xpr.mem_width = width;
        p.xpr.mem_height = height;
        p.xpr.mem_count = count;
        p.xpr.mem_volume_out = volume_out;
        p.xpr.mem_on_device = on_device;
    }
    try {
        auto recon = xpr::make_reconstruction(p);
        if (!p.recon.simulation) recon->exec();
        if (!p.recon.proj_output.empty()) recon->forward_exec();
        const std::string rec = xpr::json_dump(recon->run_record(), 0);
        if (record && record_cap) {
            std::strncpy(record, rec.c_str(), record_cap - 1);
            record[record_cap - 1] = '\0';
        }
    } catch (const std::exception &e) {
        std::cerr << "Error while creating the reconstruction object : " << e.what() << std::endl;
        return -3;                                           // RETURN_CODE_ERROR_RECONSTRUCTION
    }
    return 0;
}

} // namespace

extern "C" int xpr_reconstruct(const char *params, char *record, size_t record_cap)
{
    return run(params, nullptr, 0, 0, 0, nullptr, record, record_cap);
}

extern "C" int xpr_reconstruct_mem(const char *params, const float *projections, int64_t width,
                                   int64_t height, int64_t count, float *volume_out, char *record,
                                   size_t record_cap)
{
    if (!projections || !volume_out || width <= 0 || height <= 0 || count <= 0) return -1;
    return run(params, projections, width, height, count, volume_out, record, record_cap);
}

int xprint_fail(int code, const char *msg);

extern "C" int xpr_plan_chunks(const char *params, int64_t width, int64_t height, int64_t count,
                               int64_t *chunks_out, int64_t cap, int64_t *n_chunks)
{
    if (n_chunks) *n_chunks = 0;
    if (!params || !n_chunks || width <= 0 || height <= 0 || count <= 0)
        return xprint_fail(XPR_E_INVALID, "xpr_plan_chunks: null argument or empty projection stack");
    xpr::Parameters p;
    try {
        p = xpr::load_parameters(params);
    } catch (const std::exception &e) {
        return xprint_fail(XPR_E_INVALID, e.what());
    }
    try {
        if (!(p.xpr.device_budget_gb > 0.0))
            return xprint_fail(XPR_E_INVALID, "xpr_plan_chunks: xpr.device_budget_gb must be set");
        p.xpr.shape_only = true;
        p.xpr.quiet = true;
        p.xpr.mem_width = width;
        p.xpr.mem_height = height;
        p.xpr.mem_count = count;
        xpr::Dataset ds(p);
        const auto chunks = xpr::plan_chunks(p, ds, uint64_t(p.xpr.device_budget_gb * 1e9));
        *n_chunks = int64_t(chunks.size());
        if (!chunks_out || cap < int64_t(chunks.size()))
            return xprint_fail(XPR_E_INVALID, "xpr_plan_chunks: chunks_out holds too few chunks");
        for (size_t i = 0; i < chunks.size(); ++i) {
            const auto &c = chunks[i];
            const int64_t v[6] = {c.vOffset, c.vSize, c.iOffset, c.iSize, c.vPadTop, c.vPadBot};
            std::memcpy(chunks_out + 6 * i, v, sizeof(v));
        }
    } catch (const std::exception &e) {
        return xprint_fail(XPR_E_NOMEM, e.what());
    }
    return XPR_OK;
}

extern "C" int xpr_reconstruct_dev(const char *params, const float *projections, int64_t width,
                                   int64_t height, int64_t count, float *volume_out, char *record,
                                   size_t record_cap)
{
    if (!projections || !volume_out || width <= 0 || height <= 0 || count <= 0) return -1;
    return run(params, projections, width, height, count, volume_out, record, record_cap, true);
}

extern "C" int xpr_template_json(char *out, size_t cap)
{
    const std::string t = xpr::json_dump(xpr::parameters_to_json(xpr::Parameters{}));
    if (!out || cap <= t.size()) return -1;
    std::memcpy(out, t.c_str(), t.size() + 1);
    return 0;
}
