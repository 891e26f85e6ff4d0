// In-process entry to the XPRecon drop-in (the CLI's body, src/main.cpp:69-137),
// for callers that load the library instead of spawning the executable.
#include <cstring>
#include <iostream>
#include <string>

#include "json.hpp"
#include "params.hpp"
#include "recon.hpp"

extern "C" int xpr_reconstruct(const char *params, char *record, size_t record_cap)
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

extern "C" int xpr_template_json(char *out, size_t cap)
{
    const std::string t = xpr::json_dump(xpr::parameters_to_json(xpr::Parameters{}));
    if (!out || cap <= t.size()) return -1;
    std::memcpy(out, t.c_str(), t.size() + 1);
    return 0;
}
