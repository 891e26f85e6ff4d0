// XPRecon command line (src/main.cpp:62-137): same arguments, same exit codes.
//   XPRecon -template          print the default parameters as JSON
//   XPRecon <json_file>        run with a parameter file
//   XPRecon '{...}'            run with inline JSON
#include <fstream>
#include <iostream>
#include <string>

#include "json.hpp"
#include "params.hpp"
#include "recon.hpp"

namespace {
constexpr int RETURN_CODE_OK = 0;
constexpr int RETURN_CODE_ERROR_ARGUMENTS = -1;
constexpr int RETURN_CODE_ERROR_PARAMETERS = -2;
constexpr int RETURN_CODE_ERROR_RECONSTRUCTION = -3;

void usage()
{
    std::cout << "XPRecon -template|<json_file>|<json>\n"
              << "\t-template : Output in stdout a default JSON file\n"
              << "\t<json_file> : path to the parameters file\n"
              << "\t<json> : parameters as a JSON string (starting with a '{')" << std::endl;
}
} // namespace

int main(int argc, char *argv[])
{
    if (argc < 2) {
        usage();
        return RETURN_CODE_ERROR_ARGUMENTS;
    }
    if (std::string(argv[1]) == "-template") {
        std::cout << xpr::json_dump(xpr::parameters_to_json(xpr::Parameters{})) << std::endl;
        return RETURN_CODE_OK;
    }
    xpr::Parameters params;
    try {
        params = xpr::load_parameters(argv[1]);
    } catch (const std::exception &e) {
        std::cerr << e.what() << std::endl;
        return RETURN_CODE_ERROR_PARAMETERS;
    }
    std::unique_ptr<xpr::Reconstruction> recon;
    try {
        recon = xpr::make_reconstruction(params);
    } catch (const std::exception &e) {
        std::cerr << "Error while creating the reconstruction object : " << e.what() << std::endl;
        return RETURN_CODE_ERROR_RECONSTRUCTION;
    }
    try {
        if (!params.recon.simulation) recon->exec();
        if (!params.recon.proj_output.empty()) recon->forward_exec();
    } catch (const std::exception &e) {
        // the reference has no handler here (an exception ends the process);
        // report it with the construction error code
        std::cerr << "Error during the reconstruction : " << e.what() << std::endl;
        return RETURN_CODE_ERROR_RECONSTRUCTION;
    }
    if (!params.xpr.run_log.empty()) {
        std::ofstream f(params.xpr.run_log);
        f << xpr::json_dump(recon->run_record()) << std::endl;
    }
    recon.reset();
    return RETURN_CODE_OK;
}
