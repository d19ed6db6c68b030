// Host-only driver for the NC.inp parser (sonium_amd/csrc/numcalc.cpp) under
// AddressSanitizer and UndefinedBehaviorSanitizer: parses every file named on
// the command line with sbh_nc_parse and, when that succeeds, reads the
// project back through the info/values/string accessors.  The parser's only
// links to the rest of the library are the error string and mesh creation,
// defined here (mesh creation is never reached by parsing).  Exit status 0
// means every input returned a status; a sanitizer report aborts the run.
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "internal.hpp"

namespace sbh {
namespace {
thread_local std::string g_error;
}
void set_last_error(const std::string& msg) { g_error = msg; }
}  // namespace sbh

extern "C" const char* sbh_last_error(void) { return sbh::g_error.c_str(); }

extern "C" sbh_status sbh_mesh_create_oriented(sbh_context*, const double*, int64_t, const int64_t*, int64_t,
                                               const int32_t*, const sbh_c64*, const sbh_c64*, const uint8_t*, int,
                                               sbh_mesh**) {
    return SBH_ERR_UNSUPPORTED;
}

int main(int argc, char** argv) {
    int ok = 0, errors = 0;
    for (int a = 1; a < argc; ++a) {
        std::ifstream f(argv[a], std::ios::binary);
        std::stringstream ss;
        ss << f.rdbuf();
        const std::string text = ss.str();
        sbh_nc_project* p = nullptr;
        const sbh_status st = sbh_nc_parse(text.c_str(), ".", &p);
        if (st != SBH_OK) {
            ++errors;
            if (sbh_last_error()[0] == 0) {
                std::fprintf(stderr, "%s: status %d without a message\n", argv[a], (int)st);
                return 2;
            }
            continue;
        }
        ++ok;
        sbh_nc_info info{};
        if (sbh_nc_project_info(p, &info) != SBH_OK) return 3;
        for (int which = 0; which < 8; ++which) {
            int64_t count = 0;
            if (sbh_nc_project_values(p, which, nullptr, 0, &count) == SBH_OK && count > 0 && count < (1 << 24)) {
                std::vector<double> v((size_t)count);
                (void)sbh_nc_project_values(p, which, v.data(), count, &count);
            }
            char buf[64];
            (void)sbh_nc_project_string(p, which, 0, buf, sizeof(buf));
        }
        sbh_nc_project_destroy(p);
    }
    std::printf("parsed %d, rejected %d\n", ok, errors);
    return 0;
}
