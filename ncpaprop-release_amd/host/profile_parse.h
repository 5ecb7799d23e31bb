// The profile file reader shared by ncpa_profile_read and ncpa_profile_check (host only).
#pragma once
#include <string>
#include <vector>

namespace ncpa {

// SI node quantities of one profile (z ascending)
struct ParsedProfile {
    int n = 0;
    std::vector<double> z, u, v, T, P, rho, c0;
    bool has_z0 = false;
    double z0 = 0.0;
};

// NCPA_OK, or NCPA_ERR_ARG with the reference's message in err
int parse_profile_file(const std::string &path, const std::string &header_path, ParsedProfile &out,
                       std::string &err);

}  // namespace ncpa
