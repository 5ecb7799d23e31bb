// Atmosphere profile file reader and the node quantities ESSModeSolver::setParams derives
// (host C++, no device: the C ABI's ncpa_profile_read uploads what this returns and
// ncpa_profile_check reports its errors without a context).
//
// Format and semantics follow Atmosphere1D::read_header_from_stream / read_values_from_stream
// (src/atmosphere/Atmosphere1D.cpp:85-285): '#%' lines describe columns as
// "col, KEY, units[, value]" (column 0 = scalar, column 1 = altitude), '#' lines are
// comments, data fields are separated by spaces, tabs or commas, keys must be unique.
// Units follow src/common/units.cpp:131-178,225-283.  The node quantities are left as
// ESSModeSolver::setParams leaves them (ESSModeSolver.cpp:159-220): z [m], U/V [m/s],
// T [K], P [Pa], RHO [kg/m3], c0 = C0 column, else sqrt(1.4 P/RHO), else sqrt(1.4*287*T).
// Every malformed input ends in an error status and the reference's message, never in an
// exception or a read past an array.
#include "profile_parse.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "ncpamodes.h"

namespace {

std::string trim(const std::string &s, const char *ws = " \t\n\r") {
    const size_t b = s.find_first_not_of(ws);
    if (b == std::string::npos) return "";
    const size_t e = s.find_last_not_of(ws);
    return s.substr(b, e - b + 1);
}

std::vector<std::string> split(const std::string &in, const std::string &delims) {
    std::vector<std::string> out;
    size_t first = in.find_first_not_of(delims);
    size_t last = in.find_first_of(delims, first);
    while (first != std::string::npos) {
        out.push_back(in.substr(first, last - first));
        first = in.find_first_not_of(delims, last);
        last = in.find_first_of(delims, first);
    }
    return out;
}

std::string upper(std::string s) {
    for (auto &c : s) c = (char)std::toupper((unsigned char)c);
    return s;
}

// canonical unit tags
const std::map<std::string, std::string> &unit_table() {
    static const std::map<std::string, std::string> t = {
        {"", "none"}, {"N/A", "none"},
        {"K", "K"}, {"DEGK", "K"}, {"DEG K", "K"}, {"DEGREES K", "K"}, {"KELVIN", "K"},
        {"C", "C"}, {"DEGC", "C"}, {"DEG C", "C"}, {"DEGREES C", "C"}, {"CELSIUS", "C"}, {"CENTIGRADE", "C"},
        {"F", "F"}, {"DEGF", "F"}, {"DEG F", "F"}, {"DEGREES F", "F"}, {"FAHRENHEIT", "F"},
        {"M", "m"}, {"METERS", "m"}, {"KM", "km"}, {"KILOMETERS", "km"},
        {"M/S", "m/s"}, {"MPS", "m/s"}, {"MPERS", "m/s"}, {"M PER S", "m/s"}, {"METERS PER SECOND", "m/s"},
        {"KM/S", "km/s"}, {"KMPS", "km/s"}, {"KMPERS", "km/s"}, {"KM PER S", "km/s"},
        {"KILOMETERS PER SECOND", "km/s"},
        {"PA", "Pa"}, {"PASCAL", "Pa"}, {"PASCALS", "Pa"},
        {"MBAR", "mbar"}, {"MILLIBAR", "mbar"}, {"MILLIBARS", "mbar"},
        {"HECTOPASCALS", "hPa"}, {"HPA", "hPa"}, {"ATMOSPHERES", "atm"}, {"ATM", "atm"},
        {"KG/M3", "kg/m3"}, {"KGPM3", "kg/m3"}, {"KILOGRAMS PER CUBIC METER", "kg/m3"},
        {"G/CM3", "g/cm3"}, {"GPCM3", "g/cm3"}, {"GRAMS PER CUBIC CENTIMETER", "g/cm3"},
        {"DEGREES CLOCKWISE FROM NORTH", "degcwn"}, {"DEG CW FROM N", "degcwn"}, {"AZIMUTH", "degcwn"},
        {"DEGREES COUNTERCLOCKWISE FROM EAST", "degccwe"}, {"DEG CCW FROM E", "degccwe"},
        {"DEG", "deg"}, {"DEGREES", "deg"}, {"RAD", "rad"}, {"RADIANS", "rad"},
    };
    return t;
}

bool convert_value(double in, const std::string &from, const std::string &to, double &out) {
    if (from == to) { out = in; return true; }
    struct C { const char *f, *t; double m; double a; };
    // out = in * m + a, the table of units.cpp:131-157 (temperature offsets included)
    static const C table[] = {
        {"C", "K", 1.0, 273.15}, {"m", "km", 0.001, 0.0}, {"km", "m", 1000.0, 0.0},
        {"m/s", "km/s", 0.001, 0.0}, {"km/s", "m/s", 1000.0, 0.0},
        {"Pa", "mbar", 0.01, 0.0}, {"Pa", "hPa", 0.01, 0.0}, {"mbar", "Pa", 100.0, 0.0},
        {"mbar", "hPa", 1.0, 0.0}, {"hPa", "Pa", 100.0, 0.0}, {"hPa", "mbar", 1.0, 0.0},
        {"atm", "hPa", 1013.25, 0.0}, {"atm", "Pa", 101325.0, 0.0}, {"atm", "mbar", 1013.25, 0.0},
        {"kg/m3", "g/cm3", 0.001, 0.0}, {"g/cm3", "kg/m3", 1000.0, 0.0},
    };
    for (const C &c : table)
        if (from == c.f && to == c.t) { out = in * c.m + c.a; return true; }
    if (from == "F" && to == "K") { out = ((in - 32.0) * 5.0 / 9.0) + 273.15; return true; }
    return false;
}

struct Column {
    int col;
    std::string key, units;
    double value;
};

}  // namespace

namespace ncpa {

int parse_profile_file(const std::string &path, const std::string &header_path, ParsedProfile &pp,
                       std::string &err) {
    auto fail = [&err](const std::string &msg) {
        err = msg;
        return NCPA_ERR_ARG;
    };
    const std::string hfile = header_path.empty() ? path : header_path;
    std::vector<std::string> headerlines;
    {
        std::ifstream in(hfile);
        if (!in.good()) return fail("Atmosphere1D - Input stream not in good state: " + hfile);
        std::string line;
        while (std::getline(in, line)) {
            line = trim(line);
            if (line.size() > 1 && line[0] == '#' && line[1] == '%') headerlines.push_back(line.substr(2));
        }
    }
    if (headerlines.empty()) return fail("Atmosphere1D - No descriptive fields found.");
    std::vector<Column> cols;
    for (const std::string &h : headerlines) {
        std::vector<std::string> f = split(trim(h), " ,");
        if (f.size() != 3 && f.size() != 4)
            return fail("Atmosphere1D - Error parsing descriptive line: " + h +
                        " (must be column,key,units[,value])");
        Column c;
        try {
            c.col = std::stoi(f[0]);
        } catch (...) {
            return fail("Atmosphere1D - First field not parseable as an integer: " + h);
        }
        if (c.col < 0 || c.col > 4096) return fail("Atmosphere1D - column number out of range: " + h);
        auto it = unit_table().find(upper(trim(f[2])));
        if (it == unit_table().end() || it->second == "none")
            return fail("Unrecognized units string: " + f[2]);
        c.units = it->second;
        c.key = trim(f[1]);
        c.value = 0.0;
        if (f.size() == 4) {
            try {
                c.value = std::stod(trim(f[3]));
            } catch (...) {
                return fail("Atmosphere1D - bad scalar value: " + h);
            }
        }
        cols.push_back(c);
    }
    {
        std::vector<std::string> keys;
        for (auto &c : cols) keys.push_back(c.key);
        std::sort(keys.begin(), keys.end());
        if (std::adjacent_find(keys.begin(), keys.end()) != keys.end())
            return fail("Atmosphere1D: key values are not unique");
    }
    int ncols = 0;
    std::string zunits;
    for (auto &c : cols) {
        ncols = std::max(ncols, c.col);
        if (c.col == 1) zunits = c.units;
    }
    std::vector<std::vector<double>> data(ncols);
    {
        std::ifstream in(path);
        if (!in.good()) return fail("Atmosphere1D - cannot open " + path);
        std::string line;
        while (std::getline(in, line)) {
            line = trim(line);
            if (line.empty() || line[0] == '#') continue;
            std::vector<std::string> f = split(line, " \t,");
            if ((int)f.size() != ncols) {
                std::ostringstream o;
                o << "Error parsing data line:\n" << line << "\n" << ncols << " columns expected, " << f.size()
                  << " columns found.";
                return fail(o.str());
            }
            // std::stod as the reference (Atmosphere1D.cpp:244-262): a field with a numeric
            // prefix reads as that prefix; out-of-range and non-numeric fields are errors
            for (int i = 0; i < ncols; ++i) {
                try {
                    data[i].push_back(std::stod(f[i]));
                } catch (const std::out_of_range &) {
                    return fail("Error converting data line:\n" + line + "\nValue " + f[i] +
                                " is out of range for double precision.");
                } catch (...) {
                    return fail("Error parsing data line:\n" + line + "\nCan't parse field " + f[i] + " as a double");
                }
            }
        }
    }
    const int n = ncols > 0 ? (int)data[0].size() : 0;
    if (n < 3) return fail("Atmosphere1D - need at least 3 data rows");
    std::map<std::string, std::pair<const std::vector<double> *, std::string>> vec;
    std::map<std::string, std::pair<double, std::string>> sca;
    for (auto &c : cols) {
        if (c.col == 0) sca[c.key] = {c.value, c.units};
        else if (c.col > 1) vec[c.key] = {&data[c.col - 1], c.units};
    }
    auto get_vec = [&](const char *key, const char *units, std::vector<double> &out) -> int {
        auto it = vec.find(key);
        if (it == vec.end()) return 1;
        out.resize(n);
        for (int i = 0; i < n; ++i)
            if (!convert_value((*it->second.first)[i], it->second.second, units, out[i]))
                return fail(std::string("Undefined conversion for ") + key);
        return 0;
    };
    std::vector<double> z(n), u, v, T, P, rho, c0;
    for (int i = 0; i < n; ++i)
        if (!convert_value(data[0][i], zunits, "m", z[i]))
            return fail("altitude units must be a distance");
    int rc;
    if ((rc = get_vec("U", "m/s", u)) == 1) return fail("U wind vector not found in atmosphere file");
    if (rc) return rc;
    if ((rc = get_vec("V", "m/s", v)) == 1) return fail("V wind vector not found in atmosphere file");
    if (rc) return rc;
    const bool haveP = get_vec("P", "Pa", P) == 0;
    const bool haveR = get_vec("RHO", "kg/m3", rho) == 0;
    const bool haveT = get_vec("T", "K", T) == 0;
    if (get_vec("C0", "m/s", c0) == 0) {
        if (!haveT) {
            T.resize(n);
            for (int i = 0; i < n; ++i) T[i] = c0[i] * c0[i] / 1.4 / 287.0;
        }
    } else if (haveP && haveR) {
        c0.resize(n);
        for (int i = 0; i < n; ++i) c0[i] = std::sqrt(1.4 * P[i] / rho[i]);
    } else if (haveT) {
        c0.resize(n);
        for (int i = 0; i < n; ++i) c0[i] = std::sqrt(1.4 * 287.0 * T[i]);
    } else {
        return fail("Can't calculate static sound speed, either temperature T or "
                                            "pressure P and density RHO missing");
    }
    if (T.empty()) return fail("T temperature vector not found in atmosphere file");
    if (!haveP) return fail("P pressure vector not found in atmosphere file");
    if (!haveR) return fail("RHO density vector not found in atmosphere file");
    for (int i = 1; i < n; ++i)
        if (!(z[i] > z[i - 1])) return fail("Atmosphere1D - altitudes must increase");
    for (int i = 0; i < n; ++i)
        if (!std::isfinite(z[i]) || !std::isfinite(u[i]) || !std::isfinite(v[i]) || !std::isfinite(T[i]) ||
            !std::isfinite(P[i]) || !std::isfinite(rho[i]) || !std::isfinite(c0[i]))
            return fail("Atmosphere1D - non-finite value in the profile");
    pp.n = n;
    pp.z = std::move(z);
    pp.u = std::move(u);
    pp.v = std::move(v);
    pp.T = std::move(T);
    pp.P = std::move(P);
    pp.rho = std::move(rho);
    pp.c0 = std::move(c0);
    pp.has_z0 = false;
    pp.z0 = 0.0;
    auto it = sca.find("Z0");
    double z0;
    if (it != sca.end() && convert_value(it->second.first, it->second.second, "m", z0)) {
        pp.has_z0 = true;
        pp.z0 = z0;
    }
    return NCPA_OK;
}

}  // namespace ncpa

extern "C" int ncpa_profile_check(const char *path, const char *header_path, ncpa_profile_info_t *info, char *msg,
                                  int msg_len) {
    if (msg && msg_len > 0) msg[0] = 0;
    if (!path) return NCPA_ERR_ARG;
    ncpa::ParsedProfile pp;
    std::string err;
    const int rc = ncpa::parse_profile_file(path, header_path ? header_path : "", pp, err);
    if (rc != NCPA_OK) {
        if (msg && msg_len > 0) {
            std::strncpy(msg, err.c_str(), (size_t)msg_len - 1);
            msg[msg_len - 1] = 0;
        }
        return rc;
    }
    if (info) {
        info->n_nodes = pp.n;
        info->z_bottom = pp.z.front();
        info->z_top = pp.z.back();
        info->has_z0 = pp.has_z0 ? 1 : 0;
        info->z0 = pp.z0;
    }
    return NCPA_OK;
}
