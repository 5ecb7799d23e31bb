// Atmosphere profile reader and SI node preparation (host C++ behind the C ABI).
//
// Format and semantics follow Atmosphere1D::read_header_from_stream / read_values_from_stream
// (src/atmosphere/Atmosphere1D.cpp:85-285): '#%' lines describe columns as
// "col, KEY, units[, value]" (column 0 = scalar, column 1 = altitude), '#' lines are
// comments, data fields are separated by spaces, tabs or commas, keys must be unique.
// Units follow src/common/units.cpp:131-178,225-283.  The node quantities are left as
// ESSModeSolver::setParams leaves them (ESSModeSolver.cpp:159-220): z [m], U/V [m/s],
// T [K], P [Pa], RHO [kg/m3], c0 = C0 column, else sqrt(1.4 P/RHO), else sqrt(1.4*287*T).
#include <algorithm>
#include <cctype>
#include <cmath>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../csrc/common.h"
#include "../csrc/ctx.h"
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

using namespace ncpa;

extern "C" {

int ncpa_profile_read(ncpa_ctx_t *ctx, const char *path, const char *header_path, int *profile_id,
                      ncpa_profile_info_t *info) {
    if (!ctx || !path || !profile_id) return set_error(ctx, NCPA_ERR_ARG, "ncpa_profile_read: bad arguments");
    const std::string hfile = (header_path && header_path[0]) ? header_path : path;
    std::vector<std::string> headerlines;
    {
        std::ifstream in(hfile);
        if (!in.good()) return set_error(ctx, NCPA_ERR_ARG, "Atmosphere1D - Input stream not in good state: " + hfile);
        std::string line;
        while (std::getline(in, line)) {
            line = trim(line);
            if (line.size() > 1 && line[0] == '#' && line[1] == '%') headerlines.push_back(line.substr(2));
        }
    }
    if (headerlines.empty()) return set_error(ctx, NCPA_ERR_ARG, "Atmosphere1D - No descriptive fields found.");
    std::vector<Column> cols;
    for (const std::string &h : headerlines) {
        std::vector<std::string> f = split(trim(h), " ,");
        if (f.size() != 3 && f.size() != 4)
            return set_error(ctx, NCPA_ERR_ARG, "Atmosphere1D - Error parsing descriptive line: " + h +
                                                    " (must be column,key,units[,value])");
        Column c;
        try {
            c.col = std::stoi(f[0]);
        } catch (...) {
            return set_error(ctx, NCPA_ERR_ARG, "Atmosphere1D - First field not parseable as an integer: " + h);
        }
        auto it = unit_table().find(upper(trim(f[2])));
        if (it == unit_table().end() || it->second == "none")
            return set_error(ctx, NCPA_ERR_ARG, "Unrecognized units string: " + f[2]);
        c.units = it->second;
        c.key = trim(f[1]);
        c.value = 0.0;
        if (f.size() == 4) {
            try {
                c.value = std::stod(trim(f[3]));
            } catch (...) {
                return set_error(ctx, NCPA_ERR_ARG, "Atmosphere1D - bad scalar value: " + h);
            }
        }
        cols.push_back(c);
    }
    {
        std::vector<std::string> keys;
        for (auto &c : cols) keys.push_back(c.key);
        std::sort(keys.begin(), keys.end());
        if (std::adjacent_find(keys.begin(), keys.end()) != keys.end())
            return set_error(ctx, NCPA_ERR_ARG, "Atmosphere1D: key values are not unique");
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
        if (!in.good()) return set_error(ctx, NCPA_ERR_ARG, std::string("Atmosphere1D - cannot open ") + path);
        std::string line;
        while (std::getline(in, line)) {
            line = trim(line);
            if (line.empty() || line[0] == '#') continue;
            std::vector<std::string> f = split(line, " \t,");
            if ((int)f.size() != ncols) {
                std::ostringstream o;
                o << "Error parsing data line:\n" << line << "\n" << ncols << " columns expected, " << f.size()
                  << " columns found.";
                return set_error(ctx, NCPA_ERR_ARG, o.str());
            }
            for (int i = 0; i < ncols; ++i) {
                try {
                    data[i].push_back(std::stod(f[i]));
                } catch (...) {
                    return set_error(ctx, NCPA_ERR_ARG, "Error parsing data line:\n" + line +
                                                            "\nCan't parse field " + f[i] + " as a double");
                }
            }
        }
    }
    const int n = ncols > 0 ? (int)data[0].size() : 0;
    if (n < 3) return set_error(ctx, NCPA_ERR_ARG, "Atmosphere1D - need at least 3 data rows");
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
                return set_error(ctx, NCPA_ERR_ARG, std::string("Undefined conversion for ") + key);
        return 0;
    };
    std::vector<double> z(n), u, v, T, P, rho, c0;
    for (int i = 0; i < n; ++i)
        if (!convert_value(data[0][i], zunits, "m", z[i]))
            return set_error(ctx, NCPA_ERR_ARG, "altitude units must be a distance");
    int rc;
    if ((rc = get_vec("U", "m/s", u)) == 1) return set_error(ctx, NCPA_ERR_ARG, "U wind vector not found in atmosphere file");
    if (rc) return rc;
    if ((rc = get_vec("V", "m/s", v)) == 1) return set_error(ctx, NCPA_ERR_ARG, "V wind vector not found in atmosphere file");
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
        return set_error(ctx, NCPA_ERR_ARG, "Can't calculate static sound speed, either temperature T or "
                                            "pressure P and density RHO missing");
    }
    if (T.empty()) return set_error(ctx, NCPA_ERR_ARG, "T temperature vector not found in atmosphere file");
    if (!haveP) return set_error(ctx, NCPA_ERR_ARG, "P pressure vector not found in atmosphere file");
    if (!haveR) return set_error(ctx, NCPA_ERR_ARG, "RHO density vector not found in atmosphere file");
    rc = ncpa_profile_upload(ctx, n, z.data(), c0.data(), u.data(), v.data(), T.data(), P.data(), rho.data(),
                             profile_id);
    if (rc) return rc;
    ProfileDev &pd = ctx->profiles[*profile_id];
    {
        auto it = sca.find("Z0");
        double z0;
        if (it != sca.end() && convert_value(it->second.first, it->second.second, "m", z0)) {
            pd.has_z0 = true;
            pd.z0 = z0;
        }
    }
    if (info) {
        info->n_nodes = n;
        info->z_bottom = z[0];
        info->z_top = z[n - 1];
        info->has_z0 = pd.has_z0 ? 1 : 0;
        info->z0 = pd.z0;
    }
    return NCPA_OK;
}

int ncpa_profile_nodes(ncpa_ctx_t *ctx, int profile_id, double *z, double *c0, double *ws, double *wd_deg,
                       double *T, double *P, double *rho) {
    if (!ctx || profile_id < 0 || profile_id >= (int)ctx->profiles.size())
        return set_error(ctx, NCPA_ERR_ARG, "ncpa_profile_nodes: unknown profile");
    const ProfileDev &pd = ctx->profiles[profile_id];
    const int n = pd.n;
    double *dst[] = {z, c0, ws, wd_deg, T, P, rho};
    for (int f = 0; f < 7; ++f)
        if (dst[f]) std::copy(pd.host.begin() + size_t(f) * n, pd.host.begin() + size_t(f + 1) * n, dst[f]);
    return NCPA_OK;
}

int ncpa_grid_make(ncpa_ctx_t *ctx, int profile_id, int Nz, double maxheight_km, double zground_km,
                   ncpa_grid_t *grid) {
    if (!ctx || !grid || profile_id < 0 || profile_id >= (int)ctx->profiles.size() || Nz < 10)
        return set_error(ctx, NCPA_ERR_ARG, "ncpa_grid_make: bad arguments");
    const ProfileDev &pd = ctx->profiles[profile_id];
    const double *zn = pd.host.data();
    double z_min = zground_km * 1000.0;
    double maxheight = maxheight_km * 1000.0;
    // ESSModeSolver.cpp:144-148, 212-220
    if (maxheight >= zn[pd.n - 1]) maxheight = zn[pd.n - 1] - 1e-6;
    if (pd.has_z0 && z_min < pd.z0) z_min = pd.z0;
    grid->Nz = Nz;
    grid->z_min = z_min;
    grid->dz = (maxheight - z_min) / (Nz - 1);
    grid->maxheight = maxheight;
    return NCPA_OK;
}

int ncpa_profile_admittance(ncpa_ctx_t *ctx, int profile_id, double z_min, double *adm) {
    if (!ctx || !adm || profile_id < 0 || profile_id >= (int)ctx->profiles.size())
        return set_error(ctx, NCPA_ERR_ARG, "ncpa_profile_admittance: bad arguments");
    const ProfileDev &pd = ctx->profiles[profile_id];
    const int n = pd.n;
    const double *zn = pd.host.data();
    const double *rn = pd.host.data() + size_t(PF_RHO) * n;
    auto yf = [rn](int k) { return rn[k]; };
    // -(drho/dz)/rho/2 at the ground (ESSModeSolver.cpp:347-349)
    *adm = -steffen_deriv(zn, n, yf, z_min) / steffen_eval(zn, n, yf, z_min) / 2.0;
    return NCPA_OK;
}

int ncpa_profile_sample(ncpa_ctx_t *ctx, int profile_id, const ncpa_grid_t *grid, double *rho, double *P,
                        double *T, double *u_dummy) {
    (void)u_dummy;
    if (!ctx || !grid || profile_id < 0 || profile_id >= (int)ctx->profiles.size())
        return set_error(ctx, NCPA_ERR_ARG, "ncpa_profile_sample: bad arguments");
    const ProfileDev &pd = ctx->profiles[profile_id];
    const int n = pd.n;
    const double *zn = pd.host.data();
    for (int f = 0; f < 3; ++f) {
        double *dst = f == 0 ? rho : (f == 1 ? P : T);
        if (!dst) continue;
        const double *yn = pd.host.data() + size_t(f == 0 ? PF_RHO : (f == 1 ? PF_P : PF_T)) * n;
        auto yf = [yn](int k) { return yn[k]; };
        for (int i = 0; i < grid->Nz; ++i) dst[i] = steffen_eval(zn, n, yf, grid->z_min + i * grid->dz);
    }
    return NCPA_OK;
}

}  // extern "C"
