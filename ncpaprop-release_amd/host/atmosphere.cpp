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
#include "profile_parse.h"

using namespace ncpa;

extern "C" {

int ncpa_profile_read(ncpa_ctx_t *ctx, const char *path, const char *header_path, int *profile_id,
                      ncpa_profile_info_t *info) {
    if (!ctx || !path || !profile_id) return set_error(ctx, NCPA_ERR_ARG, "ncpa_profile_read: bad arguments");
    ParsedProfile pp;
    std::string err;
    if (parse_profile_file(path, header_path ? header_path : "", pp, err) != NCPA_OK)
        return set_error(ctx, NCPA_ERR_ARG, err);
    const int rc = ncpa_profile_upload(ctx, pp.n, pp.z.data(), pp.c0.data(), pp.u.data(), pp.v.data(), pp.T.data(),
                                       pp.P.data(), pp.rho.data(), profile_id);
    if (rc) return rc;
    ProfileDev &pd = ctx->profiles[*profile_id];
    pd.has_z0 = pp.has_z0;
    pd.z0 = pp.z0;
    if (info) {
        info->n_nodes = pp.n;
        info->z_bottom = pp.z.front();
        info->z_top = pp.z.back();
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
