// Context, profile store, caching allocator and per-kernel timing (host side).
#pragma once
#include <hip/hip_runtime.h>
#include <algorithm>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "ncpamodes.h"

namespace ncpa {

// Packed per-profile node arrays on the device:
// z, c0, ws, wd_deg, T, P, rho, alpha_user (n each).
enum ProfileField { PF_Z = 0, PF_C0, PF_WS, PF_WD, PF_T, PF_P, PF_RHO, PF_ALPHA, PF_COUNT };

struct ProfileDev {
    int n = 0;
    double *d = nullptr;         // PF_COUNT * n doubles on the device
    std::vector<double> host;    // same layout on the host
    bool user_alpha = false;
    bool has_z0 = false;
    double z0 = 0.0;             // profile ground height Z0 [m]
};

// Stream-ordered caching allocator: every use of a cached block happens on the context
// stream (a block another context used is put back only behind a wait for that use, see
// ncpa_field_stage_free), so a block returned to the pool can be handed out again at once.
// Accounting: live() = bytes handed out (requests rounded to 256 B) plus the device buffers
// the context allocated outside the pool (ncpa_dev_alloc), peak() its largest value since
// reset_peak(), cached() = bytes held in free blocks.
class DevicePool {
public:
    ~DevicePool();
    void *get(size_t bytes);
    void put(void *p);
    void release_all();
    void add_external(size_t bytes);
    void remove_external(size_t bytes);
    size_t live() const { return live_req_ + ext_; }
    size_t peak() const { return peak_; }
    size_t cached() const { return cached_; }
    void reset_peak() { peak_ = live(); }
private:
    void note() { peak_ = std::max(peak_, live()); }
    std::multimap<size_t, void *> free_;
    std::map<void *, std::pair<size_t, size_t>> live_;   // block bytes, requested bytes
    size_t live_req_ = 0, ext_ = 0, peak_ = 0, cached_ = 0;
};

}  // namespace ncpa

struct ncpa_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::string err;
    bool timing = false;
    std::map<std::string, std::vector<std::pair<hipEvent_t, hipEvent_t>>> events;
    std::vector<ncpa::ProfileDev> profiles;
    ncpa::DevicePool pool;
    std::map<void *, size_t> dev_allocs;   // ncpa_dev_alloc buffers and their sizes
    int ms_passes = 2;     // multisection passes of the eigenvalue kernel (autotuned by plans)
    bool modes_smem = true;    // eager eigenvalue sweeps read the diagonal by scalar loads
    int field_variant[8] = {0, 0, 0, 0, 0, 0, 0, 0};   // contraction variant per precision (ncpa_field_tune)
};

namespace ncpa {

int set_error(ncpa_ctx *ctx, int code, const std::string &msg);
int hip_error(ncpa_ctx *ctx, hipError_t e, const char *what);

// RAII timing scope around one kernel launch on ctx->stream
struct TimedLaunch {
    ncpa_ctx *ctx;
    const char *name;
    hipEvent_t start = nullptr, stop = nullptr;
    TimedLaunch(ncpa_ctx *c, const char *n);
    ~TimedLaunch();
};

template <class T>
T *pool_alloc(ncpa_ctx *ctx, size_t count) {
    return static_cast<T *>(ctx->pool.get(count * sizeof(T)));
}

}  // namespace ncpa

#define NCPA_HIP(ctx, call)                                          \
    do {                                                             \
        hipError_t e_ = (call);                                      \
        if (e_ != hipSuccess) return ncpa::hip_error((ctx), e_, #call); \
    } while (0)
