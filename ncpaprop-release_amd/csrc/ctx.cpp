// Context lifecycle, profile upload, allocator and timing.
#include "ctx.h"
#include "common.h"
#include "internal.h"

#include <cmath>
#include <cstring>
#include <algorithm>
#include <new>

namespace ncpa {

DevicePool::~DevicePool() { release_all(); }

void *DevicePool::get(size_t bytes) {
    if (bytes == 0) bytes = 256;
    bytes = (bytes + 255) & ~size_t(255);
    auto it = free_.lower_bound(bytes);
    if (it != free_.end() && it->first <= 2 * bytes) {
        void *p = it->second;
        live_[p] = {it->first, bytes};
        cached_ -= it->first;
        free_.erase(it);
        live_req_ += bytes;
        note();
        return p;
    }
    void *p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
        // drop cached blocks and retry once
        for (auto &kv : free_) (void)hipFree(kv.second);
        free_.clear();
        cached_ = 0;
        (void)hipGetLastError();
        if (hipMalloc(&p, bytes) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
    }
    live_[p] = {bytes, bytes};
    live_req_ += bytes;
    note();
    return p;
}

void DevicePool::put(void *p) {
    if (!p) return;
    auto it = live_.find(p);
    if (it == live_.end()) return;
    free_.emplace(it->second.first, p);
    cached_ += it->second.first;
    live_req_ -= it->second.second;
    live_.erase(it);
}

void DevicePool::add_external(size_t bytes) {
    ext_ += bytes;
    note();
}

void DevicePool::remove_external(size_t bytes) { ext_ -= std::min(ext_, bytes); }

void DevicePool::release_all() {
    for (auto &kv : free_) (void)hipFree(kv.second);
    free_.clear();
    for (auto &kv : live_) (void)hipFree(kv.first);
    live_.clear();
    live_req_ = cached_ = 0;
}

int set_error(ncpa_ctx *ctx, int code, const std::string &msg) {
    if (ctx) ctx->err = msg;
    return code;
}

int hip_error(ncpa_ctx *ctx, hipError_t e, const char *what) {
    std::string m = std::string(what) + ": " + hipGetErrorString(e);
    return set_error(ctx, NCPA_ERR_HIP, m);
}

TimedLaunch::TimedLaunch(ncpa_ctx *c, const char *n) : ctx(c), name(n) {
    if (!ctx->timing) return;
    if (hipEventCreate(&start) != hipSuccess || hipEventCreate(&stop) != hipSuccess) {
        start = stop = nullptr;
        return;
    }
    (void)hipEventRecord(start, ctx->stream);
}

TimedLaunch::~TimedLaunch() {
    if (!start) return;
    (void)hipEventRecord(stop, ctx->stream);
    ctx->events[name].emplace_back(start, stop);
}

}  // namespace ncpa

using namespace ncpa;

extern "C" {

int ncpa_ctx_create(int device, void *hip_stream, ncpa_ctx_t **out) {
    if (!out) return NCPA_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device || device < 0) {
        (void)hipGetLastError();
        return NCPA_ERR_HIP;
    }
    ncpa_ctx *ctx = new (std::nothrow) ncpa_ctx();
    if (!ctx) return NCPA_ERR_ALLOC;
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess) {
        delete ctx;
        return NCPA_ERR_HIP;
    }
    // the device address of ncpa_bb_pulse_noalloc's fixed array, once per device
    double *fixed = nullptr;
    if (ncpa::pulse_scratch_setup(ctx, &fixed) != NCPA_OK) {
        delete ctx;
        return NCPA_ERR_HIP;
    }
    if (hip_stream) {
        ctx->stream = static_cast<hipStream_t>(hip_stream);
    } else {
        if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
            delete ctx;
            return NCPA_ERR_HIP;
        }
        ctx->own_stream = true;
    }
    *out = ctx;
    return NCPA_OK;
}

void ncpa_ctx_destroy(ncpa_ctx_t *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    ncpa_ctx_reset_timing(ctx);
    for (auto &p : ctx->profiles) (void)hipFree(p.d);
    ctx->profiles.clear();
    ctx->pool.release_all();
    if (ctx->own_stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char *ncpa_ctx_error(const ncpa_ctx_t *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int ncpa_ctx_synchronize(ncpa_ctx_t *ctx) {
    if (!ctx) return NCPA_ERR_ARG;
    NCPA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return NCPA_OK;
}

void *ncpa_ctx_stream(ncpa_ctx_t *ctx) { return ctx ? static_cast<void *>(ctx->stream) : nullptr; }

int ncpa_ctx_set_timing(ncpa_ctx_t *ctx, int enable) {
    if (!ctx) return NCPA_ERR_ARG;
    ctx->timing = enable != 0;
    return NCPA_OK;
}

int ncpa_ctx_set_modes_form(ncpa_ctx_t *ctx, int form) {
    if (!ctx || form < 0 || form > 1) return NCPA_ERR_ARG;
    ctx->modes_smem = form == 1;
    return NCPA_OK;
}

int ncpa_ctx_kernel_time(ncpa_ctx_t *ctx, const char *kernel, double *total_ms, int *launches) {
    if (!ctx || !kernel) return NCPA_ERR_ARG;
    NCPA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    double tot = 0.0;
    int cnt = 0;
    auto it = ctx->events.find(kernel);
    if (it != ctx->events.end()) {
        for (auto &ev : it->second) {
            float ms = 0.f;
            NCPA_HIP(ctx, hipEventElapsedTime(&ms, ev.first, ev.second));
            tot += ms;
            ++cnt;
        }
    }
    if (total_ms) *total_ms = tot;
    if (launches) *launches = cnt;
    return NCPA_OK;
}

int ncpa_ctx_reset_timing(ncpa_ctx_t *ctx) {
    if (!ctx) return NCPA_ERR_ARG;
    (void)hipStreamSynchronize(ctx->stream);
    for (auto &kv : ctx->events)
        for (auto &ev : kv.second) {
            (void)hipEventDestroy(ev.first);
            (void)hipEventDestroy(ev.second);
        }
    ctx->events.clear();
    return NCPA_OK;
}

struct ncpa_event {
    hipEvent_t e;
};

int ncpa_event_record(ncpa_ctx_t *ctx, ncpa_event_t **ev) {
    if (!ctx || !ev) return NCPA_ERR_ARG;
    *ev = nullptr;
    NCPA_HIP(ctx, hipSetDevice(ctx->device));
    ncpa_event *x = new (std::nothrow) ncpa_event{nullptr};
    if (!x) return set_error(ctx, NCPA_ERR_ALLOC, "ncpa_event_record: allocation failed");
    hipError_t e = hipEventCreateWithFlags(&x->e, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventRecord(x->e, ctx->stream);
    if (e != hipSuccess) {
        if (x->e) (void)hipEventDestroy(x->e);
        delete x;
        return hip_error(ctx, e, "ncpa_event_record");
    }
    *ev = x;
    return NCPA_OK;
}

int ncpa_event_wait(ncpa_ctx_t *ctx, ncpa_event_t *ev) {
    if (!ctx || !ev) return set_error(ctx, NCPA_ERR_ARG, "ncpa_event_wait: bad arguments");
    NCPA_HIP(ctx, hipStreamWaitEvent(ctx->stream, ev->e, 0));
    return NCPA_OK;
}

int ncpa_event_synchronize(ncpa_event_t *ev) {
    if (!ev) return NCPA_ERR_ARG;
    return hipEventSynchronize(ev->e) == hipSuccess ? NCPA_OK : NCPA_ERR_HIP;
}

void ncpa_event_destroy(ncpa_event_t *ev) {
    if (!ev) return;
    (void)hipEventDestroy(ev->e);
    delete ev;
}

int ncpa_dev_alloc(ncpa_ctx_t *ctx, size_t bytes, void **dev_ptr) {
    if (!ctx || !dev_ptr) return NCPA_ERR_ARG;
    *dev_ptr = nullptr;
    NCPA_HIP(ctx, hipSetDevice(ctx->device));
    NCPA_HIP(ctx, hipMalloc(dev_ptr, bytes ? bytes : 1));
    ctx->dev_allocs[*dev_ptr] = bytes;
    ctx->pool.add_external(bytes);
    return NCPA_OK;
}

int ncpa_ctx_memory(ncpa_ctx_t *ctx, int reset_peak, size_t *live_bytes, size_t *peak_bytes,
                    size_t *cached_bytes) {
    if (!ctx) return NCPA_ERR_ARG;
    if (live_bytes) *live_bytes = ctx->pool.live();
    if (peak_bytes) *peak_bytes = ctx->pool.peak();
    if (cached_bytes) *cached_bytes = ctx->pool.cached();
    if (reset_peak) ctx->pool.reset_peak();
    return NCPA_OK;
}

int ncpa_dev_free(ncpa_ctx_t *ctx, void *dev_ptr) {
    if (!ctx) return NCPA_ERR_ARG;
    if (!dev_ptr) return NCPA_OK;
    NCPA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    NCPA_HIP(ctx, hipFree(dev_ptr));
    auto it = ctx->dev_allocs.find(dev_ptr);
    if (it != ctx->dev_allocs.end()) {
        ctx->pool.remove_external(it->second);
        ctx->dev_allocs.erase(it);
    }
    return NCPA_OK;
}

int ncpa_dev_upload(ncpa_ctx_t *ctx, void *dst, const void *src, size_t bytes) {
    if (!ctx || (!dst && bytes) || (!src && bytes)) return NCPA_ERR_ARG;
    NCPA_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    NCPA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return NCPA_OK;
}

int ncpa_dev_download(ncpa_ctx_t *ctx, void *dst, const void *src, size_t bytes) {
    if (!ctx || (!dst && bytes) || (!src && bytes)) return NCPA_ERR_ARG;
    NCPA_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    NCPA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return NCPA_OK;
}

int ncpa_profile_upload(ncpa_ctx_t *ctx, int n, const double *z, const double *c0,
                        const double *u, const double *v, const double *T, const double *P,
                        const double *rho, int *profile_id) {
    if (!ctx || n < 3 || !z || !c0 || !u || !v || !T || !P || !rho || !profile_id)
        return set_error(ctx, NCPA_ERR_ARG, "ncpa_profile_upload: bad arguments (need >= 3 nodes)");
    for (int i = 1; i < n; ++i)
        if (!(z[i] > z[i - 1]))
            return set_error(ctx, NCPA_ERR_ARG, "ncpa_profile_upload: altitudes must increase");
    std::vector<double> h(size_t(PF_COUNT) * n, 0.0);
    for (int i = 0; i < n; ++i) {
        h[PF_Z * n + i] = z[i];
        h[PF_C0 * n + i] = c0[i];
        // Atmosphere1D::calculate_wind_speed / calculate_wind_direction (Atmosphere1D.cpp:505-540)
        h[PF_WS * n + i] = std::sqrt(u[i] * u[i] + v[i] * v[i]);
        h[PF_WD * n + i] = (PI / 2.0 - std::atan2(v[i], u[i])) * 180.0 / PI;
        h[PF_T * n + i] = T[i];
        h[PF_P * n + i] = P[i];
        h[PF_RHO * n + i] = rho[i];
    }
    ProfileDev pd;
    pd.n = n;
    NCPA_HIP(ctx, hipSetDevice(ctx->device));
    NCPA_HIP(ctx, hipMalloc(&pd.d, h.size() * sizeof(double)));
    NCPA_HIP(ctx, hipMemcpy(pd.d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    pd.host = std::move(h);
    ctx->profiles.push_back(std::move(pd));
    *profile_id = int(ctx->profiles.size()) - 1;
    return NCPA_OK;
}

int ncpa_profile_set_attenuation(ncpa_ctx_t *ctx, int profile_id, const double *alpha) {
    if (!ctx || profile_id < 0 || profile_id >= int(ctx->profiles.size()) || !alpha)
        return set_error(ctx, NCPA_ERR_ARG, "ncpa_profile_set_attenuation: bad arguments");
    ProfileDev &pd = ctx->profiles[profile_id];
    NCPA_HIP(ctx, hipMemcpy(pd.d + size_t(PF_ALPHA) * pd.n, alpha, pd.n * sizeof(double),
                            hipMemcpyHostToDevice));
    std::copy(alpha, alpha + pd.n, pd.host.begin() + size_t(PF_ALPHA) * pd.n);
    pd.user_alpha = true;
    return NCPA_OK;
}

int ncpa_profile_clear(ncpa_ctx_t *ctx) {
    if (!ctx) return NCPA_ERR_ARG;
    (void)hipStreamSynchronize(ctx->stream);
    for (auto &p : ctx->profiles) (void)hipFree(p.d);
    ctx->profiles.clear();
    return NCPA_OK;
}

}  // extern "C"
