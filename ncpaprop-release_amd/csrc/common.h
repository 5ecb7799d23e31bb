// Shared device/host helpers for the ncpamodes HIP library (gfx950 / CDNA4).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cmath>

namespace ncpa {

// pi as src/common/units.h:63-65 and src/common/util.h:13-15 define it
constexpr double PI = 3.141592653589793;

// gsl_interp_bsearch semantics: largest k with z[k] <= x, within [0, n-2]
__host__ __device__ inline int bsearch_interval(const double *z, int n, double x) {
    int lo = 0, hi = n - 1;
    while (hi > lo + 1) {
        int mid = (hi + lo) >> 1;
        if (z[mid] > x) hi = mid; else lo = mid;
    }
    return lo;
}

__host__ __device__ inline double sgn1(double v) { return v < 0.0 ? -1.0 : 1.0; }

// Steffen (1990) monotone cubic, slopes from the node values (A&A 239, 443: eq. 11 for
// interior nodes, one-sided end slopes of sec. 2.2, Hermite coefficients eqs. 2-5).
// YF(k) returns the node value y_k.  Outside [z0, z_{n-1}] the end value is returned,
// as AtmosphericProperty1D::get does (src/atmosphere/AtmosphericProperty1D.cpp:116-126).
template <class YF>
__host__ __device__ inline double steffen_eval_at(const double *z, int n, YF yf, double x, int k);
template <class ZF, class YF>
__host__ __device__ inline double steffen_eval_zf(ZF z, int n, YF yf, double x, int k);

template <class YF>
__host__ __device__ inline double steffen_eval(const double *z, int n, YF yf, double x) {
    if (x > z[n - 1]) return yf(n - 1);
    if (x < z[0]) return yf(0);
    return steffen_eval_at(z, n, yf, x, bsearch_interval(z, n, x));
}

// node abscissae read through an accessor (global arrays, or a window staged in LDS)
struct PtrNodes {
    const double *z;
    __host__ __device__ double operator[](int k) const { return z[k]; }
};

// Hermite coefficients of Steffen's cubic on interval k: y(x) = y + dx (yp + dx (b + dx a)),
// dx = x - z_k.  Split from the evaluation so that a kernel can form them once per interval
// and evaluate many samples with the same arithmetic.
struct SteffenCoef {
    double y, yp, b, a;
};

template <class ZF, class YF>
__host__ __device__ inline SteffenCoef steffen_coef(ZF z, int n, YF yf, int k) {
    const double hk = z[k + 1] - z[k];
    const double yk = yf(k), yk1 = yf(k + 1);
    const double sk = (yk1 - yk) / hk;
    double ypk, ypk1;
    if (k == 0) {
        ypk = sk;
    } else {
        const double hm = z[k] - z[k - 1];
        const double sm = (yk - yf(k - 1)) / hm;
        const double p = (sm * hk + sk * hm) / (hm + hk);
        ypk = (sgn1(sm) + sgn1(sk)) * fmin(fabs(sm), fmin(fabs(sk), 0.5 * fabs(p)));
    }
    if (k + 1 == n - 1) {
        ypk1 = sk;
    } else {
        const double hp = z[k + 2] - z[k + 1];
        const double sp = (yf(k + 2) - yk1) / hp;
        const double p = (sk * hp + sp * hk) / (hk + hp);
        ypk1 = (sgn1(sk) + sgn1(sp)) * fmin(fabs(sk), fmin(fabs(sp), 0.5 * fabs(p)));
    }
    SteffenCoef c;
    c.y = yk;
    c.yp = ypk;
    c.a = (ypk + ypk1 - 2.0 * sk) / hk / hk;
    c.b = (3.0 * sk - 2.0 * ypk - ypk1) / hk;
    return c;
}

__host__ __device__ inline double steffen_poly(const SteffenCoef &c, double dx) {
    return c.y + dx * (c.yp + dx * (c.b + dx * c.a));
}

// the same, on interval k = bsearch_interval(z, n, x) found by the caller (x inside the nodes)
template <class ZF, class YF>
__host__ __device__ inline double steffen_eval_zf(ZF z, int n, YF yf, double x, int k) {
    return steffen_poly(steffen_coef(z, n, yf, k), x - z[k]);
}

template <class YF>
__host__ __device__ inline double steffen_eval_at(const double *z, int n, YF yf, double x, int k) {
    return steffen_eval_zf(PtrNodes{z}, n, yf, x, k);
}

// Steffen first derivative (AtmosphericProperty1D.cpp:128-136: zero outside the range)
template <class YF>
__host__ __device__ inline double steffen_deriv(const double *z, int n, YF yf, double x) {
    if (x > z[n - 1] || x < z[0]) return 0.0;
    const int k = bsearch_interval(z, n, x);
    const double hk = z[k + 1] - z[k];
    const double yk = yf(k), yk1 = yf(k + 1);
    const double sk = (yk1 - yk) / hk;
    double ypk, ypk1;
    if (k == 0) {
        ypk = sk;
    } else {
        const double hm = z[k] - z[k - 1];
        const double sm = (yk - yf(k - 1)) / hm;
        const double p = (sm * hk + sk * hm) / (hm + hk);
        ypk = (sgn1(sm) + sgn1(sk)) * fmin(fabs(sm), fmin(fabs(sk), 0.5 * fabs(p)));
    }
    if (k + 1 == n - 1) {
        ypk1 = sk;
    } else {
        const double hp = z[k + 2] - z[k + 1];
        const double sp = (yf(k + 2) - yk1) / hp;
        const double p = (sk * hp + sp * hk) / (hk + hp);
        ypk1 = (sgn1(sk) + sgn1(sp)) * fmin(fabs(sk), fmin(fabs(sp), 0.5 * fabs(p)));
    }
    const double a = (ypk + ypk1 - 2.0 * sk) / hk / hk;
    const double b = (3.0 * sk - 2.0 * ypk - ypk1) / hk;
    const double dx = x - z[k];
    return ypk + dx * (2.0 * b + dx * 3.0 * a);
}

}  // namespace ncpa
