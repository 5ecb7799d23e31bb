/* Oracle (test infrastructure only): the ESSModeSolver path of ncpaprop restated in plain C,
 * FP64, for the timed CPU baseline and as a second restatement the Python oracle checks.
 * Only tests/ and bench.py's cpu_baseline leg run it; nothing in the product links it.
 *
 * It follows the reference's loop structure (one solve per (profile, frequency, azimuth)):
 *   profile reader + units      Atmosphere1D.cpp:85-285, units.cpp:131-178
 *   Steffen splines             AtmosphericProperty1D.cpp:72-136 (Steffen 1990, A&A 239:443)
 *   setParams grid              ESSModeSolver.cpp:30-240
 *   Sutherland-Bass alpha       AtmosphericModel.cpp:86-219
 *   getModalTrace (+ WKB)       ESSModeSolver.cpp:518-645, filter :366-369
 *   eigenpairs in [kmin^2, kmax^2] of the tridiagonal of EigenEngine.cpp:57-79: Sturm-count
 *                               bisection (LAPACK dstebz's method) and inverse iteration
 *                               (dstein's), the method the manual describes (modess.tex:85)
 *   doSelect / doPerturb        ESSModeSolver.cpp:647-665, ModeSolver.cpp:263-284
 *   getTLoss1D / getTLoss2D     ModeSolver.cpp:292-381, 492-565: complex exp and sqrt per term
 *   compute_modal_sum           ModalBroadbandPropagator.cpp:245-277
 *
 *   ess_port solve  --profile P --freq F --azimuth A [--Nz N] [--maxheight_km H] [--zs_m Z]
 *                   [--zr_m Z] [--wkb] [--lamb] [--filter CMIN CMAX] [--nr N --dr M]
 *                   [--rows R --stride S --field_ranges K]
 *       one solve; prints a JSON object with the modes, the 1-D TL and stage times
 *   ess_port bench  --threads T --freq F ... --job PROFILE:AZ:FREQ ...
 *       every job on T threads (one job per thread at a time); prints per-job stage times
 *       (no result arrays)
 */
#define _GNU_SOURCE
#include <complex.h>
#include <ctype.h>
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define PI 3.141592653589793   /* units.h:63-65, util.h:13-15 */
#define GAMMA_FOR_C 1.4        /* AtmosphericModel.cpp:5-7 */
#define R_FOR_C 287.0          /* AtmosphericModel.cpp:9-11 */
#define MAX_MODES 4000         /* ESSModeSolver.cpp:17 */

static void die(const char *msg) {
    fprintf(stderr, "ess_port: %s\n", msg);
    exit(2);
}

static double now_s(void) {
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return t.tv_sec + 1e-9 * t.tv_nsec;
}

/* ------------------------------------------------------------------------------------ */
/* units (units.cpp:131-178): every quantity ends in SI                                  */
/* ------------------------------------------------------------------------------------ */
typedef enum { U_BAD, U_K, U_C, U_F, U_M, U_KM, U_MPS, U_KMPS, U_PA, U_MBAR, U_HPA, U_ATM, U_KGM3, U_GCM3 } unit_t;

static unit_t unit_from(const char *s) {
    char u[64];
    size_t n = 0;
    for (; *s && n + 1 < sizeof u; ++s) u[n++] = (char)toupper((unsigned char)*s);
    u[n] = 0;
    static const struct { const char *name; unit_t u; } tab[] = {
        {"K", U_K}, {"DEGK", U_K}, {"DEG K", U_K}, {"DEGREES K", U_K}, {"KELVIN", U_K},
        {"C", U_C}, {"DEGC", U_C}, {"DEG C", U_C}, {"DEGREES C", U_C}, {"CELSIUS", U_C}, {"CENTIGRADE", U_C},
        {"F", U_F}, {"DEGF", U_F}, {"DEG F", U_F}, {"DEGREES F", U_F}, {"FAHRENHEIT", U_F},
        {"M", U_M}, {"METERS", U_M}, {"KM", U_KM}, {"KILOMETERS", U_KM},
        {"M/S", U_MPS}, {"MPS", U_MPS}, {"MPERS", U_MPS}, {"M PER S", U_MPS}, {"METERS PER SECOND", U_MPS},
        {"KM/S", U_KMPS}, {"KMPS", U_KMPS}, {"KMPERS", U_KMPS}, {"KM PER S", U_KMPS},
        {"KILOMETERS PER SECOND", U_KMPS},
        {"PA", U_PA}, {"PASCAL", U_PA}, {"PASCALS", U_PA}, {"MBAR", U_MBAR}, {"MILLIBAR", U_MBAR},
        {"MILLIBARS", U_MBAR}, {"HECTOPASCALS", U_HPA}, {"HPA", U_HPA}, {"ATMOSPHERES", U_ATM}, {"ATM", U_ATM},
        {"KG/M3", U_KGM3}, {"KGPM3", U_KGM3}, {"KILOGRAMS PER CUBIC METER", U_KGM3},
        {"G/CM3", U_GCM3}, {"GPCM3", U_GCM3}, {"GRAMS PER CUBIC CENTIMETER", U_GCM3},
    };
    for (size_t i = 0; i < sizeof tab / sizeof tab[0]; ++i)
        if (!strcmp(tab[i].name, u)) return tab[i].u;
    return U_BAD;
}

/* to SI (K, m, m/s, Pa, kg/m3) */
static double to_si(double x, unit_t u) {
    switch (u) {
    case U_C: return x + 273.15;
    case U_F: return (x - 32.0) * 5.0 / 9.0 + 273.15;
    case U_KM: return x * 1000.0;
    case U_KMPS: return x * 1000.0;
    case U_MBAR: return x * 100.0;
    case U_HPA: return x * 100;
    case U_ATM: return x * 101325.0;
    case U_GCM3: return x * 1000.0;
    default: return x;
    }
}

/* ------------------------------------------------------------------------------------ */
/* profile (Atmosphere1D.cpp:85-285)                                                     */
/* ------------------------------------------------------------------------------------ */
typedef struct {
    int n;
    double *z, *u, *v, *T, *P, *rho, *c0, *ws, *wd;   /* SI at the nodes */
    int has_z0;
    double z0;
} Profile;

static int split(char *s, const char *delims, char **out, int max) {
    int n = 0;
    char *save = NULL;
    for (char *t = strtok_r(s, delims, &save); t && n < max; t = strtok_r(NULL, delims, &save)) out[n++] = t;
    return n;
}

static char *trim(char *s) {
    while (*s == ' ' || *s == '\t' || *s == '\r' || *s == '\n') ++s;
    size_t n = strlen(s);
    while (n && (s[n - 1] == ' ' || s[n - 1] == '\t' || s[n - 1] == '\r' || s[n - 1] == '\n')) s[--n] = 0;
    return s;
}

static void read_profile(const char *path, Profile *pr) {
    FILE *f = fopen(path, "r");
    if (!f) die("cannot open the profile");
    char line[4096];
    int cols[64], nh = 0;
    char keys[64][32];
    unit_t units[64];
    double vals[64];
    long data_cap = 4096, nrow = 0;
    int ncols = 0;
    double *data = NULL;
    while (fgets(line, sizeof line, f)) {
        char *t = trim(line);
        if (t[0] == '#' && t[1] == '%') {
            char *fld[8];
            int nf = split(t + 2, " ,", fld, 8);
            if (nf != 3 && nf != 4) die("Atmosphere1D - Error parsing descriptive line");
            if (nh >= 64) die("too many header lines");
            cols[nh] = atoi(fld[0]);
            snprintf(keys[nh], sizeof keys[nh], "%s", fld[1]);
            units[nh] = unit_from(fld[2]);
            if (units[nh] == U_BAD) die("Unrecognized units string");
            vals[nh] = nf == 4 ? strtod(fld[3], NULL) : 0.0;
            for (int k = 0; k < nh; ++k)
                if (!strcmp(keys[k], keys[nh])) die("Atmosphere1D: key values are not unique");
            if (cols[nh] > ncols) ncols = cols[nh];
            ++nh;
        } else if (t[0] && t[0] != '#') {
            if (!data) data = malloc(sizeof(double) * data_cap * ncols);
            if ((nrow + 1) * ncols > data_cap * ncols) {
                data_cap *= 2;
                data = realloc(data, sizeof(double) * data_cap * ncols);
            }
            char *fld[64];
            int nf = split(t, " \t,", fld, 64);
            if (nf != ncols) die("Error parsing data line: wrong number of columns");
            for (int c = 0; c < ncols; ++c) data[nrow * ncols + c] = strtod(fld[c], NULL);
            ++nrow;
        }
    }
    fclose(f);
    if (!nh) die("Atmosphere1D - No descriptive fields found.");
    pr->n = (int)nrow;
    int n = pr->n;
    double **dst[] = {&pr->z, &pr->u, &pr->v, &pr->T, &pr->P, &pr->rho, &pr->c0, &pr->ws, &pr->wd};
    for (size_t i = 0; i < sizeof dst / sizeof dst[0]; ++i) *dst[i] = calloc(n, sizeof(double));
    int have_c0 = 0, have_T = 0, have_P = 0, have_rho = 0, have_u = 0, have_v = 0;
    pr->has_z0 = 0;
    for (int h = 0; h < nh; ++h) {
        double *out = NULL;
        if (cols[h] == 0) {
            if (!strcmp(keys[h], "Z0")) { pr->has_z0 = 1; pr->z0 = to_si(vals[h], units[h]); }
            continue;
        }
        if (cols[h] == 1) out = pr->z;
        else if (!strcmp(keys[h], "U")) { out = pr->u; have_u = 1; }
        else if (!strcmp(keys[h], "V")) { out = pr->v; have_v = 1; }
        else if (!strcmp(keys[h], "T")) { out = pr->T; have_T = 1; }
        else if (!strcmp(keys[h], "P")) { out = pr->P; have_P = 1; }
        else if (!strcmp(keys[h], "RHO")) { out = pr->rho; have_rho = 1; }
        else if (!strcmp(keys[h], "C0")) { out = pr->c0; have_c0 = 1; }
        if (!out) continue;
        for (int i = 0; i < n; ++i) out[i] = to_si(data[(long)i * ncols + cols[h] - 1], units[h]);
    }
    free(data);
    if (!have_u || !have_v) die("U or V wind vector not found in atmosphere file");
    if (!have_P || !have_rho) die("P pressure or RHO density vector not found in atmosphere file");
    /* _C0_ priority C0 > (P, RHO) > T (ESSModeSolver.cpp:159-183) */
    for (int i = 0; i < n; ++i) {
        if (!have_c0) pr->c0[i] = sqrt(GAMMA_FOR_C * pr->P[i] / pr->rho[i]);
        if (!have_T) pr->T[i] = pr->c0[i] * pr->c0[i] / GAMMA_FOR_C / R_FOR_C;
        /* Atmosphere1D.cpp:484-544: |w|, direction CW from N in degrees */
        pr->ws[i] = sqrt(pr->u[i] * pr->u[i] + pr->v[i] * pr->v[i]);
        pr->wd[i] = (PI / 2.0 - atan2(pr->v[i], pr->u[i])) * 180.0 / PI;
    }
}

/* ------------------------------------------------------------------------------------ */
/* Steffen monotone cubic (GSL gsl_interp_steffen); clamped outside (AtmosphericProperty1D) */
/* ------------------------------------------------------------------------------------ */
typedef struct {
    int n;
    const double *x;
    double *a, *b, *c, *d, y0, yn;
} Spline;

static double sgn1(double v) { return v < 0.0 ? -1.0 : 1.0; }

static void spline_init(Spline *s, const double *x, const double *y, int n) {
    s->n = n;
    s->x = x;
    s->a = malloc(sizeof(double) * 4 * (n - 1));
    s->b = s->a + (n - 1);
    s->c = s->b + (n - 1);
    s->d = s->c + (n - 1);
    double *yp = malloc(sizeof(double) * n), *h = malloc(sizeof(double) * (n - 1)), *sl = malloc(sizeof(double) * (n - 1));
    for (int i = 0; i < n - 1; ++i) {
        h[i] = x[i + 1] - x[i];
        sl[i] = (y[i + 1] - y[i]) / h[i];
    }
    yp[0] = sl[0];
    yp[n - 1] = sl[n - 2];
    for (int i = 1; i < n - 1; ++i) {
        const double p = (sl[i - 1] * h[i] + sl[i] * h[i - 1]) / (h[i - 1] + h[i]);
        yp[i] = (sgn1(sl[i - 1]) + sgn1(sl[i])) * fmin(fabs(sl[i - 1]), fmin(fabs(sl[i]), 0.5 * fabs(p)));
    }
    for (int i = 0; i < n - 1; ++i) {
        s->a[i] = (yp[i] + yp[i + 1] - 2.0 * sl[i]) / h[i] / h[i];
        s->b[i] = (3.0 * sl[i] - 2.0 * yp[i] - yp[i + 1]) / h[i];
        s->c[i] = yp[i];
        s->d[i] = y[i];
    }
    s->y0 = y[0];
    s->yn = y[n - 1];
    free(yp);
    free(h);
    free(sl);
}

static void spline_free(Spline *s) { free(s->a); }

static int spline_index(const Spline *s, double xq) {
    int lo = 0, hi = s->n - 1;   /* x[lo] <= xq < x[hi], capped at n-2 */
    while (hi - lo > 1) {
        const int mid = (lo + hi) / 2;
        if (s->x[mid] > xq) hi = mid; else lo = mid;
    }
    return lo > s->n - 2 ? s->n - 2 : lo;
}

static double spline_get(const Spline *s, double xq) {
    if (xq > s->x[s->n - 1]) return s->yn;
    if (xq < s->x[0]) return s->y0;
    const int i = spline_index(s, xq);
    const double dx = xq - s->x[i];
    return s->d[i] + dx * (s->c[i] + dx * (s->b[i] + dx * s->a[i]));
}

static double spline_deriv(const Spline *s, double xq) {
    if (xq > s->x[s->n - 1] || xq < s->x[0]) return 0.0;
    const int i = spline_index(s, xq);
    const double dx = xq - s->x[i];
    return s->c[i] + dx * (2.0 * s->b[i] + dx * 3.0 * s->a[i]);
}

/* ------------------------------------------------------------------------------------ */
/* Sutherland-Bass attenuation [Np/m] at z [km] (AtmosphericModel.cpp:86-219)           */
/* ------------------------------------------------------------------------------------ */
static double sutherland_bass(double z, double T, double P, double D, double freq) {
    const double mu_o = 18.192E-6, T_o = 293.15, P_o = 101325.0, S = 117.0;
    const double Cv_R[4] = {5.0 / 2.0, 5.0 / 2.0, 3.0, 3.0}, Cp_R[4] = {7.0 / 2.0, 7.0 / 2.0, 4.0, 4.0};
    const double theta[4] = {2239.1, 3352.0, 915.0, 1037.0};
    const double c_snd = sqrt(GAMMA_FOR_C * P / D);
    const double mu = mu_o * sqrt(T / T_o) * ((1.0 + S / T_o) / (1.0 + S / T));
    const double nu = (8.0 * PI * freq * mu) / (3.0 * P);
    const double z2 = z * z, z3 = z2 * z, z4 = z3 * z, z5 = z4 * z;
    double X[7];
    X[0] = z > 90.0 ? pow(10.0, 49.296 - 1.5524 * z + 1.8714E-2 * z2 - 1.1069E-4 * z3 + 3.199E-7 * z4 - 3.6211E-10 * z5)
                    : pow(10.0, -0.67887);
    X[1] = z > 76.0 ? pow(10.0, 1.3972E-1 - 5.6269E-3 * z + 3.9407E-5 * z2 - 1.0737E-7 * z3) : pow(10.0, -0.10744);
    X[2] = pow(10.0, -3.3979);
    X[3] = z > 80.0 ? pow(10.0, -4.234 - 3.0975E-2 * z)
                    : pow(10.0, -19.027 + 1.3093 * z - 4.6496E-2 * z2 + 7.8543E-4 * z3 - 6.5169E-6 * z4 + 2.1343E-8 * z5);
    X[4] = z > 95.0 ? pow(10.0, -3.2456 + 4.6642E-2 * z - 2.6894E-4 * z2 + 5.264E-7 * z3)
                    : pow(10.0, -11.195 + 1.5408E-1 * z - 1.4348E-3 * z2 + 1.0166E-5 * z3);
    X[5] = pow(10.0, -53.746 + 1.5439 * z - 1.8824E-2 * z2 + 1.1587E-4 * z3 - 3.5399E-7 * z4 + 4.2609E-10 * z5);
    /* the 'z > 100' H2O branch sits under 'z <= 30' in the reference and is never taken */
    X[6] = z > 30.0 ? pow(10.0, -4.2563 + 7.6245E-2 * z - 2.1824E-3 * z2 - 2.3010E-6 * z3 + 2.4265E-7 * z4 - 1.2500E-09 * z5)
                    : pow(10.0, -1.7491 + 4.4986E-2 * z - 6.8549E-2 * z2 + 5.4639E-3 * z3 - 1.5539E-4 * z4 + 1.5063E-06 * z5);
    const double X_ON = (X[0] + X[1]) / 0.9903;
    const double Z_rot_0 = 54.1 * exp(-17.3 * pow(T, -1.0 / 3.0));
    const double Z_rot_1 = 63.3 * exp(-16.7 * pow(T, -1.0 / 3.0));
    const double Z_rot = 1.0 / ((X[1] / Z_rot_1) + (X[0] / Z_rot_0));
    const double sigma = 5.0 / sqrt(21.0);
    const double nn = (4.0 / 5.0) * sqrt(3.0 / 7.0) * Z_rot;
    const double chi = 3.0 * nn * nu / 4.0;
    const double cchi = 2.36 * chi;
    const double a_cl = (2 * PI * freq / c_snd) *
                        sqrt(0.5 * (sqrt(1 + nu * nu) - 1) * (1 + cchi * cchi) / ((1 + nu * nu) * (1 + (sigma * cchi) * (sigma * cchi))));
    const double a_rot = (2 * PI * freq / c_snd) * X_ON * ((sigma * sigma - 1) * chi / (2 * sigma)) *
                         sqrt(0.5 * (sqrt(1 + nu * nu) + 1) / ((1 + nu * nu) * (1 + cchi * cchi)));
    const double a_diff = 0.003 * a_cl;
    const double Tr = pow(T / T_o, -1.0 / 3.0) - 1.0;
    const double A1 = (X[0] + X[1]) * 24.0 * exp(-9.16 * Tr), A2 = (X[4] + X[5]) * 2400.0;
    const double B = 40400.0 * exp(10.0 * Tr), C = 0.02 * exp(-11.2 * Tr), Dd = 0.391 * exp(8.41 * Tr);
    const double E = 9.0 * exp(-19.9 * Tr), F = 60000.0, G = 28000.0 * exp(-4.17 * Tr);
    const double H = 22000.0 * exp(-7.68 * Tr), Iv = 15100.0 * exp(-10.4 * Tr), J = 11500.0 * exp(-9.17 * Tr);
    const double K = 8.48E08 * exp(9.17 * Tr), L = exp(-7.72 * Tr);
    const double ZZ = H * X[2] + Iv * (X[0] + 0.5 * X[4]) + J * (X[1] + 0.5 * X[5]) + K * (X[6] + X[3]);
    const double hu = 100.0 * (X[3] + X[6]);
    const double pf = (P / P_o) * (mu_o / mu);
    const double f_vib[4] = {pf * (A1 + A2 + B * hu * (C + hu) * (Dd + hu)), pf * (E + F * X[3] + G * X[6]), pf * ZZ,
                             pf * 1.2E5 * L};
    double a_vib = 0.0;
    for (int m = 0; m < 4; ++m) {
        const double th = theta[m] / T, ex = exp(-th);
        const double C_R = (th * th * ex) / ((1 - ex) * (1 - ex));
        const double A_max = (X[m] * (PI / 2) * C_R) / (Cp_R[m] * (Cv_R[m] + C_R));
        const double fr = freq / f_vib[m];
        a_vib += (A_max / c_snd) * ((2 * (freq * freq) / f_vib[m]) / (1 + fr * fr));
    }
    return a_cl + a_rot + a_diff + a_vib;
}

/* ------------------------------------------------------------------------------------ */
/* one solve                                                                             */
/* ------------------------------------------------------------------------------------ */
typedef struct {
    const Profile *prof;
    double freq, az, zs, zr, maxheight_km, c_min, c_max;
    int Nz, wkb, lamb, filter;
    int nr;          /* 1-D TL ranges (dr) */
    double dr;
    int rows, stride, field_ranges;   /* 2-D field: rows 0, stride, .. at field_ranges ranges */
    /* results */
    int nm, used_wkb;
    double k_min, k_max, dz;
    double *k, *kpi, *ps, *pr;         /* k (unperturbed), Im k_pert... per mode */
    double complex *kp;
    double *tl_re, *tl_im, *tl_inc;
    double t_setup, t_eig, t_tl1d, t_tl2d;
    double field_check;                /* sum |p| of the sampled field (keeps it computed) */
} Solve;

/* number of eigenvalues of the tridiagonal (a, unit e) below x: LDL^T inertia with the
 * pivot guard of dstebz (|d| < pivmin -> -pivmin); 4 shifts at once */
static void sturm4(const double *a, int n, double e2, double pivmin, const double *x, int *cnt) {
    double d0 = a[0] - x[0], d1 = a[0] - x[1], d2 = a[0] - x[2], d3 = a[0] - x[3];
    int c0 = d0 < 0, c1 = d1 < 0, c2 = d2 < 0, c3 = d3 < 0;
    for (int i = 1; i < n; ++i) {
        if (fabs(d0) < pivmin) d0 = -pivmin;
        if (fabs(d1) < pivmin) d1 = -pivmin;
        if (fabs(d2) < pivmin) d2 = -pivmin;
        if (fabs(d3) < pivmin) d3 = -pivmin;
        d0 = (a[i] - x[0]) - e2 / d0;
        d1 = (a[i] - x[1]) - e2 / d1;
        d2 = (a[i] - x[2]) - e2 / d2;
        d3 = (a[i] - x[3]) - e2 / d3;
        c0 += d0 < 0;
        c1 += d1 < 0;
        c2 += d2 < 0;
        c3 += d3 < 0;
    }
    cnt[0] = c0;
    cnt[1] = c1;
    cnt[2] = c2;
    cnt[3] = c3;
}

static int sturm1(const double *a, int n, double e2, double pivmin, double x) {
    double xs[4] = {x, x, x, x};
    int c[4];
    sturm4(a, n, e2, pivmin, xs, c);
    return c[0];
}

/* inverse iteration for eigenvalue lam of tridiag(a, e): Gaussian elimination with partial
 * pivoting of (T - lam I) (dlagtf), three solves from a fixed start, each normalised (dstein's
 * iteration); v: unit 2-norm */
static void inverse_iteration(const double *a, int n, double e, double lam, double *v, double *work) {
    double *u1 = work, *u2 = u1 + n, *u3 = u2 + n, *l = u3 + n;
    int *piv = (int *)(l + n);
    /* LU of T - lam I with row interchanges: rows i, i+1 */
    double diag = a[0] - lam, sup = e;
    for (int i = 0; i < n - 1; ++i) {
        const double sub = e, nd = a[i + 1] - lam, ns = (i + 1 < n - 1) ? e : 0.0;
        if (fabs(diag) >= fabs(sub)) {
            piv[i] = 0;
            if (diag == 0.0) diag = 1e-300;
            const double m = sub / diag;
            l[i] = m;
            u1[i] = diag;
            u2[i] = sup;
            u3[i] = 0.0;
            diag = nd - m * sup;
            sup = ns;
        } else {
            piv[i] = 1;
            const double m = diag / sub;
            l[i] = m;
            u1[i] = sub;
            u2[i] = nd;
            u3[i] = ns;
            diag = sup - m * nd;
            sup = -m * ns;
        }
    }
    u1[n - 1] = diag == 0.0 ? 1e-300 : diag;
    for (int i = 0; i < n; ++i) v[i] = 1.0 + 1e-3 * (double)((i * 7919) % 1000) / 1000.0;
    for (int it = 0; it < 3; ++it) {
        for (int i = 0; i < n - 1; ++i) {   /* forward: L^-1 with the interchanges */
            if (piv[i]) {
                const double t = v[i];
                v[i] = v[i + 1];
                v[i + 1] = t - l[i] * v[i + 1];
            } else {
                v[i + 1] -= l[i] * v[i];
            }
        }
        v[n - 1] /= u1[n - 1];   /* back: U^-1 */
        if (n > 1) v[n - 2] = (v[n - 2] - u2[n - 2] * v[n - 1]) / u1[n - 2];
        for (int i = n - 3; i >= 0; --i) v[i] = (v[i] - u2[i] * v[i + 1] - u3[i] * v[i + 2]) / u1[i];
        double s = 0.0;
        for (int i = 0; i < n; ++i) s += v[i] * v[i];
        s = 1.0 / sqrt(s);
        for (int i = 0; i < n; ++i) v[i] *= s;
    }
}

static void solve(Solve *S) {
    const double t0 = now_s();
    const Profile *pr = S->prof;
    const int n = pr->n, Nz = S->Nz;
    /* setParams grid (ESSModeSolver.cpp:144-151, 212-220) */
    double z_min = 0.0, maxheight = S->maxheight_km * 1000.0;
    if (maxheight >= pr->z[n - 1]) maxheight = pr->z[n - 1] - 1e-6;
    if (pr->has_z0 && z_min < pr->z0) z_min = pr->z0;
    const double dz = (maxheight - z_min) / (Nz - 1);
    S->dz = dz;
    const double omega = 2 * PI * S->freq;
    const double azr = S->az * PI / 180.0;
    double *ceffn = malloc(sizeof(double) * n), *alphan = malloc(sizeof(double) * n);
    for (int k = 0; k < n; ++k) {
        ceffn[k] = pr->ws[k] * cos(pr->wd[k] * PI / 180.0 - azr) + pr->c0[k];   /* Atmosphere1D.cpp:546-610 */
        alphan[k] = sutherland_bass(pr->z[k] * 0.001, pr->T[k], pr->P[k], pr->rho[k], S->freq);
    }
    Spline sc, sa, sp, sr;
    spline_init(&sc, pr->z, ceffn, n);
    spline_init(&sa, pr->z, alphan, n);
    spline_init(&sp, pr->z, pr->P, n);
    spline_init(&sr, pr->z, pr->rho, n);
    double *ceff = malloc(sizeof(double) * Nz), *diag = malloc(sizeof(double) * Nz), *w = malloc(sizeof(double) * Nz);
    for (int i = 0; i < Nz; ++i) {
        const double z = z_min + i * dz;
        ceff[i] = spline_get(&sc, z);
        const double cT = sqrt(1.4 * spline_get(&sp, z) / spline_get(&sr, z));
        w[i] = dz * (omega / cT) * spline_get(&sa, z) * 2;                 /* ModeSolver.cpp:276-278 */
        diag[i] = (omega / ceff[i]) * (omega / ceff[i]);
    }
    const double adm = S->lamb ? -spline_deriv(&sr, z_min) / spline_get(&sr, z_min) / 2.0 : 0.0;
    diag[0] = (1.0 / (dz * adm + 1)) / (dz * dz) + diag[0];
    /* getModalTrace interval (ESSModeSolver.cpp:561-621) */
    double ceffmin = ceff[0];
    for (int i = 0; i < Nz; ++i) ceffmin = fmin(ceffmin, ceff[i]);
    double k_max;
    S->used_wkb = 0;
    if (fabs(S->zs) < 1.0e-3 && fabs(S->zr) < 1.0e-3 && S->wkb) {
        S->used_wkb = 1;
        const double kmf = omega / ceffmin, kg = omega / ceff[0];
        const double dkk = (pow(kmf, 2) - pow(kg, 2)) / 100.0;
        double kk = pow(kg, 2);
        if (dkk > 1.0e-10) {
            for (kk = pow(kg, 2); kk < pow(kmf, 2); kk = kk + dkk) {
                double integral = 0.0, term = 1.0;
                int i = 0;
                while (term > dkk && i < Nz) {
                    const double ke = omega / ceff[i];
                    term = fabs(kk - pow(ke, 2));
                    integral = integral + dz * sqrt(term);
                    ++i;
                }
                if (integral >= 10.0) break;
            }
        }
        k_max = sqrt(kk);
    } else {
        k_max = omega / ceffmin;
    }
    const int top = Nz - Nz / 10;
    double k_min = omega / ceff[top + 1];
    if (S->used_wkb && k_max < k_min) k_max = omega / ceffmin;
    if (S->filter) {
        k_min = omega / S->c_max;
        k_max = omega / S->c_min;
    }
    S->k_min = k_min;
    S->k_max = k_max;
    /* the tridiagonal of doESSCalculation (EigenEngine.cpp:57-79), last row diag[Nz-2] */
    const double h2 = dz * dz, e = 1.0 / h2, e2 = e * e;
    double *a = malloc(sizeof(double) * Nz);
    for (int i = 0; i < Nz - 1; ++i) a[i] = -2.0 / h2 + diag[i];
    a[Nz - 1] = -2.0 / h2 + diag[Nz - 2];
    double tnorm = 0.0;
    for (int i = 0; i < Nz; ++i) tnorm = fmax(tnorm, fabs(a[i]) + 2 * e);
    const double pivmin = 1e-300 * fmax(1.0, e2);
    S->t_setup = now_s() - t0;
    const double t1 = now_s();
    const double lo = k_min * k_min, hi = k_max * k_max;
    const int n_lo = sturm1(a, Nz, e2, pivmin, lo), n_hi = sturm1(a, Nz, e2, pivmin, hi);
    int nm = n_hi - n_lo;
    if (nm > MAX_MODES) die("more than MAX_MODES eigenvalues in the interval");
    if (nm < 0) nm = 0;
    double *lam = malloc(sizeof(double) * (nm + 4));
    /* bisection per index (dstebz), four indices at a time; each bracket starts at [lo, hi]
     * narrowed by the previous index's eigenvalue */
    const double tol = 4.0 * 2.220446049250313e-16 * tnorm;
    double prev = lo;
    for (int j0 = 0; j0 < nm; j0 += 4) {
        double blo[4], bhi[4], x[4];
        int cnt[4], idx[4];
        for (int q = 0; q < 4; ++q) {
            idx[q] = n_lo + (j0 + q < nm ? j0 + q : nm - 1);
            blo[q] = prev;
            bhi[q] = hi;
        }
        for (int it = 0; it < 200; ++it) {
            int done = 1;
            for (int q = 0; q < 4; ++q) {
                x[q] = 0.5 * (blo[q] + bhi[q]);
                if (bhi[q] - blo[q] > tol) done = 0;
            }
            if (done) break;
            sturm4(a, Nz, e2, pivmin, x, cnt);
            for (int q = 0; q < 4; ++q) {
                if (bhi[q] - blo[q] <= tol) continue;
                if (cnt[q] <= idx[q]) blo[q] = x[q]; else bhi[q] = x[q];
            }
        }
        for (int q = 0; q < 4 && j0 + q < nm; ++q) lam[j0 + q] = 0.5 * (blo[q] + bhi[q]);
        prev = lam[j0 + (j0 + 3 < nm ? 3 : nm - 1 - j0)];
    }
    /* eigenvectors: inverse iteration; per mode the perturbation integral (doPerturb) and the
     * recorded rows (source, receiver, field rows) */
    const int i_s = (int)ceil(S->zs / dz), i_r = (int)ceil(S->zr / dz);
    double *v = malloc(sizeof(double) * Nz), *vprev = malloc(sizeof(double) * Nz);
    double *work = malloc(sizeof(double) * Nz * 5 + sizeof(int) * Nz);
    double *psi_rows = S->rows > 0 ? malloc(sizeof(double) * (size_t)S->rows * (nm + 1)) : NULL;
    S->k = malloc(sizeof(double) * (nm + 1));
    S->kp = malloc(sizeof(double complex) * (nm + 1));
    S->ps = malloc(sizeof(double) * (nm + 1));
    S->pr = malloc(sizeof(double) * (nm + 1));
    int kept = 0;
    double lam_prev = -1e300;
    const double isq = 1.0 / sqrt(dz);
    for (int j = 0; j < nm; ++j) {
        inverse_iteration(a, Nz, e, lam[j], v, work);
        if (lam[j] - lam_prev < 1e-10 * tnorm) {   /* a numerically clustered pair: orthogonalise */
            double d = 0.0, s = 0.0;
            for (int i = 0; i < Nz; ++i) d += v[i] * vprev[i];
            for (int i = 0; i < Nz; ++i) { v[i] -= d * vprev[i]; s += v[i] * v[i]; }
            s = 1.0 / sqrt(s);
            for (int i = 0; i < Nz; ++i) v[i] *= s;
        }
        memcpy(vprev, v, sizeof(double) * Nz);
        lam_prev = lam[j];
        const double kj = sqrt(lam[j]);
        if (!(kj >= k_min && kj <= k_max)) continue;   /* doSelect */
        double A = 0.0;
        for (int i = 0; i < Nz; ++i) A += (v[i] * isq) * (v[i] * isq) * w[i];
        const double sg = v[i_s] < 0 ? -1.0 : 1.0;   /* psi(zs) >= 0 */
        S->k[kept] = kj;
        S->kp[kept] = csqrt(kj * kj + I * A);
        S->ps[kept] = sg * v[i_s] * isq;
        S->pr[kept] = sg * v[i_r] * isq;
        for (int q = 0; q < S->rows; ++q) psi_rows[(size_t)q * (nm + 1) + kept] = sg * v[q * S->stride] * isq;
        ++kept;
    }
    S->nm = kept;
    S->t_eig = now_s() - t1;
    /* getTLoss1D (ModeSolver.cpp:325-356): complex exp and sqrt per term */
    const double t2 = now_s();
    const double complex expov8pi = 4 * PI * I * cexp(-I * PI * 0.25) / sqrt(8.0 * PI);
    S->tl_re = malloc(sizeof(double) * (S->nr + 1));
    S->tl_im = malloc(sizeof(double) * (S->nr + 1));
    S->tl_inc = malloc(sizeof(double) * (S->nr + 1));
    for (int i = 0; i < S->nr; ++i) {
        const double r = (i + 1) * S->dr;
        double complex sc_ = 0.0;
        double si = 0.0;
        for (int m = 0; m < kept; ++m) {
            sc_ = sc_ + S->ps[m] * S->pr[m] * cexp(I * S->kp[m] * r) / csqrt(S->kp[m]);
            si = si + pow(S->ps[m] * S->pr[m], 2) * exp(-2 * cimag(S->kp[m]) * r) / cabs(S->kp[m]);
        }
        sc_ = 1000.0 * expov8pi * sc_ / sqrt(r);
        S->tl_re[i] = creal(sc_);
        S->tl_im[i] = cimag(sc_);
        S->tl_inc[i] = 1000.0 * 4 * PI * sqrt(si) * sqrt(1. / 8. / PI / r);
    }
    S->t_tl1d = now_s() - t2;
    /* getTLoss2D (ModeSolver.cpp:524-544) over the first field_ranges ranges of the 1-D grid */
    const double t3 = now_s();
    double chk = 0.0;
    for (int i = 0; i < S->field_ranges; ++i) {
        const double r = (i + 1) * S->dr;
        for (int q = 0; q < S->rows; ++q) {
            double complex sum = 0.0;
            const double *row = psi_rows + (size_t)q * (nm + 1);
            for (int m = 0; m < kept; ++m) sum = sum + S->ps[m] * row[m] * cexp(I * S->kp[m] * r) / csqrt(S->kp[m]);
            sum = 1000.0 * expov8pi * sum / sqrt(r);
            chk += cabs(sum);
        }
    }
    S->field_check = chk;
    S->t_tl2d = now_s() - t3;
    free(psi_rows);
    free(v);
    free(vprev);
    free(work);
    free(lam);
    free(a);
    free(ceff);
    free(diag);
    free(w);
    free(ceffn);
    free(alphan);
    spline_free(&sc);
    spline_free(&sa);
    spline_free(&sp);
    spline_free(&sr);
}

static void solve_free(Solve *S) {
    free(S->k);
    free(S->kp);
    free(S->ps);
    free(S->pr);
    free(S->tl_re);
    free(S->tl_im);
    free(S->tl_inc);
}

/* ------------------------------------------------------------------------------------ */
/* command line                                                                          */
/* ------------------------------------------------------------------------------------ */
typedef struct {
    Solve *jobs;
    int n, next;
    pthread_mutex_t mu;
} Pool;

static void *worker(void *arg) {
    Pool *p = arg;
    for (;;) {
        pthread_mutex_lock(&p->mu);
        const int j = p->next++;
        pthread_mutex_unlock(&p->mu);
        if (j >= p->n) return NULL;
        solve(&p->jobs[j]);
    }
}

static void print_array(const char *name, const double *x, int n, int last) {
    printf("\"%s\": [", name);
    for (int i = 0; i < n; ++i) printf(i ? ", %.17g" : "%.17g", x[i]);
    printf(last ? "]" : "], ");
}

int main(int argc, char **argv) {
    if (argc < 2) die("usage: ess_port solve|bench [options]");
    const int bench = !strcmp(argv[1], "bench");
    Solve base;
    memset(&base, 0, sizeof base);
    base.Nz = 20000;
    base.maxheight_km = 150.0;
    base.freq = 0.1;
    base.az = 90.0;
    base.nr = 0;
    base.dr = 100.0;
    int threads = 1;
    const char *profile = NULL;
    char **jobspec = calloc(argc, sizeof(char *));
    int njobs = 0;
    for (int i = 2; i < argc; ++i) {
        const char *o = argv[i];
#define NEXT() (i + 1 < argc ? argv[++i] : (die("missing option value"), ""))
        if (!strcmp(o, "--profile")) profile = NEXT();
        else if (!strcmp(o, "--freq")) base.freq = atof(NEXT());
        else if (!strcmp(o, "--azimuth")) base.az = atof(NEXT());
        else if (!strcmp(o, "--Nz")) base.Nz = atoi(NEXT());
        else if (!strcmp(o, "--maxheight_km")) base.maxheight_km = atof(NEXT());
        else if (!strcmp(o, "--zs_m")) base.zs = atof(NEXT());
        else if (!strcmp(o, "--zr_m")) base.zr = atof(NEXT());
        else if (!strcmp(o, "--wkb")) base.wkb = 1;
        else if (!strcmp(o, "--lamb")) base.lamb = 1;
        else if (!strcmp(o, "--filter")) { base.filter = 1; base.c_min = atof(NEXT()); base.c_max = atof(NEXT()); }
        else if (!strcmp(o, "--nr")) base.nr = atoi(NEXT());
        else if (!strcmp(o, "--dr")) base.dr = atof(NEXT());
        else if (!strcmp(o, "--rows")) base.rows = atoi(NEXT());
        else if (!strcmp(o, "--stride")) base.stride = atoi(NEXT());
        else if (!strcmp(o, "--field_ranges")) base.field_ranges = atoi(NEXT());
        else if (!strcmp(o, "--threads")) threads = atoi(NEXT());
        else if (!strcmp(o, "--job")) jobspec[njobs++] = (char *)NEXT();
        else die("unknown option");
#undef NEXT
    }
    if (base.Nz < 20) die("Nz must be at least 20");
    if (base.rows > 0 && (base.stride < 1 || (long)(base.rows - 1) * base.stride >= base.Nz)) die("field rows beyond the grid");
    if (!bench) {
        if (!profile) die("--profile is required");
        Profile pr;
        read_profile(profile, &pr);
        base.prof = &pr;
        solve(&base);
        printf("{\"nm\": %d, \"used_wkb\": %d, \"k_min\": %.17g, \"k_max\": %.17g, \"dz\": %.17g, ", base.nm,
               base.used_wkb, base.k_min, base.k_max, base.dz);
        double *kr = malloc(sizeof(double) * (base.nm + 1)), *ki = malloc(sizeof(double) * (base.nm + 1));
        for (int m = 0; m < base.nm; ++m) { kr[m] = creal(base.kp[m]); ki[m] = cimag(base.kp[m]); }
        print_array("k", base.k, base.nm, 0);
        print_array("kp_re", kr, base.nm, 0);
        print_array("kp_im", ki, base.nm, 0);
        print_array("psi_s", base.ps, base.nm, 0);
        print_array("psi_r", base.pr, base.nm, 0);
        print_array("tl_re", base.tl_re, base.nr, 0);
        print_array("tl_im", base.tl_im, base.nr, 0);
        print_array("tl_inc", base.tl_inc, base.nr, 0);
        printf("\"times\": [%.6f, %.6f, %.6f, %.6f]}\n", base.t_setup, base.t_eig, base.t_tl1d, base.t_tl2d);
        solve_free(&base);
        return 0;
    }
    /* bench: jobs "profile:azimuth:freq" on `threads` threads */
    if (njobs < 1) die("bench needs at least one --job");
    Profile *profs = calloc(njobs, sizeof(Profile));
    Solve *jobs = calloc(njobs, sizeof(Solve));
    for (int j = 0; j < njobs; ++j) {
        char *spec = jobspec[j];
        char *c1 = strrchr(spec, ':');
        if (!c1) die("job: profile:azimuth:freq");
        *c1 = 0;
        char *c2 = strrchr(spec, ':');
        if (!c2) die("job: profile:azimuth:freq");
        *c2 = 0;
        read_profile(spec, &profs[j]);
        jobs[j] = base;
        jobs[j].prof = &profs[j];
        jobs[j].az = atof(c2 + 1);
        jobs[j].freq = atof(c1 + 1);
    }
    Pool pool = {jobs, njobs, 0, PTHREAD_MUTEX_INITIALIZER};
    pthread_t *th = calloc(threads, sizeof(pthread_t));
    const double t0 = now_s();
    for (int t = 0; t < threads; ++t) pthread_create(&th[t], NULL, worker, &pool);
    for (int t = 0; t < threads; ++t) pthread_join(th[t], NULL);
    const double wall = now_s() - t0;
    printf("{\"threads\": %d, \"wall_s\": %.6f, \"jobs\": [", threads, wall);
    for (int j = 0; j < njobs; ++j)
        printf("%s{\"nm\": %d, \"times\": [%.6f, %.6f, %.6f, %.6f], \"check\": %.6e}", j ? ", " : "", jobs[j].nm,
               jobs[j].t_setup, jobs[j].t_eig, jobs[j].t_tl1d, jobs[j].t_tl2d, jobs[j].field_check);
    printf("]}\n");
    return 0;
}
