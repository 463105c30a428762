/*
 * orc_p2d.c -- CPU oracle for GALAXY's 2-D polar-grid PM step.
 *
 * TEST INFRASTRUCTURE ONLY (see orc_p2d.h).  Built by oracle/Makefile with
 * gcc -O2 -ffp-contract=off into oracle/liboracle.so; loaded by
 * oracle/oracle.py.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "orc_p2d.h"

/* 2.d0 * pi of SRC/lib15/inc/pi.f */
#define TWO_PI_D (2.0 * 3.141592653589793238462643)

/* ---------------- REAL-templated particle routines ----------------
 * The FP32 logarithm of uofgr is the correctly rounded logf, taken through
 * the double logarithm: the reference's logf comes from whatever libm its
 * compiler links (unpinned; glibc's logf is within 0.82 ulp, not correctly
 * rounded), so the oracle fixes the one value every libm aims at. */
static inline float logf_cr(float v) { return (float)log((double)v); }
#define REAL float
#define SFX(name) name##_f
#define RLOG logf_cr
#define REXP expf
#define RSQRT sqrtf
#define RFMOD fmodf
#include "orc_particles.inc"
#undef REAL
#undef SFX
#undef RLOG
#undef REXP
#undef RSQRT
#undef RFMOD

#define REAL double
#define SFX(name) name##_d
#define RLOG log
#define REXP exp
#define RSQRT sqrt
#define RFMOD fmod
#include "orc_particles.inc"
#undef REAL
#undef SFX

/* alpha = 2 pi / na, stored to REAL (SRC/lib15/setgrd.f:88) */
float orc_alpha_f(int na) { return (float)(TWO_PI_D / (double)na); }
double orc_alpha_d(int na) { return TWO_PI_D / (double)na; }

/* Cubic Hermite interpolation of a tabulated radial force on r_k = k dr
 * (val, der = its derivative), with s = r * (1/dr) and t = s - k: the cubic
 * written as c0 + t (c1 + t (c2 + t c3)) with c0 = v_k, c1 = dr d_k,
 * c2 = (3 (v_k+1 - v_k) - 2 c1) - dr d_k+1, c3 = (2 (v_k - v_k+1) + c1) +
 * dr d_k+1, evaluated by fused multiply-adds; past the last entry the
 * constant val[nfr - 1].  (The Hermite form h00 v_k + h10 dr d_k + h01 v_k+1
 * + h11 dr d_k+1 regrouped by powers of t; the library evaluates the same
 * coefficients the same way, so the two agree to the bit.) */
double orc_fixforce(double r, int nfr, double dr, const double *val,
                    const double *der)
{
    double inv_dr = 1.0 / dr;
    double s = r * inv_dr;
    int k = (int)s;
    if (k >= nfr - 1) return val[nfr - 1];
    if (k < 0) k = 0;
    double t = s - (double)k;
    double v0 = val[k], v1 = val[k + 1];
    double c1 = dr * der[k];
    double c2 = (3.0 * (v1 - v0) - 2.0 * c1) - dr * der[k + 1];
    double c3 = (2.0 * (v0 - v1) + c1) + dr * der[k + 1];
    return fma(fma(fma(c3, t, c2), t, c1), t, v0);
}

/* ---------------- FFTPACK conventions, direct sums ---------------- */

/* cos / sin of 2 pi k / n with k reduced exactly modulo n */
static double cos2pi(long k, long n)
{
    k %= n;
    if (k < 0) k += n;
    return cos(TWO_PI_D * (double)k / (double)n);
}
static double sin2pi(long k, long n)
{
    k %= n;
    if (k < 0) k += n;
    return sin(TWO_PI_D * (double)k / (double)n);
}

/* forward real DFT, half-complex (stest.f:84-103):
 * y[0] = sum x; y[2k-1] = sum x cos, y[2k] = -sum x sin; y[n-1] = alt sum */
void orc_rfftf(int n, const double *x, double *y)
{
    double s = 0.0;
    for (int i = 0; i < n; i++) s += x[i];
    y[0] = s;
    for (int k = 1; 2 * k < n; k++) {
        double a = 0.0, b = 0.0;
        for (int i = 0; i < n; i++) {
            a += x[i] * cos2pi((long)k * i, n);
            b += x[i] * sin2pi((long)k * i, n);
        }
        y[2 * k - 1] = a;
        y[2 * k] = -b;
    }
    if (n % 2 == 0) {
        double a = 0.0;
        for (int i = 0; i < n; i++) a += (i % 2 ? -x[i] : x[i]);
        y[n - 1] = a;
    }
}

/* unnormalised backward real DFT (stest.f:106-122) */
void orc_rfftb(int n, const double *x, double *y)
{
    for (int i = 0; i < n; i++) {
        double s = x[0];
        for (int k = 1; 2 * k < n; k++)
            s += 2.0 * (x[2 * k - 1] * cos2pi((long)k * i, n) -
                        x[2 * k] * sin2pi((long)k * i, n));
        if (n % 2 == 0) s += (i % 2 ? -x[n - 1] : x[n - 1]);
        y[i] = s;
    }
}

/* DCT-I of n points (rcost, stest.f:169-195):
 * y_i = x_0 + (-1)^i x_{n-1} + 2 sum_{k=1}^{n-2} x_k cos(pi k i/(n-1)) */
/* first nout outputs of the DCT-I below */
static void cost_n(int n, const double *x, double *y, int nout)
{
    int m = n - 1;
    for (int i = 0; i < nout; i++) {
        double s = x[0] + (i % 2 ? -x[m] : x[m]);
        for (int k = 1; k < m; k++) s += 2.0 * x[k] * cos2pi((long)k * i, 2 * m);
        y[i] = s;
    }
}

void orc_cost(int n, const double *x, double *y) { cost_n(n, x, y, n); }

/* DST-I of n points (rsint, stest.f:138-165):
 * y_i = 2 sum_{k=1}^{n} x_k sin(pi k i/(n+1)), i = 1..n */
static void sint_n(int n, const double *x, double *y, int nout)
{
    for (int i = 1; i <= nout; i++) {
        double s = 0.0;
        for (int k = 1; k <= n; k++)
            s += 2.0 * x[k - 1] * sin2pi((long)k * i, 2 * (n + 1));
        y[i - 1] = s;
    }
}

void orc_sint(int n, const double *x, double *y) { sint_n(n, x, y, n); }

/* ---------------- Green's function kernels (FP64) ---------------- */

/* softm cubic-spline enclosed-mass fraction (SRC/lib15/softm.f:15-31) */
static double softm(double ds)
{
    if (ds <= 0.0) return 0.0;
    if (ds < 1.0)
        return 4.0 * ds * ds * ds * (1.0 / 3.0 - 0.3 * ds * ds + 0.125 * ds * ds * ds);
    if (ds < 2.0)
        return -1.0 / 15.0 + ds * ds * ds *
               (8.0 / 3.0 - 3.0 * ds + 1.2 * ds * ds - ds * ds * ds / 6.0);
    return 1.0;
}

/* radfor (SRC/lib15/radfor.f:25-38) with zf = 0 */
static double radfor(double rs, double rf, double th, double eps, int tsoft)
{
    double cth = cos(th);
    double d2 = rs * rs + rf * rf - 2.0 * rs * rf * cth;
    if (tsoft == 1) return (rs * cth - rf) / pow(d2 + eps * eps, 1.5);
    if (d2 > 0.0) return (rs * cth - rf) * softm(sqrt(d2) / eps) / pow(d2, 1.5);
    return 0.0;
}

/* azifor (SRC/lib15/azifor.f:25-40) with zf = 0 */
static double azifor(double rs, double rf, double th, double eps, int tsoft)
{
    double d2 = rs * rs + rf * rf - 2.0 * rs * rf * cos(th);
    if (tsoft == 1) return -rs * sin(th) / pow(d2 + eps * eps, 1.5);
    if (d2 > 0.0) return -rs * sin(th) * softm(sqrt(d2) / eps) / pow(d2, 1.5);
    return 0.0;
}

/* grnfun -> sftpot (SRC/lib15/grnfun.f, sftpot.f:20-40) with zf = 0 */
static double grnfun(double rs, double rf, double th, double eps, int tsoft)
{
    double d2 = rs * rs + rf * rf - 2.0 * rs * rf * cos(th);
    if (tsoft == 1) return -1.0 / sqrt(d2 + eps * eps);
    double x = sqrt(d2) / eps;
    if (x <= 0.0) return -1.4 / eps;
    if (x < 1.0) return (-1.4 + x * x * (2.0 / 3.0 - 0.3 * x * x + 0.1 * x * x * x)) / eps;
    if (x < 2.0)
        return (-1.6 + 1.0 / (15.0 * x) +
                x * x * (4.0 / 3.0 - x + 0.3 * x * x - x * x * x / 30.0)) / eps;
    return -1.0 / sqrt(d2);
}

/* p2grnm (SRC/lib15/p2grnm.f:98-203) + p2gwrt layout (p2gwrt.f:27-56). */
void orc_green(int nr, int na, int ng, int nsect, double alpha, double softl,
               int tsoft, double *G, double *D)
{
    int mmax = na / 2 + 1;
    int resyn = ng < mmax;
    size_t nn = (size_t)nr * nr;
    memset(G, 0, sizeof(double) * 3 * (size_t)ng * nn);
    double *f = malloc(sizeof(double) * mmax);
    double *g = malloc(sizeof(double) * mmax);
    for (int type = 0; type < 3; type++) {
        int azim = (type == 1);
        int it1 = azim ? 2 : 1, it2 = azim ? mmax - 1 : mmax;
        int npt = it2 - it1 + 1;
        for (int is = 1; is <= nr; is++) {
            double rs = exp(alpha * (double)(is - 1)) - 1.0;
            for (int ir = 1; ir <= nr; ir++) {
                double rf = exp(alpha * (double)(ir - 1)) - 1.0;
                for (int it = it1; it <= it2; it++) {
                    double th = TWO_PI_D * (double)(it - 1) / (double)na;
                    double v = type == 0 ? radfor(rs, rf, th, softl, tsoft)
                             : type == 1 ? azifor(rs, rf, th, softl, tsoft)
                                         : grnfun(rs, rf, th, softl, tsoft);
                    f[it - it1] = v;
                }
                /* without the direct kernels only the harmonics m < ng are
                 * needed: the rest are filtered out below */
                int nout = D ? npt : (ng - (azim ? 1 : 0) < npt ? ng - (azim ? 1 : 0) : npt);
                for (int q = nout; q < npt; q++) g[q] = 0.0;
                if (azim) sint_n(npt, f, g, nout);
                else cost_n(npt, f, g, nout);
                /* filter: harmonic m = it - 1 kept iff m < ng and m mod nsect == 0 */
                for (int it = it1; it <= it2; it++) {
                    int m = it - 1;
                    if (m >= ng || m % nsect != 0) g[it - it1] = 0.0;
                }
                int nm = azim ? (mmax - 2 < ng - 1 ? mmax - 2 : ng - 1) : ng;
                for (int q = 0; q < nm; q++) {
                    int m = azim ? q + 1 : q;
                    G[((size_t)type * ng + m) * nn + (size_t)(ir - 1) * nr + (is - 1)] =
                        g[q];
                }
                if (D) {
                    /* direct kernel at angles it = 1..mmax: filtered then
                     * re-synthesised (times 1/na), or raw if nothing filtered */
                    double *d = D + (((size_t)type * nr + (is - 1)) * nr + (ir - 1)) * mmax;
                    for (int it = 0; it < mmax; it++) d[it] = 0.0;
                    if (!resyn) {
                        for (int it = it1; it <= it2; it++) d[it - 1] = f[it - it1];
                    } else {
                        double *h = malloc(sizeof(double) * mmax);
                        if (azim) orc_sint(npt, g, h);
                        else orc_cost(npt, g, h);
                        for (int it = it1; it <= it2; it++)
                            d[it - 1] = h[it - it1] / (double)na;
                        free(h);
                    }
                }
            }
        }
    }
    free(f);
    free(g);
}

/* ---------------- field solve (FP64) ---------------- */

/* Half-complex forward transform of one ring restricted to harmonics
 * m < ng (the only ones p2conv reads); other slots are set to zero. */
static void rfftf_lim(int na, int ng, const double *tw, const double *x,
                      double *y)
{
    for (int i = 0; i < na; i++) y[i] = 0.0;
    for (int m = 0; m < ng; m++) {
        double a = 0.0, b = 0.0;
        for (int j = 0; j < na; j++) {
            int k = (int)(((long)m * j) % na);
            a += x[j] * tw[2 * k];
            b += x[j] * tw[2 * k + 1];
        }
        if (m == 0) y[0] = a;
        else if (2 * m == na) y[na - 1] = a;
        else { y[2 * m - 1] = a; y[2 * m] = -b; }
    }
}

/* Unnormalised backward transform of one ring from harmonics m < ng. */
static void rfftb_lim(int na, int ng, const double *tw, const double *x,
                      double *y)
{
    for (int j = 0; j < na; j++) {
        double s = x[0];
        for (int m = 1; m < ng; m++) {
            int k = (int)(((long)m * j) % na);
            if (2 * m < na)
                s += 2.0 * (x[2 * m - 1] * tw[2 * k] - x[2 * m] * tw[2 * k + 1]);
            else
                s += (j % 2 ? -x[na - 1] : x[na - 1]);
        }
        y[j] = s;
    }
}

/* p2manl + p2conv + p2fsyn.  The forward transform of p2manl is formed for
 * the harmonics m < ng only (p2conv never reads the others); the rings of
 * [jrs, krs] outside [jrf, krf] therefore hold zeros in the inactive slots
 * where the reference leaves all coefficients. */
void orc_fndf(int nr, int na, int ng, const int32_t *lg, const double *G,
              double *rho, int jrs, int krs, int jrf, int krf, int ntypes,
              double *fld)
{
    int mmax = na / 2 + 1;
    int mesh = nr * na;
    size_t nn = (size_t)nr * nr;
    double *row = malloc(sizeof(double) * na);
    double *tw = malloc(sizeof(double) * 2 * na);
    for (int k = 0; k < na; k++) {
        tw[2 * k] = cos2pi(k, na);
        tw[2 * k + 1] = sin2pi(k, na);
    }
    /* p2manl on rings jrs..krs in place */
    for (int ir = jrs; ir <= krs; ir++) {
        double *p = rho + (size_t)(ir - 1) * na;
        rfftf_lim(na, ng, tw, p, row);
        memcpy(p, row, sizeof(double) * na);
    }
    double *w = malloc(sizeof(double) * mesh);
    for (int type = 0; type < ntypes; type++) {
        double *F = fld + (size_t)type * mesh;
        int azim = (type == 1);
        /* blkcpy grdmss -> grdfld over the source rings (p2fndf.f:61) */
        for (int i = (jrs - 1) * na; i < krs * na; i++) F[i] = rho[i];
        /* p2conv: clear the field rings, then sum over sources */
        for (int i = (jrf - 1) * na; i < krf * na; i++) w[i] = 0.0;
        int im1 = azim ? 2 : 1;
        int im2 = azim ? (mmax - 1 < ng ? mmax - 1 : ng) : ng;
        for (int im = im1; im <= im2; im++) {
            int m = im - 1;
            if (lg[m]) continue;
            int pair = (im > 1) && (im < mmax);
            int off = (2 * (im - 1) > 1 ? 2 * (im - 1) : 1) - 1; /* max(1,2(im-1)) - 1 */
            const double *Gm = G + ((size_t)type * ng + m) * nn;
            for (int ir = jrf; ir <= krf; ir++) {
                int k = (ir - 1) * na + off;
                double sc = 0.0, ss = 0.0;
                for (int is = jrs; is <= krs; is++) {
                    double gv = Gm[(size_t)(ir - 1) * nr + (is - 1)];
                    double mc = F[(is - 1) * na + off];
                    double ms = pair ? F[(is - 1) * na + off + 1] : 0.0;
                    if (!azim) {
                        sc += gv * mc;
                        ss += gv * ms;
                    } else {
                        sc += gv * ms;
                        ss -= gv * mc;
                    }
                }
                w[k] += sc;
                if (pair) w[k + 1] += ss;
            }
        }
        for (int i = (jrf - 1) * na; i < krf * na; i++) F[i] = w[i];
        /* p2fsyn on rings jrf..krf, times 1/na */
        for (int ir = jrf; ir <= krf; ir++) {
            double *p = F + (size_t)(ir - 1) * na;
            rfftb_lim(na, ng, tw, p, row);
            for (int j = 0; j < na; j++) p[j] = row[j] / (double)na;
        }
    }
    free(w);
    free(row);
    free(tw);
}

/* polcat (SRC/lib15/polcat.f:26-40), FP64 */
void orc_polcat(int nr, int na, double *fld)
{
    int mesh = nr * na;
    for (int j = 0; j < na; j++) {
        double cp = cos2pi(j, na), sp = sin2pi(j, na);
        for (int i = 0; i < nr; i++) {
            int l = i * na + j;
            double fr = fld[l], fp = fld[mesh + l];
            fld[l] = fr * cp - fp * sp;
            fld[mesh + l] = fr * sp + fp * cp;
        }
    }
}

/* polcat in the reference's FP32 arithmetic: t = 2 pi (j-1)/na in double
 * stored to REAL, cos/sin of the REAL angle */
void orc_polcat_f(int nr, int na, float *fld)
{
    int mesh = nr * na;
    for (int j = 0; j < na; j++) {
        float t = (float)(TWO_PI_D * (double)(float)j / (double)(float)na);
        float cp = cosf(t), sp = sinf(t);
        for (int i = 0; i < nr; i++) {
            int l = i * na + j;
            float fr = fld[l], fp = fld[mesh + l];
            fld[l] = fr * cp - fp * sp;
            fld[mesh + l] = fr * sp + fp * cp;
        }
    }
}

/* Direct circular convolution with the filtered kernel D (p2chkf.f:81-128):
 * for a source at (is, js) and a field point (ir, jf), the angle offset is
 * d = jf - js (mod na); even kernels use D at min(d, na-d), odd kernels are
 * antisymmetric. */
void orc_direct_field(int nr, int na, int type, const double *D,
                      const double *rho, double *out)
{
    int mmax = na / 2 + 1;
    int mesh = nr * na;
    int azim = (type == 1);
    for (int i = 0; i < mesh; i++) out[i] = 0.0;
    for (int s = 0; s < mesh; s++) {
        double q = rho[s];
        if (q == 0.0) continue;
        int is = s / na, js = s % na;
        for (int ir = 0; ir < nr; ir++) {
            const double *d = D + (((size_t)type * nr + is) * nr + ir) * mmax;
            for (int jf = 0; jf < na; jf++) {
                int dd = ((jf - js) % na + na) % na;
                double v;
                if (dd <= na / 2) v = d[dd];
                else v = azim ? -d[na - dd] : d[na - dd];
                out[ir * na + jf] += q * v;
            }
        }
    }
}
