/*
 * orc_p3d.c -- CPU oracle for GALAXY's 3-D polar-grid (p3d) PM step.
 *
 * TEST INFRASTRUCTURE ONLY (see orc_p3d.h).  Built by oracle/Makefile with
 * gcc -O2 -ffp-contract=off into oracle/liboracle.so together with
 * orc_p2d.c; loaded by oracle/oracle3.py.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "orc_p2d.h"
#include "orc_p3d.h"

/* 2.d0 * pi of SRC/lib15/inc/pi.f */
#define TWO_PI_D (2.0 * 3.141592653589793238462643)

/* Corner cells of a 0-based p3d cell in grdacc's order
 * (SRC/lib15/grdacc.f:193-203): c, c+1 (wrapped when c+1 is a multiple of
 * na), c+na (next plane), c+1+na, then the same four one ring out
 * (+ ngz na).  Weights w1..w8 of weight() belong to these corners. */
static inline void corners3(int32_t c, int na, int ngz, int32_t ic[8])
{
    int wrap = ((c + 1) % na) == 0;
    ic[0] = c;
    ic[1] = wrap ? c + 1 - na : c + 1;
    ic[2] = c + na;
    ic[3] = ic[1] + na;
    ic[4] = c + ngz * na;
    ic[5] = wrap ? ic[4] + 1 - na : ic[4] + 1;
    ic[6] = ic[4] + na;
    ic[7] = ic[5] + na;
}

/* The 27 cells of a quadratic-interpolation particle about its central
 * cell c (massgn.f:182-195, grdacc.f:262-275): spoke offsets -1, 0, +1
 * fastest (with the azimuthal wrap), then plane -1, 0, +1 (+-na), then ring
 * -1, 0, +1 (+-ngz na), as flat indices. */
static inline void corners27(int32_t c, int na, int ngz, int32_t ic[27])
{
    int32_t m[3];
    m[1] = c;
    m[2] = c + 1;
    m[0] = c - 1;
    if ((c + 1) % na == 0) m[2] -= na;
    if (c % na == 0) m[0] += na;
    for (int k = 0; k < 3; k++)
        for (int j = 0; j < 3; j++)
            for (int i = 0; i < 3; i++)
                ic[9 * k + 3 * j + i] = m[i] + (j - 1) * na + (k - 1) * ngz * na;
}

#define REAL float
#define SFX(name) name##_f
#define RSQRT sqrtf
#define RFMOD fmodf
#include "orc_p3d_particles.inc"
#undef REAL
#undef SFX
#undef RSQRT
#undef RFMOD

#define REAL double
#define SFX(name) name##_d
#define RSQRT sqrt
#define RFMOD fmod
#include "orc_p3d_particles.inc"
#undef REAL
#undef SFX
#undef RSQRT
#undef RFMOD

/* ---------------- transforms: direct sums, exact twiddle reduction ------- */

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

/* first nout outputs of the DCT-I of n points (rcost, stest.f:169-195) */
static void cost_lim(int n, const double *x, double *y, int nout)
{
    int m = n - 1;
    for (int i = 0; i < nout; i++) {
        double s = x[0] + (i % 2 ? -x[m] : x[m]);
        for (int k = 1; k < m; k++) s += 2.0 * x[k] * cos2pi((long)k * i, 2 * m);
        y[i] = s;
    }
}

/* first nout outputs of the DST-I of n points (rsint, stest.f:138-165) */
static void sint_lim(int n, const double *x, double *y, int nout)
{
    for (int i = 1; i <= nout; i++) {
        double s = 0.0;
        for (int k = 1; k <= n; k++) s += 2.0 * x[k - 1] * sin2pi((long)k * i, 2 * (n + 1));
        y[i - 1] = s;
    }
}

/* ---------------- Green's function kernels with height zf (FP64) --------- */

static double softm3(double ds)
{
    if (ds <= 0.0) return 0.0;
    if (ds < 1.0)
        return 4.0 * ds * ds * ds * (1.0 / 3.0 - 0.3 * ds * ds + 0.125 * ds * ds * ds);
    if (ds < 2.0)
        return -1.0 / 15.0 + ds * ds * ds *
               (8.0 / 3.0 - 3.0 * ds + 1.2 * ds * ds - ds * ds * ds / 6.0);
    return 1.0;
}

/* radfor, azifor, vrtfor, grnfun -> sftpot (SRC/lib15/radfor.f:25-38,
 * azifor.f:25-40, vrtfor.f:25-40, grnfun.f, sftpot.f:20-40) */
static double kern3(int type, double rs, double rf, double th, double zf,
                    double eps, int tsoft)
{
    double cth = cos(th);
    double d2 = rs * rs + rf * rf - 2.0 * rs * rf * cth + zf * zf;
    if (type == 3) {
        if (tsoft == 1) return -1.0 / sqrt(d2 + eps * eps);
        double x = sqrt(d2) / eps;
        if (x <= 0.0) return -1.4 / eps;
        if (x < 1.0) return (-1.4 + x * x * (2.0 / 3.0 - 0.3 * x * x + 0.1 * x * x * x)) / eps;
        if (x < 2.0)
            return (-1.6 + 1.0 / (15.0 * x) +
                    x * x * (4.0 / 3.0 - x + 0.3 * x * x - x * x * x / 30.0)) / eps;
        return -1.0 / sqrt(d2);
    }
    double num = type == 0 ? rs * cth - rf : type == 1 ? -rs * sin(th) : -zf;
    if (tsoft == 1) return num / pow(d2 + eps * eps, 1.5);
    if (type == 2 ? zf != 0.0 : d2 > 0.0) return num * softm3(sqrt(d2) / eps) / pow(d2, 1.5);
    return 0.0;
}

void orc3_green(int nr, int na, int ngz, int ng, const int32_t *lg, double alpha,
                double dzg, double softl, int tsoft, double *G, double *D)
{
    const int mmax = na / 2 + 1;
    int resyn = ng < mmax;
    for (int m = 0; m < ng; m++) resyn = resyn || lg[m];
    const size_t nn = (size_t)nr * nr;
    const size_t gtype = (size_t)(ngz + 1) * ng * nn;
    memset(G, 0, sizeof(double) * 4 * gtype);
    if (D) memset(D, 0, sizeof(double) * 4 * nn * (ngz + 1) * mmax);
    double *f = malloc(sizeof(double) * mmax);
    double *g = malloc(sizeof(double) * mmax);
    double *h = malloc(sizeof(double) * mmax);
    double *col = malloc(sizeof(double) * (ngz + 1));
    double *zo = malloc(sizeof(double) * (ngz + 1));
    /* per (is, ir): angular transforms of every plane, then the vertical */
    double *T = malloc(sizeof(double) * (ngz + 1) * mmax);
    for (int type = 0; type < 4; type++) {
        if (type == 1 && ng == 1) continue;
        const int azim = type == 1, vert = type == 2;
        const int it1 = azim ? 2 : 1, it2 = azim ? mmax - 1 : mmax;
        const int npt = it2 - it1 + 1;
        const int iz1 = vert ? 2 : 1, iz2 = vert ? ngz : ngz + 1;
        const int nzp = iz2 - iz1 + 1;
        /* harmonics stored: m = 0..ng-1, or 1..min(ng, mmax-1)-1 */
        const int nm = azim ? (ng < mmax - 1 ? ng : mmax - 1) - 1 : ng;
        for (int is = 1; is <= nr; is++) {
            double rs = exp(alpha * (double)(is - 1)) - 1.0;
            for (int ir = 1; ir <= nr; ir++) {
                double rf = exp(alpha * (double)(ir - 1)) - 1.0;
                for (int iz = iz1; iz <= iz2; iz++) {
                    double zf = dzg * (double)(iz - 1);
                    for (int it = it1; it <= it2; it++)
                        f[it - it1] = kern3(type, rs, rf, alpha * (double)(it - 1), zf,
                                            softl, tsoft);
                    int nout = D ? npt : (nm < npt ? nm : npt);
                    for (int q = nout; q < npt; q++) g[q] = 0.0;
                    if (azim) sint_lim(npt, f, g, nout);
                    else cost_lim(npt, f, g, nout);
                    /* filter (p3grnm.f:156-166): harmonic it-1 >= ng or lg */
                    for (int it = it1; it <= it2; it++)
                        if (it > ng || lg[it - 1]) g[it - it1] = 0.0;
                    for (int q = 0; q < npt; q++) T[(size_t)(iz - iz1) * mmax + q] = g[q];
                    if (D) {
                        double *d = D + ((((size_t)type * nr + (is - 1)) * nr + (ir - 1)) *
                                         (ngz + 1) + (iz - 1)) * mmax;
                        if (!resyn) {
                            for (int it = it1; it <= it2; it++) d[it - 1] = f[it - it1];
                        } else {
                            if (azim) sint_lim(npt, g, h, npt);
                            else cost_lim(npt, g, h, npt);
                            for (int it = it1; it <= it2; it++)
                                d[it - 1] = h[it - it1] / (double)na;
                        }
                    }
                }
                /* vertical transform of each kept harmonic (p3grnm.f:209-243) */
                for (int q = 0; q < nm; q++) {
                    int m = azim ? q + 1 : q;
                    for (int p = 0; p < nzp; p++) col[p] = T[(size_t)p * mmax + q];
                    if (vert) sint_lim(nzp, col, zo, nzp);
                    else cost_lim(nzp, col, zo, nzp);
                    for (int p = 0; p < nzp; p++) {
                        int kz = vert ? p + 1 : p;
                        G[(size_t)type * gtype + ((size_t)kz * ng + m) * nn +
                          (size_t)(is - 1) * nr + (ir - 1)] = zo[p];
                    }
                }
            }
        }
    }
    free(f); free(g); free(h); free(col); free(zo); free(T);
}

/* ---------------- field solve ---------------- */

/* The compact component list of p3manl (p3manl.f:47-58,74-89): active
 * harmonics in order, one slot for m = 0 and m = na/2, two (cos, sin) for
 * the rest.  slot[m] = first slot of harmonic m or -1. */
static int slots3(int na, int ng, const int32_t *lg, int *slot)
{
    int mmax = na / 2 + 1, nm = 0;
    for (int m = 0; m < mmax; m++) slot[m] = -1;
    for (int im = 1; im <= ng; im++) {
        if (lg[im - 1]) continue;
        slot[im - 1] = nm;
        nm += (im == 1 || im == mmax) ? 1 : 2;
    }
    return nm;
}

void orc3_fndf(int nr, int na, int ngz, int ng, const int32_t *lg,
               const double *G, const double *rho, int jrs, int krs, int jrf,
               int krf, int ntypes, double *fld)
{
    const int mmax = na / 2 + 1, P = 2 * ngz;
    const size_t mesh = (size_t)nr * ngz * na, nn = (size_t)nr * nr;
    const size_t gtype = (size_t)(ngz + 1) * ng * nn;
    int *slot = malloc(sizeof(int) * mmax);
    int *sm = malloc(sizeof(int) * 2 * (na + 2));   /* slot -> harmonic, part */
    const int nm = slots3(na, ng, lg, slot);
    for (int m = 0; m < mmax; m++) {
        if (slot[m] < 0) continue;
        sm[2 * slot[m]] = m; sm[2 * slot[m] + 1] = 0;
        if (m > 0 && m < mmax - 1) { sm[2 * slot[m] + 2] = m; sm[2 * slot[m] + 3] = 1; }
    }
    /* p3manl: azimuthal half-complex components, then the doubled vertical
     * transform of each component column: M[ring][slot][P] */
    /* cos, sin of 2 pi k / na and of 2 pi k / P, k reduced exactly */
    double *ca = malloc(sizeof(double) * na), *sa = malloc(sizeof(double) * na);
    double *cz = malloc(sizeof(double) * P), *sz = malloc(sizeof(double) * P);
    for (int k = 0; k < na; k++) { ca[k] = cos2pi(k, na); sa[k] = sin2pi(k, na); }
    for (int k = 0; k < P; k++) { cz[k] = cos2pi(k, P); sz[k] = sin2pi(k, P); }
    double *M = calloc((size_t)nr * nm * P, sizeof(double));
    double *col = malloc(sizeof(double) * ngz);
    for (int ir = jrs; ir <= krs; ir++) {
        for (int s = 0; s < nm; s++) {
            const int m = sm[2 * s], part = sm[2 * s + 1];
            for (int p = 0; p < ngz; p++) {
                const double *row = rho + ((size_t)(ir - 1) * ngz + p) * na;
                double acc = 0.0;
                for (int j = 0; j < na; j++) {
                    int k = (int)(((long)m * j) % na);
                    acc += row[j] * (part ? -sa[k] : ca[k]);
                }
                col[p] = acc;
            }
            double *H = M + ((size_t)(ir - 1) * nm + s) * P;
            for (int k = 0; k <= ngz; k++) {
                double a = 0.0, b = 0.0;
                for (int p = 0; p < ngz; p++) {
                    int kk = (int)(((long)k * p) % P);
                    a += col[p] * cz[kk];
                    b += col[p] * sz[kk];
                }
                if (k == 0) H[0] = a;
                else if (k == ngz) H[P - 1] = a;
                else { H[2 * k - 1] = a; H[2 * k] = -b; }
            }
        }
    }
    double *W = malloc(sizeof(double) * (size_t)nr * nm * P);
    double *zc = malloc(sizeof(double) * (size_t)nm * ngz);
    for (int type = 0; type < ntypes; type++) {
        double *F = fld + (size_t)type * mesh;
        if (type == 1 && ng == 1) {
            /* p3fndf.f:77-80 */
            for (size_t i = 0; i < mesh; i++) F[i] = 0.0;
            continue;
        }
        const int jm = type == 1 ? 2 : 1;
        const int km = type == 1 ? (ng < mmax - 1 ? ng : mmax - 1) : ng;
        const int jz = type == 2 ? 2 : 1, kz = type == 2 ? ngz : ngz + 1;
        memset(W, 0, sizeof(double) * (size_t)nr * nm * P);
        for (int iz = jz; iz <= kz; iz++) {
            const int nvskip = iz > 1 && iz <= ngz;
            const int p = (2 * (iz - 1) > 1 ? 2 * (iz - 1) : 1) - 1;
            for (int im = jm; im <= km; im++) {
                if (lg[im - 1]) continue;
                const int naskip = im > 1 && im < mmax;
                const int sc = slot[im - 1], ss = sc + 1;
                const double *Gm = G + (size_t)type * gtype + ((size_t)(iz - 1) * ng + (im - 1)) * nn;
                for (int is = jrs; is <= krs; is++) {
                    const double *Ms = M + (size_t)(is - 1) * nm * P;
                    for (int ir = jrf; ir <= krf; ir++) {
                        const double g = Gm[(size_t)(is - 1) * nr + (ir - 1)];
                        double *Wf = W + (size_t)(ir - 1) * nm * P;
                        if (type == 0 || type == 3) {
                            Wf[sc * P + p] += Ms[sc * P + p] * g;
                            if (naskip) Wf[ss * P + p] += Ms[ss * P + p] * g;
                            if (nvskip) {
                                Wf[sc * P + p + 1] += Ms[sc * P + p + 1] * g;
                                if (naskip) Wf[ss * P + p + 1] += Ms[ss * P + p + 1] * g;
                            }
                        } else if (type == 1) {
                            Wf[sc * P + p] += Ms[ss * P + p] * g;
                            Wf[ss * P + p] -= Ms[sc * P + p] * g;
                            if (nvskip) {
                                Wf[sc * P + p + 1] += Ms[ss * P + p + 1] * g;
                                Wf[ss * P + p + 1] -= Ms[sc * P + p + 1] * g;
                            }
                        } else {
                            Wf[sc * P + p] += Ms[sc * P + p + 1] * g;
                            Wf[sc * P + p + 1] -= Ms[sc * P + p] * g;
                            if (naskip) {
                                Wf[ss * P + p] += Ms[ss * P + p + 1] * g;
                                Wf[ss * P + p + 1] -= Ms[ss * P + p] * g;
                            }
                        }
                    }
                }
            }
        }
        /* p3fsyn: inverse vertical transform (first ngz planes), inverse
         * azimuthal transform of the active components, 1/(2 ngz na) */
        const double scale = 1.0 / ((double)P * (double)na);
        for (int ir = jrf; ir <= krf; ir++) {
            for (int s = 0; s < nm; s++) {
                const double *H = W + ((size_t)(ir - 1) * nm + s) * P;
                for (int q = 0; q < ngz; q++) {
                    double v = H[0] + (q % 2 ? -H[P - 1] : H[P - 1]);
                    for (int k = 1; k < ngz; k++) {
                        int kk = (int)(((long)k * q) % P);
                        v += 2.0 * (H[2 * k - 1] * cz[kk] - H[2 * k] * sz[kk]);
                    }
                    zc[(size_t)s * ngz + q] = v;
                }
            }
            for (int q = 0; q < ngz; q++) {
                double *out = F + ((size_t)(ir - 1) * ngz + q) * na;
                for (int j = 0; j < na; j++) {
                    double v = 0.0;
                    for (int s = 0; s < nm; s++) {
                        const int m = sm[2 * s], part = sm[2 * s + 1];
                        const double c = zc[(size_t)s * ngz + q];
                        if (m == 0) v += c;
                        else if (m == mmax - 1) v += (j % 2 ? -c : c);
                        else if (part == 0) v += 2.0 * c * ca[((long)m * j) % na];
                        else v -= 2.0 * c * sa[((long)m * j) % na];
                    }
                    out[j] = v * scale;
                }
            }
        }
    }
    free(slot); free(sm); free(M); free(col); free(W); free(zc);
    free(ca); free(sa); free(cz); free(sz);
}

void orc3_direct_field(int nr, int na, int ngz, int type, const double *D,
                       const double *rho, double *out)
{
    const int mmax = na / 2 + 1;
    const size_t mesh = (size_t)nr * ngz * na;
    for (size_t i = 0; i < mesh; i++) out[i] = 0.0;
    for (size_t s = 0; s < mesh; s++) {
        const double q = rho[s];
        if (q == 0.0) continue;
        const int js = (int)(s % na), izs = (int)((s / na) % ngz), is = (int)(s / ((size_t)na * ngz));
        for (int ir = 0; ir < nr; ir++) {
            for (int izf = 0; izf < ngz; izf++) {
                int dz = izf - izs;
                double sv = (type == 2 && dz < 0) ? -1.0 : 1.0;
                if (dz < 0) dz = -dz;
                const double *d = D + ((((size_t)type * nr + is) * nr + ir) * (ngz + 1) + dz) * mmax;
                double *o = out + ((size_t)ir * ngz + izf) * na;
                for (int jf = 0; jf < na; jf++) {
                    int dd = ((jf - js) % na + na) % na;
                    double v;
                    if (dd <= na / 2) v = d[dd];
                    else v = type == 1 ? -d[na - dd] : d[na - dd];
                    o[jf] += q * sv * v;
                }
            }
        }
    }
}
