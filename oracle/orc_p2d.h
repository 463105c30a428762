/*
 * orc_p2d.h -- CPU oracle for GALAXY's 2-D polar-grid particle-mesh step.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (galaxy_amd/, include/)
 * may include, link or call this code; only tests/, the smoke check and the
 * CPU-baseline leg of bench.py use it, and only as the checker.
 *
 * The oracle restates the reference algorithm in plain C, once in the
 * reference's own FP32 arithmetic (suffix _f) and once promoted to FP64
 * (suffix _d).  Every function cites the reference routine it follows
 * (paths relative to the reference tree).  Field solves are always computed
 * in double precision (direct DFT sums with exactly reduced twiddles) so the
 * oracle is a high-accuracy reference for both GPU precisions.
 *
 * Conventions shared with the reference:
 *   - ring index ir = 1..nr, spoke it = 1..na; cells are stored ring-major,
 *     cell = (ir-1)*na + (it-1) (0-based here, ncl-1 in the reference);
 *   - half-complex FFT order [a0, a1, b1, ..., a_{na/2}] per ring
 *     (SRC/utils/sfftpack/stest.f:84-103);
 *   - zone z = 1..nzones; the off-grid list is zone nlists = nzones+1.
 */
#ifndef ORC_P2D_H
#define ORC_P2D_H
#include <stdint.h>

#define ORC_MAXZ 10

/* ---- geometry (SRC/lib15/setgrd.f:52-200, grofu.f, uofgr.f) ---- */
float  orc_alpha_f(int na);
double orc_alpha_d(int na);
float  orc_uofgr_f(float r, float alpha);
double orc_uofgr_d(double r, double alpha);
float  orc_grofu_f(float u, float alpha);
double orc_grofu_d(double u, double alpha);

/* ---- weight: cell numbers + CIC weights (SRC/lib15/weight.f:61-97) ----
 * fold != 0 applies the mass-assignment sector fold t = mod(t, thmax). */
void orc_weight_f(int n, const float *x, const float *y, int nr, int na,
                  float alpha, int fold, float thmax,
                  int32_t *cell, float *wt, float *rr);
void orc_weight_d(int n, const double *x, const double *y, int nr, int na,
                  double alpha, int fold, double thmax,
                  int32_t *cell, double *wt, double *rr);

/* ---- massgn: p2d scatter of 4 weights (SRC/lib15/massgn.f:68-89) ----
 * Particles with 1 <= iz <= nzones are assigned to zone iz.  Two sums are
 * formed at once: grdmss (reference order, REAL arithmetic) and mfx
 * (fixed-point int64, weight * 2^qbits, order independent).  jrad/krad are
 * 1-based ring ranges per zone, updated as min/max like the reference. */
void orc_massgn_f(int n, const float *x, const float *y, const int32_t *iz,
                  int nr, int na, float alpha, int nsect, int nzones,
                  float *grdmss, int64_t *mfx, int qbits,
                  int32_t *jrad, int32_t *krad);
void orc_massgn_d(int n, const double *x, const double *y, const int32_t *iz,
                  int nr, int na, double alpha, int nsect, int nzones,
                  double *grdmss, int64_t *mfx, int qbits,
                  int32_t *jrad, int32_t *krad);

/* ---- fixed radial force table (rigid halo, SRC/lib15/getacc.f:115-135) ----
 * Cubic Hermite interpolation of F(r) on r_k = k*dr, k = 0..nfr-1, from
 * values val[k] and derivatives der[k] (internal units). */
double orc_fixforce(double r, int nfr, double dr, const double *val,
                    const double *der);

/* ---- per-particle step: getacc + stpgrp + offgrd + rezgrp + massgn ----
 * (SRC/lib15/step.f:166-234, getacc.f:35-163, grdacc.f:40-66,
 *  stpgrp.f:27-56, addoff.f:30-70, rezgrp.f:30-50, zoneno.f, tsfac.f,
 *  chstep.f:22-31).  fx/fy are the Cartesian grid accelerations after
 *  polcat.  Returns the change in the off-grid count. */
typedef struct {
    int nr, na, nsect, nzones, mzone, offanal, fixforce;
    int nstep[ORC_MAXZ];
    double zbr[ORC_MAXZ];   /* zone inner radii (grid units), zbr[0] = 0 */
    double alpha, rgrid, thmax;
    int nfr; double fr_dr; const double *fr_val; const double *fr_der;
} orc_stepcfg;

int orc_push_f(const orc_stepcfg *c, int n, float *x, float *y, float *vx,
               float *vy, int32_t *iz, const float *fx, const float *fy,
               float *grdmss, int64_t *mfx, int qbits,
               int32_t *jrad, int32_t *krad, uint8_t *moved);
int orc_push_d(const orc_stepcfg *c, int n, double *x, double *y, double *vx,
               double *vy, int32_t *iz, const double *fx, const double *fy,
               double *grdmss, int64_t *mfx, int qbits,
               int32_t *jrad, int32_t *krad, uint8_t *moved);
/* The particle loop on every host thread: fixed-point sums only, equal to
 * orc_push_*'s (integer sums commute); no order-dependent REAL sums. */
int orc_push_par_f(const orc_stepcfg *c, int n, float *x, float *y, float *vx,
                   float *vy, int32_t *iz, const float *fx, const float *fy,
                   int64_t *mfx, int qbits, int32_t *jrad, int32_t *krad,
                   uint8_t *moved);
int orc_push_par_d(const orc_stepcfg *c, int n, double *x, double *y, double *vx,
                   double *vy, int32_t *iz, const double *fx, const double *fy,
                   int64_t *mfx, int qbits, int32_t *jrad, int32_t *krad,
                   uint8_t *moved);
void orc_tmcent_f(const orc_stepcfg *c, int n, const float *x, const float *y,
                  float *vx, float *vy, const int32_t *iz, const float *fx,
                  const float *fy, int start);
void orc_tmcent_d(const orc_stepcfg *c, int n, const double *x, const double *y,
                  double *vx, double *vy, const int32_t *iz, const double *fx,
                  const double *fy, int start);

/* ---- FFTPACK-convention transforms, direct sums in double ----
 * rfftf: forward real DFT, half-complex output (stest.f:84-103)
 * rfftb: unnormalised backward (stest.f:106-122)
 * cost : DCT-I of n points (rcost, stest.f:169-195)
 * sint : DST-I of n points (rsint, stest.f:138-165) */
void orc_rfftf(int n, const double *x, double *y);
void orc_rfftb(int n, const double *x, double *y);
void orc_cost(int n, const double *x, double *y);
void orc_sint(int n, const double *x, double *y);

/* ---- Green's tables (SRC/lib15/p2grnm.f:98-203, p2gwrt.f:27-56) ----
 * G[type][m][ir][is], type 0 = F_R (radfor, DCT), 1 = F_phi (azifor, DST),
 * 2 = potential (grnfun, DCT); m = 0..ng-1 (type 1 uses m = 1..
 * min(mmax-2, ng-1), other rows zero).  Harmonics m >= ng or m mod nsect
 * != 0 are zero.  D (optional) holds the filtered, re-synthesised direct
 * kernels D[type][is][ir][it], it = 0..mmax-1 (angles 0..pi).  tsoft 1 =
 * Plummer, 2 = cubic spline. */
void orc_green(int nr, int na, int ng, int nsect, double alpha, double softl,
               int tsoft, double *G, double *D);

/* ---- field solve (SRC/lib15/p2fndf.f:26-66, p2manl.f, p2conv.f:95-186,
 *      p2fsyn.f:40-52, polcat.f:26-40) ----
 * rho: combined, scaled mass (mesh), overwritten by its transform on rings
 * jrs..krs exactly as p2manl does.  fld[3][mesh] in/out: rings jrf..krf of
 * types 0..ntypes-1 are replaced by the solved field; rings of [jrs,krs]
 * outside [jrf,krf] hold copied coefficients, as in the reference.
 * lg[m] != 0 skips harmonic m.  polcat is applied by orc_polcat. */
void orc_fndf(int nr, int na, int ng, const int32_t *lg, const double *G,
              double *rho, int jrs, int krs, int jrf, int krf, int ntypes,
              double *fld);
void orc_polcat(int nr, int na, double *fld);
void orc_polcat_f(int nr, int na, float *fld);

/* ---- direct convolution with the filtered kernel (p2chkf.f:81-128) ----
 * out[mesh] = sum over non-zero source cells of D-kernel * rho. */
void orc_direct_field(int nr, int na, int type, const double *D,
                      const double *rho, double *out);

#endif
