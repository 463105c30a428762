/*
 * orc_p3d.h -- CPU oracle for GALAXY's 3-D polar-grid (p3d) particle-mesh
 * step: cylindrical rings x spokes x z-planes, azimuthal FFT plus a vertical
 * FFT of doubled length, per-(k_z, m) radial Green's contraction.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (galaxy_amd/, include/)
 * may include, link or call this code; only tests/, the smoke check and the
 * CPU-baseline leg of bench.py use it, and only as the checker.
 *
 * Conventions (reference paths relative to its tree):
 *   - mesh cell (ir, iz, it), all 1-based in the reference, is stored at
 *     ((ir-1)*ngz + (iz-1))*na + (it-1) (SRC/lib15/weight.f:232,
 *     p3fsyn.f:120-126);
 *   - the vertical grid coordinate is (z + zm)/dzg with
 *     zm = dzg (ngz + 1 - jmass)/2 (SRC/lib15/setgrd.f:146-150), so the
 *     planes sit at z = -zm + dzg (iz - 1) for jmass = 2;
 *   - field types 0..3 = F_R, F_phi, F_z, potential (the reference's itype
 *     1..4); after polcat 0 and 1 hold F_x, F_y.
 */
#ifndef ORC_P3D_H
#define ORC_P3D_H
#include <stdint.h>

#define ORC3_MAXZ 10

/* jmass: 2 = linear (8 corners, weight.f:195-250), 3 = quadratic (27
 * corners, weight.f:251-300); ORC3_NW(j) weights per particle. */
#define ORC3_NW(j) ((j) == 3 ? 27 : 8)

typedef struct {
    int nr, na, ngz, nsect, nzones, mzone, offanal, fixforce, jmass;
    int nstep[ORC3_MAXZ];
    double zbr[ORC3_MAXZ];   /* zone inner radii (spherical, grid units) */
    double alpha, rgrid, thmax, zm, dzg;
    int nfr; double fr_dr; const double *fr_val; const double *fr_der;
} orc3_cfg;

/* ---- weight, p3d linear interpolation (SRC/lib15/weight.f:195-250) ----
 * Returns 0-based cells (ncl - 1; for jmass 3 the central cell of the 27),
 * ORC3_NW(jmass) weights per particle in the reference's order (spoke
 * fastest, then plane, then ring), cylindrical radii. */
void orc3_weight_f(const orc3_cfg *c, int n, const float *x, const float *y,
                   const float *z, int fold, int32_t *cell, float *wt, float *rr);
void orc3_weight_d(const orc3_cfg *c, int n, const double *x, const double *y,
                   const double *z, int fold, int32_t *cell, double *wt, double *rr);

/* ---- massgn p3d (SRC/lib15/massgn.f:140-172) ----
 * grdmss: [nzones][mesh] sums in the reference's order and arithmetic;
 * mfx: the same as fixed-point int64 (weight * 2^qbits). */
void orc3_massgn_f(const orc3_cfg *c, int n, const float *x, const float *y,
                   const float *z, const int32_t *iz, float *grdmss, int64_t *mfx,
                   int qbits, int32_t *jrad, int32_t *krad);
void orc3_massgn_d(const orc3_cfg *c, int n, const double *x, const double *y,
                   const double *z, const int32_t *iz, double *grdmss, int64_t *mfx,
                   int qbits, int32_t *jrad, int32_t *krad);

/* ---- grdacc p3d (SRC/lib15/grdacc.f:189-211) + getacc's fixed spherical
 * force and zone rescaling (getacc.f:106-151), for particles of list iz
 * (no acceleration on the off-grid list nzones+1).  fld: [3][mesh]
 * Cartesian components after polcat. ---- */
void orc3_getacc_f(const orc3_cfg *c, int n, const float *x, const float *y,
                   const float *z, const int32_t *iz, const float *fld, float *acc);
void orc3_getacc_d(const orc3_cfg *c, int n, const double *x, const double *y,
                   const double *z, const int32_t *iz, const double *fld, double *acc);

/* ---- the particle loop of step (SRC/lib15/step.f:166-234) for p3d:
 * getacc, stpgrp, offgrd (|z| > zm or R > rgrid, offgrd.f:69-72), addoff,
 * rezgrp with spherical zone radii (zoneno.f:40-41, setrun.f:810), chstep
 * on three components, massgn.  Returns the change in the off-grid count.
 * ---- */
int orc3_push_f(const orc3_cfg *c, int n, float *x, float *y, float *z,
                float *vx, float *vy, float *vz, int32_t *iz, const float *fld,
                float *grdmss, int64_t *mfx, int qbits, int32_t *jrad,
                int32_t *krad, uint8_t *moved);
int orc3_push_d(const orc3_cfg *c, int n, double *x, double *y, double *z,
                double *vx, double *vy, double *vz, int32_t *iz, const double *fld,
                double *grdmss, int64_t *mfx, int qbits, int32_t *jrad,
                int32_t *krad, uint8_t *moved);
void orc3_tmcent_f(const orc3_cfg *c, int n, const float *x, const float *y,
                   const float *z, float *vx, float *vy, float *vz,
                   const int32_t *iz, const float *fld, int start);
void orc3_tmcent_d(const orc3_cfg *c, int n, const double *x, const double *y,
                   const double *z, double *vx, double *vy, double *vz,
                   const int32_t *iz, const double *fld, int start);

/* ---- Green's tables (SRC/lib15/p3grnm.f:93-263, p3gwrt.f:27-66), FP64 ----
 * G[type][kz][m][is][ir] (field ring fastest), type 0..3 = radfor, azifor,
 * vrtfor, grnfun with the field point at height zf = dzg (iz - 1);
 * kz = 0..ngz (type 2: 1..ngz-1), m = 0..ng-1 (type 1: 1..min(ng,mmax-1)-1);
 * other entries zero.  lg[m] != 0 filters harmonic m.  D (optional): the
 * direct kernels for the convolution check, D[type][is][ir][iz][it],
 * iz = 0..ngz, it = 0..mmax-1, filtered and re-synthesised in angle when
 * any harmonic is filtered. */
void orc3_green(int nr, int na, int ngz, int ng, const int32_t *lg, double alpha,
                double dzg, double softl, int tsoft, double *G, double *D);

/* ---- field solve (p3fndf.f:37-104, p3manl.f:46-104, p3conv.f:39-241,
 * p3fsyn.f:39-129), FP64 direct sums in the reference's conventions.
 * rho: combined scaled mass [mesh].  fld: [4][mesh]; rings jrf..krf of
 * types 0..ntypes-1 are replaced (polar components, no polcat). ---- */
void orc3_fndf(int nr, int na, int ngz, int ng, const int32_t *lg,
               const double *G, const double *rho, int jrs, int krs, int jrf,
               int krf, int ntypes, double *fld);

/* ---- direct convolution with the kernels D (p3chkf.f:95-163) ----
 * out[mesh] = sum over source cells of rho * D with the angle and height
 * offsets of field and source (odd in angle for type 1, in height for
 * type 2). */
void orc3_direct_field(int nr, int na, int ngz, int type, const double *D,
                       const double *rho, double *out);

#endif
