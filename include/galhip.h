/*
 * galhip.h -- C ABI of libgalhip.so, the MI355X (gfx950) implementation of
 * GALAXY's polar-grid particle-mesh step: the 2-D grid (p2d) and the 3-D
 * cylindrical grid (p3d) with vertical planes.
 *
 * Every entry point takes plain scalars and pointers and returns 0 on
 * success or a negative GALHIP_E* code; galhip_last_error() returns the
 * message of the last failure on the calling thread.  The reference has no
 * return codes (its crash() STOPs, SRC/lib15/crash.f:23-37); a caller that
 * wants the reference's fatal semantics stops on any non-zero return.
 *
 * The reference keeps its state in COMMON blocks and module arrays
 * (SRC/lib15/inc, SRC/lib15/aarrays.f); here that state lives in an
 * opaque context owned by the library, with device-resident particles and
 * grids.  Host buffers cross the boundary only at load/unload and in the
 * parity entry points.  All work of one context goes on one HIP stream in
 * submission order; a context must be used from one host thread at a time
 * (the reference is single-threaded too).
 *
 * Array conventions (as the reference's): rings ir = 1..nr, spokes
 * it = 1..na, cell index (ir-1)*na + (it-1), for p3d with planes
 * iz = 1..ngz ((ir-1)*ngz + (iz-1))*na + (it-1); zones 1..nzones, zone
 * nzones+1 = the off-grid list.  Coordinates are grid units, velocities are
 * displacements per step and accelerations velocity changes per step
 * (internal units, SRC/lib15/setrun.f:858-873).  Floating point buffers are
 * float for a GALHIP_FP32 context and double for a GALHIP_FP64 context.
 */
#ifndef GALHIP_H
#define GALHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GALHIP_FP32 4
#define GALHIP_FP64 8
#define GALHIP_MAXZONES 10
#define GALHIP_MAXSKIP 16

#define GALHIP_OK 0
#define GALHIP_EINVAL (-1)   /* invalid argument or state */
#define GALHIP_EHIP (-2)     /* HIP runtime failure */
#define GALHIP_ERANGE (-3)   /* reference invariant violated (crash() cases) */

typedef struct galhip_ctx galhip_ctx;

/* Grid, filter, zone and scale parameters: the values the reference reads
 * from its .dat cards into /gridsi/, /gridsr/, /filter/, /zones/ and
 * /scalesr/ (p2dset.f:24-80, filset.f:42-141, setrun.f:200-220,770-873). */
typedef struct {
    int32_t nr, na;            /* GRID card: rings, spokes (na forced even) */
    int32_t hash;              /* HASH card: highest active harmonic */
    int32_t nsect;             /* SECT card: sectors (1 = whole plane) */
    int32_t nskip;             /* SKIP card: number of harmonics listed */
    int32_t skip[GALHIP_MAXSKIP];
    double softl;              /* SOFT card: softening length, grid units */
    int32_t tsoft;             /* 1 Plummer, 2 cubic spline */
    int32_t nzones;            /* ZONE cards */
    int32_t nstep[GALHIP_MAXZONES];  /* step multiple of zone i (nstep[0] = 1) */
    double zbr[GALHIP_MAXZONES];     /* inner radius of zone i, grid units */
    int32_t offanal;           /* OFFG card, first flag: move off-grid particles */
    double pmass;              /* particle mass, internal units (setrun.f:866) */
    int32_t precision;         /* GALHIP_FP32 or GALHIP_FP64 */
} galhip_p2d_params;

/* p2dset + setgrd + filset + setspc: create a context for one p2d grid.
 * Replaces SRC/lib15/p2dset.f:1, SRC/lib15/setgrd.f:1, SRC/lib15/filset.f:1,
 * SRC/lib15/setspc.f:1 (grid and zone set-up; particle storage is sized by
 * galhip_restrt). */
int galhip_p2dset(const galhip_p2d_params *p, galhip_ctx **out);
int galhip_free(galhip_ctx *ctx);
const char *galhip_last_error(void);

/* Geometry constants as the reference computes them (setgrd.f:52-200):
 * out[0] = alpha, out[1] = rgrid, out[2] = thmax, out[3] = ng, out[4] = mmax,
 * out[5] = mesh. */
int galhip_geometry(const galhip_ctx *ctx, double *out);

/* greenm -> p2grnm: build the Fourier-transformed Green's tables on the GPU
 * in FP64 (SRC/lib15/greenm.f:1, SRC/lib15/p2grnm.f:1).  The tables are kept
 * on the device; nothing is written to disk. */
int galhip_greenm(galhip_ctx *ctx);
/* Copy the tables in or out as G[3][ng][nr][nr] doubles (type 0 = F_R,
 * 1 = F_phi, 2 = potential; [field ring][source ring]), the content of the
 * reference's .grt records (SRC/lib15/p2gwrt.f:1). */
int galhip_get_green(const galhip_ctx *ctx, double *G);
int galhip_set_green(galhip_ctx *ctx, const double *G);

/* Tabulated fixed radial force F(r) (internal units), cubic Hermite between
 * r_k = k*dr, k = 0..n-1, from values and derivatives.  Stands in for the
 * rigid-halo / axisymmetric force functions halfrc and axifrc
 * (SRC/lib15/halfrc.f:1, SRC/lib15/axifrc.f:1) called from getacc
 * (SRC/lib15/getacc.f:115-135).  n = 0 removes the force. */
int galhip_set_fixforce(galhip_ctx *ctx, int32_t n, double dr,
                        const double *val, const double *der);

/* restrt (particle part, SRC/lib15/restrt.f:1): load n particles from host
 * arrays; iz[i] = zone (1..nzones) or nzones+1 if off the grid.  Particles
 * are stored on the device as structure of arrays, ordered by cell. */
int galhip_restrt(galhip_ctx *ctx, int64_t n, const void *x, const void *y,
                  const void *vx, const void *vy, const int32_t *iz);
/* unload (SRC/lib15/unload.f:1): copy the particles back to host arrays in
 * their load order.  Any pointer may be NULL. */
int galhip_unload(galhip_ctx *ctx, void *x, void *y, void *vx, void *vy,
                  int32_t *iz);

/* Step counter and zone time stamps (/admin/ istep, /point/ lstep,
 * SRC/lib15/inc/admin.f:36-37).  lstep is [2][nzones] with zone 1 unused;
 * either pointer may be NULL on get.  Setting lstep also forms zoncmb's
 * zone weights for the new time (SRC/lib15/zoncmb.f:40-125), failing as
 * zoncmb does when a zone's mass array is at a useless time. */
int galhip_set_step(galhip_ctx *ctx, int32_t istep, const int32_t *lstep);
int galhip_get_step(const galhip_ctx *ctx, int32_t *istep, int32_t *mzone,
                    int32_t *lstep);

/* masset: assign every on-grid particle, then combine zones and scale
 * (SRC/lib15/masset.f:1). */
int galhip_masset(galhip_ctx *ctx);

/* findf(lpot) -> p2fndf -> polcat: solve for the Cartesian accelerations
 * (and the potential when lpot != 0) over the reference's radial ranges
 * (SRC/lib15/findf.f:1, SRC/lib15/p2fndf.f:1, SRC/lib15/polcat.f:1). */
int galhip_findf(galhip_ctx *ctx, int32_t lpot);
/* p2fndf(jzone, kzone) with an explicit zone range, then polcat
 * (SRC/lib15/p2fndf.f:1). */
int galhip_p2fndf(galhip_ctx *ctx, int32_t jzone, int32_t kzone, int32_t lpot);

/* step: advance the particles of the zones due at this step, assign their
 * mass, combine zones and scale (SRC/lib15/step.f:1 with masscl, getacc,
 * grdacc, stpgrp, rezgrp, massgn, mascmb/zoncmb and scaled). */
int galhip_step(galhip_ctx *ctx);

/* tmcent(start) for the retarded-velocity leap-frog (SRC/lib15/tmcent.f:1,
 * tmcgrp.f:26-36): every particle's velocity minus (start != 0) or plus
 * half its getacc acceleration from the current field, turning loaded
 * velocities into the half-step-back ones the leap-frog carries (as pcs2dmp
 * does before the first step) or back.  Needs a field (masset + findf);
 * calling it twice in the same direction is the reference's logic error
 * (GALHIP_ERANGE). */
int galhip_tmcent(galhip_ctx *ctx, int32_t start);

/* galaxy's inner loop: nsteps x (findf(.false.); step) on the device
 * (SRC/progs/galaxy.f:107-148 without analysis and I/O). */
int galhip_run(galhip_ctx *ctx, int32_t nsteps);

/* ---- group-level entry points (parity checks; host buffers) ---- */

/* weight(jst, old) for n positions (SRC/lib15/weight.f:1): 0-based cells
 * (ncl-1) and weights wt[n][4].  old = 0 applies the sector fold used for
 * mass assignment. */
int galhip_weight(galhip_ctx *ctx, int64_t n, const void *x, const void *y,
                  int32_t old, int32_t *cell, void *wt);
/* grdacc(jst) for n positions (SRC/lib15/grdacc.f:1): grid accelerations
 * acc[n][2] interpolated from the current Cartesian field. */
int galhip_grdacc(galhip_ctx *ctx, int64_t n, const void *x, const void *y,
                  void *acc);
/* Group-level routines for parity work, on a caller's group of n particles
 * (host arrays in the context's precision; iz the group's lists, 1..nz
 * on the grid, nz + 1 off it), with the step's own device arithmetic:
 *   galhip_getacc   getacc(jst) (SRC/lib15/getacc.f:1): grid acceleration at
 *                   the particles' positions from the current field, the
 *                   fixed radial force, the zone's step factor squared;
 *                   acc[n][2] (x, y); zero off the grid.
 *   galhip_stpgrp   stpgrp(jst) (SRC/lib15/stpgrp.f:1) with offgrd/addoff and
 *                   rezgrp(jst) (SRC/lib15/rezgrp.f:1): v += acc, x += v, the
 *                   off-grid test (with chstep to the top zone's step when
 *                   offanal) and the zone change at the current istep's
 *                   mzone; x, y, vx, vy and iz updated in place.  The
 *                   context's off-grid count is not changed.
 *   galhip_massgn   massgn(jst) (SRC/lib15/massgn.f:1): adds the group's
 *                   weights to the context's zone mass sums and ring ranges. */
int galhip_getacc(galhip_ctx *ctx, int64_t n, const void *x, const void *y, const int32_t *iz,
                  void *acc);
int galhip_stpgrp(galhip_ctx *ctx, int64_t n, void *x, void *y, void *vx, void *vy,
                  const void *acc, int32_t *iz);
int galhip_massgn(galhip_ctx *ctx, int64_t n, const void *x, const void *y, const int32_t *iz);

/* Read back grids: mass sums as fixed-point int64 [nzones][mesh] (weight x
 * 2^qbits), the saved zone masses grdmoz, and the zone ring ranges jrad,
 * krad [nzones] (/zones/).  Any pointer may be NULL. */
int galhip_get_mass(const galhip_ctx *ctx, int64_t *mfx, int64_t *mozfx,
                    int32_t *jrad, int32_t *krad, int32_t *qbits);
/* Load the zone mass arrays and ring ranges, as restrt does from a dump's
 * mass records (SRC/lib15/restrt.f:1, records of SRC/lib15/dumpp.f:57-198);
 * mfx/mozfx in the fixed-point units of galhip_get_mass.  Any pointer may be
 * NULL to leave that array unchanged. */
int galhip_set_mass(galhip_ctx *ctx, const int64_t *mfx, const int64_t *mozfx,
                    const int32_t *jrad, const int32_t *krad);
/* The combined, scaled mass of zone 1 (grdmss(:,1) before p2manl). */
int galhip_get_rho(const galhip_ctx *ctx, void *rho);
/* Field as fld[3][mesh]: F_x, F_y (after polcat) and potential. */
int galhip_get_field(const galhip_ctx *ctx, void *fld);
int galhip_set_field(galhip_ctx *ctx, const void *fld);
/* Radial ranges used by the last solve: jrs, krs, jrf, krf (1-based). */
int galhip_get_ranges(const galhip_ctx *ctx, int32_t *r4);
/* Off-grid particle count (/admin/ noff). */
int galhip_get_noff(const galhip_ctx *ctx, int64_t *noff);

/* ---- diagnostics ---- */

/* Fixed radial potential for the energy diagnostics (the reference's
 * axipot/halpot, SRC/lib15/icheck.f:47-57): a cubic Hermite table on
 * r_k = k*dr like galhip_set_fixforce's; n = 0 removes it. */
int galhip_set_fixpot(galhip_ctx *ctx, int32_t n, double dr, const double *val,
                      const double *der);
/* Global integrals of icheck (SRC/lib15/icheck.f:24-70), summed in FP64 in a
 * fixed order over all particles, internal units with velocities per zone-1
 * step, time-centred as icheck takes them ((v + a/2)/tfac with the current
 * field's getacc acceleration) when centred != 0: out[0] mass, out[1..2] mass x position, out[3..4] momentum,
 * out[5] sum m v^2 (twice the kinetic energy), out[6] angular momentum Lz,
 * out[7] sum m Phi_grid over the grid's particles (NaN unless the last
 * solve formed the potential: findf with lpot), out[8] sum m Phi_fixed.
 * Energy = out[5]/2 + out[7]/2 + out[8].  On a p3d context out has 11
 * entries: [0] mass, [1..3] m x position, [4..6] momentum, [7] sum m v^2,
 * [8] Lz, [9] sum m Phi_grid (NaN as above), [10] sum m Phi_fixed of the
 * spherical radius; energy = out[7]/2 + out[9]/2 + out[10]. */
int galhip_integrals(galhip_ctx *ctx, int32_t centred, double *out);
/* galhip_integrals without any device allocation: the same kernels and
 * results (bit for bit), with the partial sums in a static 12 MiB device
 * array that each GPU holds until the process ends.  No counterpart in the
 * reference (its icheck sums on the host); it is for callers that cannot
 * allocate device memory: code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc.  Limit: at most 3 x 2^19 partial sums, i.e.
 * 9 x (ceil(n / 256) + 1) on a p2d grid (n up to about 44.7M particles) and
 * 11 x (ceil(n / 512) + 1) on a p3d grid (about 73.2M); larger contexts get
 * GALHIP_ERANGE.  Every call on one GPU uses that GPU's one array: calls
 * that overlap on a GPU (other host threads, other contexts) race. */
int galhip_integrals_noalloc(galhip_ctx *ctx, int32_t centred, double *out);

/* Log-spiral coefficients of the particles (lgspia, SRC/lib15/lgspia.f:1,
 * normalised by the particle count as phyprp does, phyprp.f:182-197): for
 * harmonic index im = 1..nm and pitch index j = 1..np,
 *   A[j][im] = sum exp(i im (nsect theta + k_j a)) / count,
 *   a = nsect ln(R / lscale) / 4,
 * R the particle's cylindrical radius in grid units, over the particles on
 * the grid and, when offanal is set, off it (equal masses); FP64 sums in a
 * fixed order.  Index j = jc + k (jc = np/2 + 1) holds pitch k >= 0 and its
 * mirror np + 1 - j pitch -k, so k_j = j - jc for odd np; for even np index
 * np/2 is never filled and is 0.  out[2][np * nm]: the cos parts, then the
 * sin parts, each in the reference's order (j - 1) nm + (im - 1).  np * nm <= 512.  On a p3d context
 * R and theta are those of (x, y); on an ensemble out holds nmem such
 * blocks, member by member. */
int galhip_lgspia(galhip_ctx *ctx, int32_t np, int32_t nm, double lscale, double *out);

/* ---- performance controls ---- */

/* Regroup particles by cell every k steps of galhip_run (0 = never), and
 * now.  No result depends on the particle order (the reference keeps its
 * particles in zone lists in load order, SRC/lib15/gather.f:1); the order
 * only sets the locality of the device passes.  galhip_get_order copies the
 * load-order index (0-based) of the particle in each storage slot. */
int galhip_set_sort_interval(galhip_ctx *ctx, int32_t k);
int galhip_sort(galhip_ctx *ctx);
int galhip_get_order(const galhip_ctx *ctx, int32_t *ids);
/* Record HIP events around every particle-push launch on the context's
 * stream; galhip_get_profile returns the summed push time (ms), the number
 * of push launches and of particles pushed since profiling was enabled. */
int galhip_set_profile(galhip_ctx *ctx, int32_t on);
int galhip_get_profile(galhip_ctx *ctx, double *push_ms, int64_t *launches,
                       int64_t *pushes);
int galhip_sync(galhip_ctx *ctx);

/* ---- 3-D polar grid (p3d) ----
 * A p3d context takes the generic entry points above with these meanings:
 * galhip_greenm / get_green / set_green (p3grnm, tables
 * G[4][ngz+1][ng][nr][nr] doubles: type 0..3 = F_R, F_phi, F_z, potential,
 * index [type][k_z][m][source ring][field ring] as p3gwrt writes them,
 * SRC/lib15/p3gwrt.f:46-64; entries p3conv never reads are zero),
 * set_fixforce (a spherical radial force, getacc.f:106-128), set_step /
 * get_step, masset, findf and p2fndf (-> p3fndf, SRC/lib15/p3fndf.f:1, and
 * polcat over every plane), step, tmcent, run, get_mass / set_mass /
 * get_rho (mesh = nr ngz na), get_field / set_field (fld[4][mesh]: F_x,
 * F_y, F_z, potential), get_ranges, get_noff, set_sort_interval, sort,
 * get_order, set_fixpot, integrals (11 entries, see there), the profiling
 * calls and sync.  Particles go in and out through galhip_restrt3 and
 * galhip_unload3; galhip_restrt, galhip_unload, galhip_weight and
 * galhip_grdacc return GALHIP_EINVAL on a p3d context.  Zone boundaries are spherical radii (SRC/lib15/setrun.f:810,
 * zoneno.f:40-41); a particle is off the grid when |z| > zm or R > rgrid
 * (SRC/lib15/offgrd.f:69-72). */
typedef struct {
    int32_t nr, na, ngz;       /* GRID card: rings, spokes, planes */
    int32_t hash;              /* HASH card */
    int32_t nsect;             /* SECT card */
    int32_t nskip;             /* SKIP card */
    int32_t skip[GALHIP_MAXSKIP];
    double dzg;                /* Z SPACING card: plane spacing, grid units */
    double softl;              /* SOFT card (softl >= dzg) */
    int32_t tsoft;             /* 1 Plummer, 2 cubic spline */
    int32_t nzones;
    int32_t nstep[GALHIP_MAXZONES];
    double zbr[GALHIP_MAXZONES];     /* inner spherical radius of zone i */
    int32_t offanal;
    double pmass;
    int32_t precision;         /* GALHIP_FP32 or GALHIP_FP64 */
    int32_t jmass;             /* INTE card: 2 linear (default; 0 reads as 2),
                                  3 quadratic, 27 cells (weight.f:251-300) */
} galhip_p3d_params;

/* p3dset + setgrd + filset + setspc for one p3d grid, linear or quadratic
 * interpolation (jmass 2 or 3: rgrid = grofu(nr - jmass/2), zm = dzg (ngz +
 * 1 - jmass)/2, setgrd.f:135-150).  Replaces SRC/lib15/p3dset.f:1,
 * SRC/lib15/setgrd.f:1, SRC/lib15/filset.f:1, SRC/lib15/setspc.f:1. */
int galhip_p3dset(const galhip_p3d_params *p, galhip_ctx **out);
/* restrt / unload (particle part) for p3d: positions and velocities in
 * six arrays, zones as for galhip_restrt (SRC/lib15/restrt.f:1,
 * SRC/lib15/unload.f:1).  Any output pointer may be NULL. */
int galhip_restrt3(galhip_ctx *ctx, int64_t n, const void *x, const void *y,
                   const void *z, const void *vx, const void *vy, const void *vz,
                   const int32_t *iz);
int galhip_unload3(galhip_ctx *ctx, void *x, void *y, void *z, void *vx,
                   void *vy, void *vz, int32_t *iz);
/* weight(jst, old) for n p3d positions (SRC/lib15/weight.f:195-300):
 * 0-based cells (the central cell for jmass 3) and wt[n][8] (jmass 3:
 * wt[n][27]) in the reference's order. */
int galhip_weight3(galhip_ctx *ctx, int64_t n, const void *x, const void *y,
                   const void *z, int32_t old, int32_t *cell, void *wt);
/* grdacc(jst) for n p3d positions (SRC/lib15/grdacc.f:189-211): acc[n][3]
 * interpolated from the current Cartesian field. */
int galhip_grdacc3(galhip_ctx *ctx, int64_t n, const void *x, const void *y,
                   const void *z, void *acc);

/* ---- ensembles of independent p2d discs ----
 * No counterpart in the reference, which runs one model per process (its
 * up to four grids, SRC/lib15/inc/params.f:17-18, are nested grids of one
 * model).  nmem discs share one grid geometry, harmonic filter, zone set-up,
 * softening and Green's table (the p2d parameters), the fixed-force and
 * fixed-potential tables and the step counter; each member has its own
 * particles, particle mass, zone mass arrays, ring ranges and field.  One
 * launch per kernel advances every member.  A member's results are
 * bit-identical to the same disc run alone in a galhip_p2dset context with
 * the same parameters and particle mass.
 * On an ensemble context the generic entry points galhip_free,
 * galhip_geometry, greenm / get_green / set_green, set_fixforce,
 * set_fixpot, set_step / get_step, masset, findf, p2fndf, step, tmcent,
 * run, weight, set_sort_interval, sort, get_order (load-order index of the
 * concatenated members), set_graph, the profiling calls (pushes summed over
 * members) and sync act on every member; galhip_restrt, unload, grdacc,
 * get_mass, set_mass, get_rho, get_field, set_field, get_ranges, get_noff
 * and integrals return GALHIP_EINVAL: use the galhip_ens_* calls below. */
#define GALHIP_MAXMEMBERS 4096

/* An ensemble of nmem discs on the grid of p (p->pmass is ignored);
 * pmass[nmem] the particle mass of each member. */
int galhip_ensset(const galhip_p2d_params *p, int32_t nmem, const double *pmass,
                  galhip_ctx **out);
/* Load the members' particles: count[nmem] particles each, the arrays
 * holding member 0's particles, then member 1's, ... (zones as for
 * galhip_restrt; at most 2^31 - 1 particles in all). */
int galhip_ens_restrt(galhip_ctx *ctx, const int64_t *count, const void *x, const void *y,
                      const void *vx, const void *vy, const int32_t *iz);
/* Copy every member's particles back in the layout galhip_ens_restrt took. */
int galhip_ens_unload(galhip_ctx *ctx, void *x, void *y, void *vx, void *vy, int32_t *iz);
/* galhip_get_mass, get_field, set_field and get_ranges of one member. */
int galhip_ens_get_mass(const galhip_ctx *ctx, int32_t member, int64_t *mfx, int64_t *mozfx,
                        int32_t *jrad, int32_t *krad, int32_t *qbits);
int galhip_ens_get_field(const galhip_ctx *ctx, int32_t member, void *fld);
int galhip_ens_set_field(galhip_ctx *ctx, int32_t member, const void *fld);
int galhip_ens_get_ranges(const galhip_ctx *ctx, int32_t member, int32_t *r4);
/* Off-grid counts noff[nmem]. */
int galhip_ens_get_noff(const galhip_ctx *ctx, int64_t *noff);
/* galhip_integrals of every member: out[nmem][9]. */
int galhip_ens_integrals(galhip_ctx *ctx, int32_t centred, double *out);

/* Where the radial Green's contraction runs (FP32 p2d grids and
 * ensembles): mode 0 (default) in the synthesis kernel, one FP32 dot product
 * per (ring, type, harmonic) -- the arithmetic that makes an ensemble member
 * bit-identical to its lone disc; mode 1 as one GEMM per (type, harmonic)
 * over all members' cos/sin columns on the matrix cores, each operand split
 * into three bf16 words (the tables from their FP64 master), the six
 * leading cross products accumulated in FP32 -- the accuracy of the FP32
 * dot products, not bit-identical to them.  The GEMM's tile shape is chosen on the first
 * solve by timing the built variants (0: 16x16 tiles, one wavefront per
 * block; 1: 16x16, four per block; 2: 32x32 tiles); modes 2, 3, 4 select
 * variant 0, 1, 2 without tuning (variants 0 and 1 sum in the same order,
 * variant 2 in another, so a fixed variant makes runs reproducible from
 * process to process).  galhip_get_contraction returns the mode (0 or 1),
 * the variant (-1: not tuned yet) and each variant's measured time in
 * microseconds (us[3], 0 when not tuned). */
int galhip_set_contraction(galhip_ctx *ctx, int32_t mode);
int galhip_get_contraction(const galhip_ctx *ctx, int32_t *mode, int32_t *variant, double *us);

/* Compensated position updates (FP32 p2d grids and ensembles; off by
 * default, which keeps the reference's plain FP32 x <- x + v): each
 * particle carries low words of (x, y), and the update forms the new
 * (x, lo) = TwoSum(x, v + lo), so the rounding error of every drift is
 * kept instead of dropped.  Forces, cells and masses use x.  Low words start
 * at zero at load; galhip_get_position_lo copies them in load order (zeros
 * when off).  No counterpart in the reference. */
int galhip_set_compensated(galhip_ctx *ctx, int32_t on);
int galhip_get_position_lo(const galhip_ctx *ctx, void *xlo, void *ylo);

/* Sets the particle mass of a single grid (p2d or p3d, not an ensemble)
 * in place of the creation parameter's (setrun.f:866 fixes it for a
 * run), so one context can serve models that differ only in the disc
 * mass.  The mass sums are scaled by it at every field solve; the last
 * field is invalidated (findf again before stepping). */
int galhip_set_pmass(galhip_ctx *ctx, double pmass);

/* Run galhip_run's steps as HIP graphs (p2d grids and ensembles): each
 * step's three launches are captured once per distinct set of launch
 * arguments and replayed.  Results are identical to on = 0.  While
 * profiling is on, galhip_run launches directly. */
int galhip_set_graph(galhip_ctx *ctx, int32_t on);

/* ---- adjoint of the p2d step ----
 * No counterpart in the reference.  Reverse-mode derivative of the map one
 * step of galhip_run computes (findf, then step: interpolation, fixed
 * force, leap-frog, mass assignment), for one grid with one zone and one
 * sector, full steps.  The cell, ring-range and off-grid choices are held
 * fixed (the map is piecewise smooth).  Adjoints are FP64 and kept on the
 * device per particle in load order: lam[4][n] = adjoints of x, y, vx, vy.
 *
 * galhip_adj_record / galhip_adj_restore save and restore the particle
 * state (positions, velocities, grid coordinates, zones, order) with the
 * zone mass sums and ring ranges in device memory slot `slot` (checkpoints,
 * and the state before each step of a recomputed segment).  After a
 * restore, findf rebuilds that step's field exactly. */
int galhip_adj_record(galhip_ctx *ctx, int32_t slot);
int galhip_adj_restore(galhip_ctx *ctx, int32_t slot);
int galhip_adj_set(galhip_ctx *ctx, const double *lam);
int galhip_adj_get(galhip_ctx *ctx, double *lam);
/* Derivative of the fixed radial force with respect to a model parameter
 * p, tabulated like galhip_set_fixforce's table (n = 0 removes it). */
int galhip_set_fixforce_dpar(galhip_ctx *ctx, int32_t n, double dr, const double *val,
                             const double *der);
/* One backward step at the context's current positions x_n, whose field
 * the last findf formed: lam holds the adjoints of (x_{n+1}, v_{n+1}) on
 * entry and of (x_n, v_n) on return.  gpar[0] += sum over particles of
 * la . a_grid (the derivative for a factor on the particle mass, and so on
 * the grid force), gpar[1] += sum of la . (dF/dp)(r) x / r (the fixed
 * force's parameter derivative table), la being the adjoint of the step's
 * acceleration.  gpar may be NULL: the sums are kept on the device
 * (galhip_adj_get_par) and the call does not wait for the GPU. */
int galhip_adj_step(galhip_ctx *ctx, double *gpar);
/* gpar[2]: the totals of galhip_adj_step's two parameter sums over the
 * steps since the adjoints were last set (galhip_adj_set or
 * galhip_adj_seed_a2). */
int galhip_adj_get_par(galhip_ctx *ctx, double *gpar);
/* Sets the adjoints for the loss L = |A_2|^2 of the current particles, as
 * galhip_adj_set would with lam = (dL/dx, dL/dy, 0, 0), computed on the
 * device: A_2 = sum e^{-2 i phi} / count over the particles with
 * r1 <= R <= r2 (R = sqrt(x^2 + y^2) / lscale, equal masses, the m = 2
 * coefficient of densty.f's DANL analysis), the annulus membership held
 * fixed in the derivative; FP64.  out[4] (may be NULL) = L, Re A_2,
 * Im A_2, count (at least 1).  Clears the parameter sums. */
int galhip_adj_seed_a2(galhip_ctx *ctx, double lscale, double r1, double r2, double *out);
/* The transposed field solve alone: g[2][mesh] (adjoints of F_x, F_y) ->
 * rbar[mesh] (adjoint of the combined, scaled mass), with the ring ranges
 * of the last findf (FP64 tables). */
int galhip_adj_field(galhip_ctx *ctx, const double *g, double *rbar);
/* With profiling on (galhip_set_profile), galhip_adj_step records HIP
 * events around its particle kernels: ms[0] the summed scatter time, ms[1]
 * the gather time (milliseconds) over the steps since the last call. */
int galhip_adj_get_profile(galhip_ctx *ctx, double *ms, int64_t *steps);

/* ---- tangent-linear (forward mode) ---- *
 * No counterpart in the reference.  The derivative of the same step map as
 * the adjoint's (one p2d grid, one zone, one sector, cells, ranges and the
 * off-grid choice held fixed) applied to a perturbation, in FP64 along the
 * forward run: an independent reference for the adjoint's gradients
 * (<J d, lam> = <d, J^T lam>).  galhip_tl_set / galhip_tl_get: the
 * perturbation d[4][n] of (x, y, vx, vy) in load order (re-sorts leave it
 * in place).  galhip_tl_step carries it through the coming step: call it
 * after findf and before step, at the state the step starts from.
 * galhip_tl_field: the field solve's own map, rho[mesh] -> (F_x, F_y)
 * f[2][mesh] in FP64 with the ring ranges of the last findf (equal to
 * galhip_findf on the same density). */
int galhip_tl_set(galhip_ctx *ctx, const double *d);
int galhip_tl_get(galhip_ctx *ctx, double *d);
int galhip_tl_step(galhip_ctx *ctx);
int galhip_tl_field(galhip_ctx *ctx, const double *rho, double *f);

#ifdef __cplusplus
}
#endif
#endif
