// galhip_p3d.h -- argument blocks and launchers of the p3d kernels
// (galhip_p3d.hip), shared with the host side (galhip_p3d_api.hip).
#pragma once
#include "galhip_kern.h"

namespace gh {

// A particle is one record of two R4 (32 B in FP32): rec[2 i] = (x, y, z, vx),
// rec[2 i + 1] = (vy, vz, u, t), (u, t) the grid coordinates of the current
// position; a sort's gather then touches one record per particle.
template <typename R> struct P3Args {
    typename Vec4<R>::T *rec;         // 2 per particle
    uint8_t *zone;
    long long n;
    const typename Vec4<R>::T *fld;   // (F_x, F_y, F_z, potential) per cell
    long long *mfx[GH_MAXZ];
    int *jrad, *krad;
    long long *counters;              // [0] off-grid count, [1] pushes
    const double4 *fr_c;              // fixed force table (coefficients)
    double fr_inv_dr;
    int nfr, fixforce;
    int nr, na, ngz, mesh, nz, mzone, offanal, nsect, qbits;
    int jmass;                        // 2 linear (8 corners), 3 quadratic (27)
    UtConst k;
    R rgrid, thmax, zm, dzg, fac;
    R zbr[GH_MAXZ];
    int nstep[GH_MAXZ];
};

template <typename R> struct P3GatherArgs {
    const R *x, *y, *z;
    long long n;
    const typename Vec4<R>::T *fld;
    R *acc;        // [n][3] or NULL
    int *cell;     // [n] or NULL
    R *wt;         // [n][8] (jmass 3: [n][27]) or NULL
    int fold;
    P3Args<R> g;   // geometry only
};

template <typename R> struct P3SolveArgs {
    const long long *mfx[GH_MAXZ];
    const long long *moz[GH_MAXZ];
    double fnew[GH_MAXZ], fold[GH_MAXZ];
    int use_new[GH_MAXZ], use_old[GH_MAXZ];
    int *jrad, *krad, *ranges;
    R *rho;          // [mesh] combined, scaled mass
    R *A;            // [nr][nm][ngz] azimuthal components per plane
    R *M;            // [nr][nm][2 ngz] after the vertical transform
    R *W;            // [4][nr][nm][2 ngz] contracted
    R *Z;            // [4][nr][ngz][nm] after the inverse vertical transform
    const R *G;      // [4][ngz+1][ng][nr(source)][nr(field)]
    const R *tw;     // [na][2] cos, sin of 2 pi k / na
    const R *twz;    // [2 ngz][2] cos, sin of 2 pi k / (2 ngz)
    const int *lg, *slot_m, *slot_p, *slot_of;
    typename Vec4<R>::T *fld;
    double q, pmass;
    int nr, na, ngz, ng, mmax, nm, nz, mzone, lpot, jzone, kzone, explicit_zones, ntypes;
};

template <typename R> struct P3PermArgs {
    const typename Vec4<R>::T *rec_in;   // 2 per particle
    typename Vec4<R>::T *rec_out;
    const uint8_t *zin;
    uint8_t *zout;
    const unsigned *iin;
    unsigned *iout;
    const unsigned *perm;
    long long n;
};

template <typename R>
hipError_t launch_p3_key(const P3Args<R> &, int kshift, unsigned kmax, unsigned *key,
                         unsigned *perm, hipStream_t);
template <typename R> hipError_t launch_p3_permute(const P3PermArgs<R> &, hipStream_t);
template <typename R>
hipError_t launch_p3_interleave(const R *s4, const R *s2, typename Vec4<R>::T *rec,
                                unsigned *id, long long n, hipStream_t);
template <typename R>
hipError_t launch_p3_deinterleave(const typename Vec4<R>::T *rec, const uint8_t *zone,
                                  const unsigned *id, R *d4, R *d2, int *iz, long long n,
                                  hipStream_t);
template <typename R> hipError_t launch_p3_massgn(const P3Args<R> &, hipStream_t);
template <typename R> hipError_t launch_p3_push(const P3Args<R> &, hipStream_t);
template <typename R> hipError_t launch_p3_tmcent(const P3Args<R> &, hipStream_t);
template <typename R> hipError_t launch_p3_ut(const P3Args<R> &, hipStream_t);
constexpr int P3_NINT = 11;   // entries of the p3d integrals
template <typename R>
hipError_t launch_p3_integrals(const P3Args<R> &, int with_pot, const double4 *fp_c,
                               int nfp, double fp_dr, int centred,
                               double *part, double *out, hipStream_t);
template <typename R> hipError_t launch_p3_gather(const P3GatherArgs<R> &, hipStream_t);
template <typename R> hipError_t launch_p3_solve(const P3SolveArgs<R> &, hipStream_t);
template <typename R> bool p3_solve_fits(const P3SolveArgs<R> &);
hipError_t launch_p3_green(double *G, const double *twd, const double *twzd, const int *lg,
                           int nr, int na, int ngz, int ng, double alpha, double dzg,
                           double eps, int tsoft, hipStream_t s);

}  // namespace gh
