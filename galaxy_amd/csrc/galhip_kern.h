// galhip_kern.h -- device-side definitions shared by the HIP translation
// units of libgalhip: kernel argument blocks and the per-particle helpers
// that must give identical results in every kernel that uses them.
//
// All device code of the library is compiled with -ffp-contract=off: the
// cell index and CIC weights must be bit-identical to the reference's FP32
// arithmetic (SRC/lib15/weight.f:61-97), which rounds every product and sum.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define GH_MAXZ 10
#define GH_BLOCK 256
// 2.d0 * pi of SRC/lib15/inc/pi.f
#define GH_TWO_PI (2.0 * 3.141592653589793238462643)

namespace gh {

template <typename R> struct Vec2;
template <> struct Vec2<float> { typedef float2 T; };
template <> struct Vec2<double> { typedef double2 T; };

// ---- per-particle geometry ------------------------------------------------

// uofgr stdpol rule u = log(1 + r) / alpha (SRC/lib15/uofgr.f:24).  The FP32
// logarithm is the correctly rounded one, taken through double.
__device__ __forceinline__ float uofgr(float r, float alpha) {
    float one_r = 1.0f + r;
    float l = (float)log((double)one_r);
    return l / alpha;
}
__device__ __forceinline__ double uofgr(double r, double alpha) {
    return log(1.0 + r) / alpha;
}

// weight() for one p2d particle (SRC/lib15/weight.f:61-97).  Returns the
// 0-based cell (ncl - 1); w[0..3] = CIC weights of corners (c, c+1, c+na,
// c+1+na).  fold applies the mass-assignment sector fold t = mod(t, thmax).
template <typename R>
__device__ __forceinline__ int weight1(R x, R y, int nr, int na, R alpha,
                                       bool fold, R thmax, R w[4], R &rr) {
    rr = sqrt(x * x + y * y);
    R u = uofgr(rr, alpha);
    R t = (R)(atan2((double)y, (double)x) * (double)na / GH_TWO_PI);
    if (t < (R)0) t = t + (R)na;
    if (fold) t = fmod(t, thmax);
    int ju = (int)u;
    if (ju > nr - 2) ju = nr - 2;
    int jt = (int)t;
    R du = u - (R)ju;
    R dt = t - (R)jt;
    R dd = du * dt;
    w[0] = (R)1 - du - dt + dd;
    w[1] = dt - dd;
    w[2] = du - dd;
    w[3] = dd;
    return ju * na + jt;
}

// The four corners with the azimuthal wrap of massgn.f:79-85 / grdacc.f:45-50.
__device__ __forceinline__ void corners(int c, int na, int ic[4]) {
    ic[0] = c;
    ic[1] = c + 1;
    if ((c + 1) % na == 0) ic[1] -= na;
    ic[2] = c + na;
    ic[3] = ic[1] + na;
}

// Fixed-point weight: w * 2^qbits, rounded to nearest even.
template <typename R>
__device__ __forceinline__ long long quant(R w, int qbits) {
    return __double2ll_rn(ldexp((double)w, qbits));
}

// Cubic Hermite interpolation of the tabulated fixed radial force.
__device__ __forceinline__ double fixforce(double r, int nfr, double dr,
                                           const double *__restrict__ val,
                                           const double *__restrict__ der) {
    double s = r / dr;
    int k = (int)s;
    if (k >= nfr - 1) return val[nfr - 1];
    if (k < 0) k = 0;
    double t = s - (double)k;
    double t2 = t * t;
    double omt = 1.0 - t;
    double omt2 = omt * omt;
    double h00 = (1.0 + 2.0 * t) * omt2;
    double h10 = t * omt2;
    double h01 = t2 * (3.0 - 2.0 * t);
    double h11 = t2 * (t - 1.0);
    return h00 * val[k] + h10 * dr * der[k] + h01 * val[k + 1] +
           h11 * dr * der[k + 1];
}

// ---- kernel argument blocks -------------------------------------------------

struct ZonePtrs {
    long long *mfx[GH_MAXZ];
    long long *moz[GH_MAXZ];
};

template <typename R> __device__ __forceinline__ typename Vec2<R>::T zero2();
template <> __device__ __forceinline__ float2 zero2<float>() { return make_float2(0.f, 0.f); }
template <> __device__ __forceinline__ double2 zero2<double>() { return make_double2(0.0, 0.0); }

template <typename R> struct PushArgs {
    R *x, *y, *vx, *vy;
    uint8_t *zone;
    long long n;
    const typename Vec2<R>::T *fld;   // [mesh] (F_x, F_y) after polcat
    long long *mfx[GH_MAXZ];          // zone mass sums, fixed point
    int *jrad, *krad;                 // [nz] 1-based ring ranges
    long long *counters;              // [0] off-grid delta, [1] pushes
    const double *fr_val, *fr_der;
    double fr_dr;
    int nfr, fixforce;
    int nr, na, mesh, nz, mzone, offanal, nsect, qbits;
    R alpha, rgrid, thmax;
    R zbr[GH_MAXZ];
    int nstep[GH_MAXZ];
};

// grdacc-only / weight-only argument block (parity entry points)
template <typename R> struct GatherArgs {
    const R *x, *y;
    long long n;
    const typename Vec2<R>::T *fld;
    R *acc;        // [n][2] or NULL
    int *cell;     // [n] or NULL
    R *wt;         // [n][4] or NULL
    int nr, na, fold;
    R alpha, thmax;
};

template <typename R> struct AnaArgs {
    const long long *mfx[GH_MAXZ];
    const long long *moz[GH_MAXZ];
    long long *mfx_w[GH_MAXZ];        // fused masscl: zone arrays to clear
    long long *moz_w[GH_MAXZ];        // fused masscl: saved copies
    double fnew[GH_MAXZ], fold[GH_MAXZ];
    int use_new[GH_MAXZ], use_old[GH_MAXZ];
    int *jrad, *krad;
    int *ranges;                      // out: jrs, krs, jrf, krf
    R *rho;                           // [mesh]
    R *coef;                          // [nr][ncoef]
    const R *tw;                      // [na][2] cos, sin of 2 pi k / na
    const int *lg;                    // [mmax] skip flags
    double q, pmass;
    int nr, na, ng, ncoef, nz, mzone, lpot, jzone, kzone, explicit_zones;
    int masscl_z;                     // fused masscl for zones 1..masscl_z (0 = none)
};

template <typename R> struct SynArgs {
    const R *coef;
    const R *G;                       // [3][ng][nr][nr]
    const R *tw;
    const int *lg;
    const int *ranges;
    typename Vec2<R>::T *fld;
    R *pot;
    int nr, na, ng, ncoef, mmax, lpot;
};

}  // namespace gh
