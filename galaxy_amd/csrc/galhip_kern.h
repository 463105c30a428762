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

#include "galhip_math.h"

#define GH_MAXZ 10
#define GH_BLOCK 256
// the p2d push kernel's launch shape: threads per block and particles per
// thread (a block takes GH_PUSH_PER consecutive particles)
#ifndef GH_PUSH_BLOCK
#define GH_PUSH_BLOCK 256
#endif
#ifndef GH_PPT
#define GH_PPT 4
#endif
#define GH_PUSH_PER (GH_PUSH_BLOCK * GH_PPT)
// 2.d0 * pi of SRC/lib15/inc/pi.f
#define GH_TWO_PI (2.0 * 3.141592653589793238462643)

namespace gh {

template <typename R> struct Vec2;
template <> struct Vec2<float> { typedef float2 T; };
template <> struct Vec2<double> { typedef double2 T; };
template <typename R> struct Vec4;
template <> struct Vec4<float> { typedef float4 T; };
template <> struct Vec4<double> { typedef double4 T; };

__device__ __forceinline__ float2 mk2(float a, float b) { return make_float2(a, b); }
__device__ __forceinline__ double2 mk2(double a, double b) { return make_double2(a, b); }
__device__ __forceinline__ float4 mk4(float a, float b, float c, float d) { return make_float4(a, b, c, d); }
__device__ __forceinline__ double4 mk4(double a, double b, double c, double d) { return make_double4(a, b, c, d); }

// ---- wavefront reductions -------------------------------------------------
//
// min / max / sum of a 32-bit integer over the 64 lanes of a wavefront, all
// lanes active: DPP exchanges inside each 16-lane row (quad permutations
// [1,0,3,2] and [2,3,0,1], the half-row mirror, the row mirror), then the
// four rows' values read into scalar registers.  The result is uniform.
// Four VALU steps and four lane reads instead of six LDS-routed shuffles.
#define GH_DPP_QUAD_1032 0xB1
#define GH_DPP_QUAD_2301 0x4E
#define GH_DPP_ROW_HALF_MIRROR 0x141
#define GH_DPP_ROW_MIRROR 0x140

// (every lane reads a lane of its own row, so nothing falls back to the
// "old" value; mov_dpp lets the compiler fold the exchange into the min,
// max or add that consumes it)
#define GH_DPP(v, ctrl) __builtin_amdgcn_mov_dpp((v), (ctrl), 0xF, 0xF, true)

__device__ __forceinline__ int wave_min_i32(int v) {
    v = min(v, GH_DPP(v, GH_DPP_QUAD_1032));
    v = min(v, GH_DPP(v, GH_DPP_QUAD_2301));
    v = min(v, GH_DPP(v, GH_DPP_ROW_HALF_MIRROR));
    v = min(v, GH_DPP(v, GH_DPP_ROW_MIRROR));
    return min(min(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 16)),
               min(__builtin_amdgcn_readlane(v, 32), __builtin_amdgcn_readlane(v, 48)));
}
__device__ __forceinline__ int wave_max_i32(int v) {
    v = max(v, GH_DPP(v, GH_DPP_QUAD_1032));
    v = max(v, GH_DPP(v, GH_DPP_QUAD_2301));
    v = max(v, GH_DPP(v, GH_DPP_ROW_HALF_MIRROR));
    v = max(v, GH_DPP(v, GH_DPP_ROW_MIRROR));
    return max(max(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 16)),
               max(__builtin_amdgcn_readlane(v, 32), __builtin_amdgcn_readlane(v, 48)));
}
__device__ __forceinline__ int wave_sum_i32(int v) {
    v = v + GH_DPP(v, GH_DPP_QUAD_1032);
    v = v + GH_DPP(v, GH_DPP_QUAD_2301);
    v = v + GH_DPP(v, GH_DPP_ROW_HALF_MIRROR);
    v = v + GH_DPP(v, GH_DPP_ROW_MIRROR);
    return (__builtin_amdgcn_readlane(v, 0) + __builtin_amdgcn_readlane(v, 16)) +
           (__builtin_amdgcn_readlane(v, 32) + __builtin_amdgcn_readlane(v, 48));
}

// n / d for 0 <= n < 2^16 and 1 <= d < 2^16 without the integer division
// sequence: n times the single-precision reciprocal of d is within 2^-6 of
// the quotient, so its truncation is off by at most one, which two exact
// integer tests correct.
__device__ __forceinline__ int div_small(int n, int d) {
    int q = (int)((float)n * __builtin_amdgcn_rcpf((float)d));
    q -= q * d > n ? 1 : 0;
    q += (q + 1) * d <= n ? 1 : 0;
    return q;
}

// Floating-point wavefront sum by DPP row exchanges (pairs, quads,
// half-rows, rows of 16 lanes), then the four rows' sums read from lanes 0,
// 16, 32, 48; every lane gets the total.  Short dot products in the solve
// kernels were waiting on the LDS-routed shuffles of a butterfly.
template <int CTRL> __device__ __forceinline__ float dpp_mov(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, true));
}
template <int CTRL> __device__ __forceinline__ double dpp_mov(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ float lane_value(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double lane_value(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
template <typename R> __device__ __forceinline__ R wave_sum_dpp(R v) {
    v += dpp_mov<0xB1>(v);    // quad_perm 1,0,3,2
    v += dpp_mov<0x4E>(v);    // quad_perm 2,3,0,1
    v += dpp_mov<0x141>(v);   // row_half_mirror
    v += dpp_mov<0x140>(v);   // row_mirror
    return (lane_value(v, 0) + lane_value(v, 16)) + (lane_value(v, 32) + lane_value(v, 48));
}

// ---- per-particle geometry ------------------------------------------------

// Grid coordinates of a position as weight() forms them
// (SRC/lib15/weight.f:72-79, uofgr.f:24): u = log(1 + r) / alpha and
// t = atan2(y, x) na / (2 pi), plus na when negative (t is the unfolded
// angle).  FP32: the reference's FP32 logarithm and quotient, correctly
// rounded -- log through gh_log_ge1 rounded once to float, and l / alpha as
// (double)l * (1/alpha) rounded once (equal to the IEEE float quotient
// because a float quotient is never within 2^-49 of a float midpoint); the
// angle in double, rounded once.  FP64: the same formulas in double.
struct UtConst {
    double inv_alpha;   // 1 / (double)alpha
    double tscale;      // na / (2 pi)
    double alpha_d;     // alpha in the context precision, as double
    int na;
};

// tab: the math tables (GH_MATH_TAB, or a copy of it staged in LDS).
__device__ __forceinline__ void ut_of(float x, float y, float rr,
                                      const UtConst &k, float &u, float &t,
                                      const double *tab = GH_MATH_TAB) {
    float v = 1.0f + rr;
    float l = (float)gh_log_ge1((double)v, tab);
    u = (float)((double)l * k.inv_alpha);
    t = (float)(gh_atan2((double)y, (double)x, tab) * k.tscale);
    if (t < 0.0f) t = t + (float)k.na;
}
__device__ __forceinline__ void ut_of(double x, double y, double rr,
                                      const UtConst &k, double &u, double &t,
                                      const double *tab = GH_MATH_TAB) {
    double l = gh_log_ge1(1.0 + rr, tab);
    u = l / k.alpha_d;
    t = gh_atan2(y, x, tab) * (double)k.na / GH_TWO_PI;
    if (t < 0.0) t = t + (double)k.na;
}

// Cell and CIC weights from (u, t) (weight.f:84-95).  The reference's cell
// is c = ju*na + jt with jt in [0, na]; (ring, spoke) = (c div na, c mod na)
// without a division.  w[0..3] belong to the corners (ring, spoke),
// (ring, spoke+1), (ring+1, spoke), (ring+1, spoke+1), spoke+1 wrapping
// (massgn.f:79-85, grdacc.f:45-50).
template <typename R>
__device__ __forceinline__ void cic(R u, R t, int nr, int na, int &ring,
                                    int &spoke, R w[4]) {
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
    ring = ju;
    spoke = jt;
    if (jt >= na) {
        ring = ju + 1;
        spoke = jt - na;
    }
}

// Flat corner indices (row-major cell = ring*na + spoke) of a cell.
__device__ __forceinline__ void corners_rs(int ring, int spoke, int na,
                                           int ic[4]) {
    int s2 = spoke + 1 == na ? 0 : spoke + 1;
    ic[0] = ring * na + spoke;
    ic[1] = ring * na + s2;
    ic[2] = ic[0] + na;
    ic[3] = ic[1] + na;
}

// weight() for one position: cell (c = ring*na + spoke), weights, radius.
template <typename R>
__device__ __forceinline__ int weight1(R x, R y, int nr, const UtConst &k,
                                       bool fold, R thmax, R w[4], R &rr,
                                       R &u, R &t) {
    rr = sqrt(x * x + y * y);
    ut_of(x, y, rr, k, u, t);
    R tf = fold ? (R)fmod(t, thmax) : t;
    int ring, spoke;
    cic(u, tf, nr, k.na, ring, spoke, w);
    return ring * k.na + spoke;
}

// Fixed-point weight: w * 2^qbits, rounded to nearest even.
template <typename R>
__device__ __forceinline__ long long quant(R w, int qbits) {
    return __double2ll_rn(ldexp((double)w, qbits));
}
// FP32 weights, the same values: a normal w >= 0 is m 2^(e - 150) with the
// 24-bit significand m and biased exponent e, so for 0 <= e - 150 + qbits
// <= 39 the product w 2^qbits is the integer m << (e - 150 + qbits),
// exact (nothing to round) and below 2^63.  Zero, negative and very small
// weights (the shift out of range, as an unsigned compare) take the FP64
// path.
template <>
__device__ __forceinline__ long long quant<float>(float w, int qbits) {
    const unsigned b = __float_as_uint(w);
    const unsigned sh = (b >> 23) + (unsigned)(qbits - 150);
    if (sh <= 39u)
        return (long long)((unsigned long long)((b & 0x7fffffu) | 0x800000u) << sh);
    return __double2ll_rn(ldexp((double)w, qbits));
}

// The N weights of one deposit as fixed point, the values of quant, for an
// active lane (an inactive lane's values are left undefined).  FP32: the
// exact shift of quant<float>; only a wavefront in which some active lane
// has a weight outside the shift's range (zero, negative, very small, or not
// below 2) takes the FP64 rounding, behind a wave-uniform branch, so that the
// others do not issue it.
template <typename R, int N>
__device__ __forceinline__ void quant_n(const R *w, int qbits, bool active, long long *v) {
#pragma unroll
    for (int q = 0; q < N; q++) v[q] = active ? quant<R>(w[q], qbits) : 0;
}
template <int N>
__device__ __forceinline__ void quant_n_f32(const float *w, int qbits, bool active,
                                            long long *v) {
    bool slow = false;
#pragma unroll
    for (int q = 0; q < N; q++) {
        const unsigned b = __float_as_uint(w[q]);
        const unsigned sh = (b >> 23) + (unsigned)(qbits - 150);
        const unsigned long long m = (unsigned long long)((b & 0x7fffffu) | 0x800000u);
        v[q] = (long long)(m << (sh & 63u));
        slow = slow || sh > 39u;
    }
    if (__ballot(active && slow)) {
#pragma unroll
        for (int q = 0; q < N; q++) {
            const unsigned sh = (__float_as_uint(w[q]) >> 23) + (unsigned)(qbits - 150);
            if (sh > 39u) v[q] = __double2ll_rn(ldexp((double)w[q], qbits));
        }
    }
}
template <>
__device__ __forceinline__ void quant_n<float, 4>(const float *w, int qbits, bool active,
                                                  long long *v) {
    quant_n_f32<4>(w, qbits, active, v);
}
template <>
__device__ __forceinline__ void quant_n<float, 8>(const float *w, int qbits, bool active,
                                                  long long *v) {
    quant_n_f32<8>(w, qbits, active, v);
}
template <>
__device__ __forceinline__ void quant_n<float, 27>(const float *w, int qbits, bool active,
                                                   long long *v) {
    quant_n_f32<27>(w, qbits, active, v);
}

// Fixed radial tables (the halo force, its potential, the force's
// derivative in a model parameter) are the cubic Hermite interpolants of
// (val, der) on r_k = k dr, held per interval k as the coefficients of
//   f = c0 + t (c1 + t (c2 + t c3)),   t = r (1/dr) - k in [0, 1),
// c0 = v_k, c1 = dr d_k, c2 = (3 (v_k+1 - v_k) - 2 c1) - dr d_k+1,
// c3 = (2 (v_k - v_k+1) + c1) + dr d_k+1 (radtab_coeffs, on the host).
// Entry n - 1 is the constant v_n-1, which holds past the table's end
// (t = 0 there).  Evaluated with fused multiply-adds.

// Interval of radius r: k with r_k <= r < r_k+1, n - 1 past the last.
__device__ __forceinline__ int fixforce_index(double r, int n, double inv_dr) {
    const int k = (int)(r * inv_dr);
    return k >= n - 1 ? n - 1 : (k < 0 ? 0 : k);
}

// The table at r in interval k (from fixforce_index); c holds the table
// from entry k0 on (a slice of it).
__device__ __forceinline__ double fixforce_at(double r, int k, int n, double inv_dr,
                                              const double4 *__restrict__ c, int k0) {
    const double t = k == n - 1 ? 0.0 : r * inv_dr - (double)k;
    const double4 q = c[k - k0];
    return fma(fma(fma(q.w, t, q.z), t, q.y), t, q.x);
}

__device__ __forceinline__ double fixforce(double r, int n, double inv_dr,
                                           const double4 *__restrict__ c) {
    return fixforce_at(r, fixforce_index(r, n, inv_dr), n, inv_dr, c, 0);
}

// d/dr of the table at r (0 past the end, where the constant has no slope)
__device__ __forceinline__ double fixforce_deriv(double r, int n, double inv_dr,
                                                 const double4 *__restrict__ c) {
    const int k = fixforce_index(r, n, inv_dr);
    const double t = k == n - 1 ? 0.0 : r * inv_dr - (double)k;
    const double4 q = c[k];
    return fma(fma(3.0 * q.w, t, 2.0 * q.z), t, q.y) * inv_dr;
}

// ---- shared by the solve and particle kernels ---------------------------------

// Radial ranges of p2fndf (p2fndf.f:26-55, findf.f:42-54).  Returns false
// if no mass is on the grid.  kr receives krad after p2fndf's cumulative
// update.
__device__ __forceinline__ bool solve_ranges(const int *jrad, const int *krad, int nz, int nr,
                             int mzone, int lpot, int explicit_zones,
                             int jzone_in, int kzone_in, int kr[GH_MAXZ],
                             int out[4]) {
    int jrs = nr, krs = 0;
    for (int ipm = 1; ipm <= nz; ipm++) {
        jrs = min(jrs, jrad[ipm - 1]);
        krs = max(krs, krad[ipm - 1]);
        kr[ipm - 1] = krad[ipm - 1];
        if (ipm > 1) kr[ipm - 1] = max(kr[ipm - 1], krs);
    }
    if (jrs > krs || krs == 0) return false;
    int jzone, kzone;
    if (explicit_zones) {
        jzone = jzone_in;
        kzone = kzone_in;
    } else if (lpot || mzone == nz) {
        jzone = 0;
        kzone = nz + 1;
    } else {
        jzone = 1;
        kzone = mzone;
    }
    int jrf = 1;
    if (jzone > 0 && jrad[jzone - 1] < nr) jrf = jrad[jzone - 1];
    int krf = nr;
    if (kzone <= nz) {
        krf = kr[kzone - 1];
        while (krf < jrf) {
            kzone++;
            if (kzone > nz) return false;
            krf = kr[kzone - 1];
        }
    }
    out[0] = jrs;
    out[1] = krs;
    out[2] = jrf;
    out[3] = krf;
    return true;
}

// Global deposit of one corner value with Fortran's flat indexing over the
// [nz][mesh] zone arrays (grdmss(j, jz)); before the first array or past the
// last it is dropped.  mo: the grid's offset in each zone array (ensemble
// member x mesh).
__device__ __forceinline__ void deposit_global(long long *const *mfx, int nz,
                                               int mesh, int zone, int idx,
                                               long long v, long long mo = 0) {
    long long f = (long long)(zone - 1) * mesh + idx;
    if (f >= 0 && f < (long long)nz * mesh) {
        int zf = zone - 1, cf = idx;
        while (cf >= mesh) {
            cf -= mesh;
            zf++;
        }
        while (cf < 0) {
            cf += mesh;
            zf--;
        }
        atomicAdd((unsigned long long *)&mfx[zf][mo + cf], (unsigned long long)v);
    }
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
    typename Vec4<R>::T *xv;          // (x, y, vx, vy) per particle
    typename Vec2<R>::T *ut;          // (u, t) of the current position
    uint8_t *zone;
    long long n;
    const typename Vec2<R>::T *fld;   // [mesh] (F_x, F_y) after polcat
    long long *mfx[GH_MAXZ];          // zone mass sums, fixed point
    int *jrad, *krad;                 // [nz] 1-based ring ranges
    long long *counters;              // [0] off-grid delta, [1] pushes
    const double4 *fr_c;              // fixed force table (coefficients)
    double fr_inv_dr;
    int nfr, fixforce;
    int nr, na, mesh, nz, mzone, offanal, nsect, qbits;
    UtConst k;
    R rgrid, thmax;
    R zbr[GH_MAXZ];
    int nstep[GH_MAXZ];
    // ensembles (push kernel): nblocks blocks, block b covering particles
    // [blk[b].y, blk[b].y + blk[b].z) of member blk[b].x, whose grids sit
    // at member x mesh in fld and the zone arrays, member x GH_MAXZ in
    // jrad/krad and member x 4 in counters, with fixed-point scale
    // qbits_m[member].  blk = NULL: one grid.
    const int4 *blk;
    const int *qbits_m;
    int nblocks;
    // compensated positions (FP32, opt-in): the low words of (x, y), or NULL
    typename Vec2<R>::T *xlo;
};

// grdacc-only / weight-only argument block (parity entry points)
template <typename R> struct GatherArgs {
    const R *x, *y;
    long long n;
    const typename Vec2<R>::T *fld;
    R *acc;        // [n][2] or NULL
    int *cell;     // [n] or NULL
    R *wt;         // [n][4] or NULL
    int nr, fold;
    UtConst k;
    R thmax;
};

template <typename R> struct AnaArgs {
    const long long *mfx[GH_MAXZ];
    const long long *moz[GH_MAXZ];
    double fnew[GH_MAXZ], fold[GH_MAXZ];
    int use_new[GH_MAXZ], use_old[GH_MAXZ];
    int *jrad, *krad;
    int *ranges;                      // out: jrs, krs, jrf, krf
    R *rho;                           // [mesh]
    R *coef;                          // [ncoef][nr]
    const R *tw;                      // [na][2] cos, sin of 2 pi k / na
    const int *lg;                    // [mmax] skip flags
    double q, pmass;
    int nr, na, ng, ncoef, nz, mzone, lpot, jzone, kzone, explicit_zones;
    // ensembles: nmem grids (grid index blockIdx.z) with per-member qbits
    // and particle mass (NULL: q and pmass above); nmem = 1 for one grid
    int nmem;
    const int *qbits_m;
    const double *pmass_m;
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
    // masscl of the coming push (masscl.f:67-146), done here because the
    // analysis kernel has consumed the zone sums: zones 1..masscl_z cleared,
    // zones 2..masscl_z saved, their jrad/krad reset (0 = none)
    long long *mfx[GH_MAXZ];
    long long *moz[GH_MAXZ];
    int *jrad, *krad;
    int masscl_z;
    int nmem;                         // grids (blockIdx.z); 1 for one grid
    const float *wpre;                // contraction done beforehand, or NULL
};

// the ensemble's MFMA contraction (galhip_solve.hip)
struct EnsGemmArgs {
    const unsigned short *gsp;        // bf16 words of G, [3 words][3][ng][nrp][nrp]
    const float *coef;                // [nmem][ncoef][nr]
    const int *ranges;                // [nmem][4]
    const int *lg;
    float *W;                         // [nmem][3][2][ng][nr]
    int nr, nrp, ng, ncoef, nmem, mmax, ntypes;
};

}  // namespace gh
