// galhip_p3d.hip -- kernels of the 3-D polar grid (p3d): cylindrical rings
// x spokes x z-planes, linear (trilinear) interpolation (jmass = 2).
//
// Mesh cell (ring, plane, spoke) is ((ring)*ngz + plane)*na + spoke, 0-based
// (SRC/lib15/weight.f:232).  The field is held interleaved as one 4-vector
// per cell, (F_x, F_y, F_z, potential), so that an interpolation corner is
// one 16-byte (FP32) load.
//
// Particle kernels (one thread per particle):
//   p3_massgn_kernel  massgn of every on-grid particle (massgn.f:140-172)
//   p3_push_kernel    getacc (grdacc.f:189-211, getacc.f:106-151) + stpgrp
//                     + offgrd (|z| > zm or R > rgrid, offgrd.f:69-72) +
//                     addoff + rezgrp with spherical zone radii
//                     (zoneno.f:40-41) + massgn at the new position
//   p3_tmcent_kernel  tmcent's half-step velocity shift
//   p3_gather_kernel  weight / grdacc for given positions (parity checks)
// Solve (p3fndf.f:37-104), FP32 or FP64:
//   p3_analyse_kernel block (plane, ring): zoncmb + scaled, then the
//                     azimuthal components of the active harmonics
//                     (p3manl.f:60-89) as direct sums
//   p3_zfwd_kernel    block per ring: the doubled vertical transform of each
//                     component column, zero-padded to 2 ngz (p3manl.f:90-101)
//   p3_conv_kernel    block per (type, k_z, m): the radial contraction with
//                     that Green's matrix (p3conv.f:139-226); table streamed
//                     once, field ring per thread
//   p3_zinv_kernel    block per (type, ring): inverse vertical transform,
//                     first ngz planes (p3fsyn.f:59-91)
//   p3_synth_kernel   block per (plane, ring): inverse azimuthal transform,
//                     1/(2 ngz na) (p3fsyn.f:92-129) and polcat
// p3_green_kernel: p3grnm's tables in FP64 (p3grnm.f:93-263), p3gwrt's
//   per-(k_z, m) matrices with the field ring fastest (p3gwrt.f:46-64).
#include "galhip_p3d.h"

#include <climits>

namespace gh {

constexpr int P3_BLOCK = 256;    // solve kernels
constexpr int P3_PBLOCK = 512;   // particle kernels: a larger block shares one deposit window

// (u, t) of particle i: the second half of its record's second R4
template <typename R>
__device__ __forceinline__ void p3_set_ut(const P3Args<R> &a, long long i, R u, R t) {
    reinterpret_cast<typename Vec2<R>::T *>(a.rec + 2 * i + 1)[1] = mk2(u, t);
}
constexpr int P3_MAX_NA = 2048;
constexpr int P3_MAX_NM = 128;   // compact components (2 ng)
constexpr int P3_MAX_NR = 1024;

template <typename R> __device__ __forceinline__ R p3_wave_sum(R v) { return wave_sum_dpp(v); }

// (u, t) of a position as weight() forms them (weight.f:210-218): the FP32
// logarithm and angle of the p2d grid, t = 0 on the axis.
template <typename R>
__device__ __forceinline__ void ut3(R x, R y, const P3Args<R> &a, R &u, R &t) {
    const R rr = sqrt(x * x + y * y);
    ut_of(x, y, rr, a.k, u, t);
    if (!(rr > (R)0)) t = (R)0;
}

// weight() for one position given its (u, t), p3d jmass = 2
// (weight.f:219-248): returns the 0-based cell (ju ngz + jz) na + jt and the
// 8 weights in the reference's order (bit 0 = dt, bit 1 = dz, bit 2 = du).
template <typename R>
__device__ __forceinline__ int cell3(R u, R t, R z, const P3Args<R> &a, bool fold, R w[8],
                                     int &ring, int &plane, int &spoke) {
    R zz = z + a.zm;
    if (fold) t = (R)fmod(t, a.thmax);
    int ju = (int)u;
    if (ju > a.nr - 2) ju = a.nr - 2;
    int jt = (int)t;
    zz = zz / a.dzg;
    int jz = (int)zz;
    if (jz < 0) jz = 0;
    if (jz > a.ngz - 2) jz = a.ngz - 2;
    const R du = u - (R)ju, dt = t - (R)jt, dz = zz - (R)jz;
    const R one = (R)1;
    w[0] = (one - dt) * (one - dz) * (one - du);
    w[1] = dt * (one - dz) * (one - du);
    w[2] = (one - dt) * dz * (one - du);
    w[3] = dt * dz * (one - du);
    w[4] = (one - dt) * (one - dz) * du;
    w[5] = dt * (one - dz) * du;
    w[6] = (one - dt) * dz * du;
    w[7] = dt * dz * du;
    ring = ju;
    plane = jz;
    spoke = jt;
    return (ju * a.ngz + jz) * a.na + jt;
}

template <typename R>
__device__ __forceinline__ int weight3(R x, R y, R z, const P3Args<R> &a, bool fold, R w[8]) {
    R u, t;
    int rg, pl, sp;
    ut3(x, y, a, u, t);
    return cell3(u, t, z, a, fold, w, rg, pl, sp);
}

// weight() for jmass = 3 given (u, t) (weight.f:281-300): the central cell
// of the 27 (nearest node, ju >= 1, jt == na folded to spoke 0) and their
// weights, spoke offset -1, 0, +1 fastest, then plane, then ring.
template <typename R>
__device__ __forceinline__ int cell3q(R u, R t, R z, const P3Args<R> &a, bool fold, R w[27],
                                      int &ring, int &plane, int &spoke) {
    R zz = z + a.zm;
    if (fold) t = (R)fmod(t, a.thmax);
    const R h = (R)0.5;
    int ju = (int)(u + h);
    if (ju > a.nr - 2) ju = a.nr - 2;
    if (ju < 1) ju = 1;
    const int jt = (int)(t + h);
    zz = zz / a.dzg;
    int jz = (int)(zz + h);
    if (jz < 0) jz = 0;
    if (jz > a.ngz - 2) jz = a.ngz - 2;
    int c = (ju * a.ngz + jz) * a.na + jt;
    if (jt == a.na) c -= a.na;
    ring = ju;
    plane = jz;
    spoke = jt == a.na ? 0 : jt;
    R dp[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const R dd = i == 0 ? t + h - (R)jt : i == 1 ? zz + h - (R)jz : u + h - (R)ju;
        dp[0][i] = (R)1 - (R)2 * dd + dd * dd;
        dp[1][i] = (R)1 + (R)2 * dd - (R)2 * (dd * dd);
        dp[2][i] = dd * dd;
    }
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
        for (int j = 0; j < 3; j++)
#pragma unroll
            for (int i = 0; i < 3; i++)
                w[9 * k + 3 * j + i] = (R)0.125 * dp[i][0] * dp[j][1] * dp[k][2];
    return c;
}

// weights per particle of interpolation scheme J (jmass 2 or 3)
template <int J> struct P3NW { static constexpr int v = J == 3 ? 27 : 8; };

template <typename R, int J>
__device__ __forceinline__ int cellJ(R u, R t, R z, const P3Args<R> &a, bool fold, R *w,
                                     int &ring, int &plane, int &spoke) {
    if constexpr (J == 3) return cell3q<R>(u, t, z, a, fold, w, ring, plane, spoke);
    else return cell3<R>(u, t, z, a, fold, w, ring, plane, spoke);
}
template <typename R, int J>
__device__ __forceinline__ int cellJ(R u, R t, R z, const P3Args<R> &a, bool fold, R *w) {
    int rg, pl, sp;
    return cellJ<R, J>(u, t, z, a, fold, w, rg, pl, sp);
}

// grdacc's corners (grdacc.f:193-203): c, c+1 (wrapped when c+1 is a
// multiple of na), the next plane (+na), and the same one ring out.
__device__ __forceinline__ void corners3(int c, int spoke, int na, int ngz, int ic[8]) {
    const bool wrap = spoke + 1 == na;   // (c + 1) mod na == 0
    ic[0] = c;
    ic[1] = wrap ? c + 1 - na : c + 1;
    ic[2] = c + na;
    ic[3] = ic[1] + na;
    ic[4] = c + ngz * na;
    ic[5] = wrap ? ic[4] + 1 - na : ic[4] + 1;
    ic[6] = ic[4] + na;
    ic[7] = ic[5] + na;
}

// the 27 cells about a central cell c (grdacc.f:262-275, massgn.f:182-195):
// spoke -1, 0, +1 (with the azimuthal wrap) fastest, then plane -1, 0, +1
// (-+ na), then ring -1, 0, +1 (-+ ngz na), as flat indices
__device__ __forceinline__ void corners27(int c, int spoke, int na, int ngz, int ic[27]) {
    int m[3];
    m[1] = c;
    m[0] = spoke == 0 ? c - 1 + na : c - 1;          // c mod na == 0
    m[2] = spoke + 1 == na ? c + 1 - na : c + 1;     // (c + 1) mod na == 0
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
        for (int j = 0; j < 3; j++)
#pragma unroll
            for (int i = 0; i < 3; i++) ic[9 * k + 3 * j + i] = m[i] + (j - 1) * na + (k - 1) * ngz * na;
}

// spoke: the central cell's spoke as cellJ gives it (c mod na, or na for a
// linear cell at t == na)
template <int J> __device__ __forceinline__ void cornersJ(int c, int spoke, int na, int ngz,
                                                          int *ic) {
    if constexpr (J == 3) corners27(c, spoke, na, ngz, ic);
    else corners3(c, spoke, na, ngz, ic);
}

template <typename R>
__device__ __forceinline__ R tsfac3(const P3Args<R> &a, int iz) {
    return (R)a.nstep[(iz < a.nz ? iz : a.nz) - 1];
}

// zoneno with spherical boundaries (zoneno.f:40-55, setrun.f:810)
template <typename R>
__device__ __forceinline__ int zoneno3(const P3Args<R> &a, R x, R y, R z) {
    const R rad = sqrt(x * x + y * y + z * z);
    int k = 1;
    if (a.nz > 1)
        for (int i = 1; i <= a.nz; i++)
            if (rad >= a.zbr[i - 1]) k = i;
    return k;
}

// chstep, RVLF branch (chstep.f:22-31), one component
template <typename R>
__device__ __forceinline__ void chstep1(R tfac, R ox, R &ov, R &acc, R &nx, R &nv) {
    const R h = (R)0.5;
    ov = tfac * (ov + h * acc * ((R)1 - tfac));
    acc = acc * tfac * tfac;
    nv = ov + acc;
    nx = ox + nv;
}

// getacc of one particle of list z0: the three Cartesian components
// interpolated at its position (component sums in corner order), the fixed
// force of the spherical radius, the zone rescaling.
template <typename R, int J>
__device__ __forceinline__ void getacc3(const P3Args<R> &a, R ox, R oy, R oz, R u, R t,
                                        int z0, R acc[3]) {
    typedef typename Vec4<R>::T R4;
    constexpr int NW = P3NW<J>::v;
    acc[0] = acc[1] = acc[2] = (R)0;
    if (z0 != a.nz + 1) {
        R w[NW];
        int ic[NW];
        int rg, pl, sp;
        const int c = cellJ<R, J>(u, t, oz, a, false, w, rg, pl, sp);
        cornersJ<J>(c, sp, a.na, a.ngz, ic);
        R sx = (R)0, sy = (R)0, sz = (R)0;
#pragma unroll
        for (int q = 0; q < NW; q++) {
            // a corner outside the field array (the last ring with t == na;
            // for jmass 3 also the first ring's bottom plane) reads 0: cell 0
            // is read instead and its value multiplied by 0 (unconditional
            // loads, no per-corner branch)
            const bool in = ic[q] >= 0 && ic[q] < a.mesh;
            const R4 f = a.fld[in ? ic[q] : 0];
            const R wq = in ? w[q] : (R)0;
            sx = sx + f.x * wq;
            sy = sy + f.y * wq;
            sz = sz + f.z * wq;
        }
        if (a.fixforce) {
            R r2 = (R)0;
            r2 = r2 + ox * ox;
            r2 = r2 + oy * oy;
            r2 = r2 + oz * oz;
            const R rr = sqrt(r2);
            R fr = (R)fixforce((double)rr, a.nfr, a.fr_inv_dr, a.fr_c);
            if (fr != (R)0) {
                fr = fr / rr;
                sx = sx + fr * ox;
                sy = sy + fr * oy;
                sz = sz + fr * oz;
            }
        }
        acc[0] = sx;
        acc[1] = sy;
        acc[2] = sz;
    }
    if (a.nz > 1 && z0 > 1) {
        R t = tsfac3(a, z0);
        t = t * t;
        acc[0] = t * acc[0];
        acc[1] = t * acc[1];
        acc[2] = t * acc[2];
    }
}

__device__ __forceinline__ int p3_wave_min(int v) { return wave_min_i32(v); }
__device__ __forceinline__ int p3_wave_max(int v) { return wave_max_i32(v); }

#ifndef GH_P3_TILE
#define GH_P3_TILE 4096
#endif
#ifndef GH_P3_PASSES
#define GH_P3_PASSES 8
#endif
constexpr int P3_TILE = GH_P3_TILE;       // cells of a block's LDS deposit window
constexpr int P3_PASSES = GH_P3_PASSES;   // window passes before the direct fallback
constexpr int P3_SB = 13;                 // ints of the block scratch s_b

// Block-level fixed-point deposit.  The block's particles (sorted by cell,
// so a block covers a few neighbouring cells) are accumulated into an LDS
// window spanning the (ring, plane, spoke) box of their corners, then the
// window's non-zero cells are added to the zone array with one global atomic
// each (consecutive cells of a row, so the atomics of a wave coalesce).  The
// corners of a central cell are the next ring, plane and spoke (jmass 2) or
// the cells one either side (jmass 3); spokes wrap around the ring, so the
// window's spoke axis runs past na (or below 0) and the flush folds it back.
// A box larger than P3_TILE cells is clipped to at most 8 rings and 8
// planes, then as many spokes as fit, placed about a reference particle:
// the middle thread's, or the first pending one after it in thread order.
// The particles it leaves out (outliers that have moved away from the
// block's cells, or the far rows of a block in a sparse region) get the
// next pass's window over what remains, up to P3_PASSES passes; the
// reference particle is always inside its window.
// Particles whose corners leave the row structure (the plane next to the
// last or, for jmass 3, the first; a linear cell at t == na), those of
// another zone than the block's lowest, and those still left after the last
// pass add their corners to the zone array directly.  Integer sums: the
// result does not depend on the grouping.  Every thread of the block must
// call it (on = false for threads without a deposit).  s_b: P3_SB ints,
// s_b[0] = INT_MAX on entry; s_jr/s_kr receive the block's central-ring
// minimum and maximum per zone.
template <typename R, int J>
__device__ __forceinline__ void deposit3_block(const P3Args<R> &a, bool on, int z, int c,
                                               int ring, int plane, int spoke,
                                               const R *w, long long *tile, int *s_b,
                                               unsigned long long *s_ref, int *s_jr,
                                               int *s_kr) {
    constexpr int NW = P3NW<J>::v;
    constexpr int LO = J == 3 ? 1 : 0;   // corner cells below the central one
    const int tid = threadIdx.x, lane = tid & 63;
    // reference key: thread order from the middle thread, then the cell
    const unsigned long long rkey =
        ((unsigned long long)((tid - P3_PBLOCK / 2) & (P3_PBLOCK - 1)) << 40) |
        ((unsigned long long)(unsigned)ring << 26) | ((unsigned long long)(unsigned)plane << 13) |
        (unsigned long long)(unsigned)spoke;
    if (!on) ring = plane = spoke = 0;
    // the block's zone: the lowest among its deposits
    int zw = p3_wave_min(on ? z : INT_MAX);
    if (lane == 0 && zw != INT_MAX) atomicMin(&s_b[0], zw);
    if (tid == 0) {
        // two (min, max) x (ring, plane, spoke) box buffers, used by
        // alternate passes so that one can be reset while the other is read
#pragma unroll
        for (int k = 1; k < P3_SB; k++) s_b[k] = (k & 1) ? INT_MAX : -1;
        s_ref[0] = s_ref[1] = ~0ull;
    }
    __syncthreads();
    const int tz = s_b[0];
    bool pend = on && z == tz && plane - LO >= 0 && plane + 1 < a.ngz && spoke < a.na &&
                ring - LO >= 0;
    {
        // ring ranges: the block zone's by wave reduction, others directly
        const int jr = p3_wave_min(on && z == tz ? ring : INT_MAX);
        const int kr = p3_wave_max(on && z == tz ? ring : -1);
        if (lane == 0 && kr >= 0) {
            atomicMin(&s_jr[tz - 1], jr);
            atomicMax(&s_kr[tz - 1], kr);
        }
        if (on && z != tz) {
            atomicMin(&s_jr[z - 1], ring);
            atomicMax(&s_kr[z - 1], ring);
        }
    }
    long long v[NW];
    quant_n<R, NW>(w, a.qbits, on, v);
    bool direct = on && !pend;
    for (int pass = 0;; pass++) {
        int *const sb = s_b + 1 + 6 * (pass & 1);
        {
            const int r0 = p3_wave_min(pend ? ring : INT_MAX), r1 = p3_wave_max(pend ? ring : -1);
            const int p0 = p3_wave_min(pend ? plane : INT_MAX), p1 = p3_wave_max(pend ? plane : -1);
            const int q0 = p3_wave_min(pend ? spoke : INT_MAX), q1 = p3_wave_max(pend ? spoke : -1);
            // the wave's first pending lane holds its lowest reference key
            const unsigned long long bal = __ballot(pend);
            const int fl = bal ? __ffsll((long long)bal) - 1 : 0;
            const unsigned klo = __builtin_amdgcn_readlane((unsigned)rkey, fl);
            const unsigned khi = __builtin_amdgcn_readlane((unsigned)(rkey >> 32), fl);
            if (lane == 0 && r1 >= 0) {
                atomicMin(&sb[0], r0); atomicMax(&sb[1], r1);
                atomicMin(&sb[2], p0); atomicMax(&sb[3], p1);
                atomicMin(&sb[4], q0); atomicMax(&sb[5], q1);
                atomicMin(&s_ref[pass & 1], ((unsigned long long)khi << 32) | klo);
            }
        }
        __syncthreads();
        // the box of the pending corners, clipped from its low corner
        int r0 = sb[0] - LO, p0 = sb[2] - LO, q0 = sb[4] - LO;
        int NR = sb[1] - sb[0] + 2 + LO, NP = sb[3] - sb[2] + 2 + LO;
        int NQ = sb[5] - sb[4] + 2 + LO;
        const bool use = sb[1] >= 0;
        const unsigned long long kref = s_ref[pass & 1];
        if (tid == 0) {
            int *const nb = s_b + 1 + 6 * ((pass + 1) & 1);   // last read in the previous pass
#pragma unroll
            for (int k = 0; k < 6; k++) nb[k] = (k & 1) ? -1 : INT_MAX;
            s_ref[(pass + 1) & 1] = ~0ull;
        }
        if (use && NR * NP * NQ > P3_TILE) {
            // about the reference cell, inside the box
            const int rr = (int)((kref >> 26) & 0x1FFF), rp = (int)((kref >> 13) & 0x1FFF);
            const int rq = (int)(kref & 0x1FFF);
            if (NR > 8) {
                r0 = min(max(rr - LO - 3, r0), r0 + NR - 8);
                NR = 8;
            }
            if (NP > 8) {
                p0 = min(max(rp - LO - 3, p0), p0 + NP - 8);
                NP = 8;
            }
            const int mq = P3_TILE / (NR * NP);
            if (NQ > mq) {
                q0 = min(max(rq - LO - (mq - 2 - LO) / 2, q0), q0 + NQ - mq);
                NQ = mq;
            }
        }
        const int ncell = use ? NR * NP * NQ : 0;
        for (int e = tid; e < ncell; e += P3_PBLOCK) tile[e] = 0;
        __syncthreads();
        const bool inside = pend && ring - LO >= r0 && ring - r0 + 1 < NR && plane - LO >= p0 &&
                            plane - p0 + 1 < NP && spoke - LO >= q0 && spoke - q0 + 1 < NQ;
        if (inside) {
            const int b = ((ring - r0) * NP + (plane - p0)) * NQ + (spoke - q0);
#pragma unroll
            for (int q = 0; q < NW; q++) {
                int o;
                if constexpr (J == 3)
                    o = b + (q % 3 - 1) + ((q / 3) % 3 - 1) * NQ + (q / 9 - 1) * NP * NQ;
                else
                    o = b + ((q & 1) ? 1 : 0) + ((q & 2) ? NQ : 0) + ((q & 4) ? NP * NQ : 0);
                if (v[q] != 0) atomicAdd((unsigned long long *)&tile[o], (unsigned long long)v[q]);
            }
            pend = false;
        }
        __syncthreads();
        // flush: the window's cells over all threads in storage order (a
        // row's cells are consecutive spokes, so a wave's atomics coalesce),
        // one global atomic per non-zero cell; the (ring, plane, spoke)
        // offsets of cell e advanced by the block size with carries instead
        // of divided out per cell; window spokes q0 + dq lie in [-1, na],
        // folded back onto the ring
        if (use) {
            const int t1 = div_small(P3_PBLOCK, NQ), sq = P3_PBLOCK - t1 * NQ;
            const int sr = div_small(t1, NP), spl = t1 - sr * NP;
            const int u1 = div_small(tid, NQ);
            int dq = tid - u1 * NQ, dr = div_small(u1, NP);
            int dp = u1 - dr * NP;
            for (int e = tid; e < ncell; e += P3_PBLOCK) {
                const long long tv = tile[e];
                if (tv != 0) {
                    int sp = q0 + dq;
                    sp = sp >= a.na ? sp - a.na : sp < 0 ? sp + a.na : sp;
                    atomicAdd((unsigned long long *)&a.mfx[tz - 1][((r0 + dr) * a.ngz + (p0 + dp)) * a.na + sp],
                              (unsigned long long)tv);
                }
                dq += sq;
                const int c1 = dq >= NQ ? 1 : 0;
                dq -= c1 * NQ;
                dp += spl + c1;
                const int c2 = dp >= NP ? 1 : 0;
                dp -= c2 * NP;
                dr += sr + c2;
            }
        }
        if (pass + 1 >= P3_PASSES) {
            direct = direct || pend;
            break;
        }
        // every thread reads the flag of the same barrier: a uniform exit
        if (!__syncthreads_or(pend ? 1 : 0)) break;
    }
    if (direct) {
        int ic[NW];
        cornersJ<J>(c, spoke, a.na, a.ngz, ic);
#pragma unroll
        for (int q = 0; q < NW; q++)
            if (v[q] != 0) deposit_global(a.mfx, a.nz, a.mesh, z, ic[q], v[q]);
    }
}

// jrad/krad from the block's central rings: 1-based rings j+1..j+2 about a
// central ring j (massgn.f:146-149), j..j+2 for jmass 3 (massgn.f:176-179)
__device__ __forceinline__ void flush_ranges(int nz, int nr, int lo, const int *s_jr,
                                             const int *s_kr, int *jrad, int *krad) {
    const int tid = threadIdx.x;
    if (tid < nz && s_kr[tid] >= 0) {
        atomicMin(&jrad[tid], s_jr[tid] + 1 - lo);
        const int k = s_kr[tid] + 2;
        atomicMax(&krad[tid], k < nr ? k : nr);
    }
}

template <typename R, int J>
__global__ __launch_bounds__(P3_PBLOCK) void p3_massgn_kernel(P3Args<R> a) {
    __shared__ int s_jr[GH_MAXZ], s_kr[GH_MAXZ], s_b[P3_SB];
    __shared__ unsigned long long s_ref[2];
    __shared__ long long tile[P3_TILE];
    const int tid = threadIdx.x;
    if (tid < GH_MAXZ) {
        s_jr[tid] = INT_MAX;
        s_kr[tid] = -1;
    }
    if (tid == 0) s_b[0] = INT_MAX;
    __syncthreads();
    const long long i = (long long)blockIdx.x * P3_PBLOCK + tid;
    bool on = false;
    int z = 0, c = 0, rg = 0, pl = 0, sp = 0;
    R w[P3NW<J>::v];
    if (i < a.n) {
        z = a.zone[i];
        const typename Vec4<R>::T p = a.rec[2 * i];
        R u, t;
        ut3(p.x, p.y, a, u, t);
        p3_set_ut(a, i, u, t);
        if (z >= 1 && z <= a.nz) {
            c = cellJ<R, J>(u, t, p.z, a, a.nsect > 1, w, rg, pl, sp);
            on = true;
        }
    }
    deposit3_block<R, J>(a, on, z, c, rg, pl, sp, w, tile, s_b, s_ref, s_jr, s_kr);
    __syncthreads();
    flush_ranges(a.nz, a.nr, J == 3 ? 1 : 0, s_jr, s_kr, a.jrad, a.krad);
}

template <typename R, int J>
__global__ __launch_bounds__(P3_PBLOCK) void p3_push_kernel(P3Args<R> a) {
    typedef typename Vec4<R>::T R4;
    typedef typename Vec2<R>::T R2;
    __shared__ int s_jr[GH_MAXZ], s_kr[GH_MAXZ], s_b[P3_SB];
    __shared__ unsigned long long s_ref[2];
    __shared__ int s_dn, s_np;
    __shared__ long long tile[P3_TILE];
    const int tid = threadIdx.x;
    if (tid < GH_MAXZ) {
        s_jr[tid] = INT_MAX;
        s_kr[tid] = -1;
    }
    if (tid == 0) {
        s_dn = s_np = 0;
        s_b[0] = INT_MAX;
    }
    __syncthreads();
    const long long i = (long long)blockIdx.x * P3_PBLOCK + tid;
    int dn = 0, np = 0;
    bool don = false;
    int dz = 0, dc = 0, drg = 0, dpl = 0, dsp = 0;
    R dw[P3NW<J>::v];
    if (i < a.n) {
        const int nz = a.nz, nlists = nz + 1;
        const int z0 = a.zone[i];
        const bool offlist = z0 == nlists;
        const bool move = offlist ? (a.offanal && a.mzone == nz) : z0 <= a.mzone;
        if (move) {
            np = 1;
            const R4 p = a.rec[2 * i];
            const R4 q4 = a.rec[2 * i + 1];
            const R2 q = mk2(q4.x, q4.y);
            const R2 ut0 = mk2(q4.z, q4.w);
            R o[3] = {p.x, p.y, p.z};
            R ov[3] = {p.w, q.x, q.y};
            R acc[3], nv[3], nx[3];
            getacc3<R, J>(a, o[0], o[1], o[2], ut0.x, ut0.y, z0, acc);
            // stpgrp
#pragma unroll
            for (int d = 0; d < 3; d++) {
                nv[d] = ov[d] + acc[d];
                nx[d] = o[d] + nv[d];
            }
            int znew = z0;
            const R rnew = sqrt(nx[0] * nx[0] + nx[1] * nx[1]);
            const R az = nx[2] < (R)0 ? -nx[2] : nx[2];
            const bool off = (az > a.zm) || (rnew > a.rgrid);
            if (!offlist) {
                if (off) {
                    znew = nlists;
                    dn = 1;
                    if (a.offanal && z0 < nz) {
                        const R tfac = tsfac3(a, nz) / tsfac3(a, z0);
#pragma unroll
                        for (int d = 0; d < 3; d++) chstep1(tfac, o[d], ov[d], acc[d], nx[d], nv[d]);
                    }
                }
            } else if (!off) {
                dn = -1;
                znew = nz;
            }
            // rezgrp
            if (a.mzone > 1 && !offlist && znew <= a.mzone) {
                int kz = zoneno3(a, nx[0], nx[1], nx[2]);
                if (kz > a.mzone) kz = a.mzone;
                if (kz > znew) kz = zoneno3(a, o[0], o[1], o[2]);
                if (kz != znew) {
                    const R tfac = tsfac3(a, kz) / tsfac3(a, znew);
#pragma unroll
                    for (int d = 0; d < 3; d++) chstep1(tfac, o[d], ov[d], acc[d], nx[d], nv[d]);
                    znew = kz;
                }
            }
            R un, tn;
            ut3(nx[0], nx[1], a, un, tn);
            a.rec[2 * i] = mk4(nx[0], nx[1], nx[2], nv[0]);
            a.rec[2 * i + 1] = mk4(nv[1], nv[2], un, tn);
            a.zone[i] = (uint8_t)znew;
            if (znew >= 1 && znew <= nz) {
                dc = cellJ<R, J>(un, tn, nx[2], a, a.nsect > 1, dw, drg, dpl, dsp);
                dz = znew;
                don = true;
            }
        }
    }
    deposit3_block<R, J>(a, don, dz, dc, drg, dpl, dsp, dw, tile, s_b, s_ref, s_jr, s_kr);
    // block sums of the off-grid change and the pushes
    dn = wave_sum_i32(dn);
    np = wave_sum_i32(np);
    if ((tid & 63) == 0) {
        if (dn) atomicAdd(&s_dn, dn);
        if (np) atomicAdd(&s_np, np);
    }
    __syncthreads();
    flush_ranges(a.nz, a.nr, J == 3 ? 1 : 0, s_jr, s_kr, a.jrad, a.krad);
    if (tid == 0) {
        if (s_dn) atomicAdd((unsigned long long *)&a.counters[0], (unsigned long long)(long long)s_dn);
        if (s_np) atomicAdd((unsigned long long *)&a.counters[1], (unsigned long long)(long long)s_np);
    }
}

// (u, t) of every particle's position (after a load)
template <typename R>
__global__ __launch_bounds__(P3_PBLOCK) void p3_ut_kernel(P3Args<R> a) {
    const long long i = (long long)blockIdx.x * P3_PBLOCK + threadIdx.x;
    if (i >= a.n) return;
    const typename Vec4<R>::T p = a.rec[2 * i];
    R u, t;
    ut3(p.x, p.y, a, u, t);
    p3_set_ut(a, i, u, t);
}

template <typename R, int J>
__global__ __launch_bounds__(P3_PBLOCK) void p3_tmcent_kernel(P3Args<R> a) {
    const long long i = (long long)blockIdx.x * P3_PBLOCK + threadIdx.x;
    if (i >= a.n) return;
    typename Vec4<R>::T p = a.rec[2 * i];
    typename Vec4<R>::T q = a.rec[2 * i + 1];   // (vy, vz, u, t)
    R acc[3];
    getacc3<R, J>(a, p.x, p.y, p.z, q.z, q.w, a.zone[i], acc);
    p.w = p.w - a.fac * acc[0];
    q.x = q.x - a.fac * acc[1];
    q.y = q.y - a.fac * acc[2];
    a.rec[2 * i] = p;
    a.rec[2 * i + 1] = q;
}

template <typename R, int J>
__global__ __launch_bounds__(P3_PBLOCK) void p3_gather_kernel(P3GatherArgs<R> b) {
    typedef typename Vec4<R>::T R4;
    constexpr int NW = P3NW<J>::v;
    const long long i = (long long)blockIdx.x * P3_PBLOCK + threadIdx.x;
    if (i >= b.n) return;
    R w[NW];
    R u, t;
    ut3(b.x[i], b.y[i], b.g, u, t);
    int rg, pl, sp;
    const int c = cellJ<R, J>(u, t, b.z[i], b.g, b.fold != 0, w, rg, pl, sp);
    if (b.cell) b.cell[i] = c;
    if (b.wt)
        for (int q = 0; q < NW; q++) b.wt[NW * i + q] = w[q];
    if (b.acc) {
        int ic[NW];
        cornersJ<J>(c, sp, b.g.na, b.g.ngz, ic);
        R s[3] = {(R)0, (R)0, (R)0};
        for (int q = 0; q < NW; q++) {
            R4 f = (ic[q] >= 0 && ic[q] < b.g.mesh) ? b.fld[ic[q]] : mk4((R)0, (R)0, (R)0, (R)0);
            s[0] = s[0] + f.x * w[q];
            s[1] = s[1] + f.y * w[q];
            s[2] = s[2] + f.z * w[q];
        }
        for (int d = 0; d < 3; d++) b.acc[3 * i + d] = s[d];
    }
}

// icheck's global integrals for the 3-D grid (SRC/lib15/icheck.f:24-70):
// per-block FP64 sums of 11 quantities in a fixed order, then one block adds
// the block sums in order.  [0] count, [1..3] position, [4..6] velocity per
// zone-1 step ((v + a/2)/tfac when centred), [7] v^2, [8] Lz, [9] the grid
// potential interpolated with the particle's own weights (grid particles),
// [10] the fixed potential of the spherical radius.
constexpr int NINT3 = 11;

__device__ __forceinline__ double p3_wave_sumd(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

template <typename R, int J>
__global__ __launch_bounds__(P3_PBLOCK) void p3_integrals_kernel(P3Args<R> a, int with_pot,
                                                                 const double4 *fp_c, int nfp,
                                                                 double fp_dr, int centred,
                                                                 double *part) {
    typedef typename Vec4<R>::T R4;
    __shared__ double red[P3_PBLOCK / 64][NINT3];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const long long i = (long long)blockIdx.x * P3_PBLOCK + tid;
    double v[NINT3];
#pragma unroll
    for (int q = 0; q < NINT3; q++) v[q] = 0.0;
    if (i < a.n) {
        const int z = a.zone[i];
        if (z >= 1) {
            const R4 p = a.rec[2 * i];
            const R4 q4 = a.rec[2 * i + 1];
            const typename Vec2<R>::T q2 = mk2(q4.x, q4.y);
            const typename Vec2<R>::T ut = mk2(q4.z, q4.w);
            R vx = p.w, vy = q2.x, vz = q2.y;
            if (centred) {
                R acc[3];
                getacc3<R, J>(a, p.x, p.y, p.z, ut.x, ut.y, z, acc);
                vx = vx + (R)0.5 * acc[0];
                vy = vy + (R)0.5 * acc[1];
                vz = vz + (R)0.5 * acc[2];
            }
            const int zs = z <= a.nz ? z : a.nz;
            const double sf = 1.0 / (double)a.nstep[zs - 1];
            const double x = p.x, y = p.y, zz = p.z;
            const double ux = vx * sf, uy = vy * sf, uz = vz * sf;
            v[0] = 1.0;
            v[1] = x;
            v[2] = y;
            v[3] = zz;
            v[4] = ux;
            v[5] = uy;
            v[6] = uz;
            v[7] = ux * ux + uy * uy + uz * uz;
            v[8] = x * uy - y * ux;
            if (with_pot && z <= a.nz) {
                constexpr int NW = P3NW<J>::v;
                R w[NW];
                int ic[NW];
                int rg, pl, sp;
                const int c = cellJ<R, J>(ut.x, ut.y, p.z, a, false, w, rg, pl, sp);
                cornersJ<J>(c, sp, a.na, a.ngz, ic);
                double ph = 0.0;
                for (int k = 0; k < NW; k++)
                    if (ic[k] >= 0 && ic[k] < a.mesh) ph += (double)w[k] * (double)a.fld[ic[k]].w;
                v[9] = ph;
            }
            if (nfp > 0)
                v[10] = fixforce(sqrt(x * x + y * y + zz * zz), nfp, 1.0 / fp_dr, fp_c);
        }
    }
#pragma unroll
    for (int q = 0; q < NINT3; q++) {
        const double t = p3_wave_sumd(v[q]);
        if (lane == 0) red[wave][q] = t;
    }
    __syncthreads();
    if (tid < NINT3) {
        double t = 0.0;
        for (int w = 0; w < P3_PBLOCK / 64; w++) t += red[w][tid];
        part[(long long)blockIdx.x * NINT3 + tid] = t;
    }
}

__global__ __launch_bounds__(P3_BLOCK) void p3_integrals_sum_kernel(const double *part, int nb,
                                                                    double *out) {
    __shared__ double red[P3_BLOCK / 64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int q = 0; q < NINT3; q++) {
        double t = 0.0;
        for (int b = tid; b < nb; b += P3_BLOCK) t += part[(long long)b * NINT3 + q];
        t = p3_wave_sumd(t);
        if (lane == 0) red[wave] = t;
        __syncthreads();
        if (tid == 0) {
            double s = 0.0;
            for (int w = 0; w < P3_BLOCK / 64; w++) s += red[w];
            out[q] = s;
        }
        __syncthreads();
    }
}

// Sort keys: the particle's cell shifted right by kshift (kmax for the
// off-grid list, which sorts last), with the identity permutation.
template <typename R, int J>
__global__ __launch_bounds__(P3_PBLOCK) void p3_key_kernel(P3Args<R> a, int kshift,
                                                          unsigned kmax, unsigned *key,
                                                          unsigned *perm) {
    const long long i = (long long)blockIdx.x * P3_PBLOCK + threadIdx.x;
    if (i >= a.n) return;
    const int z = a.zone[i];
    unsigned k = kmax;
    if (z >= 1 && z <= a.nz) {
        const typename Vec4<R>::T p = a.rec[2 * i];
        const typename Vec4<R>::T q = a.rec[2 * i + 1];
        const typename Vec2<R>::T ut = mk2(q.z, q.w);
        R w[P3NW<J>::v];
        k = (unsigned)(cellJ<R, J>(ut.x, ut.y, p.z, a, false, w) >> kshift);
    }
    key[i] = k;
    perm[i] = (unsigned)i;
}

template <typename R>
__global__ __launch_bounds__(P3_PBLOCK) void p3_permute_kernel(P3PermArgs<R> p) {
    const long long i = (long long)blockIdx.x * P3_PBLOCK + threadIdx.x;
    if (i >= p.n) return;
    const unsigned j = p.perm[i];
    p.rec_out[2 * i] = p.rec_in[2 * (long long)j];
    p.rec_out[2 * i + 1] = p.rec_in[2 * (long long)j + 1];
    p.zout[i] = p.zin[j];
    p.iout[i] = p.iin[j];
}

// load: planar host arrays staged on the device (x, y, z, vx in s4 and vy, vz
// in s2, n values each, one after another) -> particle records ((u, t) set
// by p3_ut_kernel afterwards); ids = load order
template <typename R>
__global__ __launch_bounds__(P3_PBLOCK) void p3_interleave_kernel(const R *__restrict__ s4,
                                                                  const R *__restrict__ s2,
                                                                  typename Vec4<R>::T *__restrict__ rec,
                                                                  unsigned *__restrict__ id,
                                                                  long long n) {
    const long long i = (long long)blockIdx.x * P3_PBLOCK + threadIdx.x;
    if (i >= n) return;
    typename Vec4<R>::T a;
    a.x = s4[i];
    a.y = s4[n + i];
    a.z = s4[2 * n + i];
    a.w = s4[3 * n + i];
    rec[2 * i] = a;
    rec[2 * i + 1] = mk4(s2[i], s2[n + i], (R)0, (R)0);
    id[i] = (unsigned)i;
}

// unload: interleaved storage -> planar arrays in load order, zones widened
template <typename R>
__global__ __launch_bounds__(P3_PBLOCK) void p3_deinterleave_kernel(
    const typename Vec4<R>::T *__restrict__ rec, const uint8_t *__restrict__ zone,
    const unsigned *__restrict__ id, R *__restrict__ d4, R *__restrict__ d2,
    int *__restrict__ iz, long long n) {
    const long long i = (long long)blockIdx.x * P3_PBLOCK + threadIdx.x;
    if (i >= n) return;
    const long long p = id[i];
    const typename Vec4<R>::T a = rec[2 * i];
    const typename Vec4<R>::T b = rec[2 * i + 1];
    d4[p] = a.x;
    d4[n + p] = a.y;
    d4[2 * n + p] = a.z;
    d4[3 * n + p] = a.w;
    d2[p] = b.x;
    d2[n + p] = b.y;
    iz[p] = zone[i];
}

// ---- field solve ------------------------------------------------------------------

// grid (ngz, nr): block (plane, ring-1)
template <typename R>
__global__ __launch_bounds__(P3_BLOCK) void p3_analyse_kernel(P3SolveArgs<R> a) {
    __shared__ R rho_s[P3_MAX_NA];
    __shared__ R tws[2 * P3_MAX_NA];
    __shared__ int s_rng[4], s_ok;
    const int tid = threadIdx.x, na = a.na;
    const int plane = blockIdx.x, ir = blockIdx.y + 1;
    if (tid == 0) {
        int kr[GH_MAXZ], rg[4] = {0, 0, 0, 0};
        const bool ok = solve_ranges(a.jrad, a.krad, a.nz, a.nr, a.mzone, a.lpot,
                                     a.explicit_zones, a.jzone, a.kzone, kr, rg);
        s_ok = ok;
        for (int q = 0; q < 4; q++) s_rng[q] = rg[q];
        if (plane == 0 && ir == 1) {
            for (int q = 0; q < 4; q++) a.ranges[q] = ok ? rg[q] : -1;
            if (ok)
                for (int z = 1; z < a.nz; z++) a.krad[z] = kr[z];
        }
    }
    __syncthreads();
    // zoncmb + scaled for this ring and plane
    const size_t row = ((size_t)(ir - 1) * a.ngz + plane) * na;
    for (int j = tid; j < na; j += P3_BLOCK) {
        const size_t idx = row + j;
        double acc = (double)a.mfx[0][idx] * a.q;
        for (int z = 2; z <= a.nz; z++) {
            if (a.use_new[z - 1]) acc = acc + a.fnew[z - 1] * ((double)a.mfx[z - 1][idx] * a.q);
            if (a.use_old[z - 1]) acc = acc + a.fold[z - 1] * ((double)a.moz[z - 1][idx] * a.q);
        }
        const R v = (R)(acc * a.pmass);
        rho_s[j] = v;
        a.rho[idx] = v;
    }
    for (int j = tid; j < 2 * na; j += P3_BLOCK) tws[j] = a.tw[j];
    __syncthreads();
    if (!s_ok || ir < s_rng[0] || ir > s_rng[1]) return;
    const int wave = tid >> 6, lane = tid & 63, nw = P3_BLOCK / 64;
    for (int s = wave; s < a.nm; s += nw) {
        const int m = a.slot_m[s], part = a.slot_p[s];
        R sa = 0, sb = 0;
        int k = (int)(((long long)m * lane) % na);
        const int step = (int)(((long long)m * 64) % na);
        for (int j = lane; j < na; j += 64) {
            const R rv = rho_s[j];
            sa += rv * tws[2 * k];
            sb += rv * tws[2 * k + 1];
            k += step;
            if (k >= na) k -= na;
        }
        sa = p3_wave_sum(sa);
        sb = p3_wave_sum(sb);
        if (lane == 0) a.A[((size_t)(ir - 1) * a.nm + s) * a.ngz + plane] = part ? -sb : sa;
    }
}

// grid (nr): the doubled vertical transform of each component column
// The vertical twiddles (cos, sin of 2 pi k / P) in LDS with one pad entry
// after every 256 bytes: the transforms read them at index p k mod P across
// consecutive k (or q), a stride that puts many lanes on one bank when it is
// a multiple of a large power of two; the pad spreads those over the banks.
template <typename R> constexpr int tz_shift() { return sizeof(R) == 4 ? 5 : 4; }
template <typename R> __host__ __device__ constexpr size_t tz_words(size_t P) {
    return 2 * (P + (P >> tz_shift<R>()) + 1);
}
template <typename R> __device__ __forceinline__ int tz_at(int idx) {
    return 2 * (idx + (idx >> tz_shift<R>()));
}
template <typename R>
__device__ __forceinline__ void tz_stage(const R *twz, int P, R *tz) {
    for (int e = threadIdx.x; e < P; e += P3_BLOCK) {
        tz[tz_at<R>(e)] = twz[2 * e];
        tz[tz_at<R>(e) + 1] = twz[2 * e + 1];
    }
}

template <typename R>
__global__ __launch_bounds__(P3_BLOCK) void p3_zfwd_kernel(P3SolveArgs<R> a) {
    extern __shared__ unsigned char p3_lds[];
    R *col = (R *)p3_lds;                     // [nm][ngz]
    const int ngz = a.ngz, P = 2 * ngz, nm = a.nm;
    R *tz = col + (size_t)nm * ngz;           // [P][2]
    const int ir = blockIdx.x + 1, tid = threadIdx.x;
    const int jrs = a.ranges[0], krs = a.ranges[1];
    if (jrs < 1 || ir < jrs || ir > krs) return;
    const R *src = a.A + (size_t)(ir - 1) * nm * ngz;
    for (int e = tid; e < nm * ngz; e += P3_BLOCK) col[e] = src[e];
    tz_stage<R>(a.twz, P, tz);
    __syncthreads();
    R *dst = a.M + (size_t)(ir - 1) * nm * P;
    for (int o = tid; o < nm * (ngz + 1); o += P3_BLOCK) {
        const int s = o / (ngz + 1), k = o - s * (ngz + 1);
        const R *c = col + (size_t)s * ngz;
        R sa = 0, sb = 0;
        int idx = 0;
        for (int p = 0; p < ngz; p++) {
            const int t = tz_at<R>(idx);
            sa += c[p] * tz[t];
            sb += c[p] * tz[t + 1];
            idx += k;
            if (idx >= P) idx -= P;
        }
        R *h = dst + (size_t)s * P;
        if (k == 0) h[0] = sa;
        else if (k == ngz) h[P - 1] = sa;
        else {
            h[2 * k - 1] = sa;
            h[2 * k] = -sb;
        }
    }
}

// grid (ng, ngz + 1, ntypes): block (m, k_z, type) applies one Green's matrix
template <typename R>
__global__ __launch_bounds__(P3_BLOCK) void p3_conv_kernel(P3SolveArgs<R> a) {
    __shared__ R rhs[4][P3_MAX_NR];
    const int im = blockIdx.x + 1, iz = blockIdx.y + 1, type = blockIdx.z;
    const int ng = a.ng, ngz = a.ngz, mmax = a.mmax, nm = a.nm, nr = a.nr;
    if (type == 1 && (im < 2 || im > min(ng, mmax - 1))) return;
    if (type == 2 && (iz < 2 || iz > ngz)) return;
    if (a.lg[im - 1]) return;
    const int jrs = a.ranges[0], krs = a.ranges[1], jrf = a.ranges[2], krf = a.ranges[3];
    if (jrs < 1) return;
    const int P = 2 * ngz;
    const bool nvskip = iz > 1 && iz <= ngz;
    const bool naskip = im > 1 && im < mmax;
    const int p = max(1, 2 * (iz - 1)) - 1;
    const int sc = a.slot_of[im - 1], ss = sc + 1;
    int src[4], dst[4];
    R sgn[4];
    bool val[4];
    dst[0] = sc * P + p; dst[1] = sc * P + p + 1; dst[2] = ss * P + p; dst[3] = ss * P + p + 1;
    if (type == 0 || type == 3) {
        for (int q = 0; q < 4; q++) { src[q] = dst[q]; sgn[q] = (R)1; }
        val[0] = true; val[1] = nvskip; val[2] = naskip; val[3] = naskip && nvskip;
    } else if (type == 1) {
        src[0] = ss * P + p; src[1] = ss * P + p + 1; src[2] = sc * P + p; src[3] = sc * P + p + 1;
        sgn[0] = sgn[1] = (R)1; sgn[2] = sgn[3] = (R)-1;
        val[0] = val[2] = true; val[1] = val[3] = nvskip;
    } else {
        src[0] = sc * P + p + 1; src[1] = sc * P + p; src[2] = ss * P + p + 1; src[3] = ss * P + p;
        sgn[0] = sgn[2] = (R)1; sgn[1] = sgn[3] = (R)-1;
        val[0] = val[1] = true; val[2] = val[3] = naskip;
    }
    const int nrs = krs - jrs + 1;
    for (int e = threadIdx.x; e < 4 * nrs; e += P3_BLOCK) {
        const int q = e / nrs, is = jrs + (e - q * nrs);
        rhs[q][is - 1] = val[q] ? sgn[q] * a.M[(size_t)(is - 1) * nm * P + src[q]] : (R)0;
    }
    __syncthreads();
    const R *Gm = a.G + ((((size_t)type * (ngz + 1) + (iz - 1)) * ng + (im - 1)) * nr) * nr;
    for (int ir = jrf + threadIdx.x; ir <= krf; ir += P3_BLOCK) {
        R s0 = 0, s1 = 0, s2 = 0, s3 = 0;
        const R *g = Gm + (ir - 1);
#pragma unroll 8
        for (int is = jrs; is <= krs; is++) {
            const R gv = g[(size_t)(is - 1) * nr];
            s0 += gv * rhs[0][is - 1];
            s1 += gv * rhs[1][is - 1];
            s2 += gv * rhs[2][is - 1];
            s3 += gv * rhs[3][is - 1];
        }
        R *w = a.W + ((size_t)type * nr + (ir - 1)) * nm * P;
        if (val[0]) w[dst[0]] = s0;
        if (val[1]) w[dst[1]] = s1;
        if (val[2]) w[dst[2]] = s2;
        if (val[3]) w[dst[3]] = s3;
    }
}

// grid (nr, ntypes): inverse vertical transform of one ring's columns
template <typename R>
__global__ __launch_bounds__(P3_BLOCK) void p3_zinv_kernel(P3SolveArgs<R> a) {
    extern __shared__ unsigned char p3_lds[];
    const int ngz = a.ngz, P = 2 * ngz, nm = a.nm;
    R *h = (R *)p3_lds;                       // [nm][P]
    R *tz = h + (size_t)nm * P;               // [P][2]
    const int ir = blockIdx.x + 1, type = blockIdx.y, tid = threadIdx.x;
    const int jrs = a.ranges[0], jrf = a.ranges[2], krf = a.ranges[3];
    if (jrs < 1 || ir < jrf || ir > krf) return;
    const R *src = a.W + ((size_t)type * a.nr + (ir - 1)) * nm * P;
    for (int e = tid; e < nm * P; e += P3_BLOCK) h[e] = src[e];
    tz_stage<R>(a.twz, P, tz);
    __syncthreads();
    R *dst = a.Z + ((size_t)type * a.nr + (ir - 1)) * ngz * nm;
    for (int o = tid; o < nm * ngz; o += P3_BLOCK) {
        const int s = o / ngz, q = o - s * ngz;
        const R *c = h + (size_t)s * P;
        R v = c[0] + ((q & 1) ? -c[P - 1] : c[P - 1]);
        int idx = q;
        for (int k = 1; k < ngz; k++) {
            const int t = tz_at<R>(idx);
            v += (R)2 * (c[2 * k - 1] * tz[t] - c[2 * k] * tz[t + 1]);
            idx += q;
            if (idx >= P) idx -= P;
        }
        dst[(size_t)q * nm + s] = v;
    }
}

// grid (ngz, nr): block (plane, ring-1): inverse azimuthal transform of the
// active components, scale, polcat; interleaved field store
template <typename R>
__global__ __launch_bounds__(P3_BLOCK) void p3_synth_kernel(P3SolveArgs<R> a) {
    typedef typename Vec4<R>::T R4;
    __shared__ R zc[4][P3_MAX_NM];
    __shared__ R tws[2 * P3_MAX_NA];
    __shared__ int sm[P3_MAX_NM], sp[P3_MAX_NM];
    const int plane = blockIdx.x, ir = blockIdx.y + 1, tid = threadIdx.x;
    const int na = a.na, nm = a.nm, ngz = a.ngz, mmax = a.mmax;
    const int jrs = a.ranges[0], jrf = a.ranges[2], krf = a.ranges[3];
    if (jrs < 1 || ir < jrf || ir > krf) return;
    for (int e = tid; e < 4 * nm; e += P3_BLOCK) {
        const int t = e / nm, s = e - t * nm;
        R v = (R)0;
        if (t < a.ntypes && !(t == 1 && a.ng == 1))
            v = a.Z[(((size_t)t * a.nr + (ir - 1)) * ngz + plane) * nm + s];
        zc[t][s] = v;
    }
    for (int e = tid; e < nm; e += P3_BLOCK) {
        sm[e] = a.slot_m[e];
        sp[e] = a.slot_p[e];
    }
    for (int j = tid; j < 2 * na; j += P3_BLOCK) tws[j] = a.tw[j];
    __syncthreads();
    const R scale = (R)1 / (R)(2 * ngz * na);
    const size_t row = ((size_t)(ir - 1) * ngz + plane) * na;
    for (int j = tid; j < na; j += P3_BLOCK) {
        R F0 = 0, F1 = 0, F2 = 0, F3 = 0;
        // k = (m j) mod na, advanced with m (the slots come in ascending m)
        int mcur = 0, k = 0;
        for (int s = 0; s < nm; s++) {
            const int m = sm[s];
            if (m < mcur) {
                mcur = 0;
                k = 0;
            }
            for (; mcur < m; mcur++) {
                k += j;
                if (k >= na) k -= na;
            }
            R b;
            if (m == 0) b = (R)1;
            else if (m == mmax - 1) b = (j & 1) ? (R)-1 : (R)1;
            else b = sp[s] ? (R)-2 * tws[2 * k + 1] : (R)2 * tws[2 * k];
            F0 += zc[0][s] * b;
            F1 += zc[1][s] * b;
            F2 += zc[2][s] * b;
            F3 += zc[3][s] * b;
        }
        F0 *= scale;
        F1 *= scale;
        F2 *= scale;
        F3 *= scale;
        const R cp = tws[2 * j], sn = tws[2 * j + 1];
        R4 out;
        out.x = F0 * cp - F1 * sn;
        out.y = F0 * sn + F1 * cp;
        out.z = F2;
        out.w = a.lpot ? F3 : a.fld[row + j].w;
        a.fld[row + j] = out;
    }
}

// ---- Green's tables (FP64) --------------------------------------------------------

__device__ double p3_softm(double ds) {
    if (ds <= 0.0) return 0.0;
    if (ds < 1.0)
        return 4.0 * ds * ds * ds * (1.0 / 3.0 - 0.3 * ds * ds + 0.125 * ds * ds * ds);
    if (ds < 2.0)
        return -1.0 / 15.0 + ds * ds * ds *
               (8.0 / 3.0 - 3.0 * ds + 1.2 * ds * ds - ds * ds * ds / 6.0);
    return 1.0;
}

// radfor / azifor / vrtfor / grnfun with the field point at height zf
// (radfor.f:25-38, azifor.f:25-40, vrtfor.f:25-40, sftpot.f:20-40)
__device__ double p3_kernel(int type, double rs, double rf, double cth, double sth,
                            double zf, double eps, int tsoft) {
    const double d2 = rs * rs + rf * rf - 2.0 * rs * rf * cth + zf * zf;
    if (type == 3) {
        if (tsoft == 1) return -1.0 / sqrt(d2 + eps * eps);
        const double x = sqrt(d2) / eps;
        if (x <= 0.0) return -1.4 / eps;
        if (x < 1.0) return (-1.4 + x * x * (2.0 / 3.0 - 0.3 * x * x + 0.1 * x * x * x)) / eps;
        if (x < 2.0)
            return (-1.6 + 1.0 / (15.0 * x) +
                    x * x * (4.0 / 3.0 - x + 0.3 * x * x - x * x * x / 30.0)) / eps;
        return -1.0 / sqrt(d2);
    }
    const double num = type == 0 ? rs * cth - rf : type == 1 ? -rs * sth : -zf;
    if (tsoft == 1) {
        const double e = d2 + eps * eps;
        return num / (e * sqrt(e));
    }
    if (type == 2 ? zf != 0.0 : d2 > 0.0)
        return num * p3_softm(sqrt(d2) / eps) / (d2 * sqrt(d2));
    return 0.0;
}

constexpr int GREEN3_BLOCK = 128;
constexpr int GREEN3_CHUNK = 16;

// grid (nr, nr, 4): block (is, ir, type).  Phase 1: thread per height iz,
// the angular DCT-I / DST-I of the kernel samples for the kept harmonics
// (GREEN3_CHUNK at a time), filtered; phase 2: the vertical DCT-I (DST-I
// for F_z) of each harmonic's column.  twd: cos, sin of 2 pi k / na;
// twzd: cos, sin of 2 pi k / (2 ngz).
__global__ __launch_bounds__(GREEN3_BLOCK) void p3_green_kernel(
    double *__restrict__ G, const double *__restrict__ twd,
    const double *__restrict__ twzd, const int *__restrict__ lg, int nr, int na,
    int ngz, int ng, double alpha, double dzg, double eps, int tsoft) {
    extern __shared__ unsigned char p3_lds[];
    double *T = (double *)p3_lds;   // [nzp][nmq]
    const int is = blockIdx.x + 1, ir = blockIdx.y + 1, type = blockIdx.z;
    const int mmax = na / 2 + 1, P = 2 * ngz;
    const bool azim = type == 1, vert = type == 2;
    if (azim && ng == 1) return;
    const int it1 = azim ? 2 : 1, it2 = azim ? mmax - 1 : mmax;
    const int npt = it2 - it1 + 1;
    const int iz1 = vert ? 2 : 1, iz2 = vert ? ngz : ngz + 1;
    const int nzp = iz2 - iz1 + 1;
    const int nmq = azim ? min(ng, mmax - 1) - 1 : min(ng, mmax);
    if (nmq <= 0) return;
    const double rs = exp(alpha * (double)(is - 1)) - 1.0;
    const double rf = exp(alpha * (double)(ir - 1)) - 1.0;
    for (int iz = iz1 + threadIdx.x; iz <= iz2; iz += GREEN3_BLOCK) {
        const double zf = dzg * (double)(iz - 1);
        for (int m0 = 0; m0 < nmq; m0 += GREEN3_CHUNK) {
            double acc[GREEN3_CHUNK];
            int idx[GREEN3_CHUNK], inc[GREEN3_CHUNK];
#pragma unroll
            for (int u = 0; u < GREEN3_CHUNK; u++) {
                acc[u] = 0.0;
                // output index mo = m0 + u: harmonic mo (DCT) or mo + 1 (DST);
                // sample q has angle index it - 1 = q (DCT) or q + 1 (DST)
                const int h = azim ? m0 + u + 1 : m0 + u;
                inc[u] = h % na;
                idx[u] = azim ? inc[u] : 0;
            }
            for (int q = 0; q < npt; q++) {
                const int k = it1 - 1 + q;
                const double f = p3_kernel(type, rs, rf, twd[2 * k], twd[2 * k + 1], zf, eps, tsoft);
                const double wq = azim ? 2.0 : ((q == 0 || q == npt - 1) ? 1.0 : 2.0);
                const double fw = wq * f;
#pragma unroll
                for (int u = 0; u < GREEN3_CHUNK; u++) {
                    acc[u] += fw * (azim ? twd[2 * idx[u] + 1] : twd[2 * idx[u]]);
                    idx[u] += inc[u];
                    if (idx[u] >= na) idx[u] -= na;
                }
            }
#pragma unroll
            for (int u = 0; u < GREEN3_CHUNK; u++) {
                const int mo = m0 + u;
                if (mo >= nmq) break;
                const int h = azim ? mo + 1 : mo;
                T[(size_t)(iz - iz1) * nmq + mo] = (h >= ng || lg[h]) ? 0.0 : acc[u];
            }
        }
    }
    __syncthreads();
    const size_t nn = (size_t)nr * nr;
    for (int o = threadIdx.x; o < nzp * nmq; o += GREEN3_BLOCK) {
        const int p = o / nmq, mo = o - p * nmq;
        double v = 0.0;
        if (!vert) {
            // DCT-I over nzp = ngz + 1 points
            v = T[mo] + ((p & 1) ? -T[(size_t)(nzp - 1) * nmq + mo] : T[(size_t)(nzp - 1) * nmq + mo]);
            int id = p;
            for (int k = 1; k < nzp - 1; k++) {
                v += 2.0 * T[(size_t)k * nmq + mo] * twzd[2 * id];
                id += p;
                if (id >= P) id -= P;
            }
        } else {
            // DST-I over nzp = ngz - 1 points: output i = p + 1
            const int i = p + 1;
            int id = i;
            for (int k = 1; k <= nzp; k++) {
                v += 2.0 * T[(size_t)(k - 1) * nmq + mo] * twzd[2 * id + 1];
                id += i;
                if (id >= P) id -= P;
            }
        }
        const int kz = vert ? p + 1 : p;
        const int m = azim ? mo + 1 : mo;
        G[((((size_t)type * (ngz + 1) + kz) * ng + m) * nn) + (size_t)(is - 1) * nr + (ir - 1)] = v;
    }
}

// ---- launchers ------------------------------------------------------------------------

static unsigned p3_blocks(long long n) { return (unsigned)((n + P3_PBLOCK - 1) / P3_PBLOCK); }

template <typename R> hipError_t launch_p3_massgn(const P3Args<R> &a, hipStream_t s) {
    if (a.n <= 0) return hipSuccess;
    if (a.jmass == 3)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(p3_massgn_kernel<R, 3>), dim3(p3_blocks(a.n)), dim3(P3_PBLOCK), 0, s, a);
    else
        hipLaunchKernelGGL(HIP_KERNEL_NAME(p3_massgn_kernel<R, 2>), dim3(p3_blocks(a.n)), dim3(P3_PBLOCK), 0, s, a);
    return hipGetLastError();
}

template <typename R> hipError_t launch_p3_push(const P3Args<R> &a, hipStream_t s) {
    if (a.n <= 0) return hipSuccess;
    if (a.jmass == 3)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(p3_push_kernel<R, 3>), dim3(p3_blocks(a.n)), dim3(P3_PBLOCK), 0, s, a);
    else
        hipLaunchKernelGGL(HIP_KERNEL_NAME(p3_push_kernel<R, 2>), dim3(p3_blocks(a.n)), dim3(P3_PBLOCK), 0, s, a);
    return hipGetLastError();
}

template <typename R> hipError_t launch_p3_ut(const P3Args<R> &a, hipStream_t s) {
    if (a.n <= 0) return hipSuccess;
    hipLaunchKernelGGL(p3_ut_kernel<R>, dim3(p3_blocks(a.n)), dim3(P3_PBLOCK), 0, s, a);
    return hipGetLastError();
}

template <typename R> hipError_t launch_p3_tmcent(const P3Args<R> &a, hipStream_t s) {
    if (a.n <= 0) return hipSuccess;
    if (a.jmass == 3)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(p3_tmcent_kernel<R, 3>), dim3(p3_blocks(a.n)), dim3(P3_PBLOCK), 0, s, a);
    else
        hipLaunchKernelGGL(HIP_KERNEL_NAME(p3_tmcent_kernel<R, 2>), dim3(p3_blocks(a.n)), dim3(P3_PBLOCK), 0, s, a);
    return hipGetLastError();
}

template <typename R> hipError_t launch_p3_gather(const P3GatherArgs<R> &b, hipStream_t s) {
    if (b.n <= 0) return hipSuccess;
    if (b.g.jmass == 3)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(p3_gather_kernel<R, 3>), dim3(p3_blocks(b.n)), dim3(P3_PBLOCK), 0, s, b);
    else
        hipLaunchKernelGGL(HIP_KERNEL_NAME(p3_gather_kernel<R, 2>), dim3(p3_blocks(b.n)), dim3(P3_PBLOCK), 0, s, b);
    return hipGetLastError();
}

template <typename R>
hipError_t launch_p3_key(const P3Args<R> &a, int kshift, unsigned kmax, unsigned *key,
                         unsigned *perm, hipStream_t s) {
    if (a.n <= 0) return hipSuccess;
    if (a.jmass == 3)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(p3_key_kernel<R, 3>), dim3(p3_blocks(a.n)), dim3(P3_PBLOCK), 0, s, a,
                           kshift, kmax, key, perm);
    else
        hipLaunchKernelGGL(HIP_KERNEL_NAME(p3_key_kernel<R, 2>), dim3(p3_blocks(a.n)), dim3(P3_PBLOCK), 0, s, a,
                           kshift, kmax, key, perm);
    return hipGetLastError();
}

template <typename R>
hipError_t launch_p3_interleave(const R *s4, const R *s2, typename Vec4<R>::T *rec,
                                unsigned *id, long long n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(p3_interleave_kernel<R>, dim3(p3_blocks(n)), dim3(P3_PBLOCK), 0, s, s4,
                       s2, rec, id, n);
    return hipGetLastError();
}

template <typename R>
hipError_t launch_p3_deinterleave(const typename Vec4<R>::T *rec, const uint8_t *zone,
                                  const unsigned *id, R *d4, R *d2, int *iz, long long n,
                                  hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(p3_deinterleave_kernel<R>, dim3(p3_blocks(n)), dim3(P3_PBLOCK), 0, s,
                       rec, zone, id, d4, d2, iz, n);
    return hipGetLastError();
}

template <typename R> hipError_t launch_p3_permute(const P3PermArgs<R> &p, hipStream_t s) {
    if (p.n <= 0) return hipSuccess;
    hipLaunchKernelGGL(p3_permute_kernel<R>, dim3(p3_blocks(p.n)), dim3(P3_PBLOCK), 0, s, p);
    return hipGetLastError();
}

template <typename R>
hipError_t launch_p3_integrals(const P3Args<R> &a, int with_pot, const double4 *fp_c,
                               int nfp, double fp_dr, int centred,
                               double *part, double *out, hipStream_t s) {
    const unsigned nb = a.n > 0 ? p3_blocks(a.n) : 0;
    if (nb && a.jmass == 3)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(p3_integrals_kernel<R, 3>), dim3(nb), dim3(P3_PBLOCK), 0, s, a,
                           with_pot, fp_c, nfp, fp_dr, centred, part);
    else if (nb)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(p3_integrals_kernel<R, 2>), dim3(nb), dim3(P3_PBLOCK), 0, s, a,
                           with_pot, fp_c, nfp, fp_dr, centred, part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(p3_integrals_sum_kernel, dim3(1), dim3(P3_BLOCK), 0, s, part, (int)nb,
                       out);
    return hipGetLastError();
}

// Limits the launches below rely on (checked again by the context set-up).
template <typename R> bool p3_solve_fits(const P3SolveArgs<R> &a) {
    const size_t P = 2 * (size_t)a.ngz;
    return a.na <= P3_MAX_NA && a.nm <= P3_MAX_NM && a.nr <= P3_MAX_NR &&
           ((size_t)a.nm * a.ngz + tz_words<R>(P)) * sizeof(R) <= 48 * 1024 &&
           ((size_t)a.nm * P + tz_words<R>(P)) * sizeof(R) <= 48 * 1024;
}

template <typename R> hipError_t launch_p3_solve(const P3SolveArgs<R> &a, hipStream_t s) {
    if (!p3_solve_fits(a)) return hipErrorInvalidValue;
    const size_t P = 2 * (size_t)a.ngz;
    hipLaunchKernelGGL(p3_analyse_kernel<R>, dim3(a.ngz, a.nr), dim3(P3_BLOCK), 0, s, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(p3_zfwd_kernel<R>, dim3(a.nr), dim3(P3_BLOCK),
                       ((size_t)a.nm * a.ngz + tz_words<R>(P)) * sizeof(R), s, a);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(p3_conv_kernel<R>, dim3(a.ng, a.ngz + 1, a.ntypes), dim3(P3_BLOCK),
                       0, s, a);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(p3_zinv_kernel<R>, dim3(a.nr, a.ntypes), dim3(P3_BLOCK),
                       ((size_t)a.nm * P + tz_words<R>(P)) * sizeof(R), s, a);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(p3_synth_kernel<R>, dim3(a.ngz, a.nr), dim3(P3_BLOCK), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_p3_green(double *G, const double *twd, const double *twzd, const int *lg,
                           int nr, int na, int ngz, int ng, double alpha, double dzg,
                           double eps, int tsoft, hipStream_t s) {
    const int mmax = na / 2 + 1;
    const size_t lds = sizeof(double) * (size_t)(ngz + 1) * (ng < mmax ? ng : mmax);
    if (lds > 48 * 1024) return hipErrorInvalidValue;
    hipLaunchKernelGGL(p3_green_kernel, dim3(nr, nr, 4), dim3(GREEN3_BLOCK), lds, s, G, twd,
                       twzd, lg, nr, na, ngz, ng, alpha, dzg, eps, tsoft);
    return hipGetLastError();
}

#define P3_INST(R)                                                                     \
    template hipError_t launch_p3_massgn<R>(const P3Args<R> &, hipStream_t);          \
    template hipError_t launch_p3_push<R>(const P3Args<R> &, hipStream_t);            \
    template hipError_t launch_p3_tmcent<R>(const P3Args<R> &, hipStream_t);          \
    template hipError_t launch_p3_ut<R>(const P3Args<R> &, hipStream_t);              \
    template hipError_t launch_p3_integrals<R>(const P3Args<R> &, int, const double4 *, \
                                               int, double, int,                       \
                                               double *, double *, hipStream_t);       \
    template hipError_t launch_p3_gather<R>(const P3GatherArgs<R> &, hipStream_t);    \
    template hipError_t launch_p3_solve<R>(const P3SolveArgs<R> &, hipStream_t);      \
    template bool p3_solve_fits<R>(const P3SolveArgs<R> &);                            \
    template hipError_t launch_p3_key<R>(const P3Args<R> &, int, unsigned, unsigned *,  \
                                         unsigned *, hipStream_t);                     \
    template hipError_t launch_p3_permute<R>(const P3PermArgs<R> &, hipStream_t);          \
    template hipError_t launch_p3_interleave<R>(const R *, const R *, typename Vec4<R>::T *,  \
                                                unsigned *, long long, hipStream_t);          \
    template hipError_t launch_p3_deinterleave<R>(const typename Vec4<R>::T *, const uint8_t *, \
                                                  const unsigned *, R *, R *, int *, long long, \
                                                  hipStream_t);
P3_INST(float)
P3_INST(double)

}  // namespace gh
