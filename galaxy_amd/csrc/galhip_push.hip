// galhip_push.hip -- the fused particle kernel of the p2d step, the mass
// assignment of masset, the group-level gather kernel and the sort helpers.
//
// push_kernel does, for each particle of the zones due at this step, what
// the reference's per-group pipeline does (SRC/lib15/step.f:186-234):
//   getacc/grdacc  CIC weights at the old position, 4-corner interpolation
//                  of the Cartesian field, fixed radial force, zone scaling
//                  (getacc.f:35-163, grdacc.f:40-66)
//   stpgrp/offgrd  RVLF leap-frog and the off-grid test (stpgrp.f:27-56,
//                  offgrd.f:55-59, addoff.f:30-70)
//   rezgrp/chstep  zone changes (rezgrp.f:30-50, zoneno.f, chstep.f:22-31)
//   massgn         weights at the new position, scatter into the zone's
//                  mass array (massgn.f:68-89)
// in one pass over the particle arrays.
//
// The old position's weights are not recomputed: every particle carries the
// (u, t) grid coordinates that its last mass assignment formed from the
// same position with the same formulas (weight.f forms them twice per step,
// once for the forces and once for the masses), so the transcendental part
// runs once per push.
//
// Mass sums are 64-bit fixed point (weight x 2^qbits): integer adds
// commute, so the sums are bit-identical for any processing order and need
// no float atomics.  Each block deposits into an LDS tile that covers the
// rings x spokes window of its particles (particles are kept ordered by
// cell, so the window is small), with TILE_REP copies chosen by lane so
// that lanes hitting the same cell are spread over copies; the block then
// flushes every non-zero tile cell with one 64-bit global atomic.  Blocks
// whose window exceeds the tile, and deposits into another zone than the
// tile's, go straight to global atomics.  Global 64-bit atomics are the
// scarce resource on MI355X (about 24 G/s device-wide), LDS atomics far
// less so.
#include "galhip_kern.h"
#include <climits>

namespace gh {

// Launch shape (tuned on MI355X; see the DESIGN notes of the deposit):
#ifndef GH_TILE_CAP
#define GH_TILE_CAP 1024
#endif
#ifndef GH_TILE_REP
#define GH_TILE_REP 2
#endif

constexpr int PPT = GH_PPT;            // particles per thread
constexpr int PB = GH_PUSH_BLOCK;      // threads per block
// minimum wavefronts per SIMD of push_kernel (launch bounds), FP32 and FP64
#ifndef GH_PUSH_WAVES32
#define GH_PUSH_WAVES32 6
#endif
#ifndef GH_PUSH_WAVES64
#define GH_PUSH_WAVES64 4
#endif
constexpr int TILE_CAP = GH_TILE_CAP;  // LDS tile, cells of a ring x spoke window
constexpr int TILE_REP = GH_TILE_REP;  // replicas (lane & (REP-1)) against conflicts
constexpr int FF_CAP = 256;            // staged fixed-force table entries
#ifndef GH_PUSH_PASSES
#define GH_PUSH_PASSES 3               // deposit window passes before global atomics
#endif

// fields of a packed deposit cell (zone << 28 | ring << 16 | spoke)
__device__ __forceinline__ int dc_zone(int c) { return (int)((unsigned)c >> 28); }
__device__ __forceinline__ int dc_ring(int c) { return (c >> 16) & 0xFFF; }
__device__ __forceinline__ int dc_spoke(int c) { return c & 0xFFFF; }

template <typename R>
__device__ __forceinline__ R tsfac(const PushArgs<R> &a, int iz) {
    int i = iz < a.nz ? iz : a.nz;
    return (R)a.nstep[i - 1];
}

// zoneno for a position of known radius (zoneno.f:40-55)
template <typename R>
__device__ __forceinline__ int zone_of(const PushArgs<R> &a, R rad) {
    // the first two boundaries in uniform branches with the radii read at
    // fixed offsets; only more than three zones take the loop
    int z = 1;
    if (a.nz >= 2) {
        if (rad >= a.zbr[1]) z = 2;
        if (a.nz >= 3) {
            if (rad >= a.zbr[2]) z = 3;
            for (int i = 4; i <= a.nz; i++)
                if (rad >= a.zbr[i - 1]) z = i;
        }
    }
    return z;
}

template <typename R>
__device__ __forceinline__ void chstep(R tfac, R ox, R oy, R &ovx, R &ovy,
                                       R &ax, R &ay, R &nx, R &ny, R &nvx,
                                       R &nvy) {
    const R h = (R)0.5;
    ovx = tfac * (ovx + h * ax * ((R)1 - tfac));
    ax = ax * tfac * tfac;
    nvx = ovx + ax;
    nx = ox + nvx;
    ovy = tfac * (ovy + h * ay * ((R)1 - tfac));
    ay = ay * tfac * tfac;
    nvy = ovy + ay;
    ny = oy + nvy;
}

// Compensated position update: the position is hi + lo, and the new
// (hi, lo) = TwoSum(hi, v + lo) (Knuth's two-sum, exact without fused
// operations), so each update's rounding error is carried, not dropped.
template <typename R>
__device__ __forceinline__ R comp_add(R x, R v, R lo, R &lo_out) {
    const R y = v + lo;
    const R s = x + y;
    const R bp = s - x;
    lo_out = (x - (s - bp)) + (y - bp);
    return s;
}

__device__ __forceinline__ long long wave_sum64(long long v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ int wave_min(int v) { return wave_min_i32(v); }
__device__ __forceinline__ int wave_max(int v) { return wave_max_i32(v); }


// ENS: an ensemble launch, the block's particles and member from a.blk.
// COMP: compensated positions (a.xlo holds the low words).
template <typename R, bool ENS, bool COMP>
// Wavefronts per SIMD, measured: FP32 at least 6 (80 VGPRs at most; faster
// than 5 waves at 86 VGPRs; the LDS of six blocks per CU just fits); FP64
// at least 4 (128 VGPRs, a few spills, 4 % faster than 3 waves without).
__global__ __launch_bounds__(PB, sizeof(R) == 4 ? GH_PUSH_WAVES32 : GH_PUSH_WAVES64) void push_kernel(PushArgs<R> a) {
    typedef typename Vec2<R>::T R2;
    typedef typename Vec4<R>::T R4;
    __shared__ long long tile[TILE_REP][TILE_CAP];
    __shared__ int s_jr[GH_MAXZ], s_kr[GH_MAXZ];
    __shared__ int s_dnoff, s_pushes;
    __shared__ int s_box[8];                 // two (ring min, max, offset min, max) boxes
    __shared__ unsigned long long s_ref[2];  // two deposit-pass reference keys
    __shared__ int s_fref, s_frmin, s_frmax, s_fdmin, s_fdmax, s_kmin, s_kmax;
    // slice of the fixed-force table that the block's radii use
    __shared__ double4 s_ffc[FF_CAP];
    // the logarithm and arc-tangent tables, read per particle by ut_of
    __shared__ double s_tab[GH_MATH_TAB_N];

    const int tid = threadIdx.x, lane = tid & 63;
    // the block's particles and its grid (ensembles: member eb.x)
    const int4 eb = ENS ? a.blk[blockIdx.x] : make_int4(0, 0, 0, 0);
    const long long base = ENS ? (long long)eb.y : (long long)blockIdx.x * (PB * PPT);
    const long long mo = ENS ? (long long)eb.x * a.mesh : 0;
    const R2 *const fld = a.fld + mo;
    int *const jrad = a.jrad + (ENS ? eb.x * GH_MAXZ : 0);
    int *const krad = a.krad + (ENS ? eb.x * GH_MAXZ : 0);
    long long *const counters = a.counters + (ENS ? eb.x * 4 : 0);
    const int qbits = ENS ? a.qbits_m[eb.x] : a.qbits;
    const int nz = a.nz, nlists = a.nz + 1, na = a.na, nr = a.nr;
    for (int j = tid; j < GH_MATH_TAB_N; j += PB) s_tab[j] = GH_MATH_TAB[j];
    const double *const mtab = s_tab;
    if (tid < GH_MAXZ) {
        s_jr[tid] = INT_MAX;
        s_kr[tid] = -1;
    }
    if (tid == 0) {
        s_dnoff = 0;
        s_pushes = 0;
        s_box[0] = s_box[4] = INT_MAX;
        s_box[1] = s_box[5] = -1;
        s_box[2] = s_box[6] = INT_MAX;
        s_box[3] = s_box[7] = INT_MIN;
        s_ref[0] = s_ref[1] = ~0ull;
        s_frmin = INT_MAX;
        s_frmax = -1;
        s_fdmin = INT_MAX;
        s_fdmax = INT_MIN;
        s_kmin = INT_MAX;
        s_kmax = -1;
    }

    // the block's particles through block-based pointers and 32-bit
    // offsets (uniform base, per-lane offset: no 64-bit address arithmetic)
    R4 *const xvb = a.xv + base;
    R2 *const utb = a.ut + base;
    R2 *const lob = COMP ? a.xlo + base : nullptr;
    uint8_t *const zb = a.zone + base;
    const int nblk = ENS ? eb.z : (int)min((long long)(PB * PPT), a.n - base);
    int zin[PPT];
    bool mv[PPT];
    int any = 0;
#pragma unroll
    for (int k = 0; k < PPT; k++) {
        const int j = k * PB + tid;
        zin[k] = j < nblk ? (int)zb[j] : 0;
        bool off = zin[k] == nlists;
        mv[k] = zin[k] > 0 &&
                (off ? (a.offanal && a.mzone == nz) : (zin[k] <= a.mzone));
        any |= mv[k];
    }
    // every thread reaches this barrier; a block with nothing to move exits
    if (!__syncthreads_or(any)) return;

    // The loads of all PPT particles are issued before any is consumed:
    // state and (u, t) first, then the 4 x PPT field gathers, then the
    // fixed-force table reads.  These phases are straight-line code (indices
    // clamped, results selected), so each phase's reads are in flight
    // together instead of one particle's latency chain after another's.
    R4 st[PPT];
    R2 cu[PPT];
#pragma unroll
    for (int k = 0; k < PPT; k++) {
        const int j = min(k * PB + tid, nblk - 1);
        st[k] = xvb[j];
        cu[k] = utb[j];
    }
    // grdacc: weights of the old position from its (u, t) (off-grid and
    // idle particles weigh nothing), the field read from an LDS copy of the
    // window of rings x spokes that the block's old cells span -- the same
    // windowing as the deposit tile below, whose LDS it borrows.  A window
    // larger than that space falls back to reads from global memory.
    const int half = na >> 1;
    int gring[PPT], gspoke[PPT], gd[PPT];
    R gw[PPT][4];
#pragma unroll
    for (int k = 0; k < PPT; k++) {
        const bool grid = mv[k] && zin[k] != nlists;
        cic(grid ? cu[k].x : (R)0, grid ? cu[k].y : (R)0, nr, na, gring[k],
            gspoke[k], gw[k]);
        if (!grid) gring[k] = -1;
    }
    R rr[PPT], fr[PPT];
    int kf[PPT];
#pragma unroll
    for (int k = 0; k < PPT; k++) {
        rr[k] = sqrt(st[k].x * st[k].x + st[k].y * st[k].y);
        fr[k] = 0;
        kf[k] = a.fixforce ? fixforce_index((double)rr[k], a.nfr, a.fr_inv_dr) : 0;
    }
    if (tid == 0) {
        int r = 0;
#pragma unroll
        for (int k = PPT - 1; k >= 0; k--)
            if (gring[k] >= 0) r = gspoke[k];
        s_fref = r;
    }
    __syncthreads();
    const int fref = s_fref;
    {
        int rmin = INT_MAX, rmax = -1, dmin = INT_MAX, dmax = INT_MIN;
#pragma unroll
        for (int k = 0; k < PPT; k++) {
            int d = gspoke[k] - fref;
            if (d < -half) d += na;
            else if (d >= half) d -= na;
            gd[k] = d;
            if (gring[k] >= 0) {
                rmin = min(rmin, gring[k]);
                rmax = max(rmax, gring[k]);
                dmin = min(dmin, d);
                dmax = max(dmax, d);
            }
        }
        int kmin = INT_MAX, kmax = -1;
#pragma unroll
        for (int k = 0; k < PPT; k++)
            if (gring[k] >= 0) {
                kmin = min(kmin, kf[k]);
                kmax = max(kmax, kf[k]);
            }
        rmin = wave_min(rmin);
        rmax = wave_max(rmax);
        dmin = wave_min(dmin);
        dmax = wave_max(dmax);
        kmin = wave_min(kmin);
        kmax = wave_max(kmax);
        if (lane == 0 && rmax >= 0) {
            atomicMin(&s_frmin, rmin);
            atomicMax(&s_frmax, rmax);
            atomicMin(&s_fdmin, dmin);
            atomicMax(&s_fdmax, dmax);
            atomicMin(&s_kmin, kmin);
            atomicMax(&s_kmax, kmax);
        }
    }
    __syncthreads();
    R2 *ftile = reinterpret_cast<R2 *>(&tile[0][0]);
    constexpr int FT_CAP = (int)(sizeof(tile) / sizeof(R2));
    const int fr0 = s_frmin, fr1 = s_frmax, fd0 = s_fdmin, fd1 = s_fdmax;
    const int ftw = fd1 - fd0 + 2, fth = fr1 - fr0 + 2;
    const bool use_ft = fr1 >= 0 && fth * ftw <= FT_CAP;
    if (use_ft) {
        // rows past the last ring read as zero, as the global path does
        // (row, col) of element j advanced incrementally: one division per
        // thread instead of one per element
        const int drow = PB / ftw, dcol = PB - drow * ftw;
        int row = tid / ftw, col = tid - row * ftw;
        for (int j = tid; j < fth * ftw; j += PB) {
            const int ring = fr0 + row;
            int spoke = fref + fd0 + col;
            row += drow;
            col += dcol;
            if (col >= ftw) {
                col -= ftw;
                row++;
            }
            if (spoke < 0) spoke += na;
            else if (spoke >= na) spoke -= na;
            if (spoke >= na) spoke -= na;
            ftile[j] = ring < nr ? fld[ring * na + spoke] : zero2<R>();
        }
    }
    // the table intervals k0 .. k1 that the block's grid particles read
    const int k0 = s_kmin, nk = s_kmax + 1 - s_kmin;
    const bool use_ff = a.fixforce && s_kmax >= 0 && nk <= FF_CAP;
    if (use_ff)
        for (int j = tid; j < nk; j += PB) s_ffc[j] = a.fr_c[k0 + j];
    __syncthreads();
    R ax[PPT], ay[PPT];
#pragma unroll
    for (int k = 0; k < PPT; k++) {
        const bool grid = gring[k] >= 0;
        R sx = 0, sy = 0;
        if (use_ft) {
            const int b = grid ? (gring[k] - fr0) * ftw + (gd[k] - fd0) : 0;
            const int ti[4] = {b, b + 1, b + ftw, b + ftw + 1};
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const R2 f = ftile[ti[q]];
                const R wq = grid ? gw[k][q] : (R)0;
                sx = sx + f.x * wq;
                sy = sy + f.y * wq;
            }
        } else {
            int ic[4];
            corners_rs(grid ? gring[k] : 0, gspoke[k], na, ic);
#pragma unroll
            for (int q = 0; q < 4; q++) {
                // the last ring with t == na points past the field: 0
                const bool in = ic[q] < a.mesh;
                R2 f = fld[in ? ic[q] : 0];
                R wq = (grid && in) ? gw[k][q] : (R)0;
                sx = sx + f.x * wq;
                sy = sy + f.y * wq;
            }
        }
        ax[k] = sx;
        ay[k] = sy;
    }
    if (use_ff) {
        // the staged slice; particles off the grid or idle (their force is
        // not used) read its first interval
#pragma unroll
        for (int k = 0; k < PPT; k++) {
            const bool grid = gring[k] >= 0;
            fr[k] = (R)fixforce_at((double)rr[k], grid ? kf[k] : k0, a.nfr, a.fr_inv_dr,
                                   s_ffc, k0);
        }
    } else if (a.fixforce) {
#pragma unroll
        for (int k = 0; k < PPT; k++)
            fr[k] = (R)fixforce_at((double)rr[k], kf[k], a.nfr, a.fr_inv_dr, a.fr_c, 0);
    }

    // FP32: the states of particles 1 .. PPT-1 read once more (L2 hits):
    // held in registers across the gathers above, one of them would be
    // spilled to scratch at 6 wavefronts per SIMD, a store and a reload
    // through memory (FP64 keeps them: the re-read spills more there)
    if constexpr (sizeof(R) == 4) {
#pragma unroll
        for (int k = 1; k < PPT; k++) st[k] = xvb[min(k * PB + tid, nblk - 1)];
    }

    // deposit state kept per particle: ring, spoke, zone (0 = none), weights
    // dcell: zone << 28 | ring << 16 | spoke (zone 0: no deposit)
    int dcell[PPT];
    R dw[PPT][4];
    // off-grid list changes: +1 leaving the grid, -1 returning
    bool dn_plus[PPT], dn_minus[PPT];
#pragma unroll
    for (int k = 0; k < PPT; k++) {
        dcell[k] = 0;
        dn_plus[k] = false;
        dn_minus[k] = false;
        if (mv[k]) {
            const int j = k * PB + tid;
            const int z0 = zin[k];
            const bool offlist = z0 == nlists;
            R ox = st[k].x, oy = st[k].y, ovx = st[k].z, ovy = st[k].w;
            R ax_ = ax[k], ay_ = ay[k];
            if (!offlist && fr[k] != (R)0) {
                R f = fr[k] / rr[k];
                ax_ = ax_ + f * ox;
                ay_ = ay_ + f * oy;
            }
            if (nz > 1 && z0 > 1) {
                R t = tsfac(a, z0);
                t = t * t;
                ax_ = t * ax_;
                ay_ = t * ay_;
            }
            R nvx = ovx + ax_, nvy = ovy + ay_;
            R nx = ox + nvx, ny = oy + nvy;
            R2 lo = zero2<R>(), elo = zero2<R>();
            if (COMP) {
                lo = lob[j];
                nx = comp_add(ox, nvx, lo.x, elo.x);
                ny = comp_add(oy, nvy, lo.y, elo.y);
            }
            int znew = z0;
            R rnew = sqrt(nx * nx + ny * ny);
            bool off = rnew > a.rgrid;
            if (!offlist) {
                if (off) {
                    znew = nlists;
                    dn_plus[k] = true;
                    if (a.offanal && z0 < nz) {
                        R tfac = tsfac(a, nz) / tsfac(a, z0);
                        chstep(tfac, ox, oy, ovx, ovy, ax_, ay_, nx, ny, nvx, nvy);
                        if (COMP) {
                            nx = comp_add(ox, nvx, lo.x, elo.x);
                            ny = comp_add(oy, nvy, lo.y, elo.y);
                        }
                        rnew = sqrt(nx * nx + ny * ny);
                    }
                }
            } else if (!off) {
                dn_minus[k] = true;
                znew = nz;
            }
            if (a.mzone > 1 && !offlist && znew <= a.mzone) {
                int kz = zone_of(a, rnew);
                if (kz > a.mzone) kz = a.mzone;
                if (kz > znew) kz = zone_of(a, rr[k]);
                if (kz != znew) {
                    R tfac = tsfac(a, kz) / tsfac(a, znew);
                    chstep(tfac, ox, oy, ovx, ovy, ax_, ay_, nx, ny, nvx, nvy);
                    if (COMP) {
                        nx = comp_add(ox, nvx, lo.x, elo.x);
                        ny = comp_add(oy, nvy, lo.y, elo.y);
                    }
                    rnew = sqrt(nx * nx + ny * ny);
                    znew = kz;
                }
            }
            xvb[j] = mk4(nx, ny, nvx, nvy);
            if (COMP) lob[j] = elo;
            if (znew != z0) zb[j] = (uint8_t)znew;
            if (znew >= 1 && znew <= nz) {
                R u, t;
                ut_of(nx, ny, rnew, a.k, u, t, mtab);
                utb[j] = mk2(u, t);
                R tf = a.nsect > 1 ? (R)fmod(t, a.thmax) : t;
                int ring, spoke;
                cic(u, tf, nr, na, ring, spoke, dw[k]);
                dcell[k] = (int)(((unsigned)znew << 28) | ((unsigned)ring << 16) | (unsigned)spoke);
            }
        }
    }
    // the wave's pushes and off-grid change by ballots (scalar counts),
    // added to the block's counters by one lane
    {
        int wp = 0, wd = 0;
#pragma unroll
        for (int k = 0; k < PPT; k++) {
            wp += __popcll(__ballot(mv[k]));
            wd += __popcll(__ballot(dn_plus[k])) - __popcll(__ballot(dn_minus[k]));
        }
        if (lane == 0) {
            if (wp) atomicAdd(&s_pushes, wp);
            if (wd) atomicAdd(&s_dnoff, wd);
        }
    }
    // Deposit passes.  A pass takes a reference deposit -- the pending one
    // first in particle order from the block's middle particle -- and the
    // box of the pending deposits of the reference's zone (rings x spoke
    // offsets from the reference spoke, which wrap around the ring); a box
    // over TILE_CAP cells is clipped about the reference, which therefore
    // always fits.  The deposits inside go to the LDS tile, which is then
    // flushed with one global atomic per touched cell.  What the clip or a
    // second zone leaves pending gets the next pass; after PUSH_PASSES
    // passes (and for a cell on the last ring) the corners go to global
    // atomics directly.
    bool pend[PPT];
#pragma unroll
    for (int k = 0; k < PPT; k++) pend[k] = dcell[k] != 0 && dc_ring(dcell[k]) + 1 < nr;
    // this thread's reference candidate: the lowest rotated particle index
    // among its pending deposits; the wave's lowest goes to s_ref[slot]
    auto offer_ref = [&](int slot) {
        int myrot = INT_MAX, mypack = 0;
#pragma unroll
        for (int k = 0; k < PPT; k++) {
            const int rot = (k * PB + tid - PB * PPT / 2) & (PB * PPT - 1);
            if (pend[k] && rot < myrot) {
                myrot = rot;
                mypack = dcell[k];
            }
        }
        const int wmin = wave_min(myrot);
        if (wmin != INT_MAX) {
            const unsigned long long bal = __ballot(myrot == wmin);
            const int fl = __ffsll((long long)bal) - 1;
            const int pk = __builtin_amdgcn_readlane(mypack, fl);
            if (lane == 0)
                atomicMin(&s_ref[slot], ((unsigned long long)(unsigned)wmin << 32) | (unsigned)pk);
        }
    };
    offer_ref(0);
    __syncthreads();
#pragma unroll
    for (int pass = 0; pass < GH_PUSH_PASSES; pass++) {
        const unsigned long long kref = s_ref[pass & 1];
        if (kref == ~0ull) break;                  // nothing pending (uniform)
        const int zt = (int)((kref >> 28) & 0xF), rref = (int)((kref >> 16) & 0xFFF);
        const int sref = (int)(kref & 0xFFFF);
        int *const bx = s_box + 4 * (pass & 1);
        if (tid == 0) s_ref[(pass + 1) & 1] = ~0ull;   // last read in the previous pass
        // ring and spoke-offset box of the zone's pending deposits; dd: the
        // deposits' spoke offsets from the reference spoke
        int dd[PPT];
        {
            int rmin = INT_MAX, rmax = -1, dmin = INT_MAX, dmax = INT_MIN;
#pragma unroll
            for (int k = 0; k < PPT; k++) {
                int d = dc_spoke(dcell[k]) - sref;
                if (d < -half) d += na;
                else if (d >= half) d -= na;
                dd[k] = d;
                if (pend[k] && dc_zone(dcell[k]) == zt) {
                    const int ring = dc_ring(dcell[k]);
                    rmin = min(rmin, ring);
                    rmax = max(rmax, ring);
                    dmin = min(dmin, d);
                    dmax = max(dmax, d);
                }
            }
            rmin = wave_min(rmin);
            rmax = wave_max(rmax);
            dmin = wave_min(dmin);
            dmax = wave_max(dmax);
            if (lane == 0 && rmax >= 0) {
                atomicMin(&bx[0], rmin);
                atomicMax(&bx[1], rmax);
                atomicMin(&bx[2], dmin);
                atomicMax(&bx[3], dmax);
            }
        }
        __syncthreads();
        // tile: rings [r0, r0 + th) x spoke offsets [d0, d0 + tw) of zone zt
        int r0 = bx[0], th = bx[1] - bx[0] + 2, d0 = bx[2], tw = bx[3] - bx[2] + 2;
        if (tid == 0) {
            atomicMin(&s_jr[zt - 1], bx[0]);
            atomicMax(&s_kr[zt - 1], bx[1]);
            int *const nb = s_box + 4 * ((pass + 1) & 1);   // last read in the previous pass
            nb[0] = INT_MAX;
            nb[1] = -1;
            nb[2] = INT_MAX;
            nb[3] = INT_MIN;
        }
        if (th * tw > TILE_CAP) {
            if (th > 8) {
                r0 = min(max(rref - 3, r0), r0 + th - 8);
                th = 8;
            }
            const int mw = TILE_CAP / th;
            if (tw > mw) {
                d0 = min(max(-((mw - 2) >> 1), d0), d0 + tw - mw);
                tw = mw;
            }
        }
        const int tlen = th * tw;
        for (int j = tid; j < tlen; j += PB) {
#pragma unroll
            for (int rep = 0; rep < TILE_REP; rep++) tile[rep][j] = 0;
        }
        __syncthreads();
        long long *trep = tile[lane & (TILE_REP - 1)];
#pragma unroll
        for (int k = 0; k < PPT; k++) {
            const int ring = dc_ring(dcell[k]), d = dd[k];
            if (!pend[k] || dc_zone(dcell[k]) != zt || ring < r0 || ring - r0 + 1 >= th ||
                d < d0 || d - d0 + 1 >= tw)
                continue;
            long long qv[4];
            quant_n<R, 4>(dw[k], qbits, true, qv);
            const int b = (ring - r0) * tw + (d - d0);
            const int ti[4] = {b, b + 1, b + tw, b + tw + 1};
#pragma unroll
            for (int q = 0; q < 4; q++)
                if (qv[q] != 0)
                    atomicAdd((unsigned long long *)&trep[ti[q]], (unsigned long long)qv[q]);
            pend[k] = false;
        }
        const bool last = pass + 1 >= GH_PUSH_PASSES;
        if (!last) offer_ref((pass + 1) & 1);
        __syncthreads();
        // flush: one global atomic per touched cell, the tile's cells over
        // all threads in storage order (a row's cells are consecutive
        // spokes, so a wave's atomics coalesce)
        for (int j = tid; j < tlen; j += PB) {
            long long v = 0;
#pragma unroll
            for (int rep = 0; rep < TILE_REP; rep++) v += tile[rep][j];
            if (v != 0) {
                const int row = div_small(j, tw), col = j - row * tw;
                int spoke = sref + d0 + col;
                if (spoke < 0) spoke += na;
                else if (spoke >= na) spoke -= na;
                if (spoke >= na) spoke -= na;
                atomicAdd((unsigned long long *)&a.mfx[zt - 1][mo + (r0 + row) * na + spoke],
                          (unsigned long long)v);
            }
        }
        if (last) break;
    }
    // what no pass took (and cells on the last ring): global atomics
#pragma unroll
    for (int k = 0; k < PPT; k++) {
        const int z = dc_zone(dcell[k]), ring = dc_ring(dcell[k]);
        if (z == 0 || (ring + 1 < nr && !pend[k])) continue;
        long long qv[4];
        quant_n<R, 4>(dw[k], qbits, true, qv);
        atomicMin(&s_jr[z - 1], ring);
        atomicMax(&s_kr[z - 1], ring);
        int ic[4];
        corners_rs(ring, dc_spoke(dcell[k]), na, ic);
#pragma unroll
        for (int q = 0; q < 4; q++)
            if (qv[q] != 0) deposit_global(a.mfx, nz, a.mesh, z, ic[q], qv[q], mo);
    }
    __syncthreads();
    if (tid < nz && s_kr[tid] >= 0) {
        atomicMin(&jrad[tid], s_jr[tid] + 1);
        int k = s_kr[tid] + 2;
        atomicMax(&krad[tid], k < nr ? k : nr);
    }
    if (tid == 0) {
        if (s_dnoff) atomicAdd((unsigned long long *)&counters[0],
                               (unsigned long long)(long long)s_dnoff);
        if (s_pushes) atomicAdd((unsigned long long *)&counters[1],
                                (unsigned long long)s_pushes);
    }
}

// masset / massgn for every on-grid particle without a push: also fills the
// particles' (u, t) so the next push finds them.  Global atomics after the
// wavefront reduction.
template <typename R>
__global__ __launch_bounds__(GH_BLOCK) void massgn_kernel(PushArgs<R> a) {
    typedef typename Vec4<R>::T R4;
    __shared__ int s_jr[GH_MAXZ], s_kr[GH_MAXZ];
    const int tid = threadIdx.x, lane = tid & 63;
    if (tid < GH_MAXZ) {
        s_jr[tid] = INT_MAX;
        s_kr[tid] = -1;
    }
    __syncthreads();
    long long i = (long long)blockIdx.x * GH_BLOCK + tid;
    int z = i < a.n ? a.zone[i] : 0;
    bool on = z >= 1 && z <= a.nz;
    R w[4];
    int ring = 0, spoke = 0;
    if (on) {
        R rr, u, t;
        const R4 st = a.xv[i];
        (void)weight1<R>(st.x, st.y, a.nr, a.k, false, a.thmax, w, rr, u, t);
        a.ut[i] = mk2(u, t);
        R tf = a.nsect > 1 ? (R)fmod(t, a.thmax) : t;
        cic(u, tf, a.nr, a.na, ring, spoke, w);
        atomicMin(&s_jr[z - 1], ring);
        atomicMax(&s_kr[z - 1], ring);
    }
    __syncthreads();
    // one global min/max per zone and block (ring ranges are 1-based)
    if (tid < a.nz && s_kr[tid] >= 0) {
        atomicMin(&a.jrad[tid], s_jr[tid] + 1);
        int k = s_kr[tid] + 2;
        atomicMax(&a.krad[tid], k < a.nr ? k : a.nr);
    }
    const int key = on ? ((z - 1) * (a.nr + 1) + ring) * a.na + spoke : -1;
    long long qv[4];
    quant_n<R, 4>(w, a.qbits, on, qv);
    unsigned long long pend = __ballot(key >= 0);
    while (pend) {
        const int leader = __ffsll((long long)pend) - 1;
        const int lk = __shfl(key, leader, 64);
        const bool mine = key == lk;
        const unsigned long long mm = __ballot(mine);
        long long s[4];
        for (int q = 0; q < 4; q++) s[q] = wave_sum64(mine ? qv[q] : 0);
        if (lane == leader) {
            int ic[4];
            corners_rs(ring, spoke, a.na, ic);
            for (int q = 0; q < 4; q++)
                if (s[q] != 0) deposit_global(a.mfx, a.nz, a.mesh, z, ic[q], s[q]);
        }
        pend &= ~mm;
    }
}

// (u, t) of every particle's position (after a load)
template <typename R>
__global__ __launch_bounds__(GH_BLOCK) void ut_kernel(PushArgs<R> a) {
    typedef typename Vec4<R>::T R4;
    long long i = (long long)blockIdx.x * GH_BLOCK + threadIdx.x;
    if (i >= a.n) return;
    const R4 st = a.xv[i];
    R x = st.x, y = st.y;
    R rr = sqrt(x * x + y * y), u, t;
    ut_of(x, y, rr, a.k, u, t);
    a.ut[i] = mk2(u, t);
}

// weight(jst, old) and grdacc(jst) for arbitrary positions.
template <typename R>
__global__ __launch_bounds__(GH_BLOCK) void gather_kernel(GatherArgs<R> a) {
    typedef typename Vec2<R>::T R2;
    long long i = (long long)blockIdx.x * GH_BLOCK + threadIdx.x;
    if (i >= a.n) return;
    R w[4], rr, u, t;
    int c = weight1<R>(a.x[i], a.y[i], a.nr, a.k, a.fold != 0, a.thmax, w, rr,
                       u, t);
    if (a.cell) a.cell[i] = c;
    if (a.wt)
        for (int q = 0; q < 4; q++) a.wt[4 * i + q] = w[q];
    if (a.acc) {
        int ic[4];
        corners_rs(c / a.k.na, c % a.k.na, a.k.na, ic);
        int mesh = a.nr * a.k.na;
        R ax = 0, ay = 0;
        for (int q = 0; q < 4; q++) {
            R2 f = ic[q] < mesh ? a.fld[ic[q]] : zero2<R>();
            ax = ax + f.x * w[q];
            ay = ay + f.y * w[q];
        }
        a.acc[2 * i] = ax;
        a.acc[2 * i + 1] = ay;
    }
}

// getacc for one particle of list z0 (getacc.f:35-163): grid acceleration
// interpolated with the CIC weights of its cached (u, t) (grdacc.f:40-66),
// fixed radial force, zone rescaling; none off the grid.  The push's
// operation order.
template <typename R>
__device__ __forceinline__ void getacc1(const PushArgs<R> &a, const typename Vec4<R>::T &st,
                                        const typename Vec2<R>::T &cu, int z0, R &ax,
                                        R &ay) {
    typedef typename Vec2<R>::T R2;
    ax = 0;
    ay = 0;
    if (z0 != a.nz + 1) {
        R w[4];
        int ring, spoke, ic[4];
        cic(cu.x, cu.y, a.nr, a.na, ring, spoke, w);
        corners_rs(ring, spoke, a.na, ic);
        for (int q = 0; q < 4; q++) {
            R2 f = ic[q] < a.mesh ? a.fld[ic[q]] : zero2<R>();
            ax = ax + f.x * w[q];
            ay = ay + f.y * w[q];
        }
        if (a.fixforce) {
            const R rr = sqrt(st.x * st.x + st.y * st.y);
            R fr = (R)fixforce((double)rr, a.nfr, a.fr_inv_dr, a.fr_c);
            if (fr != (R)0) {
                fr = fr / rr;
                ax = ax + fr * st.x;
                ay = ay + fr * st.y;
            }
        }
    }
    if (a.nz > 1 && z0 > 1) {
        R t = tsfac(a, z0);
        t = t * t;
        ax = t * ax;
        ay = t * ay;
    }
}

// tmcent / tmcgrp for the retarded-velocity leap-frog (tmcent.f,
// tmcgrp.f:26-36): every particle's velocity minus fac times its getacc
// acceleration.
template <typename R>
__global__ __launch_bounds__(GH_BLOCK) void tmcent_kernel(PushArgs<R> a, R fac) {
    const long long i = (long long)blockIdx.x * GH_BLOCK + threadIdx.x;
    if (i >= a.n) return;
    typename Vec4<R>::T st = a.xv[i];
    R ax, ay;
    getacc1(a, st, a.ut[i], (int)a.zone[i], ax, ay);
    st.z = st.z - fac * ax;
    st.w = st.w - fac * ay;
    a.xv[i] = st;
}

// ---- group-level routines (parity use) --------------------------------------
//
// The lib15 routines that act on one group of particles -- getacc(jst),
// stpgrp(jst) with offgrd/addoff and rezgrp(jst) -- on a caller's group (device
// copies of its arrays), in the push kernel's operation order: the push fuses
// the same steps, and massgn, for every particle of a step.

// getacc(jst) (getacc.f:35-163) at the group's positions, their (u, t) formed
// as the push's cache holds them
template <typename R>
__global__ __launch_bounds__(GH_BLOCK) void getacc_group_kernel(PushArgs<R> a,
                                                                 typename Vec2<R>::T *acc) {
    const long long i = (long long)blockIdx.x * GH_BLOCK + threadIdx.x;
    if (i >= a.n) return;
    const typename Vec4<R>::T st = a.xv[i];
    const R rr = sqrt(st.x * st.x + st.y * st.y);
    R u, t, ax, ay;
    ut_of(st.x, st.y, rr, a.k, u, t);
    getacc1(a, st, mk2(u, t), (int)a.zone[i], ax, ay);
    acc[i] = mk2(ax, ay);
}

// stpgrp(jst) (stpgrp.f:27-56) with the given accelerations, the off-grid
// test and list change (offgrd.f:55-59, addoff.f:30-70, with chstep to the
// top zone's step when offanal), and rezgrp(jst) at the current mzone
// (rezgrp.f:30-50, zoneno.f:40-55, chstep.f:22-31); zones updated in place.
template <typename R>
__global__ __launch_bounds__(GH_BLOCK) void stpgrp_group_kernel(PushArgs<R> a,
                                                                 const typename Vec2<R>::T *acc) {
    const long long i = (long long)blockIdx.x * GH_BLOCK + threadIdx.x;
    if (i >= a.n) return;
    const typename Vec4<R>::T st = a.xv[i];
    const int nz = a.nz, nlists = nz + 1, z0 = a.zone[i];
    const bool offlist = z0 == nlists;
    R ox = st.x, oy = st.y, ovx = st.z, ovy = st.w;
    R ax_ = acc[i].x, ay_ = acc[i].y;
    R nvx = ovx + ax_, nvy = ovy + ay_;
    R nx = ox + nvx, ny = oy + nvy;
    int znew = z0;
    R rnew = sqrt(nx * nx + ny * ny);
    const bool off = rnew > a.rgrid;
    if (!offlist) {
        if (off) {
            znew = nlists;
            if (a.offanal && z0 < nz) {
                const R tfac = tsfac(a, nz) / tsfac(a, z0);
                chstep(tfac, ox, oy, ovx, ovy, ax_, ay_, nx, ny, nvx, nvy);
                rnew = sqrt(nx * nx + ny * ny);
            }
        }
    } else if (!off) {
        znew = nz;
    }
    if (a.mzone > 1 && !offlist && znew <= a.mzone) {
        int kz = zone_of(a, rnew);
        if (kz > a.mzone) kz = a.mzone;
        if (kz > znew) kz = zone_of(a, sqrt(ox * ox + oy * oy));
        if (kz != znew) {
            const R tfac = tsfac(a, kz) / tsfac(a, znew);
            chstep(tfac, ox, oy, ovx, ovy, ax_, ay_, nx, ny, nvx, nvy);
            znew = kz;
        }
    }
    a.xv[i] = mk4(nx, ny, nvx, nvy);
    a.zone[i] = (uint8_t)znew;
}

// ---- global integrals (icheck, SRC/lib15/icheck.f:24-70) ------------------
//
// Per particle, in FP64: 1, x, y, vx, vy, v^2, x vy - y vx, the grid
// potential interpolated with the CIC weights of its cached (u, t) (zones
// 1..nz, when a potential is given), the fixed potential (Hermite table,
// when given).  Velocities are per zone-1 step: the stored ones carry the
// step factor of their zone (the off-grid list that of the top zone).
// centred: the velocities time-centred as icheck takes them, v + a/2 in
// the working precision (anlgrp's (v + a/2)/tfac) before the division.  Each
// block writes its sums; a second pass adds the blocks in a fixed order, so
// the result does not depend on scheduling.
constexpr int NINT = 9;

__device__ __forceinline__ double wave_sumd(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

template <typename R>
__global__ __launch_bounds__(GH_BLOCK) void integrals_kernel(PushArgs<R> a, const R *pot,
                                                             const double4 *fp_c, int nfp,
                                                             double fp_dr, int centred,
                                                             double *part) {
    __shared__ double red[GH_BLOCK / 64][NINT];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const long long i = (long long)blockIdx.x * GH_BLOCK + tid;
    double v[NINT];
#pragma unroll
    for (int q = 0; q < NINT; q++) v[q] = 0.0;
    if (i < a.n) {
        const int z = a.zone[i];
        if (z >= 1) {
            typename Vec4<R>::T st = a.xv[i];
            if (centred) {
                R ax, ay;
                getacc1(a, st, a.ut[i], z, ax, ay);
                st.z = st.z + (R)0.5 * ax;
                st.w = st.w + (R)0.5 * ay;
            }
            const int zs = z <= a.nz ? z : a.nz;
            const double sf = 1.0 / (double)a.nstep[zs - 1];
            const double x = st.x, y = st.y, vx = st.z * sf, vy = st.w * sf;
            v[0] = 1.0;
            v[1] = x;
            v[2] = y;
            v[3] = vx;
            v[4] = vy;
            v[5] = vx * vx + vy * vy;
            v[6] = x * vy - y * vx;
            if (pot && z <= a.nz) {
                R w[4];
                int ring, spoke, ic[4];
                const typename Vec2<R>::T cu = a.ut[i];
                cic(cu.x, cu.y, a.nr, a.na, ring, spoke, w);
                corners_rs(ring, spoke, a.na, ic);
                double ph = 0.0;
                for (int q = 0; q < 4; q++)
                    if (ic[q] < a.mesh) ph += (double)w[q] * (double)pot[ic[q]];
                v[7] = ph;
            }
            if (nfp > 0)
                v[8] = fixforce(sqrt(x * x + y * y), nfp, 1.0 / fp_dr, fp_c);
        }
    }
#pragma unroll
    for (int q = 0; q < NINT; q++) {
        const double t = wave_sumd(v[q]);
        if (lane == 0) red[wave][q] = t;
    }
    __syncthreads();
    if (tid < NINT) {
        double t = 0.0;
        for (int w = 0; w < GH_BLOCK / 64; w++) t += red[w][tid];
        part[(long long)blockIdx.x * NINT + tid] = t;
    }
}

// out[q] = sum over blocks of part[b][q], blocks in order (one block)
__global__ __launch_bounds__(GH_BLOCK) void integrals_sum_kernel(const double *part, int nb,
                                                                 double *out) {
    __shared__ double red[GH_BLOCK / 64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int q = 0; q < NINT; q++) {
        double t = 0.0;
        for (int b = tid; b < nb; b += GH_BLOCK) t += part[(long long)b * NINT + q];
        t = wave_sumd(t);
        if (lane == 0) red[wave] = t;
        __syncthreads();
        if (tid == 0) {
            double s = 0.0;
            for (int w = 0; w < GH_BLOCK / 64; w++) s += red[w];
            out[q] = s;
        }
        __syncthreads();
    }
}

// sort key: the cell of the particle's (u, t) shifted right by kshift so
// that keys fit 16 bits (two radix passes); off-grid particles last.
template <typename R>
__global__ __launch_bounds__(GH_BLOCK) void key_kernel(
    const typename Vec2<R>::T *__restrict__ ut,
    const uint8_t *__restrict__ zone, long long n, int nr, int na, int nz,
    int kshift, unsigned kmax, unsigned short *__restrict__ key) {
    long long i = (long long)blockIdx.x * GH_BLOCK + threadIdx.x;
    if (i >= n) return;
    unsigned k = kmax;
    int z = zone[i];
    if (z >= 1 && z <= nz) {
        R w[4];
        int ring, spoke;
        typename Vec2<R>::T c = ut[i];
        cic(c.x, c.y, nr, na, ring, spoke, w);
        k = (unsigned)(ring * na + spoke) >> kshift;
        if (k >= kmax) k = kmax - 1;
    }
    key[i] = (unsigned short)k;
}

// Ensemble sort key: (member << kbits) | the member's cell key (off-grid
// particles last within their member), so that one stable sort groups the
// members in order and cells within each.  One block per entry of the push
// block table, which never straddles two members.
template <typename R>
__global__ __launch_bounds__(GH_BLOCK) void key_ens_kernel(
    const typename Vec2<R>::T *__restrict__ ut, const uint8_t *__restrict__ zone,
    int nr, int na, int nz, int kbits, unsigned kmax, const int4 *__restrict__ blk,
    unsigned *__restrict__ key, unsigned *__restrict__ idx) {
    // block b: the push block table's (member, first particle, count)
    const int4 b = blk[blockIdx.x];
    for (int j = threadIdx.x; j < b.z; j += GH_BLOCK) {
        const long long i = (long long)b.y + j;
        unsigned k = kmax;
        const int z = zone[i];
        if (z >= 1 && z <= nz) {
            R w[4];
            int ring, spoke;
            const typename Vec2<R>::T c = ut[i];
            cic(c.x, c.y, nr, na, ring, spoke, w);
            k = (unsigned)(ring * na + spoke);
            if (k >= kmax) k = kmax - 1;
        }
        key[i] = ((unsigned)b.x << kbits) | k;
        idx[i] = (unsigned)i;
    }
}

// one pass moving every particle array to sorted order
template <typename R> struct PermArgs {
    const typename Vec4<R>::T *xv_in;
    typename Vec4<R>::T *xv_out;
    const typename Vec2<R>::T *ut_in;
    typename Vec2<R>::T *ut_out;
    const uint8_t *zin;
    uint8_t *zout;
    const unsigned *iin;
    unsigned *iout;
    const unsigned *perm;
    long long n;
    const typename Vec2<R>::T *lo_in;   // compensated positions' low words, or NULL
    typename Vec2<R>::T *lo_out;
};

template <typename R>
__global__ __launch_bounds__(GH_BLOCK) void permute_all_kernel(PermArgs<R> p) {
    long long i = (long long)blockIdx.x * GH_BLOCK + threadIdx.x;
    if (i >= p.n) return;
    unsigned s = p.perm[i];
    p.xv_out[i] = p.xv_in[s];
    p.ut_out[i] = p.ut_in[s];
    p.zout[i] = p.zin[s];
    p.iout[i] = p.iin[s];
    if (p.lo_in) p.lo_out[i] = p.lo_in[s];
}

// load: planar host arrays (x, y, vx, vy, es*n bytes each, staged one after
// another in s) -> interleaved storage; ids = load order
template <typename R>
__global__ __launch_bounds__(GH_BLOCK) void interleave_kernel(const R *__restrict__ s,
                                                              typename Vec4<R>::T *__restrict__ xv,
                                                              unsigned *__restrict__ id,
                                                              long long n) {
    const long long i = (long long)blockIdx.x * GH_BLOCK + threadIdx.x;
    if (i >= n) return;
    typename Vec4<R>::T v;
    v.x = s[i];
    v.y = s[n + i];
    v.z = s[2 * n + i];
    v.w = s[3 * n + i];
    xv[i] = v;
    id[i] = (unsigned)i;
}

// unload: interleaved storage -> planar arrays in load order (d[k n + id]),
// zones widened to int32
template <typename R>
__global__ __launch_bounds__(GH_BLOCK) void deinterleave_kernel(const typename Vec4<R>::T *__restrict__ xv,
                                                                const uint8_t *__restrict__ zone,
                                                                const unsigned *__restrict__ id,
                                                                R *__restrict__ d,
                                                                int *__restrict__ iz, long long n) {
    const long long i = (long long)blockIdx.x * GH_BLOCK + threadIdx.x;
    if (i >= n) return;
    const long long p = id[i];
    const typename Vec4<R>::T v = xv[i];
    d[p] = v.x;
    d[n + p] = v.y;
    d[2 * n + p] = v.z;
    d[3 * n + p] = v.w;
    iz[p] = zone[i];
}

// ---- host launchers ---------------------------------------------------------

static inline unsigned nblocks(long long n, long long per) {
    return (unsigned)((n + per - 1) / per);
}

template <typename R, bool COMP>
static void launch_push_c(const PushArgs<R> &a, hipStream_t s) {
    if (a.blk) {
        if (a.nblocks > 0)
            hipLaunchKernelGGL(HIP_KERNEL_NAME(push_kernel<R, true, COMP>), dim3(a.nblocks),
                               dim3(PB), 0, s, a);
    } else {
        hipLaunchKernelGGL(HIP_KERNEL_NAME(push_kernel<R, false, COMP>),
                           dim3(nblocks(a.n, PB * PPT)), dim3(PB), 0, s, a);
    }
}

template <typename R>
hipError_t launch_push(const PushArgs<R> &a, hipStream_t s) {
    if (a.n <= 0) return hipSuccess;
    if constexpr (sizeof(R) == 4) {
        if (a.xlo) launch_push_c<R, true>(a, s);
        else launch_push_c<R, false>(a, s);
    } else {
        if (a.xlo) return hipErrorInvalidValue;   // compensated positions are FP32 only
        launch_push_c<R, false>(a, s);
    }
    return hipGetLastError();
}

template <typename R>
hipError_t launch_massgn(const PushArgs<R> &a, hipStream_t s) {
    if (a.n <= 0) return hipSuccess;
    hipLaunchKernelGGL(massgn_kernel<R>, dim3(nblocks(a.n, GH_BLOCK)),
                       dim3(GH_BLOCK), 0, s, a);
    return hipGetLastError();
}

template <typename R>
hipError_t launch_getacc_group(const PushArgs<R> &a, typename Vec2<R>::T *acc, hipStream_t s) {
    if (a.n <= 0) return hipSuccess;
    hipLaunchKernelGGL(getacc_group_kernel<R>, dim3(nblocks(a.n, GH_BLOCK)), dim3(GH_BLOCK), 0,
                       s, a, acc);
    return hipGetLastError();
}

template <typename R>
hipError_t launch_stpgrp_group(const PushArgs<R> &a, const typename Vec2<R>::T *acc,
                               hipStream_t s) {
    if (a.n <= 0) return hipSuccess;
    hipLaunchKernelGGL(stpgrp_group_kernel<R>, dim3(nblocks(a.n, GH_BLOCK)), dim3(GH_BLOCK), 0,
                       s, a, acc);
    return hipGetLastError();
}

template <typename R>
hipError_t launch_ut(const PushArgs<R> &a, hipStream_t s) {
    if (a.n <= 0) return hipSuccess;
    hipLaunchKernelGGL(ut_kernel<R>, dim3(nblocks(a.n, GH_BLOCK)),
                       dim3(GH_BLOCK), 0, s, a);
    return hipGetLastError();
}

template <typename R>
hipError_t launch_gather(const GatherArgs<R> &a, hipStream_t s) {
    if (a.n <= 0) return hipSuccess;
    hipLaunchKernelGGL(gather_kernel<R>, dim3(nblocks(a.n, GH_BLOCK)),
                       dim3(GH_BLOCK), 0, s, a);
    return hipGetLastError();
}

template <typename R>
hipError_t launch_integrals(const PushArgs<R> &a, const R *pot, const double4 *fp_c,
                            int nfp, double fp_dr, int centred,
                            double *part, double *out, hipStream_t s) {
    const unsigned nb = a.n > 0 ? nblocks(a.n, GH_BLOCK) : 0;
    if (nb)
        hipLaunchKernelGGL(integrals_kernel<R>, dim3(nb), dim3(GH_BLOCK), 0, s, a, pot,
                           fp_c, nfp, fp_dr, centred, part);
    hipLaunchKernelGGL(integrals_sum_kernel, dim3(1), dim3(GH_BLOCK), 0, s,
                       (const double *)part, (int)nb, out);
    return hipGetLastError();
}

template <typename R>
hipError_t launch_tmcent(const PushArgs<R> &a, R fac, hipStream_t s) {
    if (a.n <= 0) return hipSuccess;
    hipLaunchKernelGGL(tmcent_kernel<R>, dim3(nblocks(a.n, GH_BLOCK)),
                       dim3(GH_BLOCK), 0, s, a, fac);
    return hipGetLastError();
}

template <typename R>
hipError_t launch_key(const typename Vec2<R>::T *ut, const uint8_t *zone,
                      long long n, int nr, int na, int nz, int kshift,
                      unsigned kmax, unsigned short *key, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(key_kernel<R>, dim3(nblocks(n, GH_BLOCK)), dim3(GH_BLOCK),
                       0, s, ut, zone, n, nr, na, nz, kshift, kmax, key);
    return hipGetLastError();
}

template <typename R>
hipError_t launch_key_ens(const typename Vec2<R>::T *ut, const uint8_t *zone, int nblk,
                          int nr, int na, int nz, int kbits, unsigned kmax, const int4 *blk,
                          unsigned *key, unsigned *idx, hipStream_t s) {
    if (nblk <= 0) return hipSuccess;
    hipLaunchKernelGGL(key_ens_kernel<R>, dim3(nblk), dim3(GH_BLOCK), 0, s, ut, zone, nr, na,
                       nz, kbits, kmax, blk, key, idx);
    return hipGetLastError();
}

template <typename R>
hipError_t launch_interleave(const R *stage, typename Vec4<R>::T *xv, unsigned *id, long long n,
                             hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(interleave_kernel<R>, dim3(nblocks(n, GH_BLOCK)), dim3(GH_BLOCK), 0, s,
                       stage, xv, id, n);
    return hipGetLastError();
}

template <typename R>
hipError_t launch_deinterleave(const typename Vec4<R>::T *xv, const uint8_t *zone,
                               const unsigned *id, R *stage, int *iz, long long n,
                               hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(deinterleave_kernel<R>, dim3(nblocks(n, GH_BLOCK)), dim3(GH_BLOCK), 0, s,
                       xv, zone, id, stage, iz, n);
    return hipGetLastError();
}

template <typename R>
hipError_t launch_permute_all(const PermArgs<R> &p, hipStream_t s) {
    if (p.n <= 0) return hipSuccess;
    hipLaunchKernelGGL(permute_all_kernel<R>, dim3(nblocks(p.n, GH_BLOCK)),
                       dim3(GH_BLOCK), 0, s, p);
    return hipGetLastError();
}

#define GH_INST(R)                                                            \
    template hipError_t launch_push<R>(const PushArgs<R> &, hipStream_t);     \
    template hipError_t launch_massgn<R>(const PushArgs<R> &, hipStream_t);   \
    template hipError_t launch_ut<R>(const PushArgs<R> &, hipStream_t);       \
    template hipError_t launch_getacc_group<R>(const PushArgs<R> &,           \
                                               typename Vec2<R>::T *, hipStream_t); \
    template hipError_t launch_stpgrp_group<R>(const PushArgs<R> &,           \
                                               const typename Vec2<R>::T *,   \
                                               hipStream_t);                  \
    template hipError_t launch_gather<R>(const GatherArgs<R> &, hipStream_t); \
    template hipError_t launch_tmcent<R>(const PushArgs<R> &, R, hipStream_t); \
    template hipError_t launch_integrals<R>(const PushArgs<R> &, const R *,     \
                                            const double4 *, int, double, int,   \
                                            double *, double *, hipStream_t);    \
    template hipError_t launch_key<R>(const typename Vec2<R>::T *,           \
                                      const uint8_t *, long long, int, int,   \
                                      int, int, unsigned, unsigned short *,   \
                                      hipStream_t);                           \
    template hipError_t launch_key_ens<R>(const typename Vec2<R>::T *, const uint8_t *, \
                                          int, int, int, int, int, unsigned,          \
                                          const int4 *, unsigned *, unsigned *,       \
                                          hipStream_t);                               \
    template hipError_t launch_permute_all<R>(const PermArgs<R> &, hipStream_t); \
    template hipError_t launch_interleave<R>(const R *, typename Vec4<R>::T *, unsigned *, \
                                             long long, hipStream_t);                      \
    template hipError_t launch_deinterleave<R>(const typename Vec4<R>::T *, const uint8_t *, \
                                               const unsigned *, R *, int *, long long,      \
                                               hipStream_t);
GH_INST(float)
GH_INST(double)

}  // namespace gh
