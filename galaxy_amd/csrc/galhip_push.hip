// galhip_push.hip -- the fused particle kernel of the p2d step and the
// group-level gather kernels.
//
// push_kernel does, for each particle of the zones due at this step, what
// the reference's per-group pipeline does (SRC/lib15/step.f:186-234):
//   getacc/grdacc  weights at the old position, 4-corner interpolation of
//                  the Cartesian field, fixed radial force, zone scaling
//                  (getacc.f:35-163, grdacc.f:40-66)
//   stpgrp/offgrd  RVLF leap-frog and the off-grid test (stpgrp.f:27-56,
//                  offgrd.f:55-59, addoff.f:30-70)
//   rezgrp/chstep  zone changes (rezgrp.f:30-50, zoneno.f, chstep.f:22-31)
//   massgn         weights at the new position, scatter into the zone's
//                  mass array (massgn.f:68-89)
// in one pass over structure-of-arrays particle storage.  Mass sums are
// 64-bit fixed point (weight x 2^qbits): integer adds commute, so the
// result is bit-identical for any processing order and needs no float
// atomics.  Each block first accumulates into an LDS tile covering the
// rings its particles touch (particles are kept ordered by cell, so a
// block's particles share one or two rings) and flushes the tile with one
// 64-bit global atomic per touched cell.
#include "galhip_kern.h"
#include <climits>

namespace gh {

constexpr int PPT = 4;            // particles per thread
constexpr int TILE_CAP = 4096;    // LDS tile, cells (32 KiB of int64)

template <typename R>
__device__ __forceinline__ R tsfac(const PushArgs<R> &a, int iz) {
    int i = iz < a.nz ? iz : a.nz;
    return (R)a.nstep[i - 1];
}

template <typename R>
__device__ __forceinline__ int zoneno(const PushArgs<R> &a, R x, R y) {
    R rad = sqrt(x * x + y * y);
    int z = 1;
    if (a.nz > 1)
        for (int i = 1; i <= a.nz; i++)
            if (rad >= a.zbr[i - 1]) z = i;
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

template <typename R>
__global__ __launch_bounds__(GH_BLOCK) void push_kernel(PushArgs<R> a) {
    typedef typename Vec2<R>::T R2;
    __shared__ long long tile[TILE_CAP];
    __shared__ int s_jr[GH_MAXZ], s_kr[GH_MAXZ], s_cnt[GH_MAXZ];
    __shared__ int s_dnoff, s_pushes, s_zt, s_base, s_len;

    const int tid = threadIdx.x;
    const long long base = (long long)blockIdx.x * (GH_BLOCK * PPT);
    const int nz = a.nz, nlists = a.nz + 1, na = a.na, nr = a.nr;
    if (tid < GH_MAXZ) {
        s_jr[tid] = INT_MAX;
        s_kr[tid] = -1;
        s_cnt[tid] = 0;
    }
    if (tid == 0) {
        s_dnoff = 0;
        s_pushes = 0;
    }

    int zin[PPT];
    bool mv[PPT];
    int any = 0;
#pragma unroll
    for (int k = 0; k < PPT; k++) {
        long long i = base + k * GH_BLOCK + tid;
        zin[k] = i < a.n ? (int)a.zone[i] : 0;
        bool off = zin[k] == nlists;
        mv[k] = zin[k] > 0 &&
                (off ? (a.offanal && a.mzone == nz) : (zin[k] <= a.mzone));
        any |= mv[k];
    }
    // every thread reaches this barrier; a block with nothing to move exits
    if (!__syncthreads_or(any)) return;

    int dcell[PPT], dzone[PPT];
    R dw[PPT][4];
    int my_pushes = 0, my_dnoff = 0;
#pragma unroll
    for (int k = 0; k < PPT; k++) {
        dcell[k] = -1;
        dzone[k] = 0;
        if (!mv[k]) continue;
        const long long i = base + k * GH_BLOCK + tid;
        const int z0 = zin[k];
        const bool offlist = z0 == nlists;
        R ox = a.x[i], oy = a.y[i], ovx = a.vx[i], ovy = a.vy[i];
        R ax = 0, ay = 0;
        my_pushes++;
        if (!offlist) {
            R w[4], rr;
            int c = weight1<R>(ox, oy, nr, na, a.alpha, false, a.thmax, w, rr);
            int ic[4];
            corners(c, na, ic);
#pragma unroll
            for (int q = 0; q < 4; q++) {
                // the last ring with t == na points past the field: read 0
                R2 f = ic[q] < a.mesh ? a.fld[ic[q]] : zero2<R>();
                ax = ax + f.x * w[q];
                ay = ay + f.y * w[q];
            }
            if (a.fixforce) {
                R fr = (R)fixforce((double)rr, a.nfr, a.fr_dr, a.fr_val, a.fr_der);
                if (fr != (R)0) {
                    fr = fr / rr;
                    ax = ax + fr * ox;
                    ay = ay + fr * oy;
                }
            }
        }
        if (nz > 1 && z0 > 1) {
            R t = tsfac(a, z0);
            t = t * t;
            ax = t * ax;
            ay = t * ay;
        }
        R nvx = ovx + ax, nvy = ovy + ay;
        R nx = ox + nvx, ny = oy + nvy;
        int znew = z0;
        R rnew = sqrt(nx * nx + ny * ny);
        bool off = rnew > a.rgrid;
        if (!offlist) {
            if (off) {
                znew = nlists;
                my_dnoff++;
                if (a.offanal && z0 < nz) {
                    R tfac = tsfac(a, nz) / tsfac(a, z0);
                    chstep(tfac, ox, oy, ovx, ovy, ax, ay, nx, ny, nvx, nvy);
                }
            }
        } else if (!off) {
            my_dnoff--;
            znew = nz;
        }
        if (a.mzone > 1 && !offlist && znew <= a.mzone) {
            int kz = zoneno(a, nx, ny);
            if (kz > a.mzone) kz = a.mzone;
            if (kz > znew) kz = zoneno(a, ox, oy);
            if (kz != znew) {
                R tfac = tsfac(a, kz) / tsfac(a, znew);
                chstep(tfac, ox, oy, ovx, ovy, ax, ay, nx, ny, nvx, nvy);
                znew = kz;
            }
        }
        a.x[i] = nx;
        a.y[i] = ny;
        a.vx[i] = nvx;
        a.vy[i] = nvy;
        if (znew != z0) a.zone[i] = (uint8_t)znew;
        if (znew >= 1 && znew <= nz) {
            R rr;
            int c = weight1<R>(nx, ny, nr, na, a.alpha, a.nsect > 1, a.thmax,
                               dw[k], rr);
            dcell[k] = c;
            dzone[k] = znew;
            int ring = c / na;
            atomicMin(&s_jr[znew - 1], ring);
            atomicMax(&s_kr[znew - 1], ring);
            atomicAdd(&s_cnt[znew - 1], 1);
        }
    }
    if (my_pushes) atomicAdd(&s_pushes, my_pushes);
    if (my_dnoff) atomicAdd(&s_dnoff, my_dnoff);
    __syncthreads();

    // choose the tile: the zone with most depositing particles, rings
    // [jr, kr + 1] (corners reach the next ring)
    if (tid == 0) {
        int zt = -1, best = 0;
        for (int z = 0; z < nz; z++)
            if (s_cnt[z] > best) {
                best = s_cnt[z];
                zt = z;
            }
        int len = 0;
        if (zt >= 0) {
            len = (s_kr[zt] - s_jr[zt] + 2) * na;
            if (len > TILE_CAP || s_kr[zt] + 1 >= nr) zt = -1;
        }
        s_zt = zt;
        s_base = zt >= 0 ? s_jr[zt] * na : 0;
        s_len = zt >= 0 ? len : 0;
    }
    __syncthreads();
    const int zt = s_zt, tbase = s_base, tlen = s_len;
    for (int j = tid; j < tlen; j += GH_BLOCK) tile[j] = 0;
    __syncthreads();

    const long long total = (long long)nz * a.mesh;
#pragma unroll
    for (int k = 0; k < PPT; k++) {
        if (dcell[k] < 0) continue;
        int ic[4];
        corners(dcell[k], na, ic);
        const bool in_tile = (dzone[k] - 1) == zt;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            long long v = quant<R>(dw[k][q], a.qbits);
            if (v == 0) continue;
            if (in_tile) {
                atomicAdd((unsigned long long *)&tile[ic[q] - tbase],
                          (unsigned long long)v);
            } else {
                // flat index over the [nz][mesh] zone arrays, as Fortran's
                // grdmss(j, jz); past the last array the weight is dropped
                long long f = (long long)(dzone[k] - 1) * a.mesh + ic[q];
                if (f < total) {
                    int zf = (int)(f / a.mesh);
                    int cf = (int)(f - (long long)zf * a.mesh);
                    atomicAdd((unsigned long long *)&a.mfx[zf][cf],
                              (unsigned long long)v);
                }
            }
        }
    }
    __syncthreads();
    for (int j = tid; j < tlen; j += GH_BLOCK) {
        long long v = tile[j];
        if (v != 0)
            atomicAdd((unsigned long long *)&a.mfx[zt][tbase + j],
                      (unsigned long long)v);
    }
    if (tid < nz && s_cnt[tid] > 0) {
        atomicMin(&a.jrad[tid], s_jr[tid] + 1);
        int k = s_kr[tid] + 2;
        atomicMax(&a.krad[tid], k < nr ? k : nr);
    }
    if (tid == 0) {
        if (s_dnoff) atomicAdd((unsigned long long *)&a.counters[0],
                               (unsigned long long)(long long)s_dnoff);
        if (s_pushes) atomicAdd((unsigned long long *)&a.counters[1],
                                (unsigned long long)s_pushes);
    }
}

// masset / massgn for all on-grid particles (no push): the deposit half of
// push_kernel, through global atomics.
template <typename R>
__global__ __launch_bounds__(GH_BLOCK) void massgn_kernel(PushArgs<R> a) {
    long long i = (long long)blockIdx.x * GH_BLOCK + threadIdx.x;
    if (i >= a.n) return;
    int z = a.zone[i];
    if (z < 1 || z > a.nz) return;
    R w[4], rr;
    int c = weight1<R>(a.x[i], a.y[i], a.nr, a.na, a.alpha, a.nsect > 1,
                       a.thmax, w, rr);
    int ring = c / a.na;
    atomicMin(&a.jrad[z - 1], ring + 1);
    int k = ring + 2;
    atomicMax(&a.krad[z - 1], k < a.nr ? k : a.nr);
    int ic[4];
    corners(c, a.na, ic);
    const long long total = (long long)a.nz * a.mesh;
    for (int q = 0; q < 4; q++) {
        long long v = quant<R>(w[q], a.qbits);
        if (v == 0) continue;
        long long f = (long long)(z - 1) * a.mesh + ic[q];
        if (f < total) {
            int zf = (int)(f / a.mesh);
            int cf = (int)(f - (long long)zf * a.mesh);
            atomicAdd((unsigned long long *)&a.mfx[zf][cf], (unsigned long long)v);
        }
    }
}

// weight(jst, old) and grdacc(jst) for arbitrary positions.
template <typename R>
__global__ __launch_bounds__(GH_BLOCK) void gather_kernel(GatherArgs<R> a) {
    typedef typename Vec2<R>::T R2;
    long long i = (long long)blockIdx.x * GH_BLOCK + threadIdx.x;
    if (i >= a.n) return;
    R w[4], rr;
    int c = weight1<R>(a.x[i], a.y[i], a.nr, a.na, a.alpha, a.fold != 0,
                       a.thmax, w, rr);
    if (a.cell) a.cell[i] = c;
    if (a.wt)
        for (int q = 0; q < 4; q++) a.wt[4 * i + q] = w[q];
    if (a.acc) {
        int ic[4];
        corners(c, a.na, ic);
        int mesh = a.nr * a.na;
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

// sort key: cell of the (unfolded) position; off-grid particles last.
template <typename R>
__global__ __launch_bounds__(GH_BLOCK) void key_kernel(
    const R *__restrict__ x, const R *__restrict__ y,
    const uint8_t *__restrict__ zone, long long n, int nr, int na, int nz,
    R alpha, unsigned *__restrict__ key, unsigned *__restrict__ idx) {
    long long i = (long long)blockIdx.x * GH_BLOCK + threadIdx.x;
    if (i >= n) return;
    unsigned k = (unsigned)(nr * na);
    int z = zone[i];
    if (z >= 1 && z <= nz) {
        R w[4], rr;
        k = (unsigned)weight1<R>(x[i], y[i], nr, na, alpha, false, (R)na, w, rr);
    }
    key[i] = k;
    idx[i] = (unsigned)i;
}

template <typename T>
__global__ __launch_bounds__(GH_BLOCK) void permute_kernel(
    const T *__restrict__ in, T *__restrict__ out,
    const unsigned *__restrict__ perm, long long n) {
    long long i = (long long)blockIdx.x * GH_BLOCK + threadIdx.x;
    if (i < n) out[i] = in[perm[i]];
}

// ---- host launchers ---------------------------------------------------------

static inline unsigned nblocks(long long n, long long per) {
    return (unsigned)((n + per - 1) / per);
}

template <typename R>
hipError_t launch_push(const PushArgs<R> &a, hipStream_t s) {
    if (a.n <= 0) return hipSuccess;
    hipLaunchKernelGGL(push_kernel<R>, dim3(nblocks(a.n, GH_BLOCK * PPT)),
                       dim3(GH_BLOCK), 0, s, a);
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
hipError_t launch_gather(const GatherArgs<R> &a, hipStream_t s) {
    if (a.n <= 0) return hipSuccess;
    hipLaunchKernelGGL(gather_kernel<R>, dim3(nblocks(a.n, GH_BLOCK)),
                       dim3(GH_BLOCK), 0, s, a);
    return hipGetLastError();
}

template <typename R>
hipError_t launch_key(const R *x, const R *y, const uint8_t *zone, long long n,
                      int nr, int na, int nz, R alpha, unsigned *key,
                      unsigned *idx, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(key_kernel<R>, dim3(nblocks(n, GH_BLOCK)), dim3(GH_BLOCK),
                       0, s, x, y, zone, n, nr, na, nz, alpha, key, idx);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_permute(const T *in, T *out, const unsigned *perm,
                          long long n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(permute_kernel<T>, dim3(nblocks(n, GH_BLOCK)),
                       dim3(GH_BLOCK), 0, s, in, out, perm, n);
    return hipGetLastError();
}

template hipError_t launch_push<float>(const PushArgs<float> &, hipStream_t);
template hipError_t launch_push<double>(const PushArgs<double> &, hipStream_t);
template hipError_t launch_massgn<float>(const PushArgs<float> &, hipStream_t);
template hipError_t launch_massgn<double>(const PushArgs<double> &, hipStream_t);
template hipError_t launch_gather<float>(const GatherArgs<float> &, hipStream_t);
template hipError_t launch_gather<double>(const GatherArgs<double> &, hipStream_t);
template hipError_t launch_key<float>(const float *, const float *,
                                      const uint8_t *, long long, int, int,
                                      int, float, unsigned *, unsigned *,
                                      hipStream_t);
template hipError_t launch_key<double>(const double *, const double *,
                                       const uint8_t *, long long, int, int,
                                       int, double, unsigned *, unsigned *,
                                       hipStream_t);
template hipError_t launch_permute<float>(const float *, float *,
                                          const unsigned *, long long,
                                          hipStream_t);
template hipError_t launch_permute<double>(const double *, double *,
                                           const unsigned *, long long,
                                           hipStream_t);
template hipError_t launch_permute<uint8_t>(const uint8_t *, uint8_t *,
                                            const unsigned *, long long,
                                            hipStream_t);
template hipError_t launch_permute<unsigned>(const unsigned *, unsigned *,
                                             const unsigned *, long long,
                                             hipStream_t);

}  // namespace gh
