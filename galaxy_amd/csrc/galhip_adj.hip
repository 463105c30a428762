// galhip_adj.hip -- reverse-mode derivative (adjoint) of the p2d leap-frog
// step, for one zone and the whole grid (full steps).  The reference has no
// adjoint; this differentiates the map the forward kernels compute:
//
//   rho   = pmass sum_i w(x_i)                 mass assignment (massgn)
//   F     = K rho                              findf: transform, Green's
//                                              contraction, synthesis, polcat
//   a_i   = sum_q w_q(x_i) F(c_q) + h(r_i) x_i / r_i     grdacc + fixed force
//   v'    = v + a,  x' = x + v'                stpgrp (off-grid: a = 0)
//
// Given the adjoints (lx, lv) of (x', v'), one backward step forms
//   la = lv + lx                               (adjoint of v' and of a)
//   g  = sum_i w(x_i) la_i                     adj_scatter: transpose of grdacc
//   rb = K^T g                                 adj_wbar + adj_rbar
//   lx += J_direct^T la + pmass sum_q rb(c_q) grad w_q      adj_gather
//   lv  = la
// with J_direct the derivative of a_i in x_i through the interpolation
// weights and the fixed force.  The weights are CIC in the grid coordinates
// u = log(1 + r) / alpha, t = atan2(y, x) na / 2 pi (weight.f:61-97), so
// grad w_q = dw_q/du grad u + dw_q/dt grad t; the cell and range choices are
// held fixed (the map is piecewise smooth).  Adjoints and grids are FP64;
// the forward state (positions, (u, t), field) is the context's.
#include "galhip_ctx.h"

#include <climits>
#include <cstring>
#include <vector>

namespace gh {

template <typename R> struct AdjArgs {
    const typename Vec4<R>::T *xv;
    const typename Vec2<R>::T *ut;
    const uint8_t *zone;
    const unsigned *id;
    long long n;
    double *lam;                      // [4][n] in storage order: x, y, vx, vy
    double *gx, *gy;                  // [mesh]
    const double *rbar;               // [mesh]
    const typename Vec2<R>::T *fld;   // forward (F_x, F_y)
    const double4 *fr_c;              // fixed force table (coefficients)
    double fr_inv_dr;
    int nfr;
    const double4 *fd_c;              // its derivative in the model parameter
    double fd_inv_dr;
    int nfd;
    double pmass, inv_alpha, tscale;
    int nr, na, mesh;
    double *part;                     // [blocks][2] parameter partial sums
};

struct AdjSolveArgs {
    const double *gx, *gy;
    double *wbar;                     // [2 types][2 (cos, sin)][ng][nr]
    double *rbar;                     // [mesh]
    const double *G;                  // FP64 tables [3][ng][nr][nr]
    const double *twd;                // cos, sin of 2 pi k / na
    const int *lg, *ranges;
    int nr, na, ng, mmax;
};

// CIC weights and their derivatives in (u, t), with the cell choice of cic()
__device__ __forceinline__ void cic_grad(double u, double t, int nr, int na, int &ring,
                                         int &spoke, double w[4], double wu[4], double wt[4]) {
    int ju = (int)u;
    if (ju > nr - 2) ju = nr - 2;
    const int jt = (int)t;
    const double du = u - (double)ju, dt = t - (double)jt;
    w[0] = (1.0 - du) * (1.0 - dt);
    w[1] = dt * (1.0 - du);
    w[2] = du * (1.0 - dt);
    w[3] = du * dt;
    wu[0] = -(1.0 - dt);
    wu[1] = -dt;
    wu[2] = 1.0 - dt;
    wu[3] = dt;
    wt[0] = -(1.0 - du);
    wt[1] = 1.0 - du;
    wt[2] = -du;
    wt[3] = du;
    ring = ju;
    spoke = jt;
    if (jt >= na) {
        ring = ju + 1;
        spoke = jt - na;
    }
}

// dst[k][ids[i]] = src[k][i] (storage order -> load order), k = 0..3
__global__ __launch_bounds__(GH_BLOCK) void adj_to_load_kernel(const double *__restrict__ src,
                                                               const unsigned *__restrict__ ids,
                                                               double *__restrict__ dst,
                                                               long long n) {
    const long long i = (long long)blockIdx.x * GH_BLOCK + threadIdx.x;
    if (i >= n) return;
    const long long p = ids[i];
#pragma unroll
    for (int k = 0; k < 4; k++) dst[k * n + p] = src[k * n + i];
}

// dst[k][i] = src[k][ids[i]] (load order -> storage order)
__global__ __launch_bounds__(GH_BLOCK) void adj_to_store_kernel(const double *__restrict__ src,
                                                                const unsigned *__restrict__ ids,
                                                                double *__restrict__ dst,
                                                                long long n) {
    const long long i = (long long)blockIdx.x * GH_BLOCK + threadIdx.x;
    if (i >= n) return;
    const long long p = ids[i];
#pragma unroll
    for (int k = 0; k < 4; k++) dst[k * n + i] = src[k * n + p];
}

__device__ __forceinline__ int wave_min_i(int v) { return wave_min_i32(v); }
__device__ __forceinline__ int wave_max_i(int v) { return wave_max_i32(v); }

constexpr int ADJ_TILE = 1024;   // LDS window cells (x 2 components x 8 B)

// la = lv + lx stored as the new lv; on-grid particles scatter la with their
// weights (the transpose of grdacc's interpolation).  Particles are kept
// ordered by cell, so a block's corners fall in a small window of rings x
// spokes: the block accumulates them in LDS (FP64 LDS atomics) and adds
// each touched cell to the grid with one global atomic; a block whose
// window exceeds the tile adds every corner globally.
template <typename R>
__global__ __launch_bounds__(GH_BLOCK) void adj_scatter_kernel(AdjArgs<R> a) {
    __shared__ double tile[2][ADJ_TILE];
    __shared__ int s_first, s_sref, s_rmin, s_rmax, s_dmin, s_dmax;
    const int tid = threadIdx.x, lane = tid & 63, na = a.na, half = na >> 1;
    const long long i = (long long)blockIdx.x * GH_BLOCK + tid;
    if (tid == 0) {
        s_first = INT_MAX;
        s_rmin = INT_MAX;
        s_rmax = -1;
        s_dmin = INT_MAX;
        s_dmax = INT_MIN;
    }
    const bool in = i < a.n;
    const long long n = a.n, p = i;
    double lax = 0.0, lay = 0.0;
    if (in) {
        lax = a.lam[2 * n + p] + a.lam[p];
        lay = a.lam[3 * n + p] + a.lam[n + p];
        a.lam[2 * n + p] = lax;
        a.lam[3 * n + p] = lay;
    }
    const bool on = in && a.zone[i] == 1;
    double w[4] = {0.0, 0.0, 0.0, 0.0}, wu[4], wt[4];
    int ring = 0, spoke = 0;
    if (on) {
        const typename Vec2<R>::T cu = a.ut[i];
        cic_grad((double)cu.x, (double)cu.y, a.nr, na, ring, spoke, w, wu, wt);
    }
    __syncthreads();
    // reference spoke: the first on-grid particle's (one atomic per wave)
    {
        const unsigned long long bm = __ballot(on);
        if (bm && lane == __ffsll((long long)bm) - 1) atomicMin(&s_first, tid);
    }
    __syncthreads();
    if (tid == s_first) s_sref = spoke;
    __syncthreads();
    if (s_first == INT_MAX) return;   // no on-grid particle: nothing to add
    const int sref = s_sref;
    int d = spoke - sref;
    if (d < -half) d += na;
    else if (d >= half) d -= na;
    {
        const int rmin = wave_min_i(on ? ring : INT_MAX), rmax = wave_max_i(on ? ring : -1);
        const int dmin = wave_min_i(on ? d : INT_MAX), dmax = wave_max_i(on ? d : INT_MIN);
        if (lane == 0 && rmax >= 0) {
            atomicMin(&s_rmin, rmin);
            atomicMax(&s_rmax, rmax);
            atomicMin(&s_dmin, dmin);
            atomicMax(&s_dmax, dmax);
        }
    }
    __syncthreads();
    const int r0 = s_rmin, d0 = s_dmin;
    const int tw = s_dmax - d0 + 2, th = s_rmax - r0 + 2;
    const bool use_tile = th * tw <= ADJ_TILE;
    const int tlen = use_tile ? th * tw : 0;
    for (int j = tid; j < tlen; j += GH_BLOCK) {
        tile[0][j] = 0.0;
        tile[1][j] = 0.0;
    }
    __syncthreads();
    if (on) {
        if (use_tile) {
            const int b = (ring - r0) * tw + (d - d0);
            const int ti[4] = {b, b + 1, b + tw, b + tw + 1};
#pragma unroll
            for (int q = 0; q < 4; q++) {
                atomicAdd(&tile[0][ti[q]], w[q] * lax);
                atomicAdd(&tile[1][ti[q]], w[q] * lay);
            }
        } else {
            int ic[4];
            corners_rs(ring, spoke, na, ic);
#pragma unroll
            for (int q = 0; q < 4; q++)
                if (ic[q] < a.mesh) {
                    atomicAdd(&a.gx[ic[q]], w[q] * lax);
                    atomicAdd(&a.gy[ic[q]], w[q] * lay);
                }
        }
    }
    __syncthreads();
    for (int j = tid; j < tlen; j += GH_BLOCK) {
        const double vx = tile[0][j], vy = tile[1][j];
        if (vx == 0.0 && vy == 0.0) continue;
        const int row = j / tw, col = j - row * tw;
        const int rg = r0 + row;
        if (rg >= a.nr) continue;   // past the last ring: grdacc reads 0 there
        int sp = sref + d0 + col;
        sp %= na;
        if (sp < 0) sp += na;
        const int c = rg * na + sp;
        atomicAdd(&a.gx[c], vx);
        atomicAdd(&a.gy[c], vy);
    }
}

// lx += J_direct^T la + pmass sum_q rb(c_q) grad w_q; block sums of the
// parameter terms la . a_grid and la . (dF/dp)(r) x / r
template <typename R>
__global__ __launch_bounds__(GH_BLOCK) void adj_gather_kernel(AdjArgs<R> a) {
    __shared__ double red[GH_BLOCK / 64][2];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const long long i = (long long)blockIdx.x * GH_BLOCK + tid;
    double p0 = 0.0, p1 = 0.0;
    if (i < a.n && a.zone[i] == 1) {
        const long long n = a.n, p = i;
        const double lax = a.lam[2 * n + p], lay = a.lam[3 * n + p];
        const typename Vec4<R>::T st = a.xv[i];
        const typename Vec2<R>::T cu = a.ut[i];
        const double x = st.x, y = st.y;
        double w[4], wu[4], wt[4];
        int ring, spoke, ic[4];
        cic_grad((double)cu.x, (double)cu.y, a.nr, a.na, ring, spoke, w, wu, wt);
        corners_rs(ring, spoke, a.na, ic);
        double su = 0.0, sv = 0.0, agx = 0.0, agy = 0.0;
        for (int q = 0; q < 4; q++)
            if (ic[q] < a.mesh) {
                const typename Vec2<R>::T f = a.fld[ic[q]];
                const double fx = f.x, fy = f.y;
                const double s = fx * lax + fy * lay + a.pmass * a.rbar[ic[q]];
                su += s * wu[q];
                sv += s * wt[q];
                agx += w[q] * fx;
                agy += w[q] * fy;
            }
        const double r2 = x * x + y * y, r = sqrt(r2);
        double gx = 0.0, gy = 0.0;
        if (r > 0.0) {
            // u = log(1 + r) / alpha, t = atan2(y, x) na / (2 pi)
            const double cu_ = a.inv_alpha / (r * (1.0 + r));
            const double ct = a.tscale / r2;
            gx = su * cu_ * x - sv * ct * y;
            gy = su * cu_ * y + sv * ct * x;
            if (a.nfr > 0) {
                const double f = fixforce(r, a.nfr, a.fr_inv_dr, a.fr_c);
                const double fp = fixforce_deriv(r, a.nfr, a.fr_inv_dr, a.fr_c);
                const double xl = x * lax + y * lay;
                // d/dx of f(r) x / r, transposed, applied to la
                gx += fp * xl * x / r2 + f * (lax / r - xl * x / (r2 * r));
                gy += fp * xl * y / r2 + f * (lay / r - xl * y / (r2 * r));
            }
            if (a.nfd > 0)
                p1 = fixforce(r, a.nfd, a.fd_inv_dr, a.fd_c) *
                     (x * lax + y * lay) / r;
        }
        a.lam[p] += gx;
        a.lam[n + p] += gy;
        p0 = lax * agx + lay * agy;
    }
    for (int o = 32; o > 0; o >>= 1) {
        p0 += __shfl_xor(p0, o, 64);
        p1 += __shfl_xor(p1, o, 64);
    }
    if (lane == 0) {
        red[wave][0] = p0;
        red[wave][1] = p1;
    }
    __syncthreads();
    if (tid < 2) {
        double s = 0.0;
        for (int k = 0; k < GH_BLOCK / 64; k++) s += red[k][tid];
        a.part[2 * (long long)blockIdx.x + tid] = s;
    }
}

// block per field ring: polcat transposed, then the transpose of p2fsyn's
// synthesis for F_R and F_phi -> wbar (zero outside the field range)
constexpr int ADJ_MAX_NA = 2048;

__global__ __launch_bounds__(GH_BLOCK) void adj_wbar_kernel(AdjSolveArgs a) {
    __shared__ double gt[2][ADJ_MAX_NA];
    const int ir = blockIdx.x + 1, tid = threadIdx.x, na = a.na, nr = a.nr;
    const int lane = tid & 63, wave = tid >> 6;
    const int jrs = a.ranges[0], jrf = a.ranges[2], krf = a.ranges[3];
    const bool live = jrs >= 1 && ir >= jrf && ir <= krf;
    const size_t row = (size_t)(ir - 1) * na;
    if (live)
        for (int j = tid; j < na; j += GH_BLOCK) {
            const double c = a.twd[2 * j], s = a.twd[2 * j + 1];
            const double gx = a.gx[row + j], gy = a.gy[row + j];
            gt[0][j] = gx * c + gy * s;
            gt[1][j] = -gx * s + gy * c;
        }
    __syncthreads();
    const int maz = min(a.mmax - 2, a.ng - 1);
    for (int task = wave; task < 2 * a.ng; task += GH_BLOCK / 64) {
        const int type = task / a.ng, m = task - type * a.ng;
        const bool valid = live && !a.lg[m] && (type == 0 || (m >= 1 && m <= maz));
        double sc = 0.0, ss = 0.0;
        if (valid) {
            for (int j = lane; j < na; j += 64) {
                const int k = (int)(((long long)m * j) % na);
                const double g = gt[type][j];
                sc += g * a.twd[2 * k];
                ss += g * a.twd[2 * k + 1];
            }
            for (int o = 32; o > 0; o >>= 1) {
                sc += __shfl_xor(sc, o, 64);
                ss += __shfl_xor(ss, o, 64);
            }
        }
        if (lane == 0) {
            double wc = 0.0, ws = 0.0;
            if (valid) {
                if (m == 0) {
                    wc = sc / na;
                } else if (2 * m < na) {
                    wc = 2.0 * sc / na;
                    ws = -2.0 * ss / na;
                } else {
                    wc = sc / na;   // cos(pi j) = (-1)^j
                }
            }
            a.wbar[(((size_t)type * 2 + 0) * a.ng + m) * nr + ir - 1] = wc;
            a.wbar[(((size_t)type * 2 + 1) * a.ng + m) * nr + ir - 1] = ws;
        }
    }
}

// block per source ring: the transposed Green's contraction (p2conv) and
// the transposed forward transform (p2manl) -> rbar (zero outside the
// source range)
__global__ __launch_bounds__(GH_BLOCK) void adj_rbar_kernel(AdjSolveArgs a) {
    __shared__ double mc[64], ms[64];
    const int is = blockIdx.x + 1, tid = threadIdx.x, na = a.na, nr = a.nr, ng = a.ng;
    const int lane = tid & 63, wave = tid >> 6;
    const int jrs = a.ranges[0], krs = a.ranges[1], jrf = a.ranges[2], krf = a.ranges[3];
    const bool live = jrs >= 1 && is >= jrs && is <= krs;
    const size_t row = (size_t)(is - 1) * na;
    const int maz = min(a.mmax - 2, ng - 1);
    for (int m = wave; m < ng; m += GH_BLOCK / 64) {
        double c = 0.0, s = 0.0;
        if (live && !a.lg[m]) {
            const bool azim = m >= 1 && m <= maz;
            const double *g0 = a.G + ((size_t)0 * ng + m) * nr * nr;
            const double *g1 = a.G + ((size_t)1 * ng + m) * nr * nr;
            const double *wc0 = a.wbar + (((size_t)0 * 2 + 0) * ng + m) * nr;
            const double *ws0 = a.wbar + (((size_t)0 * 2 + 1) * ng + m) * nr;
            const double *wc1 = a.wbar + (((size_t)1 * 2 + 0) * ng + m) * nr;
            const double *ws1 = a.wbar + (((size_t)1 * 2 + 1) * ng + m) * nr;
            for (int ir = jrf + lane; ir <= krf; ir += 64) {
                const double G0 = g0[(size_t)(ir - 1) * nr + is - 1];
                c += G0 * wc0[ir - 1];
                s += G0 * ws0[ir - 1];
                if (azim) {
                    const double G1 = g1[(size_t)(ir - 1) * nr + is - 1];
                    c -= G1 * ws1[ir - 1];
                    s += G1 * wc1[ir - 1];
                }
            }
            for (int o = 32; o > 0; o >>= 1) {
                c += __shfl_xor(c, o, 64);
                s += __shfl_xor(s, o, 64);
            }
        }
        if (lane == 0) {
            mc[m] = c;
            // only the cos/sin pairs 1 <= m < mmax - 1 carry a sine part
            ms[m] = (m >= 1 && m < a.mmax - 1) ? s : 0.0;
        }
    }
    __syncthreads();
    for (int j = tid; j < na; j += GH_BLOCK) {
        double v = 0.0;
        if (live) {
            v = mc[0];
            for (int m = 1; m < ng; m++) {
                const int k = (int)(((long long)m * j) % na);
                if (2 * m < na) v += mc[m] * a.twd[2 * k] - ms[m] * a.twd[2 * k + 1];
                else v += (j & 1) ? -mc[m] : mc[m];
            }
        }
        a.rbar[row + j] = v;
    }
}

// out[q] = sum over blocks of part[b][q] (q = 0, 1), blocks in order, and
// out[2 + q] += out[q] (the sums since the adjoints were set)
__global__ __launch_bounds__(GH_BLOCK) void adj_sum_kernel(const double *part, int nb,
                                                           double *out) {
    __shared__ double red[GH_BLOCK / 64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int q = 0; q < 2; q++) {
        double t = 0.0;
        for (int b = tid; b < nb; b += GH_BLOCK) t += part[2 * (long long)b + q];
        for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
        if (lane == 0) red[wave] = t;
        __syncthreads();
        if (tid == 0) {
            double s = 0.0;
            for (int w = 0; w < GH_BLOCK / 64; w++) s += red[w];
            out[q] = s;
            out[2 + q] += s;
        }
        __syncthreads();
    }
}

// The seed of L = |A_2|^2 over r1 <= R <= r2 (R = r / lscale, equal masses):
// A_2 = (sum c - i sum s) / count with c = cos 2 phi = (x^2 - y^2) / r^2 and
// s = sin 2 phi = 2 x y / r^2, all in FP64 from the stored positions.
struct A2Terms {
    double c, s, inv, x, y;
    bool in;
};

template <typename R>
__device__ __forceinline__ A2Terms a2_terms(const typename Vec4<R>::T p, double lscale,
                                            double r1, double r2) {
    A2Terms t;
    t.x = (double)p.x;
    t.y = (double)p.y;
    const double rr = t.x * t.x + t.y * t.y;
    const double rad = sqrt(rr) / lscale;
    t.in = rad >= r1 && rad <= r2;
    t.inv = rr > 0.0 ? 1.0 / rr : 0.0;
    t.c = (t.x * t.x - t.y * t.y) * t.inv;
    t.s = 2.0 * t.x * t.y * t.inv;
    return t;
}

// part[3 b + q]: block b's sums of c, s and the count over the annulus
template <typename R>
__global__ __launch_bounds__(GH_BLOCK) void a2_part_kernel(const typename Vec4<R>::T *__restrict__ xv,
                                                           long long n, double lscale, double r1,
                                                           double r2, double *__restrict__ part) {
    __shared__ double red[3][GH_BLOCK / 64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const long long i = (long long)blockIdx.x * GH_BLOCK + tid;
    double v[3] = {0.0, 0.0, 0.0};
    if (i < n) {
        const A2Terms t = a2_terms<R>(xv[i], lscale, r1, r2);
        if (t.in) {
            v[0] = t.c;
            v[1] = t.s;
            v[2] = 1.0;
        }
    }
#pragma unroll
    for (int q = 0; q < 3; q++) {
        for (int o = 32; o > 0; o >>= 1) v[q] += __shfl_xor(v[q], o, 64);
        if (lane == 0) red[q][wave] = v[q];
    }
    __syncthreads();
    if (tid < 3) {
        double s = 0.0;
        for (int w = 0; w < GH_BLOCK / 64; w++) s += red[tid][w];
        part[3 * (long long)blockIdx.x + tid] = s;
    }
}

// tot[0..3] = L, Re A_2, Im A_2, count (at least 1) from the block sums, in order
__global__ __launch_bounds__(GH_BLOCK) void a2_total_kernel(const double *part, int nb,
                                                            double *tot) {
    __shared__ double red[3][GH_BLOCK / 64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    double v[3] = {0.0, 0.0, 0.0};
    for (int b = tid; b < nb; b += GH_BLOCK)
#pragma unroll
        for (int q = 0; q < 3; q++) v[q] += part[3 * (long long)b + q];
#pragma unroll
    for (int q = 0; q < 3; q++) {
        for (int o = 32; o > 0; o >>= 1) v[q] += __shfl_xor(v[q], o, 64);
        if (lane == 0) red[q][wave] = v[q];
    }
    __syncthreads();
    if (tid == 0) {
        double s[3];
        for (int q = 0; q < 3; q++) {
            s[q] = 0.0;
            for (int w = 0; w < GH_BLOCK / 64; w++) s[q] += red[q][w];
        }
        const double cnt = s[2] > 1.0 ? s[2] : 1.0;
        const double ar = s[0] / cnt, ai = -s[1] / cnt;
        tot[0] = ar * ar + ai * ai;
        tot[1] = ar;
        tot[2] = ai;
        tot[3] = cnt;
    }
}

// lam = (dL/dx, dL/dy, 0, 0), storage order, the annulus membership held fixed:
// dL/dphi = -4 (Re A_2 s + Im A_2 c) / count, dphi/dx = -y / r^2, dphi/dy = x / r^2
template <typename R>
__global__ __launch_bounds__(GH_BLOCK) void a2_seed_kernel(const typename Vec4<R>::T *__restrict__ xv,
                                                           long long n, double lscale, double r1,
                                                           double r2, const double *__restrict__ tot,
                                                           double *__restrict__ lam) {
    const long long i = (long long)blockIdx.x * GH_BLOCK + threadIdx.x;
    if (i >= n) return;
    const A2Terms t = a2_terms<R>(xv[i], lscale, r1, r2);
    const double ar = tot[1], ai = tot[2], cnt = tot[3];
    double dphi = t.in ? (-4.0 / cnt) * (ar * t.s + ai * t.c) : 0.0;
    dphi *= t.inv;
    lam[i] = -dphi * t.y;
    lam[n + i] = dphi * t.x;
    lam[2 * n + i] = 0.0;
    lam[3 * n + i] = 0.0;
}


// ---- tangent-linear (forward-mode) step ---------------------------------------------
//
// The derivative of the same map applied to a perturbation (dx, dv) of the
// state before the step, in FP64 along the forward run:
//   drho = pmass sum_i sum_q (grad w_q(x_i) . dx_i) [c_q]      tl_rho_kernel
//   dF   = K drho                                               tl_ana + tl_syn
//   da_i = sum_q w_q dF(c_q) + sum_q (grad w_q . dx_i) F(c_q) + J_halo dx_i
//   dv'  = dv + da,  dx' = dx + dv'                             tl_push_kernel
// with the cells, ring ranges and the off-grid choice held fixed, as the
// adjoint holds them.  The solve is built as the transpose of the adjoint's
// (adj_wbar_kernel, adj_rbar_kernel), which is itself the transpose of the
// forward solve, so  <J dz, lam> = <dz, J^T lam>  holds step by step to
// rounding.  Perturbations are kept in load order (indexed through the
// particles' ids), so re-sorts leave them in place.

template <typename R> struct TlArgs {
    const typename Vec4<R>::T *xv;
    const typename Vec2<R>::T *ut;
    const uint8_t *zone;
    const unsigned *id;
    long long n;
    double *d;                        // [4][n] load order: dx, dy, dvx, dvy
    double *rho;                      // [mesh] perturbed density
    const double *fx, *fy;            // [mesh] its field
    const typename Vec2<R>::T *fld;   // forward (F_x, F_y)
    const double4 *fr_c;
    double fr_inv_dr;
    int nfr;
    double pmass, inv_alpha, tscale;
    int nr, na, mesh;
};

// grad u . d and grad t . d at (x, y): u = log(1 + r) / alpha,
// t = atan2(y, x) na / (2 pi)
__device__ __forceinline__ void tl_dut(double x, double y, double dx, double dy,
                                       double inv_alpha, double tscale, double &du, double &dt) {
    const double r2 = x * x + y * y, r = sqrt(r2);
    du = 0.0;
    dt = 0.0;
    if (r > 0.0) {
        du = inv_alpha * (x * dx + y * dy) / (r * (1.0 + r));
        dt = tscale * (x * dy - y * dx) / r2;
    }
}

template <typename R>
__global__ __launch_bounds__(GH_BLOCK) void tl_rho_kernel(TlArgs<R> a) {
    const long long i = (long long)blockIdx.x * GH_BLOCK + threadIdx.x;
    if (i >= a.n || a.zone[i] != 1) return;
    const long long n = a.n, p = a.id[i];
    const typename Vec4<R>::T st = a.xv[i];
    const typename Vec2<R>::T cu = a.ut[i];
    double w[4], wu[4], wt[4], du, dt;
    int ring, spoke, ic[4];
    cic_grad((double)cu.x, (double)cu.y, a.nr, a.na, ring, spoke, w, wu, wt);
    corners_rs(ring, spoke, a.na, ic);
    tl_dut((double)st.x, (double)st.y, a.d[p], a.d[n + p], a.inv_alpha, a.tscale, du, dt);
#pragma unroll
    for (int q = 0; q < 4; q++)
        if (ic[q] < a.mesh) atomicAdd(&a.rho[ic[q]], a.pmass * (wu[q] * du + wt[q] * dt));
}

struct TlSolveArgs {
    const double *rho;                // [mesh]
    double *m;                        // [2 (cos, sin)][ng][nr]
    double *fx, *fy;                  // [mesh]
    const double *G;                  // FP64 tables [3][ng][nr][nr]
    const double *twd;
    const int *lg, *ranges;
    int nr, na, ng, mmax;
};

// block per source ring: the forward transform of the perturbed density
// over the active harmonics (the transpose of adj_rbar_kernel's synthesis):
// Mc[m] = sum_j rho_j C_m(j), Ms[m] = -sum_j rho_j sin(2 pi m j / na), C_m
// = 1 (m = 0), cos (2m < na), (-1)^j (2m = na); zero off the source range
__global__ __launch_bounds__(GH_BLOCK) void tl_ana_kernel(TlSolveArgs a) {
    __shared__ double row[ADJ_MAX_NA];
    const int is = blockIdx.x + 1, tid = threadIdx.x, na = a.na, nr = a.nr, ng = a.ng;
    const int lane = tid & 63, wave = tid >> 6;
    const int jrs = a.ranges[0], krs = a.ranges[1];
    const bool live = jrs >= 1 && is >= jrs && is <= krs;
    if (live)
        for (int j = tid; j < na; j += GH_BLOCK) row[j] = a.rho[(size_t)(is - 1) * na + j];
    __syncthreads();
    for (int m = wave; m < ng; m += GH_BLOCK / 64) {
        double c = 0.0, s = 0.0;
        if (live && !a.lg[m]) {
            for (int j = lane; j < na; j += 64) {
                const int k = (int)(((long long)m * j) % na);
                c += row[j] * a.twd[2 * k];
                s += row[j] * a.twd[2 * k + 1];
            }
            for (int o = 32; o > 0; o >>= 1) {
                c += __shfl_xor(c, o, 64);
                s += __shfl_xor(s, o, 64);
            }
        }
        if (lane == 0) {
            const bool sine = m >= 1 && m < a.mmax - 1 && 2 * m < na;
            a.m[(size_t)m * nr + is - 1] = c;
            a.m[((size_t)ng + m) * nr + is - 1] = sine ? -s : 0.0;
        }
    }
}

// block per field ring: the Green's contraction of the coefficients (F_R:
// G_R (Mc, Ms), F_phi: G_phi (Ms, -Mc)), the synthesis over the active
// harmonics and polcat's rotation to (F_x, F_y) -- the transposes of
// adj_rbar_kernel's contraction and adj_wbar_kernel; zero off the field range
__global__ __launch_bounds__(GH_BLOCK) void tl_syn_kernel(TlSolveArgs a) {
    __shared__ double wc[2][64], ws[2][64];
    const int ir = blockIdx.x + 1, tid = threadIdx.x, na = a.na, nr = a.nr, ng = a.ng;
    const int lane = tid & 63, wave = tid >> 6;
    const int jrs = a.ranges[0], krs = a.ranges[1], jrf = a.ranges[2], krf = a.ranges[3];
    const bool live = jrs >= 1 && ir >= jrf && ir <= krf;
    const int maz = min(a.mmax - 2, ng - 1);
    for (int task = wave; task < 2 * ng; task += GH_BLOCK / 64) {
        const int type = task / ng, m = task - type * ng;
        const bool valid = live && !a.lg[m] && (type == 0 || (m >= 1 && m <= maz));
        double c = 0.0, s = 0.0;
        if (valid) {
            const double *g = a.G + (((size_t)type * ng + m) * nr + (ir - 1)) * nr;
            const double *mc = a.m + (size_t)m * nr, *ms = a.m + ((size_t)ng + m) * nr;
            for (int is = jrs + lane; is <= krs; is += 64) {
                c += g[is - 1] * mc[is - 1];
                s += g[is - 1] * ms[is - 1];
            }
            for (int o = 32; o > 0; o >>= 1) {
                c += __shfl_xor(c, o, 64);
                s += __shfl_xor(s, o, 64);
            }
        }
        if (lane == 0) {
            // F_R: (wc, ws) = G (Mc, Ms); F_phi: (wc, ws) = G (Ms, -Mc)
            double vc = type == 0 ? c : s, vs = type == 0 ? s : -c;
            // synthesis weights: 1/na (m = 0 or 2m = na), 2/na (cos) and
            // -2/na (sin) otherwise
            const double al = (m == 0 || 2 * m >= na) ? 1.0 / na : 2.0 / na;
            const double be = (m >= 1 && 2 * m < na) ? -2.0 / na : 0.0;
            wc[type][m] = valid ? al * vc : 0.0;
            ws[type][m] = valid ? be * vs : 0.0;
        }
    }
    __syncthreads();
    for (int j = tid; j < na; j += GH_BLOCK) {
        double g0 = 0.0, g1 = 0.0;
        if (live)
            for (int m = 0; m < ng; m++) {
                const int k = (int)(((long long)m * j) % na);
                const double cm = a.twd[2 * k], sm = a.twd[2 * k + 1];
                g0 += wc[0][m] * cm + ws[0][m] * sm;
                g1 += wc[1][m] * cm + ws[1][m] * sm;
            }
        const double cj = a.twd[2 * j], sj = a.twd[2 * j + 1];
        a.fx[(size_t)(ir - 1) * na + j] = g0 * cj - g1 * sj;
        a.fy[(size_t)(ir - 1) * na + j] = g0 * sj + g1 * cj;
    }
}

// da from the perturbed field, the weights' derivative against the forward
// field and the halo's Jacobian; dv += da, dx += dv (off-grid: da = 0)
template <typename R>
__global__ __launch_bounds__(GH_BLOCK) void tl_push_kernel(TlArgs<R> a) {
    const long long i = (long long)blockIdx.x * GH_BLOCK + threadIdx.x;
    if (i >= a.n) return;
    const long long n = a.n, p = a.id[i];
    double dx = a.d[p], dy = a.d[n + p], dvx = a.d[2 * n + p], dvy = a.d[3 * n + p];
    if (a.zone[i] == 1) {
        const typename Vec4<R>::T st = a.xv[i];
        const typename Vec2<R>::T cu = a.ut[i];
        const double x = st.x, y = st.y;
        double w[4], wu[4], wt[4], du, dt;
        int ring, spoke, ic[4];
        cic_grad((double)cu.x, (double)cu.y, a.nr, a.na, ring, spoke, w, wu, wt);
        corners_rs(ring, spoke, a.na, ic);
        tl_dut(x, y, dx, dy, a.inv_alpha, a.tscale, du, dt);
        double ax = 0.0, ay = 0.0;
#pragma unroll
        for (int q = 0; q < 4; q++)
            if (ic[q] < a.mesh) {
                const typename Vec2<R>::T f = a.fld[ic[q]];
                const double dw = wu[q] * du + wt[q] * dt;
                ax += w[q] * a.fx[ic[q]] + dw * (double)f.x;
                ay += w[q] * a.fy[ic[q]] + dw * (double)f.y;
            }
        const double r2 = x * x + y * y, r = sqrt(r2);
        if (a.nfr > 0 && r > 0.0) {
            const double f = fixforce(r, a.nfr, a.fr_inv_dr, a.fr_c);
            const double fp = fixforce_deriv(r, a.nfr, a.fr_inv_dr, a.fr_c);
            const double xd = x * dx + y * dy;
            // d/dx of f(r) x / r applied to (dx, dy)
            ax += fp * xd * x / r2 + f * (dx / r - xd * x / (r2 * r));
            ay += fp * xd * y / r2 + f * (dy / r - xd * y / (r2 * r));
        }
        dvx += ax;
        dvy += ay;
    }
    a.d[p] = dx + dvx;
    a.d[n + p] = dy + dvy;
    a.d[2 * n + p] = dvx;
    a.d[3 * n + p] = dvy;
}

}  // namespace gh

using namespace gh;

// ---- host side ----------------------------------------------------------------------

void adj_free(galhip_ctx *c) {
    dfree(c->adj_lam);
    dfree(c->adj_g);
    dfree(c->adj_rbar);
    dfree(c->adj_wbar);
    dfree(c->adj_part);
    dfree(c->fd_c);
    dfree(c->adj_tmp);
    dfree(c->adj_id);
    dfree(c->tl_d);
    dfree(c->tl_rho);
    dfree(c->tl_f);
    dfree(c->tl_m);
    c->tl_n = 0;
    for (auto &s : c->adj_slots) {
        dfree(s.xv);
        dfree(s.ut);
        dfree(s.zone);
        dfree(s.id);
        dfree(s.counters);
        dfree(s.mass);
        dfree(s.rng);
    }
    c->adj_slots.clear();
    for (auto e : c->adj_ev) (void)hipEventDestroy(e);
    c->adj_ev.clear();
}

static hipEvent_t adj_event(galhip_ctx *c) {
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    c->adj_ev.push_back(e);
    return e;
}

// after the per-block partial sums (2 per block for a step, 3 for the A_2 seed):
// [0..1] the last step's parameter sums, [2..3] their totals since the adjoints
// were set, [4..7] the A_2 seed's L, Re A_2, Im A_2, count
static double *adj_tail(galhip_ctx *c) {
    const long long nb = (c->n + GH_BLOCK - 1) / GH_BLOCK;
    return c->adj_part + 3 * (nb > 0 ? nb : 1);
}

static int adj_check(galhip_ctx *c) {
    if (!c) return gh_fail(GALHIP_EINVAL, "null context");
    if (c->p3d || c->ens) return gh_fail(GALHIP_EINVAL, "the adjoint is built for one p2d grid");
    if (c->nz != 1 || c->nsect != 1)
        return gh_fail(GALHIP_EINVAL, "the adjoint needs one zone and one sector");
    if (c->comp) return gh_fail(GALHIP_EINVAL, "the adjoint differentiates the plain position update");
    return 0;
}

// per-particle adjoints and grid work arrays, sized for the loaded particles
static int adj_ensure(galhip_ctx *c) {
    if (c->adj_lam && c->adj_n == c->n) return 0;
    dfree(c->adj_lam);
    dfree(c->adj_part);
    dfree(c->adj_tmp);
    dfree(c->adj_id);
    int rc = dalloc(&c->adj_lam, sizeof(double) * 4 * (size_t)(c->n > 0 ? c->n : 1));
    rc |= dalloc(&c->adj_tmp, sizeof(double) * 4 * (size_t)(c->n > 0 ? c->n : 1));
    rc |= dalloc(&c->adj_id, sizeof(unsigned) * (size_t)(c->n > 0 ? c->n : 1));
    const long long nb = (c->n + GH_BLOCK - 1) / GH_BLOCK;
    rc |= dalloc(&c->adj_part, sizeof(double) * (3 * (size_t)(nb > 0 ? nb : 1) + 8));
    if (!c->adj_g) {
        rc |= dalloc(&c->adj_g, sizeof(double) * 2 * (size_t)c->mesh);
        rc |= dalloc(&c->adj_rbar, sizeof(double) * (size_t)c->mesh);
        rc |= dalloc(&c->adj_wbar, sizeof(double) * 4 * (size_t)c->ng * c->nr);
    }
    if (rc) return rc;
    c->adj_n = c->n;
    HIPCHK(hipMemset(c->adj_lam, 0, sizeof(double) * 4 * (size_t)(c->n > 0 ? c->n : 1)));
    HIPCHK(hipMemset(adj_tail(c), 0, sizeof(double) * 4));
    return 0;
}

extern "C" int galhip_adj_record(galhip_ctx *c, int32_t slot) {
    int rc = adj_check(c);
    if (rc) return rc;
    if (slot < 0 || slot > 4096) return gh_fail(GALHIP_EINVAL, "slot %d out of range", slot);
    if ((size_t)slot >= c->adj_slots.size()) c->adj_slots.resize(slot + 1);
    galhip_ctx::AdjSlot &s = c->adj_slots[slot];
    const size_t es = esize(c), n = (size_t)c->n;
    const size_t cells = (size_t)c->mesh;
    if (s.n != c->n || !s.xv) {
        dfree(s.xv);
        dfree(s.ut);
        dfree(s.zone);
        dfree(s.id);
        dfree(s.counters);
        dfree(s.mass);
        dfree(s.rng);
        rc = dalloc(&s.xv, 4 * es * (n ? n : 1)) | dalloc(&s.ut, 2 * es * (n ? n : 1)) |
             dalloc(&s.zone, n ? n : 1) | dalloc(&s.id, sizeof(unsigned) * (n ? n : 1)) |
             dalloc(&s.counters, sizeof(long long) * 4) |
             dalloc(&s.mass, sizeof(long long) * cells * c->nz) |
             dalloc(&s.rng, sizeof(int) * 2 * GH_MAXZ);
        if (rc) return rc;
        s.n = c->n;
    }
    for (int z = 0; z < c->nz; z++)
        HIPCHK(hipMemcpyAsync(s.mass + z * cells, c->mfx[z], sizeof(long long) * cells,
                              hipMemcpyDeviceToDevice, c->s));
    HIPCHK(hipMemcpyAsync(s.rng, c->jrad, sizeof(int) * GH_MAXZ, hipMemcpyDeviceToDevice, c->s));
    HIPCHK(hipMemcpyAsync(s.rng + GH_MAXZ, c->krad, sizeof(int) * GH_MAXZ,
                          hipMemcpyDeviceToDevice, c->s));
    if (n) {
        HIPCHK(hipMemcpyAsync(s.xv, c->xv, 4 * es * n, hipMemcpyDeviceToDevice, c->s));
        HIPCHK(hipMemcpyAsync(s.ut, c->ut, 2 * es * n, hipMemcpyDeviceToDevice, c->s));
        HIPCHK(hipMemcpyAsync(s.zone, c->zone, n, hipMemcpyDeviceToDevice, c->s));
        HIPCHK(hipMemcpyAsync(s.id, c->id, sizeof(unsigned) * n, hipMemcpyDeviceToDevice, c->s));
    }
    HIPCHK(hipMemcpyAsync(s.counters, c->counters, sizeof(long long) * 4,
                          hipMemcpyDeviceToDevice, c->s));
    s.istep = c->istep;
    s.order_ver = c->order_ver;
    return 0;
}

extern "C" int galhip_adj_restore(galhip_ctx *c, int32_t slot) {
    int rc = adj_check(c);
    if (rc) return rc;
    if (slot < 0 || (size_t)slot >= c->adj_slots.size() || !c->adj_slots[slot].xv)
        return gh_fail(GALHIP_EINVAL, "slot %d was not recorded", slot);
    galhip_ctx::AdjSlot &s = c->adj_slots[slot];
    if (s.n != c->n) return gh_fail(GALHIP_EINVAL, "slot %d holds %lld particles, not %lld",
                                    slot, s.n, c->n);
    const size_t es = esize(c), n = (size_t)c->n;
    if (n) {
        HIPCHK(hipMemcpyAsync(c->xv, s.xv, 4 * es * n, hipMemcpyDeviceToDevice, c->s));
        HIPCHK(hipMemcpyAsync(c->ut, s.ut, 2 * es * n, hipMemcpyDeviceToDevice, c->s));
        HIPCHK(hipMemcpyAsync(c->zone, s.zone, n, hipMemcpyDeviceToDevice, c->s));
        HIPCHK(hipMemcpyAsync(c->id, s.id, sizeof(unsigned) * n, hipMemcpyDeviceToDevice, c->s));
    }
    HIPCHK(hipMemcpyAsync(c->counters, s.counters, sizeof(long long) * 4,
                          hipMemcpyDeviceToDevice, c->s));
    const size_t cells = (size_t)c->mesh;
    for (int z = 0; z < c->nz; z++)
        HIPCHK(hipMemcpyAsync(c->mfx[z], s.mass + z * cells, sizeof(long long) * cells,
                              hipMemcpyDeviceToDevice, c->s));
    HIPCHK(hipMemcpyAsync(c->jrad, s.rng, sizeof(int) * GH_MAXZ, hipMemcpyDeviceToDevice, c->s));
    HIPCHK(hipMemcpyAsync(c->krad, s.rng + GH_MAXZ, sizeof(int) * GH_MAXZ,
                          hipMemcpyDeviceToDevice, c->s));
    c->istep = s.istep;
    c->order_ver = s.order_ver;
    c->have_field = 0;
    c->masscl_done = 0;
    c->steps_since_sort = 0;
    return 0;
}

extern "C" int galhip_adj_set(galhip_ctx *c, const double *lam) {
    int rc = adj_check(c);
    if (rc) return rc;
    if (!lam) return gh_fail(GALHIP_EINVAL, "null argument");
    rc = adj_ensure(c);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(c->s));
    const long long n = c->n;
    if (n) {
        HIPCHK(hipMemcpy(c->adj_tmp, lam, sizeof(double) * 4 * (size_t)n, hipMemcpyHostToDevice));
        const unsigned nb = (unsigned)((n + GH_BLOCK - 1) / GH_BLOCK);
        hipLaunchKernelGGL(adj_to_store_kernel, dim3(nb), dim3(GH_BLOCK), 0, c->s,
                           (const double *)c->adj_tmp, (const unsigned *)c->id, c->adj_lam, n);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(c->adj_id, c->id, sizeof(unsigned) * n, hipMemcpyDeviceToDevice, c->s));
    }
    HIPCHK(hipMemsetAsync(adj_tail(c), 0, sizeof(double) * 4, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    c->adj_ver = c->order_ver;
    return 0;
}

template <typename R> static int adj_seed_a2_impl(galhip_ctx *c, double lscale, double r1,
                                                  double r2, double *out) {
    const long long n = c->n;
    const long long nb = (n + GH_BLOCK - 1) / GH_BLOCK;
    double *tail = adj_tail(c);
    const typename Vec4<R>::T *xv = (const typename Vec4<R>::T *)c->xv;
    if (nb > 0)
        hipLaunchKernelGGL(a2_part_kernel<R>, dim3((unsigned)nb), dim3(GH_BLOCK), 0, c->s, xv, n,
                           lscale, r1, r2, c->adj_part);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(a2_total_kernel, dim3(1), dim3(GH_BLOCK), 0, c->s,
                       (const double *)c->adj_part, (int)nb, tail + 4);
    HIPCHK(hipGetLastError());
    if (nb > 0) {
        hipLaunchKernelGGL(a2_seed_kernel<R>, dim3((unsigned)nb), dim3(GH_BLOCK), 0, c->s, xv, n,
                           lscale, r1, r2, (const double *)(tail + 4), c->adj_lam);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(c->adj_id, c->id, sizeof(unsigned) * n, hipMemcpyDeviceToDevice, c->s));
    }
    HIPCHK(hipMemsetAsync(tail, 0, sizeof(double) * 4, c->s));
    double h[4];
    HIPCHK(hipMemcpyAsync(h, tail + 4, sizeof h, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    c->adj_ver = c->order_ver;
    if (out) memcpy(out, h, sizeof h);
    return 0;
}

extern "C" int galhip_adj_seed_a2(galhip_ctx *c, double lscale, double r1, double r2,
                                  double *out) {
    int rc = adj_check(c);
    if (rc) return rc;
    if (!(lscale > 0.0) || !(r1 >= 0.0) || !(r2 >= r1))
        return gh_fail(GALHIP_EINVAL, "need lscale > 0 and 0 <= r1 <= r2");
    rc = adj_ensure(c);
    if (rc) return rc;
    return c->prec == GALHIP_FP32 ? adj_seed_a2_impl<float>(c, lscale, r1, r2, out)
                                  : adj_seed_a2_impl<double>(c, lscale, r1, r2, out);
}

extern "C" int galhip_adj_get_par(galhip_ctx *c, double *gpar) {
    int rc = adj_check(c);
    if (rc) return rc;
    if (!gpar) return gh_fail(GALHIP_EINVAL, "null argument");
    if (!c->adj_lam || c->adj_n != c->n) return gh_fail(GALHIP_EINVAL, "no adjoint set");
    HIPCHK(hipMemcpyAsync(gpar, adj_tail(c) + 2, sizeof(double) * 2, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    return 0;
}

// bring the adjoints to the current storage order
static int adj_relayout(galhip_ctx *c) {
    if (c->adj_ver == c->order_ver || c->n == 0) return 0;
    const long long n = c->n;
    const unsigned nb = (unsigned)((n + GH_BLOCK - 1) / GH_BLOCK);
    hipLaunchKernelGGL(adj_to_load_kernel, dim3(nb), dim3(GH_BLOCK), 0, c->s,
                       (const double *)c->adj_lam, (const unsigned *)c->adj_id, c->adj_tmp, n);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(adj_to_store_kernel, dim3(nb), dim3(GH_BLOCK), 0, c->s,
                       (const double *)c->adj_tmp, (const unsigned *)c->id, c->adj_lam, n);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(c->adj_id, c->id, sizeof(unsigned) * n, hipMemcpyDeviceToDevice, c->s));
    c->adj_ver = c->order_ver;
    return 0;
}

extern "C" int galhip_adj_get(galhip_ctx *c, double *lam) {
    int rc = adj_check(c);
    if (rc) return rc;
    if (!lam) return gh_fail(GALHIP_EINVAL, "null argument");
    if (!c->adj_lam || c->adj_n != c->n) return gh_fail(GALHIP_EINVAL, "no adjoint set");
    const long long n = c->n;
    if (n) {
        const unsigned nb = (unsigned)((n + GH_BLOCK - 1) / GH_BLOCK);
        hipLaunchKernelGGL(adj_to_load_kernel, dim3(nb), dim3(GH_BLOCK), 0, c->s,
                           (const double *)c->adj_lam, (const unsigned *)c->adj_id, c->adj_tmp, n);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(c->s));
        HIPCHK(hipMemcpy(lam, c->adj_tmp, sizeof(double) * 4 * (size_t)n, hipMemcpyDeviceToHost));
    }
    return 0;
}

extern "C" int galhip_set_fixforce_dpar(galhip_ctx *c, int32_t n, double dr, const double *val,
                                        const double *der) {
    if (!c) return gh_fail(GALHIP_EINVAL, "null context");
    if (n != 0 && (n < 2 || !(dr > 0.0) || !val || !der))
        return gh_fail(GALHIP_EINVAL, "galhip_set_fixforce_dpar: bad table");
    HIPCHK(hipStreamSynchronize(c->s));
    dfree(c->fd_c);
    c->nfd = 0;
    if (n == 0) return 0;
    const std::vector<double4> cf = radtab_coeffs(n, dr, val, der);
    int rc = dalloc(&c->fd_c, sizeof(double4) * n);
    if (rc) return rc;
    HIPCHK(hipMemcpy(c->fd_c, cf.data(), sizeof(double4) * n, hipMemcpyHostToDevice));
    c->nfd = n;
    c->fd_dr = dr;
    return 0;
}

static AdjSolveArgs adj_solve_args(galhip_ctx *c) {
    AdjSolveArgs a;
    a.gx = c->adj_g;
    a.gy = c->adj_g + c->mesh;
    a.wbar = c->adj_wbar;
    a.rbar = c->adj_rbar;
    a.G = c->Gd;
    a.twd = c->twd;
    a.lg = c->lg_d;
    a.ranges = c->ranges;
    a.nr = c->nr;
    a.na = c->na;
    a.ng = c->ng;
    a.mmax = c->mmax;
    return a;
}

// rbar = K^T (gx, gy), with the ranges of the last solve
static int adj_solve(galhip_ctx *c) {
    if (c->na > ADJ_MAX_NA || c->ng > 64) return gh_fail(GALHIP_EINVAL, "grid too large for the adjoint");
    AdjSolveArgs a = adj_solve_args(c);
    hipLaunchKernelGGL(adj_wbar_kernel, dim3(c->nr), dim3(GH_BLOCK), 0, c->s, a);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(adj_rbar_kernel, dim3(c->nr), dim3(GH_BLOCK), 0, c->s, a);
    HIPCHK(hipGetLastError());
    return 0;
}

template <typename R> static AdjArgs<R> adj_args(galhip_ctx *c) {
    AdjArgs<R> a;
    memset(&a, 0, sizeof a);
    a.xv = (const typename Vec4<R>::T *)c->xv;
    a.ut = (const typename Vec2<R>::T *)c->ut;
    a.zone = c->zone;
    a.id = c->id;
    a.n = c->n;
    a.lam = c->adj_lam;
    a.gx = c->adj_g;
    a.gy = c->adj_g + c->mesh;
    a.rbar = c->adj_rbar;
    a.fld = (const typename Vec2<R>::T *)c->fld2;
    a.fr_c = c->fr_c;
    a.fr_inv_dr = c->nfr > 0 ? 1.0 / c->fr_dr : 0.0;
    a.nfr = c->nfr;
    a.fd_c = c->fd_c;
    a.fd_inv_dr = c->nfd > 0 ? 1.0 / c->fd_dr : 0.0;
    a.nfd = c->nfd;
    a.pmass = c->pmass;
    a.inv_alpha = 1.0 / c->alpha;
    a.tscale = (double)c->na / (2.0 * 3.141592653589793238462643);
    a.nr = c->nr;
    a.na = c->na;
    a.mesh = c->mesh;
    a.part = c->adj_part;
    return a;
}

template <typename R> static int adj_step_impl(galhip_ctx *c, double *gpar) {
    const long long nb = (c->n + GH_BLOCK - 1) / GH_BLOCK;
    int rc0 = adj_relayout(c);
    if (rc0) return rc0;
    HIPCHK(hipMemsetAsync(c->adj_g, 0, sizeof(double) * 2 * (size_t)c->mesh, c->s));
    AdjArgs<R> a = adj_args<R>(c);
    const bool prof = c->prof && nb > 0;
    if (prof) HIPCHK(hipEventRecord(adj_event(c), c->s));
    if (nb > 0) {
        hipLaunchKernelGGL(adj_scatter_kernel<R>, dim3((unsigned)nb), dim3(GH_BLOCK), 0, c->s, a);
        HIPCHK(hipGetLastError());
    }
    if (prof) HIPCHK(hipEventRecord(adj_event(c), c->s));
    int rc = adj_solve(c);
    if (rc) return rc;
    if (prof) HIPCHK(hipEventRecord(adj_event(c), c->s));
    if (nb > 0) {
        hipLaunchKernelGGL(adj_gather_kernel<R>, dim3((unsigned)nb), dim3(GH_BLOCK), 0, c->s, a);
        HIPCHK(hipGetLastError());
    }
    if (prof) HIPCHK(hipEventRecord(adj_event(c), c->s));
    double *dout = adj_tail(c);
    hipLaunchKernelGGL(adj_sum_kernel, dim3(1), dim3(GH_BLOCK), 0, c->s,
                       (const double *)c->adj_part, (int)nb, dout);
    HIPCHK(hipGetLastError());
    if (gpar) {
        double h[2];
        HIPCHK(hipMemcpyAsync(h, dout, sizeof h, hipMemcpyDeviceToHost, c->s));
        HIPCHK(hipStreamSynchronize(c->s));
        gpar[0] += h[0];
        gpar[1] += h[1];
    }
    return 0;
}

extern "C" int galhip_adj_step(galhip_ctx *c, double *gpar) {
    int rc = adj_check(c);
    if (rc) return rc;
    if (!c->have_field) return gh_fail(GALHIP_ERANGE, "adjoint step needs the step's field (findf)");
    if (!c->have_green) return gh_fail(GALHIP_EINVAL, "Green function not set");
    if (!c->adj_lam || c->adj_n != c->n) return gh_fail(GALHIP_EINVAL, "no adjoint set");
    return c->prec == GALHIP_FP32 ? adj_step_impl<float>(c, gpar) : adj_step_impl<double>(c, gpar);
}

extern "C" int galhip_adj_field(galhip_ctx *c, const double *g, double *rbar) {
    int rc = adj_check(c);
    if (rc) return rc;
    if (!g || !rbar) return gh_fail(GALHIP_EINVAL, "null argument");
    if (!c->have_green) return gh_fail(GALHIP_EINVAL, "Green function not set");
    rc = adj_ensure(c);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(c->s));
    HIPCHK(hipMemcpy(c->adj_g, g, sizeof(double) * 2 * (size_t)c->mesh, hipMemcpyHostToDevice));
    rc = adj_solve(c);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(c->s));
    HIPCHK(hipMemcpy(rbar, c->adj_rbar, sizeof(double) * (size_t)c->mesh, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int galhip_adj_get_profile(galhip_ctx *c, double *ms, int64_t *steps) {
    int rc = adj_check(c);
    if (rc) return rc;
    if (!ms) return gh_fail(GALHIP_EINVAL, "null argument");
    HIPCHK(hipStreamSynchronize(c->s));
    double t[2] = {0.0, 0.0};
    const size_t k = c->adj_ev.size() / 4;
    for (size_t i = 0; i < k; i++)
        for (int q = 0; q < 2; q++) {
            float v = 0.f;
            HIPCHK(hipEventElapsedTime(&v, c->adj_ev[4 * i + 2 * q], c->adj_ev[4 * i + 2 * q + 1]));
            t[q] += v;
        }
    ms[0] = t[0];
    ms[1] = t[1];
    if (steps) *steps = (int64_t)k;
    for (auto e : c->adj_ev) (void)hipEventDestroy(e);
    c->adj_ev.clear();
    return 0;
}

// ---- tangent-linear host side -------------------------------------------------------

static int tl_ensure(galhip_ctx *c) {
    if (c->na > ADJ_MAX_NA || c->ng > 64)
        return gh_fail(GALHIP_EINVAL, "grid too large for the tangent-linear solve");
    if (!c->tl_rho) {
        int rc = dalloc(&c->tl_rho, sizeof(double) * (size_t)c->mesh) |
                 dalloc(&c->tl_f, sizeof(double) * 2 * (size_t)c->mesh) |
                 dalloc(&c->tl_m, sizeof(double) * 2 * (size_t)c->ng * c->nr);
        if (rc) return rc;
    }
    if (c->tl_d && c->tl_n == c->n) return 0;
    dfree(c->tl_d);
    c->tl_n = 0;
    int rc = dalloc(&c->tl_d, sizeof(double) * 4 * (size_t)(c->n > 0 ? c->n : 1));
    if (rc) return rc;
    c->tl_n = c->n;
    HIPCHK(hipMemset(c->tl_d, 0, sizeof(double) * 4 * (size_t)(c->n > 0 ? c->n : 1)));
    return 0;
}

static TlSolveArgs tl_solve_args(galhip_ctx *c) {
    TlSolveArgs a;
    a.rho = c->tl_rho;
    a.m = c->tl_m;
    a.fx = c->tl_f;
    a.fy = c->tl_f + c->mesh;
    a.G = c->Gd;
    a.twd = c->twd;
    a.lg = c->lg_d;
    a.ranges = c->ranges;
    a.nr = c->nr;
    a.na = c->na;
    a.ng = c->ng;
    a.mmax = c->mmax;
    return a;
}

// (F_x, F_y) = K rho in FP64, with the ranges of the last solve
static int tl_solve(galhip_ctx *c) {
    const TlSolveArgs a = tl_solve_args(c);
    hipLaunchKernelGGL(tl_ana_kernel, dim3(c->nr), dim3(GH_BLOCK), 0, c->s, a);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(tl_syn_kernel, dim3(c->nr), dim3(GH_BLOCK), 0, c->s, a);
    HIPCHK(hipGetLastError());
    return 0;
}

template <typename R> static TlArgs<R> tl_args(galhip_ctx *c) {
    TlArgs<R> a;
    memset(&a, 0, sizeof a);
    a.xv = (const typename Vec4<R>::T *)c->xv;
    a.ut = (const typename Vec2<R>::T *)c->ut;
    a.zone = c->zone;
    a.id = c->id;
    a.n = c->n;
    a.d = c->tl_d;
    a.rho = c->tl_rho;
    a.fx = c->tl_f;
    a.fy = c->tl_f + c->mesh;
    a.fld = (const typename Vec2<R>::T *)c->fld2;
    a.fr_c = c->fr_c;
    a.fr_inv_dr = c->nfr > 0 ? 1.0 / c->fr_dr : 0.0;
    a.nfr = c->nfr;
    a.pmass = c->pmass;
    a.inv_alpha = 1.0 / c->alpha;
    a.tscale = (double)c->na / (2.0 * 3.141592653589793238462643);
    a.nr = c->nr;
    a.na = c->na;
    a.mesh = c->mesh;
    return a;
}

extern "C" int galhip_tl_set(galhip_ctx *c, const double *d) {
    int rc = adj_check(c);
    if (rc) return rc;
    if (!d) return gh_fail(GALHIP_EINVAL, "null argument");
    rc = tl_ensure(c);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(c->s));
    if (c->n) HIPCHK(hipMemcpy(c->tl_d, d, sizeof(double) * 4 * (size_t)c->n, hipMemcpyHostToDevice));
    return 0;
}

extern "C" int galhip_tl_get(galhip_ctx *c, double *d) {
    int rc = adj_check(c);
    if (rc) return rc;
    if (!d) return gh_fail(GALHIP_EINVAL, "null argument");
    if (!c->tl_d || c->tl_n != c->n) return gh_fail(GALHIP_EINVAL, "no tangent-linear state set");
    HIPCHK(hipStreamSynchronize(c->s));
    if (c->n) HIPCHK(hipMemcpy(d, c->tl_d, sizeof(double) * 4 * (size_t)c->n, hipMemcpyDeviceToHost));
    return 0;
}

template <typename R> static int tl_step_impl(galhip_ctx *c) {
    const long long nb = (c->n + GH_BLOCK - 1) / GH_BLOCK;
    HIPCHK(hipMemsetAsync(c->tl_rho, 0, sizeof(double) * (size_t)c->mesh, c->s));
    const TlArgs<R> a = tl_args<R>(c);
    if (nb > 0) {
        hipLaunchKernelGGL(tl_rho_kernel<R>, dim3((unsigned)nb), dim3(GH_BLOCK), 0, c->s, a);
        HIPCHK(hipGetLastError());
    }
    int rc = tl_solve(c);
    if (rc) return rc;
    if (nb > 0) {
        hipLaunchKernelGGL(tl_push_kernel<R>, dim3((unsigned)nb), dim3(GH_BLOCK), 0, c->s, a);
        HIPCHK(hipGetLastError());
    }
    return 0;
}

extern "C" int galhip_tl_step(galhip_ctx *c) {
    int rc = adj_check(c);
    if (rc) return rc;
    if (!c->have_field) return gh_fail(GALHIP_ERANGE, "tangent-linear step needs the step's field (findf)");
    if (!c->have_green) return gh_fail(GALHIP_EINVAL, "Green function not set");
    if (!c->tl_d || c->tl_n != c->n) return gh_fail(GALHIP_EINVAL, "no tangent-linear state set");
    return c->prec == GALHIP_FP32 ? tl_step_impl<float>(c) : tl_step_impl<double>(c);
}

extern "C" int galhip_tl_field(galhip_ctx *c, const double *rho, double *f) {
    int rc = adj_check(c);
    if (rc) return rc;
    if (!rho || !f) return gh_fail(GALHIP_EINVAL, "null argument");
    if (!c->have_green) return gh_fail(GALHIP_EINVAL, "Green function not set");
    rc = tl_ensure(c);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(c->s));
    HIPCHK(hipMemcpy(c->tl_rho, rho, sizeof(double) * (size_t)c->mesh, hipMemcpyHostToDevice));
    rc = tl_solve(c);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(c->s));
    HIPCHK(hipMemcpy(f, c->tl_f, sizeof(double) * 2 * (size_t)c->mesh, hipMemcpyDeviceToHost));
    return 0;
}
