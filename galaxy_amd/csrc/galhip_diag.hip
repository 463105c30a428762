// galhip_diag.hip -- analysis-step diagnostics of the particles: the
// log-spiral coefficients of lgspia (SRC/lib15/lgspia.f:44-110), normalised
// by the population as phyprp does (SRC/lib15/phyprp.f:182-197).
//
// For harmonic index im = 1..nm and pitch index j = 1..np,
//   A[j][im] = sum over particles of exp(i im (nsect theta + k_j a)) / count,
//   a = nsect ln(R / lscale) / 4,
// over the particles on the grid (and off it when offanal), equal masses.
// Each particle forms e^{i nsect theta} from (x + i y) / R and e^{i a}, then
// the powers and pitch factors by complex products in FP64, as the
// reference's recurrences do: index j = jc + k (jc = np/2 + 1, k >= 0)
// takes e^{i k im a}, its mirror np + 1 - j the conjugate (for even np
// index np/2 is left at 0, as in the reference).
//
// Sums are deterministic: every wavefront reduces each coefficient with
// shuffles and its lane 0 adds the result to the wave's own LDS row in
// particle order; the rows are added in wave order into one partial per
// block, and a second kernel adds the block partials in block order.
#include "galhip_ctx.h"

#include <algorithm>

namespace gh {

constexpr int LG_BLOCK = 256;
constexpr int LG_WAVES = LG_BLOCK / 64;
constexpr int LG_MAXQ = 512;    // np * nm coefficients at most (LDS: 4 x 2 x 512 doubles)

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

template <typename R>
__global__ __launch_bounds__(LG_BLOCK) void lgsp_part_kernel(
    const typename Vec4<R>::T *__restrict__ xv, int ps, const uint8_t *__restrict__ zone,
    long long n, int nz, int offanal, int nsect, int np, int nm, double inv_lscale,
    double *__restrict__ part) {
    __shared__ double acc[LG_WAVES][2 * LG_MAXQ];
    __shared__ double cnt_w[LG_WAVES];
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    const int nq = np * nm;
    for (int i = tid; i < LG_WAVES * 2 * LG_MAXQ; i += LG_BLOCK) (&acc[0][0])[i] = 0.0;
    if (tid < LG_WAVES) cnt_w[tid] = 0.0;
    __syncthreads();
    const int jc = np / 2 + 1;           // 1-based centre (k = 0)
    const int kmax = np - jc;
    const long long stride = (long long)gridDim.x * LG_BLOCK;
    double cnt = 0.0;
    for (long long base = (long long)blockIdx.x * LG_BLOCK; base < n; base += stride) {
        const long long i = base + tid;
        double c1 = 0.0, s1 = 0.0, ca = 1.0, sa = 0.0;
        if (i < n) {
            const int z = zone[i];
            const bool inc = z >= 1 && (z <= nz || (offanal && z == nz + 1));
            const typename Vec4<R>::T p = xv[i * ps];   // x, y lead the particle's data
            const double x = (double)p.x, y = (double)p.y;
            const double rc = sqrt(x * x + y * y);
            if (inc && rc > 0.0) {
                const double cx = x / rc, sy = y / rc;
                c1 = cx;
                s1 = sy;
                for (int m = 2; m <= nsect; m++) {
                    const double t = c1 * cx - s1 * sy;
                    s1 = s1 * cx + c1 * sy;
                    c1 = t;
                }
                const double a = 0.25 * (double)nsect * log(rc * inv_lscale);
                sincos(a, &sa, &ca);
                cnt += 1.0;
            }
        }
        // powers: z^im (harmonic), w^im (pitch step of this harmonic)
        double zr = 1.0, zi = 0.0, wr = 1.0, wi = 0.0;
        for (int im = 1; im <= nm; im++) {
            double t = zr * c1 - zi * s1;
            zi = zi * c1 + zr * s1;
            zr = t;
            t = wr * ca - wi * sa;
            wi = wi * ca + wr * sa;
            wr = t;
            const int q0 = (jc - 1) * nm + im - 1;
            double vr = wave_sum(zr), vi = wave_sum(zi);
            if (lane == 0) {
                acc[w][2 * q0] += vr;
                acc[w][2 * q0 + 1] += vi;
            }
            double gr = 1.0, gi = 0.0;
            for (int k = 1; k <= kmax; k++) {
                t = gr * wr - gi * wi;
                gi = gi * wr + gr * wi;
                gr = t;
                // pitch jc + k: g z; its mirror np + 1 - (jc + k): conj(g) z
                const int qp = (jc + k - 1) * nm + im - 1;
                const int qm = (np - jc - k) * nm + im - 1;
                const double pr = wave_sum(gr * zr - gi * zi);
                const double pi = wave_sum(gi * zr + gr * zi);
                const double mr = wave_sum(gr * zr + gi * zi);
                const double mi = wave_sum(gr * zi - gi * zr);
                if (lane == 0) {
                    acc[w][2 * qp] += pr;
                    acc[w][2 * qp + 1] += pi;
                    acc[w][2 * qm] += mr;
                    acc[w][2 * qm + 1] += mi;
                }
            }
        }
    }
    cnt = wave_sum(cnt);
    if (lane == 0) cnt_w[w] = cnt;
    __syncthreads();
    double *out = part + (size_t)blockIdx.x * (2 * nq + 1);
    for (int q = tid; q < 2 * nq; q += LG_BLOCK) {
        double v = 0.0;
        for (int ww = 0; ww < LG_WAVES; ww++) v += acc[ww][q];
        // [cos parts][sin parts]
        out[(q & 1) * nq + (q >> 1)] = v;
    }
    if (tid == 0) {
        double v = 0.0;
        for (int ww = 0; ww < LG_WAVES; ww++) v += cnt_w[ww];
        out[2 * nq] = v;
    }
}

// out[2 * nq] = (sum over blocks in order) / count
__global__ __launch_bounds__(LG_BLOCK) void lgsp_sum_kernel(const double *__restrict__ part,
                                                             int nb, int nq,
                                                             double *__restrict__ out) {
    const int stride = 2 * nq + 1;
    double cnt = 0.0;
    for (int b = 0; b < nb; b++) cnt += part[(size_t)b * stride + 2 * nq];
    const double inv = cnt > 0.0 ? 1.0 / cnt : 0.0;
    for (int q = threadIdx.x; q < 2 * nq; q += LG_BLOCK) {
        double v = 0.0;
        for (int b = 0; b < nb; b++) v += part[(size_t)b * stride + q];
        out[q] = v * inv;
    }
}

template <typename R>
hipError_t launch_lgspia(const void *xv, int ps, const uint8_t *zone, long long n, int nz,
                         int offanal, int nsect, int np, int nm, double lscale, double *part,
                         int nb, double *out, hipStream_t s) {
    if (np * nm > LG_MAXQ) return hipErrorInvalidValue;
    hipLaunchKernelGGL(lgsp_part_kernel<R>, dim3(nb), dim3(LG_BLOCK), 0, s,
                       (const typename Vec4<R>::T *)xv, ps, zone, n, nz, offanal, nsect, np, nm,
                       1.0 / lscale, part);
    hipLaunchKernelGGL(lgsp_sum_kernel, dim3(1), dim3(LG_BLOCK), 0, s, (const double *)part, nb,
                       np * nm, out);
    return hipGetLastError();
}

}  // namespace gh

using namespace gh;

// Blocks of the partial-sum pass: enough to fill the device, each striding
// over the particles (the sums' order depends only on n and this count).
static int lgsp_blocks(long long n) {
    long long nb = (n + LG_BLOCK - 1) / LG_BLOCK;
    if (nb > 2048) nb = 2048;
    return nb < 1 ? 1 : (int)nb;
}

extern "C" int galhip_lgspia(galhip_ctx *c, int32_t np, int32_t nm, double lscale,
                             double *out) {
    if (!c || !out) return gh_fail(GALHIP_EINVAL, "null argument");
    if (np < 1 || nm < 1 || (long long)np * nm > LG_MAXQ)
        return gh_fail(GALHIP_EINVAL, "LGSPIA: np x nm = %d x %d out of range (at most %d)", np,
                       nm, LG_MAXQ);
    if (!(lscale > 0.0)) return gh_fail(GALHIP_EINVAL, "LGSPIA: lscale must be positive");
    const int nq = np * nm;
    double *part = nullptr, *dout = nullptr;
    int nbmax = 1;
    for (int m = 0; m < c->nmem; m++) {
        const long long nm_ = c->ens ? c->m_off[m + 1] - c->m_off[m] : c->n;
        nbmax = std::max(nbmax, lgsp_blocks(nm_));
    }
    int rc = dalloc(&part, sizeof(double) * (size_t)nbmax * (2 * nq + 1)) |
             dalloc(&dout, sizeof(double) * (size_t)c->nmem * 2 * nq);
    if (!rc) {
        hipError_t e = hipSuccess;
        const size_t esz = c->prec == GALHIP_FP32 ? sizeof(float4) : sizeof(double4);
        // 4-vectors per particle: (x, y, vx, vy) on p2d, a two-vector record on p3d
        const int ps = c->p3d ? 2 : 1;
        for (int m = 0; m < c->nmem && e == hipSuccess; m++) {
            const long long off = c->ens ? c->m_off[m] : 0;
            const long long cnt = c->ens ? c->m_off[m + 1] - off : c->n;
            const void *xv = (const char *)c->xv + (size_t)off * ps * esz;
            const uint8_t *zn = c->zone + off;
            e = c->prec == GALHIP_FP32
                    ? launch_lgspia<float>(xv, ps, zn, cnt, c->nz, c->offanal, c->nsect, np,
                                           nm, lscale, part, lgsp_blocks(cnt),
                                           dout + 2 * nq * m, c->s)
                    : launch_lgspia<double>(xv, ps, zn, cnt, c->nz, c->offanal, c->nsect, np,
                                            nm, lscale, part, lgsp_blocks(cnt),
                                            dout + 2 * nq * m, c->s);
        }
        if (e == hipSuccess)
            e = hipMemcpyAsync(out, dout, sizeof(double) * (size_t)c->nmem * 2 * nq,
                               hipMemcpyDeviceToHost, c->s);
        if (e == hipSuccess) e = hipStreamSynchronize(c->s);
        if (e != hipSuccess) rc = gh_fail(GALHIP_EHIP, "lgspia: %s", hipGetErrorString(e));
    }
    dfree(part);
    dfree(dout);
    return rc;
}
