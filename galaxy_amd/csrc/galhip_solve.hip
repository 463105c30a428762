// galhip_solve.hip -- field solve of the p2d polar grid.
//
// analyse_kernel (one block per ring): zoncmb + scaled (SRC/lib15/zoncmb.f:
//   40-125, scaled.f:91) from the fixed-point zone sums, then the forward
//   azimuthal transform of p2manl (p2manl.f:35-43) restricted to the active
//   harmonics m < ng, as a direct real DFT against a twiddle table held in
//   L2: for the harmonic counts used here (ng <= 33 of na/2+1) this costs
//   about as much as a radix FFT and keeps exact FFTPACK conventions
//   (half-complex a_m = sum x cos, b_m = -sum x sin).  Optionally performs
//   the masscl of the next push (masscl.f:67-146) on its ring.
// synth_kernel (one block per field ring): p2conv's radial contraction with
//   the per-harmonic Green's matrices (p2conv.f:95-186), p2fsyn's inverse
//   transform and 1/na (p2fsyn.f:40-52) and polcat (polcat.f:26-40), writing
//   interleaved (F_x, F_y) and the potential.
// green_kernel: p2grnm's table construction in FP64 (p2grnm.f:98-203):
//   kernel samples (radfor/azifor/grnfun), DCT-I / DST-I over the half
//   circle, harmonic filter, p2gwrt's [m][field ring][source ring] layout.
#include "galhip_kern.h"

namespace gh {

constexpr int MAX_NA = 2048;
constexpr int MAX_NG = 64;

template <typename R> __device__ __forceinline__ R wave_sum(R v) { return wave_sum_dpp(v); }

// One block of SOLVE_BLOCK threads (16 wavefronts) per ring: the dot
// products over rings and spokes are short, so many wavefronts per CU are
// what hides the L2 latency of their loads.
constexpr int SOLVE_BLOCK = 1024;
constexpr int ANA_SPLIT = 1;   // blocks per ring in the analysis (harmonics)
constexpr int SYN_SPLIT = 1;   // blocks per ring in the synthesis (spokes)
constexpr int SYN_TPW = 3;     // contraction tasks per wavefront in flight together

// grid (nr, ANA_SPLIT, nmem): block (ir, h, mem) forms the ring's combined
// mass and the harmonics m = h, h + ANA_SPLIT, ... of grid mem (ensemble
// member; 0 for a single grid).
template <typename R>
__global__ __launch_bounds__(SOLVE_BLOCK) void analyse_kernel(AnaArgs<R> a) {
    __shared__ R rho_s[MAX_NA];
    __shared__ R tws[2 * MAX_NA];
    __shared__ int s_rng[4], s_ok;
    const int tid = threadIdx.x, na = a.na, ir = blockIdx.x + 1;
    const int h = blockIdx.y, mem = blockIdx.z;
    // member mem's grids
    const size_t mo = (size_t)mem * a.nr * na;
    int *const jrad = a.jrad + mem * GH_MAXZ;
    int *const krad = a.krad + mem * GH_MAXZ;
    int *const ranges = a.ranges + mem * 4;
    R *const rho = a.rho + mo;
    R *const coef = a.coef + (size_t)mem * a.ncoef * a.nr;
    const double q = a.qbits_m ? ldexp(1.0, -a.qbits_m[mem]) : a.q;
    const double pmass = a.pmass_m ? a.pmass_m[mem] : a.pmass;
    if (tid == 0) {
        int kr[GH_MAXZ], rg[4] = {0, 0, 0, 0};
        bool ok = solve_ranges(jrad, krad, a.nz, a.nr, a.mzone, a.lpot,
                               a.explicit_zones, a.jzone, a.kzone, kr, rg);
        s_ok = ok;
        for (int q = 0; q < 4; q++) s_rng[q] = rg[q];
        if (blockIdx.x == 0 && h == 0) {
            for (int q = 0; q < 4; q++) ranges[q] = ok ? rg[q] : -1;
            if (ok)
                for (int z = 1; z < a.nz; z++) krad[z] = kr[z];
        }
    }
    __syncthreads();
    // zoncmb + scaled for this ring
    const size_t row = mo + (size_t)(ir - 1) * na;
    for (int j = tid; j < na; j += SOLVE_BLOCK) {
        const size_t idx = row + j;
        double acc = (double)a.mfx[0][idx] * q;
        for (int z = 2; z <= a.nz; z++) {
            if (a.use_new[z - 1])
                acc = acc + a.fnew[z - 1] * ((double)a.mfx[z - 1][idx] * q);
            if (a.use_old[z - 1])
                acc = acc + a.fold[z - 1] * ((double)a.moz[z - 1][idx] * q);
        }
        R v = (R)(acc * pmass);
        rho_s[j] = v;
        if (h == 0) rho[idx - mo] = v;
    }
    for (int j = tid; j < 2 * na; j += SOLVE_BLOCK) tws[j] = a.tw[j];
    __syncthreads();
    if (!s_ok || ir < s_rng[0] || ir > s_rng[1]) return;
    // forward transform of the active harmonics: one wave per harmonic
    const int wave = tid >> 6, lane = tid & 63, nw = SOLVE_BLOCK / 64;
    for (int m = h + ANA_SPLIT * wave; m < a.ng; m += ANA_SPLIT * nw) {
        R sa = 0, sb = 0;
        if (!a.lg[m]) {
            int k = (m * lane) % na;
            const int step = (m * 64) % na;
            for (int j = lane; j < na; j += 64) {
                R rv = rho_s[j];
                sa += rv * tws[2 * k];
                sb += rv * tws[2 * k + 1];
                k += step;
                if (k >= na) k -= na;
            }
            sa = wave_sum(sa);
            sb = wave_sum(sb);
        }
        if (lane == 0) {
            if (m == 0) {
                coef[ir - 1] = sa;
            } else if (2 * m == na) {
                coef[(size_t)(2 * m - 1) * a.nr + ir - 1] = sa;
            } else {
                coef[(size_t)(2 * m - 1) * a.nr + ir - 1] = sa;
                coef[(size_t)(2 * m) * a.nr + ir - 1] = -sb;
            }
        }
    }
}

// grid (nr, SYN_SPLIT, nmem): block (ir, q, mem) forms the ring's
// contraction (small) and synthesises spokes [q na/SYN_SPLIT,
// (q+1) na/SYN_SPLIT) of grid mem.
// PRE: the contraction was done beforehand (the ensemble's MFMA GEMM):
// W is read from a.wpre[mem][type][cos/sin][m][ring] instead.
template <typename R, bool PRE>
__global__ __launch_bounds__(SOLVE_BLOCK) void synth_kernel(SynArgs<R> a) {
    typedef typename Vec2<R>::T R2;
    __shared__ R W[3][2][MAX_NG];
    __shared__ R tws[2 * MAX_NA];
    const int tid = threadIdx.x, na = a.na, nr = a.nr, ir = blockIdx.x + 1;
    const int q = blockIdx.y, mem = blockIdx.z;
    const int jlo = (q * na) / SYN_SPLIT, jhi = ((q + 1) * na) / SYN_SPLIT;
    const size_t mo = (size_t)mem * nr * na;
    const int *const ranges = a.ranges + mem * 4;
    const R *const coef = a.coef + (size_t)mem * a.ncoef * nr;
    // masscl of the coming push, on this block's cells
    if (a.masscl_z > 0) {
        const size_t row = mo + (size_t)(ir - 1) * na;
        for (int j = jlo + tid; j < jhi; j += SOLVE_BLOCK) {
            const size_t idx = row + j;
            for (int z = 1; z <= a.masscl_z; z++) {
                if (z >= 2) a.moz[z - 1][idx] = a.mfx[z - 1][idx];
                a.mfx[z - 1][idx] = 0;
            }
        }
        if (blockIdx.x == 0 && q == 0 && tid < a.masscl_z) {
            a.jrad[mem * GH_MAXZ + tid] = nr;
            a.krad[mem * GH_MAXZ + tid] = 0;
        }
    }
    const int jrs = ranges[0], krs = ranges[1];
    const int jrf = ranges[2], krf = ranges[3];
    if (jrs < 1 || ir < jrf || ir > krf) return;
    for (int j = tid; j < 2 * na; j += SOLVE_BLOCK) tws[j] = a.tw[j];
    const int ntypes = a.lpot ? 3 : 2;
    const int wave = tid >> 6, lane = tid & 63, nw = SOLVE_BLOCK / 64;
    const int mmax = a.mmax;
    const int maz = min(mmax - 2, a.ng - 1);
    if (PRE) {
        for (int t = tid; t < ntypes * 2 * a.ng; t += SOLVE_BLOCK) {
            const int type = t / (2 * a.ng), r = t - type * 2 * a.ng, cs = r / a.ng;
            const int m = r - cs * a.ng;
            W[type][cs][m] = (R)a.wpre[((((size_t)mem * 3 + type) * 2 + cs) * a.ng + m) * nr +
                                       ir - 1];
        }
    }
    // The contraction: one wavefront per (type, harmonic) dot product over
    // the source rings, SYN_TPW of them per wavefront at a time with all
    // their loads issued before the sums (the load latency, not the
    // arithmetic, bounds this phase); each lane's terms are added in source
    // order as one task at a time would.
    const int ntask = PRE ? 0 : ntypes * a.ng;
    for (int t0 = wave; t0 < ntask; t0 += SYN_TPW * nw) {
        R sc[SYN_TPW], ss[SYN_TPW];
        bool ok[SYN_TPW], azim[SYN_TPW], pair[SYN_TPW];
        const R *g[SYN_TPW], *mc[SYN_TPW], *ms[SYN_TPW];
#pragma unroll
        for (int u = 0; u < SYN_TPW; u++) {
            const int task = t0 + u * nw;
            const int type = task / a.ng, m = task - type * a.ng;
            azim[u] = type == 1;
            ok[u] = task < ntask && !a.lg[m] && (!azim[u] || (m >= 1 && m <= maz));
            pair[u] = m >= 1 && m < mmax - 1;
            const int off = m == 0 ? 0 : 2 * m - 1;
            g[u] = a.G + (((size_t)type * a.ng + m) * nr + (ir - 1)) * nr;
            mc[u] = coef + (size_t)off * nr;
            ms[u] = coef + (size_t)(off + 1) * nr;
            sc[u] = 0;
            ss[u] = 0;
        }
        for (int base = jrs; base <= krs; base += 4 * 64) {
            R gv[SYN_TPW][4], cv[SYN_TPW][4], sv[SYN_TPW][4];
#pragma unroll
            for (int u = 0; u < SYN_TPW; u++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int is = base + lane + 64 * r;
                    const bool in = ok[u] && is <= krs;
                    gv[u][r] = in ? g[u][is - 1] : (R)0;
                    cv[u][r] = in ? mc[u][is - 1] : (R)0;
                    sv[u][r] = (in && pair[u]) ? ms[u][is - 1] : (R)0;
                }
#pragma unroll
            for (int u = 0; u < SYN_TPW; u++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    if (base + lane + 64 * r > krs) continue;
                    if (!azim[u]) {
                        sc[u] += gv[u][r] * cv[u][r];
                        ss[u] += gv[u][r] * sv[u][r];
                    } else {
                        sc[u] += gv[u][r] * sv[u][r];
                        ss[u] -= gv[u][r] * cv[u][r];
                    }
                }
        }
#pragma unroll
        for (int u = 0; u < SYN_TPW; u++) {
            const int task = t0 + u * nw;
            if (task >= ntask) break;                 // wave-uniform
            const R c = wave_sum(sc[u]), sn = wave_sum(ss[u]);
            if (lane == 0) {
                const int type = task / a.ng, m = task - type * a.ng;
                W[type][0][m] = ok[u] ? c : (R)0;
                W[type][1][m] = ok[u] ? sn : (R)0;
            }
        }
    }
    __syncthreads();
    const R inv = (R)1 / (R)na;
    for (int j = jlo + tid; j < jhi; j += SOLVE_BLOCK) {
        R F[3] = {0, 0, 0};
        for (int type = 0; type < ntypes; type++) {
            R s = W[type][0][0];
            int k = j;
            for (int m = 1; m < a.ng; m++) {
                if (2 * m < na) {
                    s += (R)2 * (W[type][0][m] * tws[2 * k] - W[type][1][m] * tws[2 * k + 1]);
                } else {
                    s += (j & 1) ? -W[type][0][m] : W[type][0][m];
                }
                k += j;
                if (k >= na) k -= na;
            }
            F[type] = s * inv;
        }
        const R cp = tws[2 * j], sp = tws[2 * j + 1];
        R2 out;
        out.x = F[0] * cp - F[1] * sp;
        out.y = F[0] * sp + F[1] * cp;
        const size_t idx = mo + (size_t)(ir - 1) * na + j;
        a.fld[idx] = out;
        if (a.lpot) a.pot[idx] = F[2];
    }
}

// ---- Green's tables (FP64) ----

__device__ double softm(double ds) {
    if (ds <= 0.0) return 0.0;
    if (ds < 1.0)
        return 4.0 * ds * ds * ds * (1.0 / 3.0 - 0.3 * ds * ds + 0.125 * ds * ds * ds);
    if (ds < 2.0)
        return -1.0 / 15.0 + ds * ds * ds *
               (8.0 / 3.0 - 3.0 * ds + 1.2 * ds * ds - ds * ds * ds / 6.0);
    return 1.0;
}

__device__ double kernel_value(int type, double rs, double rf, double cth,
                               double sth, double eps, int tsoft) {
    double d2 = rs * rs + rf * rf - 2.0 * rs * rf * cth;
    if (type == 0) {       // radfor (radfor.f:25-38)
        if (tsoft == 1) return (rs * cth - rf) / pow(d2 + eps * eps, 1.5);
        return d2 > 0.0 ? (rs * cth - rf) * softm(sqrt(d2) / eps) / pow(d2, 1.5) : 0.0;
    }
    if (type == 1) {       // azifor (azifor.f:25-40)
        if (tsoft == 1) return -rs * sth / pow(d2 + eps * eps, 1.5);
        return d2 > 0.0 ? -rs * sth * softm(sqrt(d2) / eps) / pow(d2, 1.5) : 0.0;
    }
    // grnfun -> sftpot (sftpot.f:20-40)
    if (tsoft == 1) return -1.0 / sqrt(d2 + eps * eps);
    double x = sqrt(d2) / eps;
    if (x <= 0.0) return -1.4 / eps;
    if (x < 1.0) return (-1.4 + x * x * (2.0 / 3.0 - 0.3 * x * x + 0.1 * x * x * x)) / eps;
    if (x < 2.0)
        return (-1.6 + 1.0 / (15.0 * x) +
                x * x * (4.0 / 3.0 - x + 0.3 * x * x - x * x * x / 30.0)) / eps;
    return -1.0 / sqrt(d2);
}

// block (is, ir, type): samples over the half circle, then the transform
// of every kept harmonic.  twd = cos, sin of 2 pi k / na in double.
__global__ __launch_bounds__(GH_BLOCK) void green_kernel(
    double *__restrict__ G, const double *__restrict__ twd, int nr, int na,
    int ng, int nsect, double alpha, double eps, int tsoft) {
    __shared__ double f[MAX_NA / 2 + 1];
    const int is = blockIdx.x + 1, ir = blockIdx.y + 1, type = blockIdx.z;
    const int mmax = na / 2 + 1;
    const bool azim = type == 1;
    const double rs = exp(alpha * (double)(is - 1)) - 1.0;
    const double rf = exp(alpha * (double)(ir - 1)) - 1.0;
    for (int k = threadIdx.x; k < mmax; k += GH_BLOCK)
        f[k] = kernel_value(type, rs, rf, twd[2 * k], twd[2 * k + 1], eps, tsoft);
    __syncthreads();
    const int maz = min(mmax - 2, ng - 1);
    for (int m = threadIdx.x; m < ng; m += GH_BLOCK) {
        double v = 0.0;
        if (m % nsect == 0) {
            if (!azim) {
                // DCT-I over mmax points (rcost): f0 + (-1)^m f_n/2 + 2 sum
                v = f[0] + ((m & 1) ? -f[mmax - 1] : f[mmax - 1]);
                int k = m;
                for (int j = 1; j < mmax - 1; j++) {
                    v += 2.0 * f[j] * twd[2 * k];
                    k += m;
                    if (k >= na) k -= na;
                }
            } else if (m >= 1 && m <= maz) {
                // DST-I over the na/2-1 interior points (rsint)
                int k = m;
                for (int j = 1; j < mmax - 1; j++) {
                    v += 2.0 * f[j] * twd[2 * k + 1];
                    k += m;
                    if (k >= na) k -= na;
                }
            }
        }
        G[(((size_t)type * ng + m) * nr + (ir - 1)) * nr + (is - 1)] = v;
    }
}

template <typename S, typename D>
__global__ void convert_kernel(const S *__restrict__ in, D *__restrict__ out,
                               long long n) {
    long long i = (long long)blockIdx.x * GH_BLOCK + threadIdx.x;
    if (i < n) out[i] = (D)in[i];
}

template <typename T>
__global__ void fill_kernel(T *__restrict__ p, T v, long long n) {
    long long i = (long long)blockIdx.x * GH_BLOCK + threadIdx.x;
    if (i < n) p[i] = v;
}

// masscl (masscl.f:67-146) as a kernel: save zones 2..mz, clear zones 1..mz,
// reset jrad/krad of those zones, for the ncell = nmem x mesh cells of nmem
// grids (ring ranges GH_MAXZ apart).
__global__ void masscl_kernel(ZonePtrs zp, int mz, long long ncell, int *jrad,
                              int *krad, int nr, int nmem) {
    long long i = (long long)blockIdx.x * GH_BLOCK + threadIdx.x;
    if (i < ncell) {
        for (int z = 1; z <= mz; z++) {
            if (z >= 2) zp.moz[z - 1][i] = zp.mfx[z - 1][i];
            zp.mfx[z - 1][i] = 0;
        }
    }
    if (i < (long long)mz * nmem) {
        const int m = (int)(i / mz), z = (int)(i - (long long)m * mz);
        jrad[m * GH_MAXZ + z] = nr;
        krad[m * GH_MAXZ + z] = 0;
    }
}

// ---- host launchers ----

template <typename R>
hipError_t launch_analyse(const AnaArgs<R> &a, hipStream_t s) {
    if (a.na > MAX_NA) return hipErrorInvalidValue;
    hipLaunchKernelGGL(analyse_kernel<R>, dim3(a.nr, ANA_SPLIT, a.nmem), dim3(SOLVE_BLOCK), 0,
                       s, a);
    return hipGetLastError();
}

template <typename R>
hipError_t launch_synth(const SynArgs<R> &a, hipStream_t s) {
    if (a.ng > MAX_NG) return hipErrorInvalidValue;
    if (a.wpre)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(synth_kernel<R, true>), dim3(a.nr, SYN_SPLIT, a.nmem),
                           dim3(SOLVE_BLOCK), 0, s, a);
    else
        hipLaunchKernelGGL(HIP_KERNEL_NAME(synth_kernel<R, false>), dim3(a.nr, SYN_SPLIT, a.nmem),
                           dim3(SOLVE_BLOCK), 0, s, a);
    return hipGetLastError();
}

// ---- the ensemble's radial contraction on MFMA (split BF16, FP32 sums) ----
//
// For every (type, harmonic m) the contraction of all members is one GEMM:
//   W[field ring][column] = sum over source rings of G[ring][source] B[source][column],
// columns = (member, cos/sin) pairs, B = the members' coefficients in the
// arrangement p2conv uses for the type (F_R and the potential: (Mc, Ms);
// F_phi: (Ms, -Mc)), zero outside each member's source range.  Each operand
// x is split into three bf16 words, x1 = rn(x), x2 = rn(x - x1),
// x3 = rn(x - x1 - x2), which hold x to about 2^-24 of itself (G's split is
// taken from the FP64 master table, B's from the FP32 coefficients, which
// it holds exactly).  The product keeps the six terms whose word orders sum
// to at most 4 (x1 y1, x1 y2, x2 y1, x1 y3, x2 y2, x3 y1); each is exact in
// FP32 and the matrix cores accumulate them in FP32.  The dropped terms are
// below 2^-24 of the product, so the result has the error of an FP32 dot
// product.  The instructions are gfx950's K = 32 (16 x 16 tiles) and
// K = 16 (32 x 32 tiles) bf16 forms.

typedef __bf16 gh_v8bf __attribute__((ext_vector_type(8)));
typedef float gh_v4f __attribute__((ext_vector_type(4)));
typedef float gh_v16f __attribute__((ext_vector_type(16)));

__device__ __forceinline__ unsigned short bf16_rn(float x) {
    unsigned u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf16_to_f(unsigned short h) {
    return __uint_as_float((unsigned)h << 16);
}
__device__ __forceinline__ __bf16 bf16_bits(unsigned short h) {
    return __builtin_bit_cast(__bf16, h);
}

// G (FP64 master [3][ng][nr][nr]) -> three bf16 words [3][3][ng][nrp][nrp]
// (word-major), zero padded to nrp rings
__global__ void split_bf16_kernel(const double *__restrict__ G, unsigned short *__restrict__ w,
                                  int ng, int nr, int nrp) {
    const long long i = (long long)blockIdx.x * GH_BLOCK + threadIdx.x;
    const long long tot = 3LL * ng * nrp * nrp;
    if (i >= tot) return;
    const int k = (int)(i % nrp);
    const long long r1 = i / nrp;
    const int row = (int)(r1 % nrp);
    const long long tm = r1 / nrp;
    double g = 0.0;
    if (row < nr && k < nr) g = G[(tm * nr + row) * nr + k];
    const unsigned short h1 = bf16_rn((float)g);
    const double g1 = g - (double)bf16_to_f(h1);
    const unsigned short h2 = bf16_rn((float)g1);
    const double g2 = g1 - (double)bf16_to_f(h2);
    w[i] = h1;
    w[tot + i] = h2;
    w[2 * tot + i] = bf16_rn((float)g2);
}

// element (source ring k, column col) of the (type, m) right-hand side
__device__ __forceinline__ float ens_rhs(const EnsGemmArgs &a, int type, int m, bool pair,
                                         int off, int k, int col) {
    const int mem = col >> 1, cs = col & 1;
    if (mem >= a.nmem || k >= a.nr) return 0.f;
    const int jrs = a.ranges[4 * mem], krs = a.ranges[4 * mem + 1];
    if (jrs < 1 || k + 1 < jrs || k + 1 > krs) return 0.f;
    const float *cf = a.coef + (size_t)mem * a.ncoef * a.nr;
    const float mc = cf[(size_t)off * a.nr + k];
    const float ms = pair ? cf[(size_t)(off + 1) * a.nr + k] : 0.f;
    if (type == 1) return cs == 0 ? ms : -mc;
    return cs == 0 ? mc : ms;
}

__device__ __forceinline__ bool ens_valid(const EnsGemmArgs &a, int type, int m) {
    const int maz = min(a.mmax - 2, a.ng - 1);
    return !a.lg[m] && (type != 1 || (m >= 1 && m <= maz));
}

__device__ __forceinline__ void ens_store(const EnsGemmArgs &a, int type, int m, int row,
                                          int col, float v) {
    if (row >= a.nr || col >= 2 * a.nmem) return;
    const int mem = col >> 1, cs = col & 1;
    a.W[((((size_t)mem * 3 + type) * 2 + cs) * a.ng + m) * a.nr + row] = v;
}

// One lane's eight consecutive k of a G row (three words) and of a B column
// (split here), starting at k0 + kq.
struct SplitOps {
    gh_v8bf a1, a2, a3, b1, b2, b3;
};

__device__ __forceinline__ void load_ops(const EnsGemmArgs &a, SplitOps &o, size_t rowoff,
                                         size_t wstride, int type, int m, bool pair, int off,
                                         int k, int col) {
#pragma unroll
    for (int j = 0; j < 8; j++) {
        o.a1[j] = bf16_bits(a.gsp[rowoff + k + j]);
        o.a2[j] = bf16_bits(a.gsp[wstride + rowoff + k + j]);
        o.a3[j] = bf16_bits(a.gsp[2 * wstride + rowoff + k + j]);
        const float b = ens_rhs(a, type, m, pair, off, k + j, col);
        const unsigned short h1 = bf16_rn(b);
        const float r1 = b - bf16_to_f(h1);
        const unsigned short h2 = bf16_rn(r1);
        o.b1[j] = bf16_bits(h1);
        o.b2[j] = bf16_bits(h2);
        o.b3[j] = bf16_bits(bf16_rn(r1 - bf16_to_f(h2)));
    }
}

// 16 x 16 output tiles, K = 32 per instruction (v_mfma_f32_16x16x32_bf16):
// lane l holds A row l%16 and B column l%16 at eight k of its lane group
// l/16 (the same k for A and B), and D column l%16 at rows 4(l/16)..+3.
// WPB wavefronts per block, each its own tile along the columns.
// grid (column tiles / WPB, nrp/16, ntypes*ng).
template <int WPB>
__global__ __launch_bounds__(64 * WPB) void ens_gemm16_kernel(EnsGemmArgs a) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int tn = blockIdx.x * WPB + w, tm = blockIdx.y;
    const int type = blockIdx.z / a.ng, m = blockIdx.z - type * a.ng;
    if (tn * 16 >= 2 * a.nmem) return;
    const bool valid = ens_valid(a, type, m);
    const bool pair = m >= 1 && m < a.mmax - 1;
    const int off = m == 0 ? 0 : 2 * m - 1;
    gh_v4f acc = {0.f, 0.f, 0.f, 0.f}, lo = acc;   // x1 y1 terms; the five corrections
    const int col = tn * 16 + (lane & 15), kq = 8 * (lane >> 4);
    if (valid) {
        const size_t wstride = 3ull * a.ng * a.nrp * a.nrp;
        const size_t rowoff = ((size_t)(type * a.ng + m) * a.nrp + tm * 16 + (lane & 15)) * a.nrp;
        for (int k0 = 0; k0 < a.nrp; k0 += 32) {
            SplitOps o;
            load_ops(a, o, rowoff, wstride, type, m, pair, off, k0 + kq, col);
            lo = __builtin_amdgcn_mfma_f32_16x16x32_bf16(o.a3, o.b1, lo, 0, 0, 0);
            lo = __builtin_amdgcn_mfma_f32_16x16x32_bf16(o.a2, o.b2, lo, 0, 0, 0);
            lo = __builtin_amdgcn_mfma_f32_16x16x32_bf16(o.a1, o.b3, lo, 0, 0, 0);
            lo = __builtin_amdgcn_mfma_f32_16x16x32_bf16(o.a2, o.b1, lo, 0, 0, 0);
            lo = __builtin_amdgcn_mfma_f32_16x16x32_bf16(o.a1, o.b2, lo, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(o.a1, o.b1, acc, 0, 0, 0);
        }
    }
#pragma unroll
    for (int i = 0; i < 4; i++)
        ens_store(a, type, m, tm * 16 + 4 * (lane >> 4) + i, col, acc[i] + lo[i]);
}

// 32 x 32 output tiles, K = 16 per instruction (v_mfma_f32_32x32x16_bf16):
// lane l holds A row l%32 and B column l%32 at eight k of its lane group
// l/32, and D column l%32 at rows 8(i/4) + 4(l/32) + i%4, i = 0..15.
// grid (column tiles, nrp/32, ntypes*ng), one wavefront per block.
__global__ __launch_bounds__(64) void ens_gemm32_kernel(EnsGemmArgs a) {
    const int lane = threadIdx.x & 63;
    const int tn = blockIdx.x, tm = blockIdx.y;
    const int type = blockIdx.z / a.ng, m = blockIdx.z - type * a.ng;
    const bool valid = ens_valid(a, type, m);
    const bool pair = m >= 1 && m < a.mmax - 1;
    const int off = m == 0 ? 0 : 2 * m - 1;
    gh_v16f acc, lo;   // x1 y1 terms; the five corrections
#pragma unroll
    for (int i = 0; i < 16; i++) acc[i] = lo[i] = 0.f;
    const int col = tn * 32 + (lane & 31), kq = 8 * (lane >> 5);
    if (valid) {
        const size_t wstride = 3ull * a.ng * a.nrp * a.nrp;
        const size_t rowoff = ((size_t)(type * a.ng + m) * a.nrp + tm * 32 + (lane & 31)) * a.nrp;
        for (int k0 = 0; k0 < a.nrp; k0 += 16) {
            SplitOps o;
            load_ops(a, o, rowoff, wstride, type, m, pair, off, k0 + kq, col);
            lo = __builtin_amdgcn_mfma_f32_32x32x16_bf16(o.a3, o.b1, lo, 0, 0, 0);
            lo = __builtin_amdgcn_mfma_f32_32x32x16_bf16(o.a2, o.b2, lo, 0, 0, 0);
            lo = __builtin_amdgcn_mfma_f32_32x32x16_bf16(o.a1, o.b3, lo, 0, 0, 0);
            lo = __builtin_amdgcn_mfma_f32_32x32x16_bf16(o.a2, o.b1, lo, 0, 0, 0);
            lo = __builtin_amdgcn_mfma_f32_32x32x16_bf16(o.a1, o.b2, lo, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(o.a1, o.b1, acc, 0, 0, 0);
        }
    }
#pragma unroll
    for (int i = 0; i < 16; i++)
        ens_store(a, type, m, tm * 32 + 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3), col,
                  acc[i] + lo[i]);
}

hipError_t launch_split_bf16(const double *G, unsigned short *w, int ng, int nr, int nrp,
                             hipStream_t s) {
    const long long tot = 3LL * ng * nrp * nrp;
    hipLaunchKernelGGL(split_bf16_kernel, dim3((unsigned)((tot + GH_BLOCK - 1) / GH_BLOCK)),
                       dim3(GH_BLOCK), 0, s, G, w, ng, nr, nrp);
    return hipGetLastError();
}

// variant: 0 = 16x16 tiles, one wavefront per block; 1 = 16x16, four
// wavefronts per block; 2 = 32x32 tiles
hipError_t launch_ens_gemm(const EnsGemmArgs &a, int variant, hipStream_t s) {
    const int ncol = 2 * a.nmem;
    const unsigned z = (unsigned)(a.ntypes * a.ng);
    if (variant == 2) {
        hipLaunchKernelGGL(ens_gemm32_kernel, dim3((ncol + 31) / 32, a.nrp / 32, z), dim3(64), 0,
                           s, a);
    } else if (variant == 1) {
        const int tn = (ncol + 15) / 16;
        hipLaunchKernelGGL(HIP_KERNEL_NAME(ens_gemm16_kernel<4>), dim3((tn + 3) / 4, a.nrp / 16, z),
                           dim3(256), 0, s, a);
    } else {
        hipLaunchKernelGGL(HIP_KERNEL_NAME(ens_gemm16_kernel<1>),
                           dim3((ncol + 15) / 16, a.nrp / 16, z), dim3(64), 0, s, a);
    }
    return hipGetLastError();
}

hipError_t launch_green(double *G, const double *twd, int nr, int na, int ng,
                        int nsect, double alpha, double eps, int tsoft,
                        hipStream_t s) {
    if (na > MAX_NA) return hipErrorInvalidValue;
    hipLaunchKernelGGL(green_kernel, dim3(nr, nr, 3), dim3(GH_BLOCK), 0, s, G,
                       twd, nr, na, ng, nsect, alpha, eps, tsoft);
    return hipGetLastError();
}

hipError_t launch_d2f(const double *in, float *out, long long n, hipStream_t s) {
    hipLaunchKernelGGL((convert_kernel<double, float>),
                       dim3((unsigned)((n + GH_BLOCK - 1) / GH_BLOCK)),
                       dim3(GH_BLOCK), 0, s, in, out, n);
    return hipGetLastError();
}

hipError_t launch_fill_i64(long long *p, long long v, long long n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(fill_kernel<long long>,
                       dim3((unsigned)((n + GH_BLOCK - 1) / GH_BLOCK)),
                       dim3(GH_BLOCK), 0, s, p, v, n);
    return hipGetLastError();
}

hipError_t launch_fill_i32(int *p, int v, long long n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(fill_kernel<int>,
                       dim3((unsigned)((n + GH_BLOCK - 1) / GH_BLOCK)),
                       dim3(GH_BLOCK), 0, s, p, v, n);
    return hipGetLastError();
}

hipError_t launch_masscl(const ZonePtrs &zp, int mz, long long mesh, int *jrad,
                         int *krad, int nr, int nmem, hipStream_t s) {
    const long long ncell = mesh * nmem, nz = (long long)mz * nmem;
    long long n = ncell > nz ? ncell : nz;
    hipLaunchKernelGGL(masscl_kernel, dim3((unsigned)((n + GH_BLOCK - 1) / GH_BLOCK)),
                       dim3(GH_BLOCK), 0, s, zp, mz, ncell, jrad, krad, nr, nmem);
    return hipGetLastError();
}

template hipError_t launch_analyse<float>(const AnaArgs<float> &, hipStream_t);
template hipError_t launch_analyse<double>(const AnaArgs<double> &, hipStream_t);
template hipError_t launch_synth<float>(const SynArgs<float> &, hipStream_t);
template hipError_t launch_synth<double>(const SynArgs<double> &, hipStream_t);

}  // namespace gh
