// galhip_api.hip -- C ABI of libgalhip (include/galhip.h): context set-up,
// device storage, the reference's zone/time bookkeeping on the host, and
// the launch sequence of findf and step.
//
// Host-side state mirrors the reference's /zones/ and /point/ commons
// (SRC/lib15/inc/admin.f:36-71): istep, mzone, lstep, mstep are advanced by
// the host exactly as masscl/zoncmb advance them (masscl.f:31-98,
// zoncmb.f:40-125); jrad/krad live on the device because the particle kernel
// produces them, and the solve kernels read them there.
#include "galhip_ctx.h"

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

namespace gh {
template <typename R> hipError_t launch_push(const PushArgs<R> &, hipStream_t);
template <typename R> hipError_t launch_massgn(const PushArgs<R> &, hipStream_t);
template <typename R> hipError_t launch_gather(const GatherArgs<R> &, hipStream_t);
template <typename R> hipError_t launch_ut(const PushArgs<R> &, hipStream_t);
template <typename R>
hipError_t launch_getacc_group(const PushArgs<R> &, typename Vec2<R>::T *, hipStream_t);
template <typename R>
hipError_t launch_stpgrp_group(const PushArgs<R> &, const typename Vec2<R>::T *, hipStream_t);
template <typename R> hipError_t launch_tmcent(const PushArgs<R> &, R, hipStream_t);
template <typename R>
hipError_t launch_integrals(const PushArgs<R> &, const R *, const double4 *, int, double, int,
                            double *, double *, hipStream_t);
template <typename R>
hipError_t launch_key(const typename Vec2<R>::T *, const uint8_t *, long long,
                      int, int, int, int, unsigned, unsigned short *, hipStream_t);
template <typename R>
hipError_t launch_key_ens(const typename Vec2<R>::T *, const uint8_t *, int, int, int, int, int,
                          unsigned, const int4 *, unsigned *, unsigned *, hipStream_t);
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
    const typename Vec2<R>::T *lo_in;
    typename Vec2<R>::T *lo_out;
};
template <typename R>
hipError_t launch_permute_all(const PermArgs<R> &, hipStream_t);
template <typename R>
hipError_t launch_interleave(const R *, typename Vec4<R>::T *, unsigned *, long long, hipStream_t);
template <typename R>
hipError_t launch_deinterleave(const typename Vec4<R>::T *, const uint8_t *, const unsigned *, R *,
                               int *, long long, hipStream_t);
template <typename R> hipError_t launch_analyse(const AnaArgs<R> &, hipStream_t);
template <typename R> hipError_t launch_synth(const SynArgs<R> &, hipStream_t);
hipError_t launch_green(double *, const double *, int, int, int, int, double,
                        double, int, hipStream_t);
hipError_t launch_split_bf16(const double *, unsigned short *, int, int, int,
                             hipStream_t);
hipError_t launch_ens_gemm(const EnsGemmArgs &, int, hipStream_t);
hipError_t launch_d2f(const double *, float *, long long, hipStream_t);
hipError_t launch_fill_i64(long long *, long long, long long, hipStream_t);
hipError_t launch_fill_i32(int *, int, long long, hipStream_t);
hipError_t launch_masscl(const ZonePtrs &, int, long long, int *, int *, int, int,
                         hipStream_t);
}  // namespace gh

using namespace gh;

static thread_local std::string g_err;

int gh_fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

static void free_particles(galhip_ctx *c) {
    dfree(c->xlo); dfree(c->xlo2);
    dfree(c->xv); dfree(c->ut); dfree(c->zone); dfree(c->id);
    dfree(c->xv2); dfree(c->ut2); dfree(c->zone2); dfree(c->id2);
    dfree(c->key); dfree(c->key2); dfree(c->perm); dfree(c->perm2); dfree(c->sort_tmp);
    c->sort_tmp_bytes = 0;
    c->n = c->cap = 0;
}

extern "C" const char *galhip_last_error(void) { return g_err.c_str(); }

// ---- set-up ------------------------------------------------------------------

int qbits_for(long long n) {
    int b = 0;
    while (b < 62 && (1LL << b) < n + 1) b++;
    if (b < 1) b = 1;
    return 62 - b;
}

// p2dset for one grid (nmem = 1, pmass from the parameters) or an ensemble
// of nmem grids with per-member particle masses.
static int p2d_create(const galhip_p2d_params *p, int nmem, const double *pmass,
                      galhip_ctx **out) {
    if (!p || !out) return gh_fail(GALHIP_EINVAL, "galhip_p2dset: null argument");
    *out = nullptr;
    if (p->precision != GALHIP_FP32 && p->precision != GALHIP_FP64)
        return gh_fail(GALHIP_EINVAL, "precision must be GALHIP_FP32 or GALHIP_FP64");
    int na = 2 * (p->na / 2);                       // setgrd.f:58
    // nr <= 4096: the push packs a deposit's zone, ring and spoke in one word
    if (p->nr < 3 || p->nr > 4096 || na < 4 || na > 2048)
        return gh_fail(GALHIP_EINVAL, "grid %d x %d out of range", p->nr, p->na);
    if (p->nzones < 1 || p->nzones > GALHIP_MAXZONES)
        return gh_fail(GALHIP_EINVAL, "nzones %d out of range", p->nzones);
    if (p->nsect < 1) return gh_fail(GALHIP_EINVAL, "nsect not set");
    if (!(p->softl > 0.0)) return gh_fail(GALHIP_EINVAL, "Softening length must be greater than 0");
    if (p->tsoft != 1 && p->tsoft != 2) return gh_fail(GALHIP_EINVAL, "Unrecognized softening rule");
    if (p->nskip < 0 || p->nskip > GALHIP_MAXSKIP) return gh_fail(GALHIP_EINVAL, "nskip out of range");
    galhip_ctx *c = new galhip_ctx();   // value-initialised: scalars zeroed
    c->nmem = nmem;
    c->m_pmass.assign(nmem, p->pmass);
    if (pmass)
        for (int m = 0; m < nmem; m++) c->m_pmass[m] = pmass[m];
    c->m_qbits.assign(nmem, qbits_for(1));
    c->m_off.assign(nmem + 1, 0);
    c->prec = p->precision;
    c->nr = p->nr;
    c->na = na;
    c->mmax = na / 2 + 1;
    c->ng = p->hash + 1 < c->mmax ? p->hash + 1 : c->mmax;   // p2dset.f:40, setgrd.f:89
    if (c->ng < 1) c->ng = 1;
    if (c->ng > 64) { delete c; return gh_fail(GALHIP_EINVAL, "at most 64 harmonics supported"); }
    c->nsect = p->nsect;
    c->mesh = c->nr * na;
    c->tsoft = p->tsoft;
    c->nz = p->nzones;
    c->offanal = p->offanal != 0;
    c->softl = p->softl;
    c->pmass = p->pmass;
    for (int i = 0; i < c->nz; i++) {
        c->nstep[i] = p->nstep[i] > 0 ? p->nstep[i] : 1;
        c->zbr[i] = p->zbr[i];
    }
    // filter (filset.f:118-141): lg[m] != 0 skips harmonic m
    c->lg.assign(c->mmax, 1);
    c->lg[0] = 0;
    for (int m = 1; m < c->ng; m++) c->lg[m] = (c->nsect == 1 || m % c->nsect == 0) ? 0 : 1;
    for (int i = 0; i < p->nskip; i++)
        if (p->skip[i] >= 0 && p->skip[i] < c->mmax) c->lg[p->skip[i]] = 1;
    // alpha, rgrid = grofu(nr - 1), thmax (setgrd.f:88,127-128,172-176)
    const double two_pi = 2.0 * 3.141592653589793238462643;
    if (c->prec == GALHIP_FP32) {
        float a = (float)(two_pi / (double)na);
        c->alpha = a;
        c->rgrid = expf(a * (float)(c->nr - 1)) - 1.0f;
        c->thmax = (float)na / (float)c->nsect;
    } else {
        c->alpha = two_pi / (double)na;
        c->rgrid = exp(c->alpha * (double)(c->nr - 1)) - 1.0;
        c->thmax = (double)na / (double)c->nsect;
    }
    // outermost cell against the softening length (setgrd.f:181-199): the
    // reference stops when the radial size of the last cell exceeds 5 softl
    {
        double aw;
        if (c->prec == GALHIP_FP32) {
            float a = (float)c->alpha;
            aw = (double)((float)c->rgrid - (expf(a * (float)(c->nr - 2)) - 1.0f));
        } else {
            aw = c->rgrid - (exp(c->alpha * (double)(c->nr - 2)) - 1.0);
        }
        if (aw / c->softl > 5.0) {
            delete c;
            return gh_fail(GALHIP_ERANGE, "SETGRD: Unsuitable choice of grid parameters "
                        "(outermost cell %.4g times the softening length)", aw / p->softl);
        }
    }
    c->ncoef = 2 * c->ng;
    c->istep = 0;
    c->mzone = c->nz;
    int rc = 0;
    // a blocking stream: the synchronous hipMemcpy calls of the load/unload
    // and parity entry points (null stream) are ordered with its kernels
    hipError_t e = hipStreamCreate(&c->s);
    if (e != hipSuccess) { delete c; return gh_fail(GALHIP_EHIP, "hipStreamCreate: %s", hipGetErrorString(e)); }
    size_t es = esize(c);
    const size_t nm = (size_t)nmem, cells = nm * c->mesh;
    for (int z = 0; z < c->nz; z++) {
        rc |= dalloc(&c->mfx[z], sizeof(long long) * cells);
        rc |= dalloc(&c->moz[z], sizeof(long long) * cells);
    }
    rc |= dalloc(&c->jrad, sizeof(int) * GH_MAXZ * nm);
    rc |= dalloc(&c->krad, sizeof(int) * GH_MAXZ * nm);
    rc |= dalloc(&c->ranges, sizeof(int) * 4 * nm);
    rc |= dalloc(&c->lg_d, sizeof(int) * c->mmax);
    rc |= dalloc(&c->counters, sizeof(long long) * 4 * nm);
    rc |= dalloc(&c->rho, es * cells);
    rc |= dalloc(&c->coef, es * c->ncoef * c->nr * nm);
    rc |= dalloc(&c->fld2, 2 * es * cells);
    rc |= dalloc(&c->pot, es * cells);
    rc |= dalloc(&c->m_qbits_d, sizeof(int) * nm);
    rc |= dalloc(&c->m_pmass_d, sizeof(double) * nm);
    rc |= dalloc(&c->tw, 2 * es * na);
    rc |= dalloc(&c->twd, 2 * sizeof(double) * na);
    rc |= dalloc(&c->Gd, sizeof(double) * 3 * (size_t)c->ng * c->nr * c->nr);
    if (c->prec == GALHIP_FP32)
        rc |= dalloc(&c->G, sizeof(float) * 3 * (size_t)c->ng * c->nr * c->nr);
    else
        c->G = c->Gd;
    if (rc) { galhip_free(c); return rc; }
    // twiddles: cos, sin of 2 pi k / na, k reduced exactly
    std::vector<double> twd(2 * na);
    for (int k = 0; k < na; k++) {
        twd[2 * k] = cos(two_pi * (double)k / (double)na);
        twd[2 * k + 1] = sin(two_pi * (double)k / (double)na);
    }
    HIPCHK(hipMemcpy(c->twd, twd.data(), sizeof(double) * 2 * na, hipMemcpyHostToDevice));
    if (c->prec == GALHIP_FP32) {
        std::vector<float> twf(2 * na);
        for (int i = 0; i < 2 * na; i++) twf[i] = (float)twd[i];
        HIPCHK(hipMemcpy(c->tw, twf.data(), sizeof(float) * 2 * na, hipMemcpyHostToDevice));
    } else {
        HIPCHK(hipMemcpy(c->tw, twd.data(), sizeof(double) * 2 * na, hipMemcpyHostToDevice));
    }
    HIPCHK(hipMemcpy(c->lg_d, c->lg.data(), sizeof(int) * c->mmax, hipMemcpyHostToDevice));
    std::vector<int> jr(GH_MAXZ * nm, c->nr), kr(GH_MAXZ * nm, 0);
    HIPCHK(hipMemcpy(c->jrad, jr.data(), sizeof(int) * GH_MAXZ * nm, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->krad, kr.data(), sizeof(int) * GH_MAXZ * nm, hipMemcpyHostToDevice));
    HIPCHK(hipMemset(c->counters, 0, sizeof(long long) * 4 * nm));
    HIPCHK(hipMemset(c->fld2, 0, 2 * es * cells));
    HIPCHK(hipMemset(c->pot, 0, es * cells));
    HIPCHK(hipMemset(c->rho, 0, es * cells));
    for (int z = 0; z < c->nz; z++) {
        HIPCHK(hipMemset(c->mfx[z], 0, sizeof(long long) * cells));
        HIPCHK(hipMemset(c->moz[z], 0, sizeof(long long) * cells));
    }
    HIPCHK(hipMemcpy(c->m_qbits_d, c->m_qbits.data(), sizeof(int) * nm, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->m_pmass_d, c->m_pmass.data(), sizeof(double) * nm, hipMemcpyHostToDevice));
    c->sort_interval = 0;
    c->qbits = qbits_for(1);
    *out = c;
    return 0;
}

extern "C" int galhip_p2dset(const galhip_p2d_params *p, galhip_ctx **out) {
    return p2d_create(p, 1, nullptr, out);
}

extern "C" int galhip_ensset(const galhip_p2d_params *p, int32_t nmem, const double *pmass,
                             galhip_ctx **out) {
    if (!p || !out || !pmass) return gh_fail(GALHIP_EINVAL, "galhip_ensset: null argument");
    *out = nullptr;
    if (nmem < 1 || nmem > GALHIP_MAXMEMBERS)
        return gh_fail(GALHIP_EINVAL, "nmem %d outside 1..%d", nmem, GALHIP_MAXMEMBERS);
    for (int m = 0; m < nmem; m++)
        if (!(pmass[m] > 0.0)) return gh_fail(GALHIP_EINVAL, "member %d: particle mass not positive", m);
    int rc = p2d_create(p, nmem, pmass, out);
    if (rc) return rc;
    (*out)->ens = 1;
    return 0;
}

static void drop_graphs(galhip_ctx *c) {
    for (auto &g : c->graphs) (void)hipGraphExecDestroy(g.second);
    c->graphs.clear();
}

extern "C" int galhip_free(galhip_ctx *c) {
    if (!c) return 0;
    (void)hipStreamSynchronize(c->s);
    if (c->p3d) p3_free(c);
    free_particles(c);
    for (int z = 0; z < GH_MAXZ; z++) { dfree(c->mfx[z]); dfree(c->moz[z]); }
    dfree(c->jrad); dfree(c->krad); dfree(c->ranges); dfree(c->lg_d); dfree(c->counters);
    dfree(c->fp_c);
    dfree(c->rho); dfree(c->coef); dfree(c->fld2); dfree(c->pot); dfree(c->tw); dfree(c->twd);
    if (c->G != (void *)c->Gd) dfree(c->G);
    dfree(c->Gd);
    dfree(c->fr_c);
    dfree(c->m_qbits_d); dfree(c->m_pmass_d); dfree(c->blk_d);
    dfree(c->gsp); dfree(c->wpre);
    drop_graphs(c);
    adj_free(c);
    for (auto e : c->ev_free) (void)hipEventDestroy(e);
    for (auto e : c->ev_used) (void)hipEventDestroy(e);
    if (c->s) (void)hipStreamDestroy(c->s);
    delete c;
    return 0;
}

extern "C" int galhip_geometry(const galhip_ctx *c, double *out) {
    if (!c || !out) return gh_fail(GALHIP_EINVAL, "null argument");
    out[0] = c->alpha;
    out[1] = c->rgrid;
    out[2] = c->thmax;
    out[3] = c->ng;
    out[4] = c->mmax;
    out[5] = c->mesh;
    return 0;
}

// ---- Green's tables ------------------------------------------------------------

static int green_to_work(galhip_ctx *c) {
    c->gsplit_valid = 0;
    if (c->prec == GALHIP_FP32) {
        long long n = 3LL * c->ng * c->nr * c->nr;
        HIPCHK(launch_d2f(c->Gd, (float *)c->G, n, c->s));
    }
    c->have_green = 1;
    return 0;
}

extern "C" int galhip_greenm(galhip_ctx *c) {
    if (!c) return gh_fail(GALHIP_EINVAL, "null context");
    if (c->p3d) return p3_greenm(c);
    double alpha_d = 2.0 * 3.141592653589793238462643 / (double)c->na;
    HIPCHK(launch_green(c->Gd, c->twd, c->nr, c->na, c->ng, c->nsect, alpha_d,
                        c->softl, c->tsoft, c->s));
    int rc = green_to_work(c);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(c->s));
    return 0;
}

extern "C" int galhip_get_green(const galhip_ctx *c, double *G) {
    if (c && G && c->p3d) return p3_get_green(c, G);
    if (!c || !G) return gh_fail(GALHIP_EINVAL, "null argument");
    HIPCHK(hipStreamSynchronize(c->s));
    HIPCHK(hipMemcpy(G, c->Gd, sizeof(double) * 3 * (size_t)c->ng * c->nr * c->nr,
                     hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int galhip_set_green(galhip_ctx *c, const double *G) {
    if (c && G && c->p3d) return p3_set_green(c, G);
    if (!c || !G) return gh_fail(GALHIP_EINVAL, "null argument");
    HIPCHK(hipMemcpy(c->Gd, G, sizeof(double) * 3 * (size_t)c->ng * c->nr * c->nr,
                     hipMemcpyHostToDevice));
    int rc = green_to_work(c);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(c->s));
    return 0;
}

extern "C" int galhip_set_fixforce(galhip_ctx *c, int32_t n, double dr,
                                   const double *val, const double *der) {
    if (!c) return gh_fail(GALHIP_EINVAL, "null context");
    HIPCHK(hipStreamSynchronize(c->s));
    dfree(c->fr_c);
    c->nfr = 0;
    if (n <= 0) return 0;
    if (n < 2 || !(dr > 0.0) || !val || !der)
        return gh_fail(GALHIP_EINVAL, "fixed force table needs n >= 2, dr > 0");
    const std::vector<double4> cf = radtab_coeffs(n, dr, val, der);
    int rc = dalloc(&c->fr_c, sizeof(double4) * n);
    if (rc) return rc;
    HIPCHK(hipMemcpy(c->fr_c, cf.data(), sizeof(double4) * n, hipMemcpyHostToDevice));
    c->nfr = n;
    c->fr_dr = dr;
    return 0;
}

// Fixed potential table for the energy diagnostics, as galhip_set_fixforce.
extern "C" int galhip_set_fixpot(galhip_ctx *c, int32_t n, double dr,
                                 const double *val, const double *der) {
    if (!c) return gh_fail(GALHIP_EINVAL, "null context");
    HIPCHK(hipStreamSynchronize(c->s));
    dfree(c->fp_c);
    c->nfp = 0;
    if (n <= 0) return 0;
    if (n < 2 || !(dr > 0.0) || !val || !der)
        return gh_fail(GALHIP_EINVAL, "fixed potential table needs n >= 2, dr > 0");
    const std::vector<double4> cf = radtab_coeffs(n, dr, val, der);
    int rc = dalloc(&c->fp_c, sizeof(double4) * n);
    if (rc) return rc;
    HIPCHK(hipMemcpy(c->fp_c, cf.data(), sizeof(double4) * n, hipMemcpyHostToDevice));
    c->nfp = n;
    c->fp_dr = dr;
    return 0;
}

// ---- particles -------------------------------------------------------------------

static int ensure_particles(galhip_ctx *c, long long n) {
    if (n <= c->cap) { c->n = n; return 0; }
    free_particles(c);
    size_t es = esize(c);
    int rc = 0;
    rc |= dalloc(&c->xv, 4 * es * n); rc |= dalloc(&c->ut, 2 * es * n);
    rc |= dalloc(&c->xv2, 4 * es * n); rc |= dalloc(&c->ut2, 2 * es * n);
    rc |= dalloc(&c->zone, n); rc |= dalloc(&c->id, sizeof(unsigned) * n);
    rc |= dalloc(&c->zone2, n); rc |= dalloc(&c->id2, sizeof(unsigned) * n);
    rc |= dalloc(&c->key, sizeof(unsigned) * n); rc |= dalloc(&c->key2, sizeof(unsigned) * n);
    rc |= dalloc(&c->perm, sizeof(unsigned) * n); rc |= dalloc(&c->perm2, sizeof(unsigned) * n);
    if (rc) { free_particles(c); return rc; }
    // scratch of the two sorts: 32-bit keys with stored values (ensembles),
    // 16-bit keys with the identity as values (one grid)
    size_t tmp = 0, tmp16 = 0;
    hipError_t e = rocprim::radix_sort_pairs(
        nullptr, tmp, c->key, c->key2, c->perm, c->perm2, (int)n, 0, 32, c->s);
    if (e == hipSuccess)
        e = rocprim::radix_sort_pairs(nullptr, tmp16, (unsigned short *)c->key,
                                      (unsigned short *)c->key2,
                                      rocprim::counting_iterator<unsigned>(0u), c->perm2,
                                      (int)n, 0, 16, c->s);
    if (e != hipSuccess) return gh_fail(GALHIP_EHIP, "radix sort sizing: %s", hipGetErrorString(e));
    if (tmp16 > tmp) tmp = tmp16;
    rc = dalloc(&c->sort_tmp, tmp);
    if (c->comp) rc |= dalloc(&c->xlo, 2 * es * n) | dalloc(&c->xlo2, 2 * es * n);
    if (rc) { free_particles(c); return rc; }
    c->sort_tmp_bytes = tmp;
    c->cap = c->n = n;
    return 0;
}

template <typename R> static int sort_impl(galhip_ctx *c) {
    long long n = c->n;
    if (n <= 1) return 0;
    // keys of at most 16 bits: cells shifted right until they fit, off-grid last
    int kshift = 0;
    while (((long long)c->mesh >> kshift) > 65534) kshift++;
    unsigned kmax = (unsigned)((c->mesh - 1) >> kshift) + 1;
    int end_bit = 1;
    while ((1LL << end_bit) <= (long long)kmax) end_bit++;
    if (c->ens) {
        // member-major keys: the cell (unshifted, off-grid = mesh) below the
        // member index
        kmax = (unsigned)c->mesh;
        int kbits = 1;
        while ((1LL << kbits) <= (long long)kmax) kbits++;
        int mbits = 0;
        while ((1LL << mbits) < (long long)c->nmem) mbits++;
        if (kbits + mbits > 32) return gh_fail(GALHIP_EINVAL, "ensemble sort keys exceed 32 bits");
        HIPCHK(launch_key_ens<R>((const typename Vec2<R>::T *)c->ut, c->zone, c->nblk, c->nr,
                                 c->na, c->nz, kbits, kmax, c->blk_d, c->key, c->perm, c->s));
        end_bit = kbits + mbits;
    } else {
        // one grid: 16-bit keys, values the identity (no index array)
        HIPCHK(launch_key<R>((const typename Vec2<R>::T *)c->ut, c->zone, n, c->nr,
                             c->na, c->nz, kshift, kmax, (unsigned short *)c->key, c->s));
    }
    size_t tmp = c->sort_tmp_bytes;
    if (c->ens)
        HIPCHK(rocprim::radix_sort_pairs(c->sort_tmp, tmp, c->key, c->key2, c->perm, c->perm2,
                                         (int)n, 0, end_bit, c->s));
    else
        HIPCHK(rocprim::radix_sort_pairs(c->sort_tmp, tmp, (unsigned short *)c->key,
                                         (unsigned short *)c->key2,
                                         rocprim::counting_iterator<unsigned>(0u), c->perm2,
                                         (int)n, 0, end_bit, c->s));
    PermArgs<R> p;
    p.xv_in = (const typename Vec4<R>::T *)c->xv;
    p.xv_out = (typename Vec4<R>::T *)c->xv2;
    p.ut_in = (const typename Vec2<R>::T *)c->ut;
    p.ut_out = (typename Vec2<R>::T *)c->ut2;
    p.zin = c->zone;
    p.zout = c->zone2;
    p.iin = c->id;
    p.iout = c->id2;
    p.perm = c->perm2;
    p.n = n;
    p.lo_in = (const typename Vec2<R>::T *)c->xlo;
    p.lo_out = (typename Vec2<R>::T *)c->xlo2;
    HIPCHK(launch_permute_all<R>(p, c->s));
    if (c->xlo) std::swap(c->xlo, c->xlo2);
    std::swap(c->xv, c->xv2);
    std::swap(c->ut, c->ut2);
    std::swap(c->zone, c->zone2);
    std::swap(c->id, c->id2);
    c->order_ver++;
    c->steps_since_sort = 0;
    return 0;
}

extern "C" int galhip_sort(galhip_ctx *c) {
    if (!c) return gh_fail(GALHIP_EINVAL, "null context");
    if (c->p3d) return p3_sort(c);
    return c->prec == GALHIP_FP32 ? sort_impl<float>(c) : sort_impl<double>(c);
}

extern "C" int galhip_get_order(const galhip_ctx *c, int32_t *ids) {
    if (!c || (!ids && c->n > 0)) return gh_fail(GALHIP_EINVAL, "null argument");
    HIPCHK(hipStreamSynchronize(c->s));
    if (c->n > 0)
        HIPCHK(hipMemcpy(ids, c->id, sizeof(unsigned) * c->n, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int galhip_set_sort_interval(galhip_ctx *c, int32_t k) {
    if (!c) return gh_fail(GALHIP_EINVAL, "null context");
    c->sort_interval = k < 0 ? 0 : k;
    return 0;
}

extern "C" int galhip_set_pmass(galhip_ctx *c, double pmass) {
    if (!c) return gh_fail(GALHIP_EINVAL, "null context");
    if (c->ens) return gh_fail(GALHIP_EINVAL, "ensemble members keep the masses of galhip_ensset");
    if (!(pmass > 0.0)) return gh_fail(GALHIP_EINVAL, "particle mass not positive");
    c->pmass = pmass;
    if (!c->m_pmass.empty()) c->m_pmass[0] = pmass;
    c->have_field = 0;
    c->pot_valid = 0;
    return 0;
}

static int compute_ut(galhip_ctx *c);

// the four planar host arrays through the second particle buffer (free
// outside a sort) into interleaved storage, and the load-order ids
static int load_xv(galhip_ctx *c, long long n, const void *x, const void *y, const void *vx,
                   const void *vy) {
    const size_t es = esize(c);
    unsigned char *stage = (unsigned char *)c->xv2;
    const void *src[4] = {x, y, vx, vy};
    for (int k = 0; k < 4; k++)
        HIPCHK(hipMemcpy(stage + k * es * n, src[k], es * n, hipMemcpyHostToDevice));
    if (c->prec == GALHIP_FP32)
        HIPCHK(launch_interleave<float>((const float *)stage, (float4 *)c->xv, c->id, n, c->s));
    else
        HIPCHK(launch_interleave<double>((const double *)stage, (double4 *)c->xv, c->id, n, c->s));
    return 0;
}

extern "C" int galhip_restrt(galhip_ctx *c, int64_t n, const void *x,
                             const void *y, const void *vx, const void *vy,
                             const int32_t *iz) {
    if (!c || n < 0 || (n > 0 && (!x || !y || !vx || !vy || !iz)))
        return gh_fail(GALHIP_EINVAL, "galhip_restrt: bad arguments");
    if (c->p3d) return gh_fail(GALHIP_EINVAL, "p3d context: use galhip_restrt3");
    if (c->ens) return gh_fail(GALHIP_EINVAL, "ensemble context: use galhip_ens_restrt");
    if (n >= (1LL << 31)) return gh_fail(GALHIP_EINVAL, "at most 2^31-1 particles");
    HIPCHK(hipStreamSynchronize(c->s));
    c->tcent = 0;   // loaded velocities: not (yet) time-centred
    int rc = ensure_particles(c, n);
    if (rc) return rc;
    size_t es = esize(c);
    std::vector<uint8_t> z(n);
    long long cnt[4] = {0, 0, 0, 0};   // counters: [0] noff, [1] pushes
    for (long long i = 0; i < n; i++) {
        int v = iz[i];
        if (v < 1 || v > c->nz + 1)
            return gh_fail(GALHIP_EINVAL, "particle %lld has zone %d outside 1..%d", i, v, c->nz + 1);
        z[i] = (uint8_t)v;
        if (v == c->nz + 1) cnt[0]++;    // /admin/ noff: the off-grid list
    }
    HIPCHK(hipMemcpy(c->counters, cnt, sizeof cnt, hipMemcpyHostToDevice));
    c->qbits = qbits_for(n);
    c->order_ver++;
    if (n == 0) return 0;
    rc = load_xv(c, n, x, y, vx, vy);
    if (rc) return rc;
    HIPCHK(hipMemcpy(c->zone, z.data(), n, hipMemcpyHostToDevice));
    if (c->xlo) HIPCHK(hipMemset(c->xlo, 0, 2 * es * n));
    c->qbits = qbits_for(n);
    rc = compute_ut(c);
    if (rc) return rc;
    rc = galhip_sort(c);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(c->s));
    return 0;
}

static int unload_all(galhip_ctx *c, void *x, void *y, void *vx, void *vy, int32_t *iz);

extern "C" int galhip_unload(galhip_ctx *c, void *x, void *y, void *vx,
                             void *vy, int32_t *iz) {
    if (!c) return gh_fail(GALHIP_EINVAL, "null context");
    if (c->ens) return gh_fail(GALHIP_EINVAL, "ensemble context: use galhip_ens_unload");
    if (c->p3d) return p3_unload(c, x, y, vx, vy, iz);
    return unload_all(c, x, y, vx, vy, iz);
}

// every particle back to host arrays in load order (ensembles: the members
// one after another)
static int unload_all(galhip_ctx *c, void *x, void *y, void *vx, void *vy, int32_t *iz) {
    HIPCHK(hipStreamSynchronize(c->s));
    long long n = c->n;
    if (n <= 0) return 0;
    // planar, in load order, into the second particle buffer and the sort
    // keys (both free outside a sort), then straight to the caller's arrays
    void *dst[4] = {x, y, vx, vy};
    const size_t es = esize(c);
    unsigned char *stage = (unsigned char *)c->xv2;
    int *zs = (int *)c->key;
    if (c->prec == GALHIP_FP32)
        HIPCHK(launch_deinterleave<float>((const float4 *)c->xv, c->zone, c->id, (float *)stage,
                                          zs, n, c->s));
    else
        HIPCHK(launch_deinterleave<double>((const double4 *)c->xv, c->zone, c->id,
                                           (double *)stage, zs, n, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    for (int k = 0; k < 4; k++)
        if (dst[k]) HIPCHK(hipMemcpy(dst[k], stage + k * es * n, es * n, hipMemcpyDeviceToHost));
    if (iz) HIPCHK(hipMemcpy(iz, zs, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
    return 0;
}

// ---- zone bookkeeping (host) -----------------------------------------------------

void set_mzone(galhip_ctx *c) {
    int m = 1;
    if (c->nz > 1)
        for (int i = 1; i <= c->nz; i++)
            if (c->istep % c->nstep[i - 1] == 0) m = i;
    c->mzone = m;
}

// masscl's time stamps (masscl.f:31-40,92-96)
void masscl_host(galhip_ctx *c, bool normal) {
    for (int z = 1; z <= c->mzone; z++) {
        int ns = c->nstep[z - 1];
        if (normal) c->mstep[z - 1] = c->istep + ns;
        else c->mstep[z - 1] = c->istep <= 0 ? 0 : ((c->istep - 1) / ns + 1) * ns;
    }
    for (int z = 2; z <= c->mzone; z++) {
        c->lstep[0][z - 1] = c->lstep[1][z - 1];
        c->lstep[1][z - 1] = c->mstep[z - 1];
    }
}

// zoncmb's fractions for the current istep (zoncmb.f:40-125)
int zoncmb_host(galhip_ctx *c) {
    set_mzone(c);
    for (int z = 2; z <= c->nz; z++) {
        int ns = c->nstep[z - 1];
        int i = c->lstep[1][z - 1] - c->istep;
        if (i < 0 || i > ns)
            return gh_fail(GALHIP_ERANGE, "ZONCMB: New mass array at a useless time (istep %d, lstep %d)",
                        c->istep, c->lstep[1][z - 1]);
        c->use_new[z - 1] = i != ns;
        c->fnew[z - 1] = 1.0 - (double)i / (double)ns;
        c->use_old[z - 1] = 0;
        c->fold[z - 1] = 0.0;
        if (i != 0) {
            int j = c->istep - c->lstep[0][z - 1];
            if (c->lstep[0][z - 1] < 0 || j < 0 || j > ns)
                return gh_fail(GALHIP_ERANGE, "ZONCMB: Old mass array at a useless time (istep %d, lstep %d)",
                            c->istep, c->lstep[0][z - 1]);
            c->use_old[z - 1] = 1;
            c->fold[z - 1] = 1.0 - (double)j / (double)ns;
        }
    }
    return 0;
}

extern "C" int galhip_set_step(galhip_ctx *c, int32_t istep, const int32_t *lstep) {
    if (!c) return gh_fail(GALHIP_EINVAL, "null context");
    c->istep = istep;
    if (lstep)
        for (int z = 0; z < c->nz; z++) {
            c->lstep[0][z] = lstep[z];
            c->lstep[1][z] = lstep[c->nz + z];
        }
    set_mzone(c);
    c->masscl_done = 0;
    // with time stamps given, zoncmb's fractions for the new time too, so a
    // solve may follow mass arrays loaded by galhip_set_mass
    if (lstep) return zoncmb_host(c);
    return 0;
}

extern "C" int galhip_get_step(const galhip_ctx *c, int32_t *istep, int32_t *mzone,
                               int32_t *lstep) {
    if (!c) return gh_fail(GALHIP_EINVAL, "null context");
    if (istep) *istep = c->istep;
    if (mzone) *mzone = c->mzone;
    if (lstep)
        for (int z = 0; z < c->nz; z++) {
            lstep[z] = c->lstep[0][z];
            lstep[c->nz + z] = c->lstep[1][z];
        }
    return 0;
}

// ---- kernel argument assembly -----------------------------------------------------

static UtConst ut_const(const galhip_ctx *c) {
    UtConst k;
    k.alpha_d = c->alpha;                 // alpha in the context precision
    k.inv_alpha = 1.0 / c->alpha;
    k.tscale = (double)c->na / (2.0 * 3.141592653589793238462643);
    k.na = c->na;
    return k;
}

template <typename R> static PushArgs<R> push_args(galhip_ctx *c) {
    PushArgs<R> a;
    memset(&a, 0, sizeof a);
    a.xv = (typename Vec4<R>::T *)c->xv;
    a.ut = (typename Vec2<R>::T *)c->ut;
    a.zone = c->zone;
    a.n = c->n;
    a.fld = (const typename Vec2<R>::T *)c->fld2;
    for (int z = 0; z < c->nz; z++) a.mfx[z] = c->mfx[z];
    a.jrad = c->jrad;
    a.krad = c->krad;
    a.counters = c->counters;
    a.fr_c = c->fr_c;
    a.fr_inv_dr = c->nfr > 0 ? 1.0 / c->fr_dr : 0.0;
    a.nfr = c->nfr;
    a.fixforce = c->nfr > 0;
    a.nr = c->nr; a.na = c->na; a.mesh = c->mesh; a.nz = c->nz;
    a.mzone = c->mzone; a.offanal = c->offanal; a.nsect = c->nsect; a.qbits = c->qbits;
    a.k = ut_const(c);
    a.rgrid = (R)c->rgrid; a.thmax = (R)c->thmax;
    for (int z = 0; z < c->nz; z++) {
        a.zbr[z] = (R)c->zbr[z];
        a.nstep[z] = c->nstep[z];
    }
    if (c->ens) {
        a.blk = c->blk_d;
        a.qbits_m = c->m_qbits_d;
        a.nblocks = c->nblk;
    }
    a.xlo = (typename Vec2<R>::T *)c->xlo;
    return a;
}

// The argument block of one ensemble member as a single grid: its particle
// range, grids and fixed-point scale (one-grid kernels on member m).
template <typename R> static PushArgs<R> member_args(galhip_ctx *c, int m) {
    PushArgs<R> a = push_args<R>(c);
    if (!c->ens) return a;
    a.blk = nullptr;
    a.qbits_m = nullptr;
    a.nblocks = 0;
    const long long o = c->m_off[m];
    const size_t mo = (size_t)m * c->mesh;
    if (a.xlo) a.xlo += o;
    a.xv = (typename Vec4<R>::T *)c->xv + o;
    a.ut = (typename Vec2<R>::T *)c->ut + o;
    a.zone = c->zone + o;
    a.n = c->m_off[m + 1] - o;
    a.fld = (const typename Vec2<R>::T *)c->fld2 + mo;
    for (int z = 0; z < c->nz; z++) a.mfx[z] = c->mfx[z] + mo;
    a.jrad = c->jrad + m * GH_MAXZ;
    a.krad = c->krad + m * GH_MAXZ;
    a.counters = c->counters + m * 4;
    a.qbits = c->m_qbits[m];
    return a;
}

static int compute_ut(galhip_ctx *c) {
    if (c->prec == GALHIP_FP32) HIPCHK(launch_ut<float>(push_args<float>(c), c->s));
    else HIPCHK(launch_ut<double>(push_args<double>(c), c->s));
    return 0;
}

template <typename R>
static AnaArgs<R> ana_args(galhip_ctx *c, int lpot, int explicit_zones,
                           int jzone, int kzone) {
    AnaArgs<R> a;
    memset(&a, 0, sizeof a);
    for (int z = 0; z < c->nz; z++) {
        a.mfx[z] = c->mfx[z];
        a.moz[z] = c->moz[z];
        a.fnew[z] = c->fnew[z];
        a.fold[z] = c->fold[z];
        a.use_new[z] = c->use_new[z];
        a.use_old[z] = c->use_old[z];
    }
    a.jrad = c->jrad; a.krad = c->krad; a.ranges = c->ranges;
    a.rho = (R *)c->rho; a.coef = (R *)c->coef; a.tw = (const R *)c->tw; a.lg = c->lg_d;
    a.q = ldexp(1.0, -c->qbits);
    a.pmass = c->pmass;
    a.nr = c->nr; a.na = c->na; a.ng = c->ng; a.ncoef = c->ncoef; a.nz = c->nz;
    a.mzone = c->mzone; a.lpot = lpot; a.jzone = jzone; a.kzone = kzone;
    a.explicit_zones = explicit_zones;
    a.nmem = c->nmem;
    if (c->ens) {
        a.qbits_m = c->m_qbits_d;
        a.pmass_m = c->m_pmass_d;
    }
    return a;
}

template <typename R> static SynArgs<R> syn_args(galhip_ctx *c, int lpot,
                                                 int masscl_z) {
    SynArgs<R> a;
    memset(&a, 0, sizeof a);
    for (int z = 0; z < c->nz; z++) {
        a.mfx[z] = c->mfx[z];
        a.moz[z] = c->moz[z];
    }
    a.jrad = c->jrad;
    a.krad = c->krad;
    a.masscl_z = masscl_z;
    a.coef = (const R *)c->coef;
    a.G = (const R *)c->G;
    a.tw = (const R *)c->tw;
    a.lg = c->lg_d;
    a.ranges = c->ranges;
    a.fld = (typename Vec2<R>::T *)c->fld2;
    a.pot = (R *)c->pot;
    a.nr = c->nr; a.na = c->na; a.ng = c->ng; a.ncoef = c->ncoef; a.mmax = c->mmax;
    a.lpot = lpot;
    a.nmem = c->nmem;
    return a;
}

// ---- masset / findf / step -----------------------------------------------------------

template <typename R> static int masset_impl(galhip_ctx *c) {
    c->mzone = c->nz;
    masscl_host(c, false);
    ZonePtrs zp;
    memset(&zp, 0, sizeof zp);
    for (int z = 0; z < c->nz; z++) { zp.mfx[z] = c->mfx[z]; zp.moz[z] = c->moz[z]; }
    HIPCHK(launch_masscl(zp, c->nz, c->mesh, c->jrad, c->krad, c->nr, c->nmem, c->s));
    for (int m = 0; m < c->nmem; m++) HIPCHK(launch_massgn<R>(member_args<R>(c, m), c->s));
    int rc = zoncmb_host(c);
    if (rc) return rc;
    c->masscl_done = 0;
    HIPCHK(hipStreamSynchronize(c->s));
    return 0;
}

extern "C" int galhip_masset(galhip_ctx *c) {
    if (!c) return gh_fail(GALHIP_EINVAL, "null context");
    if (c->p3d) return p3_masset(c);
    return c->prec == GALHIP_FP32 ? masset_impl<float>(c) : masset_impl<double>(c);
}

// The MFMA contraction's arguments for a solve with ntypes field types.
static EnsGemmArgs gemm_args(galhip_ctx *c, int ntypes) {
    EnsGemmArgs g;
    memset(&g, 0, sizeof g);
    g.gsp = c->gsp;
    g.coef = (const float *)c->coef;
    g.ranges = c->ranges;
    g.lg = c->lg_d;
    g.W = c->wpre;
    g.nr = c->nr;
    g.nrp = c->nrp;
    g.ng = c->ng;
    g.ncoef = c->ncoef;
    g.nmem = c->nmem;
    g.mmax = c->mmax;
    g.ntypes = ntypes;
    return g;
}

// Before an MFMA solve: the bf16 split of the current tables, and the GEMM
// variant -- tuned once per context by timing each variant on this
// context's shapes (HIP events, 5 launches each after one warm-up).
static int mfma_prepare(galhip_ctx *c) {
    if (!c->gsplit_valid) {
        HIPCHK(launch_split_bf16(c->Gd, c->gsp, c->ng, c->nr, c->nrp, c->s));
        c->gsplit_valid = 1;
    }
    if (c->gemm_variant < 0) {
        const EnsGemmArgs g = gemm_args(c, 2);
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        int best = 0;
        for (int v = 0; v < 3; v++) {
            HIPCHK(launch_ens_gemm(g, v, c->s));
            HIPCHK(hipEventRecord(e0, c->s));
            for (int r = 0; r < 5; r++) HIPCHK(launch_ens_gemm(g, v, c->s));
            HIPCHK(hipEventRecord(e1, c->s));
            HIPCHK(hipEventSynchronize(e1));
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, e0, e1));
            c->gemm_us[v] = 1e3 * ms / 5.0;
            if (c->gemm_us[v] < c->gemm_us[best]) best = v;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        c->gemm_variant = best;
    }
    return 0;
}

// analyse, [MFMA contraction], synthesis
template <typename R>
static int launch_solve(galhip_ctx *c, const AnaArgs<R> &a, SynArgs<R> &b, int lpot) {
    HIPCHK(launch_analyse<R>(a, c->s));
    if (c->contract == 1) {
        HIPCHK(launch_ens_gemm(gemm_args(c, lpot ? 3 : 2), c->gemm_variant, c->s));
        b.wpre = c->wpre;
    }
    HIPCHK(launch_synth<R>(b, c->s));
    return 0;
}

template <typename R>
static int solve_impl(galhip_ctx *c, int lpot, int explicit_zones, int jzone,
                      int kzone, int masscl_z) {
    if (!c->have_green) return gh_fail(GALHIP_EINVAL, "P2CONV: Green function not set (call galhip_greenm)");
    if (c->contract == 1) {
        int rc = mfma_prepare(c);
        if (rc) return rc;
    }
    AnaArgs<R> a = ana_args<R>(c, lpot, explicit_zones, jzone, kzone);
    SynArgs<R> b = syn_args<R>(c, lpot, masscl_z);
    int rc = launch_solve<R>(c, a, b, lpot);
    if (rc) return rc;
    c->have_field = 1;
    c->pot_valid = lpot != 0;
    return 0;
}

extern "C" int galhip_set_contraction(galhip_ctx *c, int32_t mode) {
    if (!c) return gh_fail(GALHIP_EINVAL, "null context");
    if (mode < 0 || mode > 4) return gh_fail(GALHIP_EINVAL, "contraction mode %d", mode);
    if (mode >= 1 && (c->p3d || c->prec != GALHIP_FP32))
        return gh_fail(GALHIP_EINVAL, "the MFMA contraction is built for FP32 p2d grids");
    HIPCHK(hipStreamSynchronize(c->s));
    if (mode >= 1 && !c->gsp) {
        c->nrp = (c->nr + 31) / 32 * 32;
        const size_t ng = (size_t)3 * c->ng * c->nrp * c->nrp;
        int rc = dalloc(&c->gsp, 3 * sizeof(unsigned short) * ng) |
                 dalloc(&c->wpre, sizeof(float) * (size_t)c->nmem * 3 * 2 * c->ng * c->nr);
        if (rc) return rc;
        HIPCHK(hipMemset(c->wpre, 0, sizeof(float) * (size_t)c->nmem * 3 * 2 * c->ng * c->nr));
        c->gsplit_valid = 0;
        c->gemm_variant = -1;
    }
    c->contract = mode >= 1 ? 1 : 0;
    if (mode >= 2) c->gemm_variant = mode - 2;   // a fixed variant: no tuning
    drop_graphs(c);
    return 0;
}

extern "C" int galhip_get_contraction(const galhip_ctx *c, int32_t *mode, int32_t *variant,
                                      double *us) {
    if (!c) return gh_fail(GALHIP_EINVAL, "null context");
    if (mode) *mode = c->contract;
    if (variant) *variant = c->gemm_variant;
    if (us)
        for (int v = 0; v < 3; v++) us[v] = c->gemm_variant >= 0 ? c->gemm_us[v] : 0.0;
    return 0;
}

extern "C" int galhip_findf(galhip_ctx *c, int32_t lpot) {
    if (!c) return gh_fail(GALHIP_EINVAL, "null context");
    if (c->mzone < 1 || c->mzone > c->nz) return gh_fail(GALHIP_ERANGE, "FINDF: invalid mzone");
    if (c->p3d) return p3_solve(c, lpot != 0, 0, 0, 0);
    return c->prec == GALHIP_FP32 ? solve_impl<float>(c, lpot != 0, 0, 0, 0, 0)
                                  : solve_impl<double>(c, lpot != 0, 0, 0, 0, 0);
}

extern "C" int galhip_p2fndf(galhip_ctx *c, int32_t jzone, int32_t kzone, int32_t lpot) {
    if (!c) return gh_fail(GALHIP_EINVAL, "null context");
    if (c->p3d) return p3_solve(c, lpot != 0, 1, jzone, kzone);
    return c->prec == GALHIP_FP32 ? solve_impl<float>(c, lpot != 0, 1, jzone, kzone, 0)
                                  : solve_impl<double>(c, lpot != 0, 1, jzone, kzone, 0);
}

hipEvent_t take_event(galhip_ctx *c) {
    hipEvent_t e;
    if (!c->ev_free.empty()) {
        e = c->ev_free.back();
        c->ev_free.pop_back();
    } else {
        (void)hipEventCreate(&e);
    }
    c->ev_used.push_back(e);
    return e;
}

template <typename R> static int step_impl(galhip_ctx *c) {
    set_mzone(c);
    masscl_host(c, true);
    if (!c->masscl_done) {
        ZonePtrs zp;
        memset(&zp, 0, sizeof zp);
        for (int z = 0; z < c->nz; z++) { zp.mfx[z] = c->mfx[z]; zp.moz[z] = c->moz[z]; }
        HIPCHK(launch_masscl(zp, c->mzone, c->mesh, c->jrad, c->krad, c->nr, c->nmem, c->s));
    }
    c->masscl_done = 0;
    PushArgs<R> a = push_args<R>(c);
    if (c->prof) HIPCHK(hipEventRecord(take_event(c), c->s));
    HIPCHK(launch_push<R>(a, c->s));
    if (c->prof) {
        HIPCHK(hipEventRecord(take_event(c), c->s));
        c->prof_launch++;
    }
    c->istep += 1;
    return zoncmb_host(c);
}

extern "C" int galhip_step(galhip_ctx *c) {
    if (!c) return gh_fail(GALHIP_EINVAL, "null context");
    if (!c->have_field) return gh_fail(GALHIP_ERANGE, "GETACC: Accelerations at useless time");
    if (c->p3d) return p3_step(c);
    return c->prec == GALHIP_FP32 ? step_impl<float>(c) : step_impl<double>(c);
}

template <typename R> static int tmcent_impl(galhip_ctx *c, int start) {
    for (int m = 0; m < c->nmem; m++)
        HIPCHK(launch_tmcent<R>(member_args<R>(c, m), start ? (R)0.5 : (R)-0.5, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    c->tcent = start != 0;
    return 0;
}

extern "C" int galhip_tmcent(galhip_ctx *c, int32_t start) {
    if (!c) return gh_fail(GALHIP_EINVAL, "null context");
    if ((start != 0) == (c->tcent != 0))
        return gh_fail(GALHIP_ERANGE, "TMCENT: Logic error in call (tcent flag already %s)",
                    c->tcent ? "up" : "down");
    if (!c->have_field) return gh_fail(GALHIP_ERANGE, "GETACC: Accelerations at useless time");
    if (c->p3d) return p3_tmcent(c, start);
    return c->prec == GALHIP_FP32 ? tmcent_impl<float>(c, start) : tmcent_impl<double>(c, start);
}

// Partial-sum rows of the p2d integrals kernels: 9 per block of the
// largest grid, then the 9 totals of every grid.
static long long integrals_nb(const galhip_ctx *c) {
    long long nbmax = 1;
    for (int m = 0; m < c->nmem; m++) {
        const long long nm = c->ens ? c->m_off[m + 1] - c->m_off[m] : c->n;
        nbmax = std::max(nbmax, (nm + GH_BLOCK - 1) / GH_BLOCK);
    }
    return nbmax;
}

// out[nmem][9]: each grid's integrals (one grid: nmem = 1), member by
// member with the one-grid kernels, so that a member's sums are those of
// the same disc in its own context.  buf: a device buffer of
// 9 * (integrals_nb(c) + nmem) doubles, or null to allocate one here.
template <typename R> static int integrals_impl(galhip_ctx *c, int centred, double *out,
                                                double *buf) {
    const long long nbmax = integrals_nb(c);
    double *part = buf, *dout = buf ? buf + 9 * nbmax : nullptr;
    int rc = 0;
    if (!buf)
        rc = dalloc(&part, sizeof(double) * 9 * nbmax) |
             dalloc(&dout, sizeof(double) * 9 * c->nmem);
    if (!rc) {
        hipError_t e = hipSuccess;
        for (int m = 0; m < c->nmem && e == hipSuccess; m++) {
            PushArgs<R> a = member_args<R>(c, m);
            const R *pot = c->pot_valid ? (const R *)c->pot + (size_t)m * c->mesh : nullptr;
            e = launch_integrals<R>(a, pot, c->fp_c, c->nfp, c->fp_dr, centred,
                                    part, dout + 9 * m, c->s);
        }
        if (e == hipSuccess) e = hipMemcpyAsync(out, dout, sizeof(double) * 9 * c->nmem,
                                                hipMemcpyDeviceToHost, c->s);
        if (e == hipSuccess) e = hipStreamSynchronize(c->s);
        if (e != hipSuccess) rc = gh_fail(GALHIP_EHIP, "integrals: %s", hipGetErrorString(e));
    }
    if (!buf) {
        dfree(part);
        dfree(dout);
    }
    if (rc) return rc;
    // particle sums times the particle mass; no potential: NaN
    for (int m = 0; m < c->nmem; m++) {
        const double pm = c->ens ? c->m_pmass[m] : c->pmass;
        for (int q = 0; q < 9; q++) out[9 * m + q] *= pm;
        if (!c->pot_valid) out[9 * m + 7] = NAN;
    }
    return 0;
}

static int integrals_check(const galhip_ctx *c, int32_t centred, const double *out) {
    if (!c || !out) return gh_fail(GALHIP_EINVAL, "null argument");
    if (centred && !c->have_field)
        return gh_fail(GALHIP_ERANGE, "GETACC: Accelerations at useless time");
    if (c->ens) return gh_fail(GALHIP_EINVAL, "ensemble context: use galhip_ens_integrals");
    return 0;
}

extern "C" int galhip_integrals(galhip_ctx *c, int32_t centred, double *out) {
    int rc = integrals_check(c, centred, out);
    if (rc) return rc;
    if (c->p3d) return p3_integrals(c, centred != 0, out);
    return c->prec == GALHIP_FP32 ? integrals_impl<float>(c, centred != 0, out, nullptr)
                                  : integrals_impl<double>(c, centred != 0, out, nullptr);
}

// ---- the integrals without allocation ------------------------------------------
//
// galhip_integrals_noalloc keeps its partial sums in `scratch`, a file-scope
// device array of GH_SCRATCH_DOUBLES doubles (12 MiB) that every GPU the
// library runs on holds until the process ends.  In host code the array's
// name is only a stand-in for the symbol, not a device address: the address
// comes from hipGetSymbolAddress, taken once per device (with that device
// current) on the entry point's first call there, and only that pointer is
// handed to kernels.  Every call on one GPU shares that GPU's array, so calls
// that overlap on one GPU (two host threads, or two contexts' streams) race.
#define GH_SCRATCH_DOUBLES (3 << 19)
#define GH_SCRATCH_MAXDEV 64
static __device__ __align__(256) double scratch[GH_SCRATCH_DOUBLES];

static std::mutex g_scratch_mu;
static double *g_scratch[GH_SCRATCH_MAXDEV];

static int scratch_on_device(double **out) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= GH_SCRATCH_MAXDEV)
        return gh_fail(GALHIP_EINVAL, "device %d: at most %d devices", dev, GH_SCRATCH_MAXDEV);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return gh_fail(GALHIP_EHIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev,
                           hipGetErrorString(e));
        g_scratch[dev] = (double *)p;
    }
    *out = g_scratch[dev];
    return 0;
}

extern "C" int galhip_integrals_noalloc(galhip_ctx *c, int32_t centred, double *out) {
    int rc = integrals_check(c, centred, out);
    if (rc) return rc;
    const long long need = c->p3d ? p3_integrals_doubles(c) : 9 * (integrals_nb(c) + 1);
    if (need > GH_SCRATCH_DOUBLES)
        return gh_fail(GALHIP_ERANGE,
                       "integrals_noalloc: %lld partial sums exceed the %d of the static "
                       "scratch (12 MiB; use galhip_integrals)", need, GH_SCRATCH_DOUBLES);
    double *buf = nullptr;
    rc = scratch_on_device(&buf);
    if (rc) return rc;
    if (c->p3d) return p3_integrals(c, centred != 0, out, buf);
    return c->prec == GALHIP_FP32 ? integrals_impl<float>(c, centred != 0, out, buf)
                                  : integrals_impl<double>(c, centred != 0, out, buf);
}

// One step of galhip_run -- findf with the fused masscl, then step -- as a
// HIP graph of its three launches, with the host bookkeeping of
// solve_impl and step_impl in the same order.  A graph is fully determined
// by its launch arguments, so graphs are cached under the bytes of the
// argument blocks and replayed whenever a step repeats them (the zone
// schedule is periodic; a sort swaps the particle arrays).
template <typename R> static int graph_step(galhip_ctx *c) {
    if (!c->have_green) return gh_fail(GALHIP_EINVAL, "P2CONV: Green function not set (call galhip_greenm)");
    if (c->contract == 1) {
        int rc = mfma_prepare(c);
        if (rc) return rc;
    }
    const AnaArgs<R> a = ana_args<R>(c, 0, 0, 0, 0);
    SynArgs<R> b = syn_args<R>(c, 0, c->mzone);
    if (c->contract == 1) b.wpre = c->wpre;
    c->have_field = 1;
    c->pot_valid = 0;
    set_mzone(c);
    masscl_host(c, true);
    c->masscl_done = 0;
    const PushArgs<R> p = push_args<R>(c);
    std::string key((const char *)&a, sizeof a);
    key.append((const char *)&b, sizeof b);
    key.append((const char *)&p, sizeof p);
    key.append((const char *)&c->contract, sizeof c->contract);
    key.append((const char *)&c->gemm_variant, sizeof c->gemm_variant);
    hipGraphExec_t ex = nullptr;
    for (auto &g : c->graphs)
        if (g.first == key) {
            ex = g.second;
            break;
        }
    if (!ex) {
        if (c->graphs.size() >= 64) {
            HIPCHK(hipStreamSynchronize(c->s));
            drop_graphs(c);
        }
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamBeginCapture(c->s, hipStreamCaptureModeThreadLocal));
        hipError_t e = launch_analyse<R>(a, c->s);
        if (e == hipSuccess && c->contract == 1)
            e = launch_ens_gemm(gemm_args(c, 2), c->gemm_variant, c->s);
        if (e == hipSuccess) e = launch_synth<R>(b, c->s);
        if (e == hipSuccess) e = launch_push<R>(p, c->s);
        hipError_t e2 = hipStreamEndCapture(c->s, &g);
        if (e == hipSuccess) e = e2;
        if (e == hipSuccess) e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        if (g) (void)hipGraphDestroy(g);
        if (e != hipSuccess) return gh_fail(GALHIP_EHIP, "step graph: %s", hipGetErrorString(e));
        c->graphs.emplace_back(key, ex);
    }
    HIPCHK(hipGraphLaunch(ex, c->s));
    c->istep += 1;
    return zoncmb_host(c);
}

extern "C" int galhip_set_graph(galhip_ctx *c, int32_t on) {
    if (!c) return gh_fail(GALHIP_EINVAL, "null context");
    if (c->p3d) return gh_fail(GALHIP_EINVAL, "graphs are built for p2d grids and ensembles");
    HIPCHK(hipStreamSynchronize(c->s));
    c->graph = on != 0;
    if (!c->graph) drop_graphs(c);
    return 0;
}

extern "C" int galhip_run(galhip_ctx *c, int32_t nsteps) {
    if (!c) return gh_fail(GALHIP_EINVAL, "null context");
    if (c->p3d) {
        for (int k = 0; k < nsteps; k++) {
            set_mzone(c);
            int rc = p3_solve(c, 0, 0, 0, 0);
            if (rc) return rc;
            rc = galhip_step(c);
            if (rc) return rc;
            if (c->sort_interval > 0 && ++c->steps_since_sort >= c->sort_interval) {
                rc = p3_sort(c);
                if (rc) return rc;
            }
        }
        return 0;
    }
    for (int k = 0; k < nsteps; k++) {
        set_mzone(c);
        int rc;
        if (c->graph && !c->prof) {
            rc = c->prec == GALHIP_FP32 ? graph_step<float>(c) : graph_step<double>(c);
            if (rc) return rc;
        } else {
            // findf with the masscl of this step (zone sums saved and cleared,
            // jrad/krad reset) fused into the synthesis kernel
            rc = c->prec == GALHIP_FP32 ? solve_impl<float>(c, 0, 0, 0, 0, c->mzone)
                                        : solve_impl<double>(c, 0, 0, 0, 0, c->mzone);
            if (rc) return rc;
            c->masscl_done = 1;
            rc = galhip_step(c);
            if (rc) return rc;
        }
        if (c->sort_interval > 0 && ++c->steps_since_sort >= c->sort_interval) {
            rc = galhip_sort(c);
            if (rc) return rc;
        }
    }
    return 0;
}

// ---- group-level parity entry points ------------------------------------------------

template <typename R>
static int gather_impl(galhip_ctx *c, int64_t n, const void *x, const void *y,
                       int fold, int32_t *cell, void *wt, void *acc) {
    if (n <= 0) return 0;
    size_t es = sizeof(R);
    R *dx, *dy, *dwt = nullptr, *dacc = nullptr;
    int *dcell = nullptr;
    int rc = dalloc(&dx, es * n) | dalloc(&dy, es * n);
    if (cell) rc |= dalloc(&dcell, sizeof(int) * n);
    if (wt) rc |= dalloc(&dwt, 4 * es * n);
    if (acc) rc |= dalloc(&dacc, 2 * es * n);
    if (!rc) {
        (void)hipMemcpy(dx, x, es * n, hipMemcpyHostToDevice);
        (void)hipMemcpy(dy, y, es * n, hipMemcpyHostToDevice);
        GatherArgs<R> a;
        a.x = dx; a.y = dy; a.n = n;
        a.fld = (const typename Vec2<R>::T *)c->fld2;
        a.acc = dacc; a.cell = dcell; a.wt = dwt;
        a.nr = c->nr; a.fold = fold && c->nsect > 1;
        a.k = ut_const(c);
        a.thmax = (R)c->thmax;
        hipError_t e = launch_gather<R>(a, c->s);
        if (e == hipSuccess) e = hipStreamSynchronize(c->s);
        if (e != hipSuccess) rc = gh_fail(GALHIP_EHIP, "gather kernel: %s", hipGetErrorString(e));
        if (!rc && cell) (void)hipMemcpy(cell, dcell, sizeof(int) * n, hipMemcpyDeviceToHost);
        if (!rc && wt) (void)hipMemcpy(wt, dwt, 4 * es * n, hipMemcpyDeviceToHost);
        if (!rc && acc) (void)hipMemcpy(acc, dacc, 2 * es * n, hipMemcpyDeviceToHost);
    }
    dfree(dx); dfree(dy); dfree(dcell); dfree(dwt); dfree(dacc);
    return rc;
}

extern "C" int galhip_weight(galhip_ctx *c, int64_t n, const void *x, const void *y,
                             int32_t old, int32_t *cell, void *wt) {
    if (c && c->p3d) return gh_fail(GALHIP_EINVAL, "p3d context: use the p3d entry point");
    if (!c || n < 0) return gh_fail(GALHIP_EINVAL, "bad arguments");
    HIPCHK(hipStreamSynchronize(c->s));
    return c->prec == GALHIP_FP32 ? gather_impl<float>(c, n, x, y, !old, cell, wt, nullptr)
                                  : gather_impl<double>(c, n, x, y, !old, cell, wt, nullptr);
}

extern "C" int galhip_grdacc(galhip_ctx *c, int64_t n, const void *x, const void *y,
                             void *acc) {
    if (c && c->p3d) return gh_fail(GALHIP_EINVAL, "p3d context: use the p3d entry point");
    if (c && c->ens) return gh_fail(GALHIP_EINVAL, "ensemble context: no single field");
    if (!c || n < 0 || !acc) return gh_fail(GALHIP_EINVAL, "bad arguments");
    HIPCHK(hipStreamSynchronize(c->s));
    return c->prec == GALHIP_FP32 ? gather_impl<float>(c, n, x, y, 0, nullptr, nullptr, acc)
                                  : gather_impl<double>(c, n, x, y, 0, nullptr, nullptr, acc);
}

// ---- group-level routines: getacc(jst), stpgrp(jst) + rezgrp(jst), massgn(jst) ----
//
// For parity work on groups of particles given by the caller: the group is
// copied to device buffers (positions and velocities interleaved as the
// context stores them, zones as bytes), run through the same device code as
// the step, and copied back.  They act on the context's field, zone
// schedule and zone mass sums as the reference's routines act on its commons.

enum { GRP_GETACC, GRP_STPGRP, GRP_MASSGN };

template <typename R>
static int group_impl(galhip_ctx *c, int kind, int64_t n, void *x, void *y, void *vx, void *vy,
                      void *acc, int32_t *iz) {
    typedef typename Vec4<R>::T R4;
    typedef typename Vec2<R>::T R2;
    if (n <= 0) return 0;
    for (int64_t i = 0; i < n; i++)
        if (iz[i] < 1 || iz[i] > c->nz + 1)
            return gh_fail(GALHIP_EINVAL, "group particle %lld: zone %d out of 1..%d",
                           (long long)i, iz[i], c->nz + 1);
    std::vector<R4> hx((size_t)n);
    std::vector<uint8_t> hz((size_t)n);
    const R *px = (const R *)x, *py = (const R *)y, *pvx = (const R *)vx, *pvy = (const R *)vy;
    for (int64_t i = 0; i < n; i++) {
        hx[i].x = px[i];
        hx[i].y = py[i];
        hx[i].z = pvx ? pvx[i] : (R)0;
        hx[i].w = pvy ? pvy[i] : (R)0;
        hz[i] = (uint8_t)iz[i];
    }
    std::vector<R2> ha;
    if (kind == GRP_STPGRP) {
        ha.resize((size_t)n);
        const R *pa = (const R *)acc;
        for (int64_t i = 0; i < n; i++) {
            ha[i].x = pa[2 * i];
            ha[i].y = pa[2 * i + 1];
        }
    }
    R4 *dxv = nullptr;
    R2 *dut = nullptr, *dacc = nullptr;
    uint8_t *dz = nullptr;
    int rc = dalloc(&dxv, sizeof(R4) * n) | dalloc(&dz, (size_t)n) |
             dalloc(&dut, sizeof(R2) * n) | dalloc(&dacc, sizeof(R2) * n);
    if (!rc) {
        hipError_t e = hipMemcpy(dxv, hx.data(), sizeof(R4) * n, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(dz, hz.data(), (size_t)n, hipMemcpyHostToDevice);
        if (e == hipSuccess && kind == GRP_STPGRP)
            e = hipMemcpy(dacc, ha.data(), sizeof(R2) * n, hipMemcpyHostToDevice);
        PushArgs<R> a = push_args<R>(c);
        a.xv = dxv;
        a.ut = dut;
        a.zone = dz;
        a.n = n;
        a.xlo = nullptr;
        if (e == hipSuccess) {
            if (kind == GRP_GETACC) e = launch_getacc_group<R>(a, dacc, c->s);
            else if (kind == GRP_STPGRP) e = launch_stpgrp_group<R>(a, dacc, c->s);
            else e = launch_massgn<R>(a, c->s);
        }
        if (e == hipSuccess) e = hipStreamSynchronize(c->s);
        if (e == hipSuccess && kind == GRP_GETACC) {
            ha.resize((size_t)n);
            e = hipMemcpy(ha.data(), dacc, sizeof(R2) * n, hipMemcpyDeviceToHost);
            R *pa = (R *)acc;
            for (int64_t i = 0; e == hipSuccess && i < n; i++) {
                pa[2 * i] = ha[i].x;
                pa[2 * i + 1] = ha[i].y;
            }
        }
        if (e == hipSuccess && kind == GRP_STPGRP) {
            e = hipMemcpy(hx.data(), dxv, sizeof(R4) * n, hipMemcpyDeviceToHost);
            if (e == hipSuccess) e = hipMemcpy(hz.data(), dz, (size_t)n, hipMemcpyDeviceToHost);
            R *wx = (R *)x, *wy = (R *)y, *wvx = (R *)vx, *wvy = (R *)vy;
            for (int64_t i = 0; e == hipSuccess && i < n; i++) {
                wx[i] = hx[i].x;
                wy[i] = hx[i].y;
                wvx[i] = hx[i].z;
                wvy[i] = hx[i].w;
                iz[i] = hz[i];
            }
        }
        if (e != hipSuccess) rc = gh_fail(GALHIP_EHIP, "group kernel: %s", hipGetErrorString(e));
    }
    dfree(dxv);
    dfree(dz);
    dfree(dut);
    dfree(dacc);
    return rc;
}

static int group_check(galhip_ctx *c, int64_t n, const void *x, const void *y,
                       const int32_t *iz) {
    if (!c) return gh_fail(GALHIP_EINVAL, "null context");
    if (c->p3d || c->ens) return gh_fail(GALHIP_EINVAL, "group routines: one p2d grid");
    if (n < 0 || (n > 0 && (!x || !y || !iz))) return gh_fail(GALHIP_EINVAL, "bad arguments");
    return 0;
}

extern "C" int galhip_getacc(galhip_ctx *c, int64_t n, const void *x, const void *y,
                             const int32_t *iz, void *acc) {
    int rc = group_check(c, n, x, y, iz);
    if (rc) return rc;
    if (n > 0 && !acc) return gh_fail(GALHIP_EINVAL, "bad arguments");
    if (!c->have_field) return gh_fail(GALHIP_ERANGE, "GETACC: Accelerations at useless time");
    HIPCHK(hipStreamSynchronize(c->s));
    int32_t *zz = const_cast<int32_t *>(iz);   // read only for GETACC
    return c->prec == GALHIP_FP32
               ? group_impl<float>(c, GRP_GETACC, n, (void *)x, (void *)y, nullptr, nullptr, acc, zz)
               : group_impl<double>(c, GRP_GETACC, n, (void *)x, (void *)y, nullptr, nullptr, acc,
                                    zz);
}

extern "C" int galhip_stpgrp(galhip_ctx *c, int64_t n, void *x, void *y, void *vx, void *vy,
                             const void *acc, int32_t *iz) {
    int rc = group_check(c, n, x, y, iz);
    if (rc) return rc;
    if (n > 0 && (!vx || !vy || !acc)) return gh_fail(GALHIP_EINVAL, "bad arguments");
    HIPCHK(hipStreamSynchronize(c->s));
    set_mzone(c);
    return c->prec == GALHIP_FP32
               ? group_impl<float>(c, GRP_STPGRP, n, x, y, vx, vy, (void *)acc, iz)
               : group_impl<double>(c, GRP_STPGRP, n, x, y, vx, vy, (void *)acc, iz);
}

extern "C" int galhip_massgn(galhip_ctx *c, int64_t n, const void *x, const void *y,
                             const int32_t *iz) {
    int rc = group_check(c, n, x, y, iz);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(c->s));
    int32_t *zz = const_cast<int32_t *>(iz);
    return c->prec == GALHIP_FP32
               ? group_impl<float>(c, GRP_MASSGN, n, (void *)x, (void *)y, nullptr, nullptr,
                                   nullptr, zz)
               : group_impl<double>(c, GRP_MASSGN, n, (void *)x, (void *)y, nullptr, nullptr,
                                    nullptr, zz);
}

// grid m's zone masses, ring ranges and fixed-point scale
static int get_mass_m(const galhip_ctx *c, int m, int64_t *mfx, int64_t *mozfx, int32_t *jrad,
                      int32_t *krad, int32_t *qbits) {
    HIPCHK(hipStreamSynchronize(c->s));
    const size_t mo = (size_t)m * c->mesh;
    for (int z = 0; z < c->nz; z++) {
        if (mfx) HIPCHK(hipMemcpy(mfx + (size_t)z * c->mesh, c->mfx[z] + mo, 8 * c->mesh, hipMemcpyDeviceToHost));
        if (mozfx) HIPCHK(hipMemcpy(mozfx + (size_t)z * c->mesh, c->moz[z] + mo, 8 * c->mesh, hipMemcpyDeviceToHost));
    }
    if (jrad) HIPCHK(hipMemcpy(jrad, c->jrad + m * GH_MAXZ, sizeof(int) * c->nz, hipMemcpyDeviceToHost));
    if (krad) HIPCHK(hipMemcpy(krad, c->krad + m * GH_MAXZ, sizeof(int) * c->nz, hipMemcpyDeviceToHost));
    if (qbits) *qbits = c->ens ? c->m_qbits[m] : c->qbits;
    return 0;
}

extern "C" int galhip_get_mass(const galhip_ctx *c, int64_t *mfx, int64_t *mozfx,
                               int32_t *jrad, int32_t *krad, int32_t *qbits) {
    if (!c) return gh_fail(GALHIP_EINVAL, "null context");
    if (c->ens) return gh_fail(GALHIP_EINVAL, "ensemble context: use the galhip_ens_* calls");
    return get_mass_m(c, 0, mfx, mozfx, jrad, krad, qbits);
}

extern "C" int galhip_set_mass(galhip_ctx *c, const int64_t *mfx,
                               const int64_t *mozfx, const int32_t *jrad,
                               const int32_t *krad) {
    if (!c) return gh_fail(GALHIP_EINVAL, "null context");
    if (c->ens) return gh_fail(GALHIP_EINVAL, "ensemble context: use the galhip_ens_* calls");
    HIPCHK(hipStreamSynchronize(c->s));
    for (int z = 0; z < c->nz; z++) {
        if (mfx) HIPCHK(hipMemcpy(c->mfx[z], mfx + (size_t)z * c->mesh, 8 * c->mesh, hipMemcpyHostToDevice));
        if (mozfx) HIPCHK(hipMemcpy(c->moz[z], mozfx + (size_t)z * c->mesh, 8 * c->mesh, hipMemcpyHostToDevice));
    }
    if (jrad) HIPCHK(hipMemcpy(c->jrad, jrad, sizeof(int) * c->nz, hipMemcpyHostToDevice));
    if (krad) HIPCHK(hipMemcpy(c->krad, krad, sizeof(int) * c->nz, hipMemcpyHostToDevice));
    return 0;
}

extern "C" int galhip_get_rho(const galhip_ctx *c, void *rho) {
    if (!c || !rho) return gh_fail(GALHIP_EINVAL, "null argument");
    if (c->ens) return gh_fail(GALHIP_EINVAL, "ensemble context: use the galhip_ens_* calls");
    HIPCHK(hipStreamSynchronize(c->s));
    HIPCHK(hipMemcpy(rho, c->rho, esize(c) * c->mesh, hipMemcpyDeviceToHost));
    return 0;
}

// grid g's field as [3][mesh] (F_x, F_y, potential)
static int get_field_m(const galhip_ctx *c, int g, void *fld) {
    HIPCHK(hipStreamSynchronize(c->s));
    size_t es = esize(c), m = c->mesh;
    std::vector<unsigned char> tmp(2 * es * m);
    HIPCHK(hipMemcpy(tmp.data(), (const unsigned char *)c->fld2 + 2 * es * m * g, 2 * es * m,
                     hipMemcpyDeviceToHost));
    unsigned char *o = (unsigned char *)fld;
    for (size_t i = 0; i < m; i++) {
        memcpy(o + i * es, tmp.data() + 2 * i * es, es);
        memcpy(o + (m + i) * es, tmp.data() + (2 * i + 1) * es, es);
    }
    HIPCHK(hipMemcpy(o + 2 * m * es, (const unsigned char *)c->pot + es * m * g, es * m,
                     hipMemcpyDeviceToHost));
    return 0;
}

static int set_field_m(galhip_ctx *c, int g, const void *fld) {
    HIPCHK(hipStreamSynchronize(c->s));
    size_t es = esize(c), m = c->mesh;
    std::vector<unsigned char> tmp(2 * es * m);
    const unsigned char *in = (const unsigned char *)fld;
    for (size_t i = 0; i < m; i++) {
        memcpy(tmp.data() + 2 * i * es, in + i * es, es);
        memcpy(tmp.data() + (2 * i + 1) * es, in + (m + i) * es, es);
    }
    HIPCHK(hipMemcpy((unsigned char *)c->fld2 + 2 * es * m * g, tmp.data(), 2 * es * m,
                     hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy((unsigned char *)c->pot + es * m * g, in + 2 * m * es, es * m,
                     hipMemcpyHostToDevice));
    return 0;
}

extern "C" int galhip_get_field(const galhip_ctx *c, void *fld) {
    if (!c || !fld) return gh_fail(GALHIP_EINVAL, "null argument");
    if (c->p3d) return p3_get_field(c, fld);
    if (c->ens) return gh_fail(GALHIP_EINVAL, "ensemble context: use the galhip_ens_* calls");
    return get_field_m(c, 0, fld);
}

extern "C" int galhip_set_field(galhip_ctx *c, const void *fld) {
    if (!c || !fld) return gh_fail(GALHIP_EINVAL, "null argument");
    if (c->p3d) return p3_set_field(c, fld);
    if (c->ens) return gh_fail(GALHIP_EINVAL, "ensemble context: use the galhip_ens_* calls");
    int rc = set_field_m(c, 0, fld);
    if (rc) return rc;
    c->have_field = 1;
    return 0;
}

extern "C" int galhip_get_ranges(const galhip_ctx *c, int32_t *r4) {
    if (!c || !r4) return gh_fail(GALHIP_EINVAL, "null argument");
    if (c->ens) return gh_fail(GALHIP_EINVAL, "ensemble context: use the galhip_ens_* calls");
    HIPCHK(hipStreamSynchronize(c->s));
    HIPCHK(hipMemcpy(r4, c->ranges, sizeof(int) * 4, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int galhip_get_noff(const galhip_ctx *c, int64_t *noff) {
    if (!c || !noff) return gh_fail(GALHIP_EINVAL, "null argument");
    if (c->ens) return gh_fail(GALHIP_EINVAL, "ensemble context: use the galhip_ens_* calls");
    HIPCHK(hipStreamSynchronize(c->s));
    long long v[4];
    HIPCHK(hipMemcpy(v, c->counters, sizeof v, hipMemcpyDeviceToHost));
    *noff = v[0];
    return 0;
}

// ---- profiling ----------------------------------------------------------------------

// pushes counted so far over every grid of the context
static int pushes_total(const galhip_ctx *c, long long *out) {
    const int nm = c->nmem > 0 ? c->nmem : 1;
    std::vector<long long> v(4 * (size_t)nm);
    HIPCHK(hipMemcpy(v.data(), c->counters, sizeof(long long) * v.size(), hipMemcpyDeviceToHost));
    long long t = 0;
    for (int m = 0; m < nm; m++) t += v[4 * m + 1];
    *out = t;
    return 0;
}

extern "C" int galhip_set_profile(galhip_ctx *c, int32_t on) {
    if (!c) return gh_fail(GALHIP_EINVAL, "null context");
    HIPCHK(hipStreamSynchronize(c->s));
    for (auto e : c->ev_used) c->ev_free.push_back(e);
    c->ev_used.clear();
    c->prof = on != 0;
    c->prof_ms = 0.0;
    c->prof_launch = 0;
    return pushes_total(c, &c->prof_push0);
}

extern "C" int galhip_get_profile(galhip_ctx *c, double *push_ms, int64_t *launches,
                                  int64_t *pushes) {
    if (!c) return gh_fail(GALHIP_EINVAL, "null context");
    HIPCHK(hipStreamSynchronize(c->s));
    double ms = 0.0;
    for (size_t i = 0; i + 1 < c->ev_used.size(); i += 2) {
        float t = 0.f;
        HIPCHK(hipEventElapsedTime(&t, c->ev_used[i], c->ev_used[i + 1]));
        ms += t;
    }
    long long tot = 0;
    int rc = pushes_total(c, &tot);
    if (rc) return rc;
    if (push_ms) *push_ms = ms;
    if (launches) *launches = c->prof_launch;
    if (pushes) *pushes = tot - c->prof_push0;
    return 0;
}

extern "C" int galhip_sync(galhip_ctx *c) {
    if (!c) return gh_fail(GALHIP_EINVAL, "null context");
    HIPCHK(hipStreamSynchronize(c->s));
    return 0;
}

// ---- ensembles of p2d discs -------------------------------------------------------

extern "C" int galhip_ens_restrt(galhip_ctx *c, const int64_t *count, const void *x,
                                 const void *y, const void *vx, const void *vy,
                                 const int32_t *iz) {
    if (!c || !count) return gh_fail(GALHIP_EINVAL, "galhip_ens_restrt: bad arguments");
    if (!c->ens) return gh_fail(GALHIP_EINVAL, "not an ensemble context");
    long long n = 0;
    for (int m = 0; m < c->nmem; m++) {
        if (count[m] < 0) return gh_fail(GALHIP_EINVAL, "member %d: negative count", m);
        n += count[m];
    }
    if (n > 0 && (!x || !y || !vx || !vy || !iz))
        return gh_fail(GALHIP_EINVAL, "galhip_ens_restrt: null particle array");
    if (n >= (1LL << 31)) return gh_fail(GALHIP_EINVAL, "at most 2^31-1 particles in an ensemble");
    HIPCHK(hipStreamSynchronize(c->s));
    c->tcent = 0;
    int rc = ensure_particles(c, n);
    if (rc) return rc;
    // member ranges, push blocks of at most GH_PUSH_PER particles that never
    // straddle two members, fixed-point scales as a lone disc of that size
    // would have them
    const int per = GH_PUSH_PER;
    std::vector<int4> blk;
    c->m_off.assign(c->nmem + 1, 0);
    for (int m = 0; m < c->nmem; m++) {
        c->m_off[m + 1] = c->m_off[m] + count[m];
        for (long long s0 = c->m_off[m]; s0 < c->m_off[m + 1]; s0 += per)
            blk.push_back(make_int4(m, (int)s0, (int)std::min<long long>(per, c->m_off[m + 1] - s0), 0));
        c->m_qbits[m] = qbits_for(count[m]);
    }
    dfree(c->blk_d);
    rc = dalloc(&c->blk_d, sizeof(int4) * std::max<size_t>(blk.size(), 1));
    if (rc) return rc;
    c->nblk = (int)blk.size();
    if (!blk.empty())
        HIPCHK(hipMemcpy(c->blk_d, blk.data(), sizeof(int4) * blk.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->m_qbits_d, c->m_qbits.data(), sizeof(int) * c->nmem, hipMemcpyHostToDevice));
    size_t es = esize(c);
    std::vector<uint8_t> z(n);
    std::vector<long long> cnt(4 * (size_t)c->nmem, 0);   // per member: [0] noff
    for (int m = 0; m < c->nmem; m++)
        for (long long i = c->m_off[m]; i < c->m_off[m + 1]; i++) {
            int v = iz[i];
            if (v < 1 || v > c->nz + 1)
                return gh_fail(GALHIP_EINVAL, "particle %lld has zone %d outside 1..%d", i, v, c->nz + 1);
            z[i] = (uint8_t)v;
            if (v == c->nz + 1) cnt[4 * m]++;
        }
    HIPCHK(hipMemcpy(c->counters, cnt.data(), sizeof(long long) * cnt.size(), hipMemcpyHostToDevice));
    drop_graphs(c);
    if (n == 0) return 0;
    rc = load_xv(c, n, x, y, vx, vy);
    if (rc) return rc;
    HIPCHK(hipMemcpy(c->zone, z.data(), n, hipMemcpyHostToDevice));
    if (c->xlo) HIPCHK(hipMemset(c->xlo, 0, 2 * es * n));
    rc = compute_ut(c);
    if (rc) return rc;
    rc = galhip_sort(c);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(c->s));
    return 0;
}

extern "C" int galhip_ens_unload(galhip_ctx *c, void *x, void *y, void *vx, void *vy,
                                 int32_t *iz) {
    if (!c) return gh_fail(GALHIP_EINVAL, "null context");
    if (!c->ens) return gh_fail(GALHIP_EINVAL, "not an ensemble context");
    return unload_all(c, x, y, vx, vy, iz);
}

static int ens_member_ok(const galhip_ctx *c, int32_t m) {
    if (!c) return gh_fail(GALHIP_EINVAL, "null context");
    if (!c->ens) return gh_fail(GALHIP_EINVAL, "not an ensemble context");
    if (m < 0 || m >= c->nmem) return gh_fail(GALHIP_EINVAL, "member %d outside 0..%d", m, c->nmem - 1);
    return 0;
}

extern "C" int galhip_ens_get_mass(const galhip_ctx *c, int32_t member, int64_t *mfx,
                                   int64_t *mozfx, int32_t *jrad, int32_t *krad,
                                   int32_t *qbits) {
    int rc = ens_member_ok(c, member);
    return rc ? rc : get_mass_m(c, member, mfx, mozfx, jrad, krad, qbits);
}

extern "C" int galhip_ens_get_field(const galhip_ctx *c, int32_t member, void *fld) {
    int rc = ens_member_ok(c, member);
    if (rc) return rc;
    if (!fld) return gh_fail(GALHIP_EINVAL, "null argument");
    return get_field_m(c, member, fld);
}

extern "C" int galhip_ens_set_field(galhip_ctx *c, int32_t member, const void *fld) {
    int rc = ens_member_ok(c, member);
    if (rc) return rc;
    if (!fld) return gh_fail(GALHIP_EINVAL, "null argument");
    rc = set_field_m(c, member, fld);
    if (rc) return rc;
    c->have_field = 1;
    return 0;
}

extern "C" int galhip_ens_get_ranges(const galhip_ctx *c, int32_t member, int32_t *r4) {
    int rc = ens_member_ok(c, member);
    if (rc) return rc;
    if (!r4) return gh_fail(GALHIP_EINVAL, "null argument");
    HIPCHK(hipStreamSynchronize(c->s));
    HIPCHK(hipMemcpy(r4, c->ranges + 4 * member, sizeof(int) * 4, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int galhip_ens_get_noff(const galhip_ctx *c, int64_t *noff) {
    if (!c || !noff) return gh_fail(GALHIP_EINVAL, "null argument");
    if (!c->ens) return gh_fail(GALHIP_EINVAL, "not an ensemble context");
    HIPCHK(hipStreamSynchronize(c->s));
    std::vector<long long> v(4 * (size_t)c->nmem);
    HIPCHK(hipMemcpy(v.data(), c->counters, sizeof(long long) * v.size(), hipMemcpyDeviceToHost));
    for (int m = 0; m < c->nmem; m++) noff[m] = v[4 * m];
    return 0;
}

extern "C" int galhip_ens_integrals(galhip_ctx *c, int32_t centred, double *out) {
    if (!c || !out) return gh_fail(GALHIP_EINVAL, "null argument");
    if (!c->ens) return gh_fail(GALHIP_EINVAL, "not an ensemble context");
    if (centred && !c->have_field)
        return gh_fail(GALHIP_ERANGE, "GETACC: Accelerations at useless time");
    return c->prec == GALHIP_FP32 ? integrals_impl<float>(c, centred != 0, out, nullptr)
                                  : integrals_impl<double>(c, centred != 0, out, nullptr);
}

// ---- compensated positions ----------------------------------------------------------

extern "C" int galhip_set_compensated(galhip_ctx *c, int32_t on) {
    if (!c) return gh_fail(GALHIP_EINVAL, "null context");
    if (on && (c->p3d || c->prec != GALHIP_FP32))
        return gh_fail(GALHIP_EINVAL, "compensated positions are built for FP32 p2d grids");
    HIPCHK(hipStreamSynchronize(c->s));
    if (on && !c->xlo && c->cap > 0) {
        const size_t b = 2 * sizeof(float) * (size_t)c->cap;
        int rc = dalloc(&c->xlo, b) | dalloc(&c->xlo2, b);
        if (rc) return rc;
        HIPCHK(hipMemset(c->xlo, 0, b));
    }
    if (!on) {
        dfree(c->xlo);
        dfree(c->xlo2);
    }
    c->comp = on != 0;
    drop_graphs(c);
    return 0;
}

extern "C" int galhip_get_position_lo(const galhip_ctx *c, void *xlo, void *ylo) {
    if (!c || !xlo || !ylo) return gh_fail(GALHIP_EINVAL, "null argument");
    if (c->p3d) return gh_fail(GALHIP_EINVAL, "p2d grids only");
    HIPCHK(hipStreamSynchronize(c->s));
    const long long n = c->n;
    const size_t es = esize(c);
    if (n <= 0) return 0;
    if (!c->xlo) {
        memset(xlo, 0, es * n);
        memset(ylo, 0, es * n);
        return 0;
    }
    std::vector<unsigned> ids(n);
    std::vector<float> lo(2 * n);
    HIPCHK(hipMemcpy(ids.data(), c->id, sizeof(unsigned) * n, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(lo.data(), c->xlo, sizeof(float) * 2 * n, hipMemcpyDeviceToHost));
    float *xo = (float *)xlo, *yo = (float *)ylo;
    for (long long i = 0; i < n; i++) {
        xo[ids[i]] = lo[2 * i];
        yo[ids[i]] = lo[2 * i + 1];
    }
    return 0;
}
