// galhip_p3d_api.hip -- host side of the 3-D polar grid (p3d): context
// set-up (p3dset, setgrd, filset), particle load/unload, and the launch
// sequences behind the generic entry points (greenm, masset, findf, step,
// tmcent) when the context holds a p3d grid.  The zone and time bookkeeping
// is the p2d host code's (masscl, zoncmb, mzone), which the reference shares
// between the two grids (SRC/lib15/masscl.f, zoncmb.f, step.f).
#include "galhip_ctx.h"
#include "galhip_p3d.h"

#include <rocprim/device/device_radix_sort.hpp>

#include <cmath>
#include <cstring>
#include <vector>

using namespace gh;

// ---- set-up ------------------------------------------------------------------

extern "C" int galhip_p3dset(const galhip_p3d_params *p, galhip_ctx **out) {
    if (!p || !out) return gh_fail(GALHIP_EINVAL, "galhip_p3dset: null argument");
    *out = nullptr;
    if (p->precision != GALHIP_FP32 && p->precision != GALHIP_FP64)
        return gh_fail(GALHIP_EINVAL, "precision must be GALHIP_FP32 or GALHIP_FP64");
    const int na = 2 * (p->na / 2);                     // setgrd.f:58
    if (p->nr < 3 || p->nr > 1024 || na < 4 || na > 2048 || p->ngz < 3)
        return gh_fail(GALHIP_EINVAL, "grid %d x %d x %d out of range", p->nr, p->na, p->ngz);
    if (p->nzones < 1 || p->nzones > GALHIP_MAXZONES)
        return gh_fail(GALHIP_EINVAL, "nzones %d out of range", p->nzones);
    if (p->nsect < 1) return gh_fail(GALHIP_EINVAL, "nsect not set");
    if (!(p->dzg > 0.0)) return gh_fail(GALHIP_EINVAL, "z spacing must be positive");
    if (!(p->softl > 0.0))
        return gh_fail(GALHIP_EINVAL, "P3DSET: Softening length must be greater than 0");
    if (p->softl < p->dzg)   // p3dset.f:99-100
        return gh_fail(GALHIP_ERANGE,
                       "P3DSET: Softening length cannot be less than z-spacing of grid planes");
    if (p->tsoft != 1 && p->tsoft != 2) return gh_fail(GALHIP_EINVAL, "Unrecognized softening rule");
    if (p->nskip < 0 || p->nskip > GALHIP_MAXSKIP) return gh_fail(GALHIP_EINVAL, "nskip out of range");
    const int jmass = p->jmass == 0 ? 2 : p->jmass;
    if (jmass != 2 && jmass != 3)   // p3dset.f:80-83
        return gh_fail(GALHIP_EINVAL, "P3DSET: Interpolation scheme must be 2 or 3");
    galhip_ctx *c = new galhip_ctx();
    c->p3d = 1;
    c->nmem = 1;
    c->jmass = jmass;
    c->prec = p->precision;
    c->nr = p->nr;
    c->na = na;
    c->ngz = p->ngz;
    c->mmax = na / 2 + 1;
    c->ng = p->hash + 1 < c->mmax ? p->hash + 1 : c->mmax;
    if (c->ng < 1) c->ng = 1;
    if (c->ng > 64) { delete c; return gh_fail(GALHIP_EINVAL, "at most 64 harmonics supported"); }
    c->nsect = p->nsect;
    c->mesh = c->nr * c->ngz * na;
    c->tsoft = p->tsoft;
    c->nz = p->nzones;
    c->offanal = p->offanal != 0;
    c->softl = p->softl;
    c->pmass = p->pmass;
    for (int i = 0; i < c->nz; i++) {
        c->nstep[i] = p->nstep[i] > 0 ? p->nstep[i] : 1;
        c->zbr[i] = p->zbr[i];
    }
    // filter (filset.f:118-141)
    c->lg.assign(c->mmax, 1);
    c->lg[0] = 0;
    for (int m = 1; m < c->ng; m++) c->lg[m] = (c->nsect == 1 || m % c->nsect == 0) ? 0 : 1;
    for (int i = 0; i < p->nskip; i++)
        if (p->skip[i] >= 0 && p->skip[i] < c->mmax) c->lg[p->skip[i]] = 1;
    // compact component list of p3manl (p3manl.f:47-58): one slot for m = 0
    // and m = na/2, a (cos, sin) pair for the others
    std::vector<int> slot_m, slot_p, slot_of(c->mmax, -1);
    for (int m = 0; m < c->ng; m++) {
        if (c->lg[m]) continue;
        slot_of[m] = (int)slot_m.size();
        slot_m.push_back(m);
        slot_p.push_back(0);
        if (m > 0 && m < c->mmax - 1) {
            slot_m.push_back(m);
            slot_p.push_back(1);
        }
    }
    c->nm3 = (int)slot_m.size();
    const double two_pi = 2.0 * 3.141592653589793238462643;
    double aw;
    if (c->prec == GALHIP_FP32) {
        const float a = (float)(two_pi / (double)na);
        c->alpha = a;
        // rgrid = grofu(nr - jmass/2), zm = dzg (ngz + 1 - jmass)/2
        // (setgrd.f:135-136,146-150)
        c->rgrid = expf(a * ((float)c->nr - 0.5f * (float)jmass)) - 1.0f;
        c->thmax = (float)na / (float)c->nsect;
        const float dz = (float)p->dzg;
        c->dzg = dz;
        c->zm = (0.5f * dz) * (float)(c->ngz + 1 - jmass);
        aw = (double)((float)c->rgrid - (expf(a * (float)(c->nr - 2)) - 1.0f));
    } else {
        c->alpha = two_pi / (double)na;
        c->rgrid = exp(c->alpha * ((double)c->nr - 0.5 * (double)jmass)) - 1.0;
        c->thmax = (double)na / (double)c->nsect;
        c->dzg = p->dzg;
        c->zm = 0.5 * p->dzg * (double)(c->ngz + 1 - jmass);
        aw = c->rgrid - (exp(c->alpha * (double)(c->nr - 2)) - 1.0);
    }
    if (aw / c->softl > 5.0) {   // setgrd.f:181-199
        delete c;
        return gh_fail(GALHIP_ERANGE, "SETGRD: Unsuitable choice of grid parameters "
                       "(outermost cell %.4g times the softening length)", aw / p->softl);
    }
    c->istep = 0;
    c->mzone = c->nz;
    hipError_t e = hipStreamCreate(&c->s);
    if (e != hipSuccess) {
        delete c;
        return gh_fail(GALHIP_EHIP, "hipStreamCreate: %s", hipGetErrorString(e));
    }
    const size_t es = esize(c), P = 2 * (size_t)c->ngz, nm = c->nm3;
    {
        // the solve's launch limits
        P3SolveArgs<float> probe;
        memset(&probe, 0, sizeof probe);
        probe.na = na; probe.nm = c->nm3; probe.nr = c->nr; probe.ngz = c->ngz;
        const bool fits32 = p3_solve_fits(probe);
        P3SolveArgs<double> probe64;
        memset(&probe64, 0, sizeof probe64);
        probe64.na = na; probe64.nm = c->nm3; probe64.nr = c->nr; probe64.ngz = c->ngz;
        const bool fits = c->prec == GALHIP_FP32 ? fits32 : p3_solve_fits(probe64);
        if (!fits || sizeof(double) * (size_t)(c->ngz + 1) * c->ng > 48 * 1024) {
            galhip_free(c);
            return gh_fail(GALHIP_EINVAL, "p3d grid too large for the solve kernels "
                           "(%d components x %d planes)", c->nm3, c->ngz);
        }
    }
    int rc = 0;
    for (int z = 0; z < c->nz; z++) {
        rc |= dalloc(&c->mfx[z], sizeof(long long) * c->mesh);
        rc |= dalloc(&c->moz[z], sizeof(long long) * c->mesh);
    }
    rc |= dalloc(&c->jrad, sizeof(int) * GH_MAXZ);
    rc |= dalloc(&c->krad, sizeof(int) * GH_MAXZ);
    rc |= dalloc(&c->ranges, sizeof(int) * 4);
    rc |= dalloc(&c->lg_d, sizeof(int) * c->mmax);
    rc |= dalloc(&c->slot_m_d, sizeof(int) * (nm > 0 ? nm : 1));
    rc |= dalloc(&c->slot_p_d, sizeof(int) * (nm > 0 ? nm : 1));
    rc |= dalloc(&c->slot_of_d, sizeof(int) * c->mmax);
    rc |= dalloc(&c->counters, sizeof(long long) * 4);
    rc |= dalloc(&c->rho, es * c->mesh);
    rc |= dalloc(&c->A3, es * c->nr * nm * c->ngz);
    rc |= dalloc(&c->M3, es * c->nr * nm * P);
    rc |= dalloc(&c->W3, es * 4 * c->nr * nm * P);
    rc |= dalloc(&c->Z3, es * 4 * c->nr * c->ngz * nm);
    rc |= dalloc(&c->fld4, 4 * es * c->mesh);
    rc |= dalloc(&c->tw, 2 * es * na);
    rc |= dalloc(&c->twd, 2 * sizeof(double) * na);
    rc |= dalloc(&c->twz, 2 * es * P);
    rc |= dalloc(&c->twzd, 2 * sizeof(double) * P);
    c->gsize3 = 4 * (size_t)(c->ngz + 1) * c->ng * c->nr * c->nr;
    rc |= dalloc(&c->Gd, sizeof(double) * c->gsize3);
    if (c->prec == GALHIP_FP32) rc |= dalloc(&c->G, sizeof(float) * c->gsize3);
    else c->G = c->Gd;
    if (rc) { galhip_free(c); return rc; }
    std::vector<double> twd(2 * na), twzd(2 * P);
    for (int k = 0; k < na; k++) {
        twd[2 * k] = cos(two_pi * (double)k / (double)na);
        twd[2 * k + 1] = sin(two_pi * (double)k / (double)na);
    }
    for (size_t k = 0; k < P; k++) {
        twzd[2 * k] = cos(two_pi * (double)k / (double)P);
        twzd[2 * k + 1] = sin(two_pi * (double)k / (double)P);
    }
    HIPCHK(hipMemcpy(c->twd, twd.data(), sizeof(double) * 2 * na, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->twzd, twzd.data(), sizeof(double) * 2 * P, hipMemcpyHostToDevice));
    if (c->prec == GALHIP_FP32) {
        std::vector<float> f(2 * na), fz(2 * P);
        for (int i = 0; i < 2 * na; i++) f[i] = (float)twd[i];
        for (size_t i = 0; i < 2 * P; i++) fz[i] = (float)twzd[i];
        HIPCHK(hipMemcpy(c->tw, f.data(), sizeof(float) * 2 * na, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(c->twz, fz.data(), sizeof(float) * 2 * P, hipMemcpyHostToDevice));
    } else {
        HIPCHK(hipMemcpy(c->tw, twd.data(), sizeof(double) * 2 * na, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(c->twz, twzd.data(), sizeof(double) * 2 * P, hipMemcpyHostToDevice));
    }
    HIPCHK(hipMemcpy(c->lg_d, c->lg.data(), sizeof(int) * c->mmax, hipMemcpyHostToDevice));
    if (nm > 0) {
        HIPCHK(hipMemcpy(c->slot_m_d, slot_m.data(), sizeof(int) * nm, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(c->slot_p_d, slot_p.data(), sizeof(int) * nm, hipMemcpyHostToDevice));
    }
    HIPCHK(hipMemcpy(c->slot_of_d, slot_of.data(), sizeof(int) * c->mmax, hipMemcpyHostToDevice));
    std::vector<int> jr(GH_MAXZ, c->nr), kr(GH_MAXZ, 0);
    HIPCHK(hipMemcpy(c->jrad, jr.data(), sizeof(int) * GH_MAXZ, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->krad, kr.data(), sizeof(int) * GH_MAXZ, hipMemcpyHostToDevice));
    HIPCHK(hipMemset(c->counters, 0, sizeof(long long) * 4));
    HIPCHK(hipMemset(c->fld4, 0, 4 * es * c->mesh));
    HIPCHK(hipMemset(c->rho, 0, es * c->mesh));
    // entries of W no contraction writes (e.g. F_z at k_z = 0) stay zero
    HIPCHK(hipMemset(c->W3, 0, es * 4 * c->nr * nm * P));
    HIPCHK(hipMemset(c->Z3, 0, es * 4 * c->nr * c->ngz * nm));
    for (int z = 0; z < c->nz; z++) {
        HIPCHK(hipMemset(c->mfx[z], 0, sizeof(long long) * c->mesh));
        HIPCHK(hipMemset(c->moz[z], 0, sizeof(long long) * c->mesh));
    }
    c->qbits = qbits_for(1);
    *out = c;
    return 0;
}

void p3_free(galhip_ctx *c) {
    dfree(c->A3); dfree(c->M3); dfree(c->W3); dfree(c->Z3); dfree(c->fld4);
    dfree(c->twz); dfree(c->twzd);
    dfree(c->slot_m_d); dfree(c->slot_p_d); dfree(c->slot_of_d);
}

// ---- Green's tables ---------------------------------------------------------------

static int p3_green_to_work(galhip_ctx *c) {
    if (c->prec == GALHIP_FP32) HIPCHK(launch_d2f(c->Gd, (float *)c->G, (long long)c->gsize3, c->s));
    c->have_green = 1;
    return 0;
}

int p3_greenm(galhip_ctx *c) {
    const double alpha_d = 2.0 * 3.141592653589793238462643 / (double)c->na;
    HIPCHK(hipMemsetAsync(c->Gd, 0, sizeof(double) * c->gsize3, c->s));
    HIPCHK(launch_p3_green(c->Gd, c->twd, c->twzd, c->lg_d, c->nr, c->na, c->ngz, c->ng,
                           alpha_d, c->prec == GALHIP_FP32 ? (double)(float)c->dzg : c->dzg,
                           c->softl, c->tsoft, c->s));
    int rc = p3_green_to_work(c);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(c->s));
    return 0;
}

int p3_get_green(const galhip_ctx *c, double *G) {
    HIPCHK(hipStreamSynchronize(c->s));
    HIPCHK(hipMemcpy(G, c->Gd, sizeof(double) * c->gsize3, hipMemcpyDeviceToHost));
    return 0;
}

int p3_set_green(galhip_ctx *c, const double *G) {
    HIPCHK(hipStreamSynchronize(c->s));
    HIPCHK(hipMemcpy(c->Gd, G, sizeof(double) * c->gsize3, hipMemcpyHostToDevice));
    int rc = p3_green_to_work(c);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(c->s));
    return 0;
}

// ---- particles ----------------------------------------------------------------------

template <typename R> static P3Args<R> p3_args(galhip_ctx *c);

static void p3_free_particles(galhip_ctx *c) {
    dfree(c->xv); dfree(c->ut); dfree(c->zone); dfree(c->id);
    dfree(c->xv2); dfree(c->ut2); dfree(c->zone2); dfree(c->id2);
    dfree(c->key); dfree(c->key2); dfree(c->perm); dfree(c->perm2); dfree(c->sort_tmp);
    c->sort_tmp_bytes = 0;
    c->n = c->cap = 0;
}

// Particle storage plus the sort's second copies, keys and scratch.
static int p3_ensure(galhip_ctx *c, long long n) {
    if (n <= c->cap) { c->n = n; return 0; }
    p3_free_particles(c);
    const size_t es = esize(c), m = n > 0 ? n : 1;
    // particle records, 8 values each (see P3Args), and the sort's second copy
    int rc = dalloc(&c->xv, 8 * es * m) | dalloc(&c->zone, m) |
             dalloc(&c->id, sizeof(unsigned) * m) | dalloc(&c->xv2, 8 * es * m) |
             dalloc(&c->zone2, m) |
             dalloc(&c->id2, sizeof(unsigned) * m) | dalloc(&c->key, sizeof(unsigned) * m) |
             dalloc(&c->key2, sizeof(unsigned) * m) | dalloc(&c->perm, sizeof(unsigned) * m) |
             dalloc(&c->perm2, sizeof(unsigned) * m);
    if (rc) { p3_free_particles(c); return rc; }
    size_t tmp = 0;
    hipError_t e = rocprim::radix_sort_pairs(nullptr, tmp, c->key, c->key2, c->perm,
                                                      c->perm2, (int)m, 0, 32, c->s);
    if (e != hipSuccess) {
        p3_free_particles(c);
        return gh_fail(GALHIP_EHIP, "radix sort sizing: %s", hipGetErrorString(e));
    }
    rc = dalloc(&c->sort_tmp, tmp);
    if (rc) { p3_free_particles(c); return rc; }
    c->sort_tmp_bytes = tmp;
    c->n = c->cap = n;
    return 0;
}

extern "C" int galhip_restrt3(galhip_ctx *c, int64_t n, const void *x, const void *y,
                              const void *z, const void *vx, const void *vy, const void *vz,
                              const int32_t *iz) {
    if (!c || n < 0 || (n > 0 && (!x || !y || !z || !vx || !vy || !vz || !iz)))
        return gh_fail(GALHIP_EINVAL, "galhip_restrt3: bad arguments");
    if (!c->p3d) return gh_fail(GALHIP_EINVAL, "galhip_restrt3: not a p3d context");
    if (n >= (1LL << 31)) return gh_fail(GALHIP_EINVAL, "at most 2^31-1 particles");
    HIPCHK(hipStreamSynchronize(c->s));
    c->tcent = 0;
    int rc = p3_ensure(c, n);
    if (rc) return rc;
    const size_t es = esize(c);
    std::vector<uint8_t> zz(n);
    long long cnt[4] = {0, 0, 0, 0};
    for (long long i = 0; i < n; i++) {
        const int v = iz[i];
        if (v < 1 || v > c->nz + 1)
            return gh_fail(GALHIP_EINVAL, "particle %lld has zone %d outside 1..%d", i, v, c->nz + 1);
        zz[i] = (uint8_t)v;
        if (v == c->nz + 1) cnt[0]++;
    }
    HIPCHK(hipMemcpy(c->counters, cnt, sizeof cnt, hipMemcpyHostToDevice));
    if (n > 0) {
        // the planar arrays through the second record buffer (free outside a
        // sort, 8 n values), interleaved on the device
        unsigned char *t4 = (unsigned char *)c->xv2, *t2 = t4 + 4 * es * n;
        const void *s4[4] = {x, y, z, vx};
        const void *s2[2] = {vy, vz};
        for (int k = 0; k < 4; k++)
            HIPCHK(hipMemcpy(t4 + k * es * n, s4[k], es * n, hipMemcpyHostToDevice));
        for (int k = 0; k < 2; k++)
            HIPCHK(hipMemcpy(t2 + k * es * n, s2[k], es * n, hipMemcpyHostToDevice));
        if (c->prec == GALHIP_FP32)
            HIPCHK(launch_p3_interleave<float>((const float *)t4, (const float *)t2,
                                               (float4 *)c->xv, c->id, n, c->s));
        else
            HIPCHK(launch_p3_interleave<double>((const double *)t4, (const double *)t2,
                                                (double4 *)c->xv, c->id, n, c->s));
        HIPCHK(hipMemcpy(c->zone, zz.data(), n, hipMemcpyHostToDevice));
    }
    c->qbits = qbits_for(n);
    if (c->prec == GALHIP_FP32) HIPCHK(launch_p3_ut<float>(p3_args<float>(c), c->s));
    else HIPCHK(launch_p3_ut<double>(p3_args<double>(c), c->s));
    int rc2 = p3_sort(c);
    if (rc2) return rc2;
    HIPCHK(hipStreamSynchronize(c->s));
    return 0;
}

extern "C" int galhip_unload3(galhip_ctx *c, void *x, void *y, void *z, void *vx,
                              void *vy, void *vz, int32_t *iz) {
    if (!c) return gh_fail(GALHIP_EINVAL, "null context");
    if (!c->p3d) return gh_fail(GALHIP_EINVAL, "galhip_unload3: not a p3d context");
    HIPCHK(hipStreamSynchronize(c->s));
    const long long n = c->n;
    if (n == 0) return 0;
    // planar, in load order, into the second particle buffers and the sort
    // keys (all free outside a sort), then straight to the caller's arrays
    const size_t es = esize(c);
    unsigned char *t4 = (unsigned char *)c->xv2, *t2 = t4 + 4 * es * n;
    int *zs = (int *)c->key;
    if (c->prec == GALHIP_FP32)
        HIPCHK(launch_p3_deinterleave<float>((const float4 *)c->xv, c->zone,
                                             c->id, (float *)t4, (float *)t2, zs, n, c->s));
    else
        HIPCHK(launch_p3_deinterleave<double>((const double4 *)c->xv, c->zone, c->id,
                                              (double *)t4, (double *)t2, zs, n, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    void *d4[4] = {x, y, z, vx};
    void *d2[2] = {vy, vz};
    for (int k = 0; k < 4; k++)
        if (d4[k]) HIPCHK(hipMemcpy(d4[k], t4 + k * es * n, es * n, hipMemcpyDeviceToHost));
    for (int k = 0; k < 2; k++)
        if (d2[k]) HIPCHK(hipMemcpy(d2[k], t2 + k * es * n, es * n, hipMemcpyDeviceToHost));
    if (iz) HIPCHK(hipMemcpy(iz, zs, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
    return 0;
}

// Regroup the particles by cell (the storage order sets the locality of the
// field gathers and lets the deposit combine the particles of a cell within
// a wave; no result depends on it): the key is the cell itself, off-grid
// particles last.
template <typename R> static int p3_sort_impl(galhip_ctx *c) {
    const long long n = c->n;
    if (n <= 1) return 0;
    const int kshift = 0;
    const unsigned kmax = (unsigned)c->mesh;
    HIPCHK(launch_p3_key<R>(p3_args<R>(c), kshift, kmax, c->key, c->perm, c->s));
    int end_bit = 1;
    while ((1LL << end_bit) <= (long long)kmax) end_bit++;
    size_t tmp = c->sort_tmp_bytes;
    HIPCHK(rocprim::radix_sort_pairs(c->sort_tmp, tmp, c->key, c->key2, c->perm,
                                              c->perm2, (int)n, 0, end_bit, c->s));
    P3PermArgs<R> p;
    p.rec_in = (const typename Vec4<R>::T *)c->xv;
    p.rec_out = (typename Vec4<R>::T *)c->xv2;
    p.zin = c->zone;
    p.zout = c->zone2;
    p.iin = c->id;
    p.iout = c->id2;
    p.perm = c->perm2;
    p.n = n;
    HIPCHK(launch_p3_permute<R>(p, c->s));
    std::swap(c->xv, c->xv2);
    std::swap(c->zone, c->zone2);
    std::swap(c->id, c->id2);
    c->steps_since_sort = 0;
    return 0;
}

int p3_sort(galhip_ctx *c) {
    return c->prec == GALHIP_FP32 ? p3_sort_impl<float>(c) : p3_sort_impl<double>(c);
}

int p3_unload(galhip_ctx *, void *, void *, void *, void *, int32_t *) {
    return gh_fail(GALHIP_EINVAL, "p3d context: use galhip_unload3");
}

// ---- argument assembly ------------------------------------------------------------------

template <typename R> static P3Args<R> p3_args(galhip_ctx *c) {
    P3Args<R> a;
    memset(&a, 0, sizeof a);
    a.rec = (typename Vec4<R>::T *)c->xv;
    a.zone = c->zone;
    a.n = c->n;
    a.fld = (const typename Vec4<R>::T *)c->fld4;
    for (int z = 0; z < c->nz; z++) a.mfx[z] = c->mfx[z];
    a.jrad = c->jrad;
    a.krad = c->krad;
    a.counters = c->counters;
    a.fr_c = c->fr_c;
    a.fr_inv_dr = c->nfr > 0 ? 1.0 / c->fr_dr : 0.0;
    a.nfr = c->nfr;
    a.fixforce = c->nfr > 0;
    a.nr = c->nr; a.na = c->na; a.ngz = c->ngz; a.mesh = c->mesh; a.nz = c->nz;
    a.mzone = c->mzone; a.offanal = c->offanal; a.nsect = c->nsect; a.qbits = c->qbits;
    a.jmass = c->jmass;
    a.k.alpha_d = c->alpha;
    a.k.inv_alpha = 1.0 / c->alpha;
    a.k.tscale = (double)c->na / (2.0 * 3.141592653589793238462643);
    a.k.na = c->na;
    a.rgrid = (R)c->rgrid; a.thmax = (R)c->thmax; a.zm = (R)c->zm; a.dzg = (R)c->dzg;
    for (int z = 0; z < c->nz; z++) {
        a.zbr[z] = (R)c->zbr[z];
        a.nstep[z] = c->nstep[z];
    }
    return a;
}

template <typename R>
static P3SolveArgs<R> p3_solve_args(galhip_ctx *c, int lpot, int explicit_zones, int jzone,
                                    int kzone) {
    P3SolveArgs<R> a;
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
    a.rho = (R *)c->rho; a.A = (R *)c->A3; a.M = (R *)c->M3; a.W = (R *)c->W3; a.Z = (R *)c->Z3;
    a.G = (const R *)c->G; a.tw = (const R *)c->tw; a.twz = (const R *)c->twz;
    a.lg = c->lg_d; a.slot_m = c->slot_m_d; a.slot_p = c->slot_p_d; a.slot_of = c->slot_of_d;
    a.fld = (typename Vec4<R>::T *)c->fld4;
    a.q = ldexp(1.0, -c->qbits);
    a.pmass = c->pmass;
    a.nr = c->nr; a.na = c->na; a.ngz = c->ngz; a.ng = c->ng; a.mmax = c->mmax; a.nm = c->nm3;
    a.nz = c->nz; a.mzone = c->mzone; a.lpot = lpot; a.jzone = jzone; a.kzone = kzone;
    a.explicit_zones = explicit_zones;
    a.ntypes = lpot ? 4 : 3;
    return a;
}

// ---- masset / findf / step / tmcent --------------------------------------------------------

template <typename R> static int p3_masset_impl(galhip_ctx *c) {
    c->mzone = c->nz;
    masscl_host(c, false);
    ZonePtrs zp;
    memset(&zp, 0, sizeof zp);
    for (int z = 0; z < c->nz; z++) { zp.mfx[z] = c->mfx[z]; zp.moz[z] = c->moz[z]; }
    HIPCHK(launch_masscl(zp, c->nz, c->mesh, c->jrad, c->krad, c->nr, 1, c->s));
    HIPCHK(launch_p3_massgn<R>(p3_args<R>(c), c->s));
    int rc = zoncmb_host(c);
    if (rc) return rc;
    c->masscl_done = 0;
    HIPCHK(hipStreamSynchronize(c->s));
    return 0;
}

int p3_masset(galhip_ctx *c) {
    return c->prec == GALHIP_FP32 ? p3_masset_impl<float>(c) : p3_masset_impl<double>(c);
}

template <typename R>
static int p3_solve_impl(galhip_ctx *c, int lpot, int explicit_zones, int jzone, int kzone) {
    if (!c->have_green) return gh_fail(GALHIP_EINVAL, "P3CONV: Green function not set (call galhip_greenm)");
    HIPCHK(launch_p3_solve<R>(p3_solve_args<R>(c, lpot, explicit_zones, jzone, kzone), c->s));
    c->have_field = 1;
    c->pot_valid = lpot != 0;
    return 0;
}

int p3_solve(galhip_ctx *c, int lpot, int explicit_zones, int jzone, int kzone) {
    return c->prec == GALHIP_FP32 ? p3_solve_impl<float>(c, lpot, explicit_zones, jzone, kzone)
                                  : p3_solve_impl<double>(c, lpot, explicit_zones, jzone, kzone);
}

template <typename R> static int p3_step_impl(galhip_ctx *c) {
    set_mzone(c);
    masscl_host(c, true);
    ZonePtrs zp;
    memset(&zp, 0, sizeof zp);
    for (int z = 0; z < c->nz; z++) { zp.mfx[z] = c->mfx[z]; zp.moz[z] = c->moz[z]; }
    HIPCHK(launch_masscl(zp, c->mzone, c->mesh, c->jrad, c->krad, c->nr, 1, c->s));
    c->masscl_done = 0;
    const P3Args<R> a = p3_args<R>(c);
    if (c->prof) HIPCHK(hipEventRecord(take_event(c), c->s));
    HIPCHK(launch_p3_push<R>(a, c->s));
    if (c->prof) {
        HIPCHK(hipEventRecord(take_event(c), c->s));
        c->prof_launch++;
    }
    c->istep += 1;
    return zoncmb_host(c);
}

int p3_step(galhip_ctx *c) {
    return c->prec == GALHIP_FP32 ? p3_step_impl<float>(c) : p3_step_impl<double>(c);
}

template <typename R> static int p3_tmcent_impl(galhip_ctx *c, int start) {
    P3Args<R> a = p3_args<R>(c);
    a.fac = start ? (R)0.5 : (R)-0.5;
    HIPCHK(launch_p3_tmcent<R>(a, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    c->tcent = start != 0;
    return 0;
}

int p3_tmcent(galhip_ctx *c, int start) {
    return c->prec == GALHIP_FP32 ? p3_tmcent_impl<float>(c, start) : p3_tmcent_impl<double>(c, start);
}

// ---- global integrals ------------------------------------------------------------------------

// part and dout: device buffers of p3_integrals_doubles(c) doubles in all
// (partials, then the P3_NINT totals), or null to allocate them here.
template <typename R>
static int p3_integrals_impl(galhip_ctx *c, int centred, double *out, double *buf) {
    const long long nb = c->n > 0 ? (c->n + 511) / 512 : 0;
    double *part = buf, *dout = buf ? buf + P3_NINT * (nb > 0 ? nb : 1) : nullptr;
    int rc = 0;
    if (!buf)
        rc = dalloc(&part, sizeof(double) * P3_NINT * (nb > 0 ? nb : 1)) |
             dalloc(&dout, sizeof(double) * P3_NINT);
    if (!rc) {
        hipError_t e = launch_p3_integrals<R>(p3_args<R>(c), c->pot_valid, c->fp_c,
                                              c->nfp, c->fp_dr, centred, part, dout, c->s);
        if (e == hipSuccess)
            e = hipMemcpyAsync(out, dout, sizeof(double) * P3_NINT, hipMemcpyDeviceToHost, c->s);
        if (e == hipSuccess) e = hipStreamSynchronize(c->s);
        if (e != hipSuccess) rc = gh_fail(GALHIP_EHIP, "integrals: %s", hipGetErrorString(e));
    }
    if (!buf) {
        dfree(part);
        dfree(dout);
    }
    if (rc) return rc;
    for (int q = 0; q < P3_NINT; q++) out[q] *= c->pmass;
    if (!c->pot_valid) out[9] = NAN;
    return 0;
}

long long p3_integrals_doubles(const galhip_ctx *c) {
    const long long nb = c->n > 0 ? (c->n + 511) / 512 : 0;
    return P3_NINT * ((nb > 0 ? nb : 1) + 1);
}

int p3_integrals(galhip_ctx *c, int centred, double *out, double *buf) {
    return c->prec == GALHIP_FP32 ? p3_integrals_impl<float>(c, centred, out, buf)
                                  : p3_integrals_impl<double>(c, centred, out, buf);
}

// ---- fields ------------------------------------------------------------------------------------

int p3_get_field(const galhip_ctx *c, void *fld) {
    HIPCHK(hipStreamSynchronize(c->s));
    const size_t es = esize(c), m = c->mesh;
    std::vector<unsigned char> tmp(4 * es * m);
    HIPCHK(hipMemcpy(tmp.data(), c->fld4, 4 * es * m, hipMemcpyDeviceToHost));
    unsigned char *o = (unsigned char *)fld;
    for (size_t i = 0; i < m; i++)
        for (int k = 0; k < 4; k++) memcpy(o + (k * m + i) * es, tmp.data() + (4 * i + k) * es, es);
    return 0;
}

int p3_set_field(galhip_ctx *c, const void *fld) {
    HIPCHK(hipStreamSynchronize(c->s));
    const size_t es = esize(c), m = c->mesh;
    std::vector<unsigned char> tmp(4 * es * m);
    const unsigned char *in = (const unsigned char *)fld;
    for (size_t i = 0; i < m; i++)
        for (int k = 0; k < 4; k++) memcpy(tmp.data() + (4 * i + k) * es, in + (k * m + i) * es, es);
    HIPCHK(hipMemcpy(c->fld4, tmp.data(), 4 * es * m, hipMemcpyHostToDevice));
    c->have_field = 1;
    return 0;
}

// ---- group-level parity entry points -----------------------------------------------------------

template <typename R>
static int p3_gather_impl(galhip_ctx *c, int64_t n, const void *x, const void *y, const void *z,
                          int fold, int32_t *cell, void *wt, void *acc) {
    if (n <= 0) return 0;
    const size_t es = sizeof(R);
    R *dx = nullptr, *dy = nullptr, *dz = nullptr, *dwt = nullptr, *dacc = nullptr;
    int *dcell = nullptr;
    int rc = dalloc(&dx, es * n) | dalloc(&dy, es * n) | dalloc(&dz, es * n);
    if (cell) rc |= dalloc(&dcell, sizeof(int) * n);
    const int nw = c->jmass == 3 ? 27 : 8;
    if (wt) rc |= dalloc(&dwt, nw * es * n);
    if (acc) rc |= dalloc(&dacc, 3 * es * n);
    hipError_t e = hipSuccess;
    if (!rc) {
        e = hipMemcpy(dx, x, es * n, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(dy, y, es * n, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(dz, z, es * n, hipMemcpyHostToDevice);
        P3GatherArgs<R> b;
        memset(&b, 0, sizeof b);
        b.x = dx; b.y = dy; b.z = dz; b.n = n;
        b.fld = (const typename Vec4<R>::T *)c->fld4;
        b.acc = dacc; b.cell = dcell; b.wt = dwt; b.fold = fold;
        b.g = p3_args<R>(c);
        if (e == hipSuccess) e = launch_p3_gather<R>(b, c->s);
        if (e == hipSuccess) e = hipStreamSynchronize(c->s);
        if (e == hipSuccess && cell) e = hipMemcpy(cell, dcell, sizeof(int) * n, hipMemcpyDeviceToHost);
        if (e == hipSuccess && wt) e = hipMemcpy(wt, dwt, nw * es * n, hipMemcpyDeviceToHost);
        if (e == hipSuccess && acc) e = hipMemcpy(acc, dacc, 3 * es * n, hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = gh_fail(GALHIP_EHIP, "p3d gather: %s", hipGetErrorString(e));
    }
    dfree(dx); dfree(dy); dfree(dz); dfree(dcell); dfree(dwt); dfree(dacc);
    return rc;
}

extern "C" int galhip_weight3(galhip_ctx *c, int64_t n, const void *x, const void *y,
                              const void *z, int32_t old, int32_t *cell, void *wt) {
    if (!c || (n > 0 && (!x || !y || !z || !cell || !wt)))
        return gh_fail(GALHIP_EINVAL, "galhip_weight3: bad arguments");
    if (!c->p3d) return gh_fail(GALHIP_EINVAL, "galhip_weight3: not a p3d context");
    HIPCHK(hipStreamSynchronize(c->s));
    const int fold = !old && c->nsect > 1;
    return c->prec == GALHIP_FP32 ? p3_gather_impl<float>(c, n, x, y, z, fold, cell, wt, nullptr)
                                  : p3_gather_impl<double>(c, n, x, y, z, fold, cell, wt, nullptr);
}

extern "C" int galhip_grdacc3(galhip_ctx *c, int64_t n, const void *x, const void *y,
                              const void *z, void *acc) {
    if (!c || (n > 0 && (!x || !y || !z || !acc)))
        return gh_fail(GALHIP_EINVAL, "galhip_grdacc3: bad arguments");
    if (!c->p3d) return gh_fail(GALHIP_EINVAL, "galhip_grdacc3: not a p3d context");
    if (!c->have_field) return gh_fail(GALHIP_ERANGE, "GETACC: Accelerations at useless time");
    HIPCHK(hipStreamSynchronize(c->s));
    return c->prec == GALHIP_FP32 ? p3_gather_impl<float>(c, n, x, y, z, 0, nullptr, nullptr, acc)
                                  : p3_gather_impl<double>(c, n, x, y, z, 0, nullptr, nullptr, acc);
}
