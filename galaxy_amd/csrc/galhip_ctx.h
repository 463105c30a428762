// galhip_ctx.h -- the library context (the opaque galhip_ctx of galhip.h)
// and the host helpers shared by the translation units of libgalhip.
#pragma once
#include "galhip_kern.h"
#include "../../include/galhip.h"

#include <string>
#include <vector>

// Sets the message galhip_last_error returns; returns code.
int gh_fail(int code, const char *fmt, ...);

#define HIPCHK(expr)                                                          \
    do {                                                                      \
        hipError_t e_ = (expr);                                               \
        if (e_ != hipSuccess)                                                 \
            return gh_fail(GALHIP_EHIP, "%s: %s (%s:%d)", #expr,                 \
                        hipGetErrorString(e_), __FILE__, __LINE__);           \
    } while (0)

struct galhip_ctx {
    int prec;
    int nr, na, ng, mmax, nsect, mesh, tsoft, nz, offanal;
    int nstep[GH_MAXZ];
    double zbr[GH_MAXZ];
    double softl, pmass;
    std::vector<int> lg;
    double alpha, rgrid, thmax;   // values in the context's precision
    // host time/zone bookkeeping
    int istep, mzone;
    int lstep[2][GH_MAXZ], mstep[GH_MAXZ];
    double fnew[GH_MAXZ], fold[GH_MAXZ];
    int use_new[GH_MAXZ], use_old[GH_MAXZ];
    int masscl_done;              // masscl of this istep already on the device
    int tcent;                    // velocities time-centred (tmcent)
    int have_green, have_field;
    // device storage
    hipStream_t s;
    long long n, cap;
    void *xv, *ut;                     // (x, y, vx, vy) and (u, t) per particle; a p3d
                                       // context keeps its particle records in xv (ut unused)
    uint8_t *zone;
    unsigned *id;
    void *xv2, *ut2;
    uint8_t *zone2;
    unsigned *id2;
    unsigned *key, *key2, *perm, *perm2;
    void *sort_tmp;
    size_t sort_tmp_bytes;
    long long *mfx[GH_MAXZ], *moz[GH_MAXZ];
    int *jrad, *krad, *ranges, *lg_d;
    long long *counters;
    void *rho, *coef, *fld2, *pot, *G, *tw;
    double *Gd, *twd;
    // fixed radial tables as cubic coefficients per interval (radtab_coeffs):
    // the force, the potential (diagnostics)
    double4 *fr_c;
    double fr_dr;
    double4 *fp_c;
    double fp_dr;
    int nfp;
    int pot_valid;                    // the last solve formed the potential
    int nfr, qbits, ncoef;
    int sort_interval, steps_since_sort;
    // profiling
    int prof;
    std::vector<hipEvent_t> ev_free, ev_used;
    double prof_ms;
    long long prof_launch, prof_push0;
    // 3-D polar grid (p3d); p3d = 0 for the 2-D grid
    int p3d, ngz, nm3, jmass;
    double dzg, zm;                   // in the context's precision
    void *A3, *M3, *W3, *Z3;          // solve work arrays
    void *fld4;                       // [mesh] (F_x, F_y, F_z, potential)
    void *twz;                        // cos, sin of 2 pi k / (2 ngz)
    double *twzd;                     // the same in double
    int *slot_m_d, *slot_p_d, *slot_of_d;
    size_t gsize3;                    // entries of the p3d Green's table
    // Ensemble of nmem independent p2d discs (ens = 1) sharing the grid,
    // filter, zones and Green's table; nmem = 1 for a single grid.  Member
    // m's part of the per-grid arrays above sits at offset m*mesh (mass
    // sums, rho, field, potential), m*ncoef*nr (coefficients), m*GH_MAXZ
    // (jrad, krad) and m*4 (ranges, counters).
    int ens, nmem;
    std::vector<long long> m_off;     // particle offsets [nmem + 1]
    std::vector<double> m_pmass;      // particle mass per member
    std::vector<int> m_qbits;         // fixed-point scale per member
    int *m_qbits_d;
    double *m_pmass_d;
    int4 *blk_d;                      // push and key blocks: (member, first, count, 0)
    int nblk;
    // HIP graphs of galhip_run's steps, keyed by their launch arguments
    int graph;
    std::vector<std::pair<std::string, hipGraphExec_t>> graphs;
    // adjoint (galhip_adj.hip): per-particle adjoints [4][n] in load order,
    // FP64 grid work arrays, the fixed force's parameter derivative table
    // and recorded particle states
    double *adj_lam, *adj_g, *adj_rbar, *adj_wbar, *adj_part;
    double *adj_tmp;                  // [4][n] work copy for re-layouts
    long long adj_n;
    // particle order: order_ver changes whenever the storage order does;
    // the adjoints are kept in the storage order of version adj_ver,
    // whose load-order index is adj_id
    long long order_ver, adj_ver;
    unsigned *adj_id;
    double4 *fd_c;                    // parameter derivative of the fixed force
    double fd_dr;
    int nfd;
    // tangent-linear (forward-mode) state: perturbations [4][n] of
    // (x, y, vx, vy) in load order, the perturbed density [mesh], its
    // field [2][mesh] and the solve's coefficients [2][ng][nr]
    double *tl_d, *tl_rho, *tl_f, *tl_m;
    long long tl_n;
    struct AdjSlot {
        void *xv, *ut;
        uint8_t *zone;
        unsigned *id;
        long long *counters;
        long long *mass;              // zone sums [nz][mesh]
        int *rng;                     // jrad, krad [2][GH_MAXZ]
        long long n, order_ver;
        int istep;
    };
    std::vector<AdjSlot> adj_slots;
    std::vector<hipEvent_t> adj_ev;   // profiling: scatter and gather start/end per step
    // radial contraction: 0 in the synthesis kernel (default), 1 on MFMA
    // with split-BF16 operands (FP32 grids); the bf16 split of G, padded to
    // nrp rings; the GEMM output; the tuned GEMM variant (-1: not yet)
    // compensated positions (FP32, opt-in): low words of (x, y), sort copy
    int comp;
    void *xlo, *xlo2;
    int contract, nrp, gsplit_valid, gemm_variant;
    unsigned short *gsp;          // three bf16 words of G (FP64 master), word-major
    float *wpre;
    double gemm_us[3];
};

static size_t esize(const galhip_ctx *c) { return c->prec == GALHIP_FP64 ? 8 : 4; }

// The cubic coefficients of the Hermite interpolant of (val, der) on
// r_k = k dr, interval by interval (the device's fixed-table layout; see
// fixforce_at), with the constant val[n - 1] as entry n - 1.
static std::vector<double4> radtab_coeffs(int n, double dr, const double *val,
                                          const double *der) {
    std::vector<double4> c((size_t)n);
    for (int k = 0; k + 1 < n; k++) {
        const double v0 = val[k], v1 = val[k + 1];
        const double c1 = dr * der[k];
        const double c2 = (3.0 * (v1 - v0) - 2.0 * c1) - dr * der[k + 1];
        const double c3 = (2.0 * (v0 - v1) + c1) + dr * der[k + 1];
        c[k] = make_double4(v0, c1, c2, c3);
    }
    c[n - 1] = make_double4(val[n - 1], 0.0, 0.0, 0.0);
    return c;
}

template <typename T> static int dalloc(T **p, size_t bytes) {
    *p = nullptr;
    if (bytes == 0) return 0;
    hipError_t e = hipMalloc((void **)p, bytes);
    if (e != hipSuccess)
        return gh_fail(GALHIP_EHIP, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    return 0;
}

template <typename T> static void dfree(T *&p) {
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}


// zone/time bookkeeping of the reference (galhip_api.hip)
int qbits_for(long long n);
void set_mzone(galhip_ctx *c);
void masscl_host(galhip_ctx *c, bool normal);
int zoncmb_host(galhip_ctx *c);
hipEvent_t take_event(galhip_ctx *c);

namespace gh {
hipError_t launch_d2f(const double *, float *, long long, hipStream_t);
hipError_t launch_fill_i64(long long *, long long, long long, hipStream_t);
hipError_t launch_fill_i32(int *, int, long long, hipStream_t);
hipError_t launch_masscl(const ZonePtrs &, int, long long, int *, int *, int, int,
                         hipStream_t);
}  // namespace gh

// adjoint storage (galhip_adj.hip)
void adj_free(galhip_ctx *c);

// p3d implementations of the generic entry points (galhip_p3d_api.hip)
void p3_free(galhip_ctx *c);
int p3_greenm(galhip_ctx *c);
int p3_get_green(const galhip_ctx *c, double *G);
int p3_set_green(galhip_ctx *c, const double *G);
int p3_masset(galhip_ctx *c);
int p3_solve(galhip_ctx *c, int lpot, int explicit_zones, int jzone, int kzone);
int p3_step(galhip_ctx *c);
int p3_tmcent(galhip_ctx *c, int start);
int p3_get_field(const galhip_ctx *c, void *fld);
int p3_set_field(galhip_ctx *c, const void *fld);
int p3_unload(galhip_ctx *c, void *x, void *y, void *vx, void *vy, int32_t *iz);
int p3_sort(galhip_ctx *c);
int p3_integrals(galhip_ctx *c, int centred, double *out, double *buf = nullptr);
long long p3_integrals_doubles(const galhip_ctx *c);
