// galhip_galaxy.cpp -- a native host driver of libgalhip through its C ABI,
// in the order of GALAXY's programs: pcs2dmp's start (SRC/progs/pcs2dmp.f:
// 94-103: load the particles, masset, findf, tmcent) and galaxy's main loop
// (SRC/progs/galaxy.f:107-148: findf + step per step, here galhip_run), for
// the p2d grid.  It reads a .pcs particle file, converts to internal units
// as loadup does (SRC/lib15/loadup.f:99,168-169: x lscale, v / gvfac with
// gvfac = 1/(lscale ts), FP32), assigns zones (zoneno.f:40-55, off-grid
// beyond rgrid) and the zone step factors (adjvls.f), runs, and writes the
// end state as dmp2pcs does (SRC/progs/dmp2pcs.f:89-100: masset, findf,
// then unload with its tmcent(.false.), unload.f:45, so the velocities are
// synchronous again) as a .pcs file in model units (velocities over their
// zone's factor, times gvfac) with the time advanced.
//
// usage: galhip_galaxy in.pcs out.pcs nr=N na=N hash=N soft=S lscale=L ts=T
//        steps=K [nstep=1,2,4] [zbr=0,r2,r3] [sort=K] [tsoft=1] [halo=file]
//        [precision=32|64]
//   soft and zbr are in grid units; halo is a binary file of n (int32), dr
//   (float64), val[n], der[n] (float64): the fixed radial force table of
//   galhip_set_fixforce in internal units.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "galhip.h"

namespace {

int fail(const char *what, int rc) {
    std::fprintf(stderr, "galhip_galaxy: %s failed (%d): %s\n", what, rc, galhip_last_error());
    return 1;
}

#define CALL(expr)                                \
    do {                                          \
        int rc_ = (expr);                         \
        if (rc_ != 0) return fail(#expr, rc_);    \
    } while (0)

struct Pcs {
    int32_t nsp[3] = {0, 0, 0}, ncoor = 0, block = 5000;
    float t = 0.f, pm = 0.f;
    int32_t pertbn = 0;
    std::vector<char> perturber;
    std::vector<float> coords;   // [n][ncoor]
};

// Fortran unformatted sequential records: 4-byte length, payload, length
bool read_records(const char *path, std::vector<std::vector<char>> &recs) {
    FILE *f = std::fopen(path, "rb");
    if (!f) return false;
    for (;;) {
        int32_t n = 0, m = 0;
        if (std::fread(&n, 4, 1, f) != 1) break;
        std::vector<char> r(n > 0 ? n : 0);
        if (n > 0 && std::fread(r.data(), 1, n, f) != (size_t)n) { std::fclose(f); return false; }
        if (std::fread(&m, 4, 1, f) != 1 || m != n) { std::fclose(f); return false; }
        recs.push_back(std::move(r));
    }
    std::fclose(f);
    return true;
}

bool read_pcs(const char *path, Pcs &p) {
    std::vector<std::vector<char>> recs;
    if (!read_records(path, recs) || recs.empty() || recs[0].size() < 32) return false;
    const char *h = recs[0].data();
    std::memcpy(p.nsp, h, 12);
    std::memcpy(&p.ncoor, h + 12, 4);
    std::memcpy(&p.block, h + 16, 4);
    std::memcpy(&p.t, h + 20, 4);
    std::memcpy(&p.pm, h + 24, 4);
    std::memcpy(&p.pertbn, h + 28, 4);
    size_t k = 1;
    if (p.pertbn) {
        if (recs.size() < 2) return false;
        p.perturber = recs[1];
        k = 2;
    }
    for (; k < recs.size(); k++) {
        const size_t nw = recs[k].size() / 4;
        const float *w = (const float *)recs[k].data();
        p.coords.insert(p.coords.end(), w, w + nw);
    }
    const long long n = (long long)p.nsp[0] + p.nsp[1] + p.nsp[2];
    return p.ncoor > 0 && (long long)p.coords.size() == n * p.ncoor;
}

void put_record(FILE *f, const void *data, int32_t n) {
    std::fwrite(&n, 4, 1, f);
    if (n) std::fwrite(data, 1, n, f);
    std::fwrite(&n, 4, 1, f);
}

bool write_pcs(const char *path, const Pcs &p) {
    FILE *f = std::fopen(path, "wb");
    if (!f) return false;
    char h[32];
    std::memcpy(h, p.nsp, 12);
    std::memcpy(h + 12, &p.ncoor, 4);
    std::memcpy(h + 16, &p.block, 4);
    std::memcpy(h + 20, &p.t, 4);
    std::memcpy(h + 24, &p.pm, 4);
    std::memcpy(h + 28, &p.pertbn, 4);
    put_record(f, h, 32);
    if (p.pertbn) put_record(f, p.perturber.data(), (int32_t)p.perturber.size());
    const long long n = (long long)p.coords.size() / p.ncoor;
    for (long long s = 0; s < n; s += p.block) {
        const long long m = (n - s < p.block) ? n - s : p.block;
        put_record(f, p.coords.data() + s * p.ncoor, (int32_t)(m * p.ncoor * 4));
    }
    return std::fclose(f) == 0;
}

std::vector<double> list_arg(const std::string &v) {
    std::vector<double> out;
    size_t s = 0;
    while (s <= v.size()) {
        size_t e = v.find(',', s);
        if (e == std::string::npos) e = v.size();
        if (e > s) out.push_back(std::atof(v.substr(s, e - s).c_str()));
        s = e + 1;
    }
    return out;
}

// the end state back to model units and .pcs layout; R = the context's
// floating point type
template <typename R>
void to_model(const std::vector<R> &x, const std::vector<R> &y, const std::vector<R> &vx,
              const std::vector<R> &vy, const std::vector<int32_t> &iz,
              const std::vector<int> &nstep, R ls, R gvfac, Pcs &out) {
    const size_t n = x.size();
    const int nz = (int)nstep.size();
    for (size_t i = 0; i < n; i++) {
        const R fac = (iz[i] >= 1 && iz[i] <= nz) ? (R)nstep[iz[i] - 1] : (R)1;
        out.coords[4 * i + 0] = (float)(x[i] / ls);
        out.coords[4 * i + 1] = (float)(y[i] / ls);
        out.coords[4 * i + 2] = (float)(vx[i] / fac * gvfac);
        out.coords[4 * i + 3] = (float)(vy[i] / fac * gvfac);
    }
}

template <typename R>
int run(const Pcs &in, Pcs &out, galhip_p2d_params &p, double lscale, double ts, int steps,
        int sort, const std::vector<int> &nstep, const std::vector<double> &zbr,
        const std::vector<double> &halo_val, const std::vector<double> &halo_der, double halo_dr) {
    const long long n = (long long)in.coords.size() / 4;
    // loadup's conversions in the working precision
    const R ls = (R)lscale, tsr = (R)ts;
    const R gvfac = (R)1 / (ls * tsr);
    p.pmass = (double)((R)in.pm * (ls * ls * ls) * (tsr * tsr));
    galhip_ctx *c = nullptr;
    CALL(galhip_p2dset(&p, &c));
    double geo[6];
    CALL(galhip_geometry(c, geo));
    const R rgrid = (R)geo[1];
    std::vector<R> x(n), y(n), vx(n), vy(n);
    std::vector<int32_t> iz(n);
    const int nz = (int)nstep.size();
    for (long long i = 0; i < n; i++) {
        x[i] = (R)in.coords[4 * i] * ls;
        y[i] = (R)in.coords[4 * i + 1] * ls;
        vx[i] = (R)in.coords[4 * i + 2] / gvfac;
        vy[i] = (R)in.coords[4 * i + 3] / gvfac;
        const R r = std::sqrt(x[i] * x[i] + y[i] * y[i]);
        int z = 1;
        for (int k = 2; k <= nz; k++)
            if (r >= (R)zbr[k - 1]) z = k;
        if (r > rgrid) z = nz + 1;
        iz[i] = z;
        if (z <= nz) {
            vx[i] = vx[i] * (R)nstep[z - 1];
            vy[i] = vy[i] * (R)nstep[z - 1];
        }
    }
    CALL(galhip_greenm(c));
    if (!halo_val.empty())
        CALL(galhip_set_fixforce(c, (int32_t)halo_val.size(), halo_dr, halo_val.data(),
                                 halo_der.data()));
    CALL(galhip_restrt(c, n, x.data(), y.data(), vx.data(), vy.data(), iz.data()));
    CALL(galhip_set_sort_interval(c, sort));
    // pcs2dmp: masset, a full findf, time-centring
    CALL(galhip_masset(c));
    CALL(galhip_findf(c, 0));
    CALL(galhip_tmcent(c, 1));
    // galaxy's loop: findf + step, progress as galaxy prints it
    const int chunk = 100;
    for (int s = 0; s < steps; s += chunk) {
        const int k = steps - s < chunk ? steps - s : chunk;
        std::printf("Starting step %d\n", s);
        std::fflush(stdout);
        CALL(galhip_run(c, k));
    }
    // dmp2pcs's end (dmp2pcs.f:89-100): masset, a full findf, then unload,
    // whose tmcent(.false.) (unload.f:45) turns the leap-frog's half-step
    // velocities back into synchronous ones before they are written
    CALL(galhip_masset(c));
    CALL(galhip_findf(c, 1));
    CALL(galhip_tmcent(c, 0));
    CALL(galhip_sync(c));
    CALL(galhip_unload(c, x.data(), y.data(), vx.data(), vy.data(), iz.data()));
    int64_t noff = 0;
    CALL(galhip_get_noff(c, &noff));
    CALL(galhip_free(c));
    out = in;
    out.t = (float)(in.t + steps * ts);
    to_model<R>(x, y, vx, vy, iz, nstep, ls, gvfac, out);
    std::printf("done: %d steps, %lld particles, %lld off the grid\n", steps, n, (long long)noff);
    return 0;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s in.pcs out.pcs key=value ...\n", argv[0]);
        return 2;
    }
    galhip_p2d_params p;
    std::memset(&p, 0, sizeof p);
    p.tsoft = 1;
    p.nsect = 1;
    p.offanal = 1;
    p.precision = GALHIP_FP32;
    double lscale = 0.0, ts = 0.0;
    int steps = 0, sort = 10;
    std::vector<int> nstep = {1};
    std::vector<double> zbr = {0.0};
    std::string halo;
    for (int a = 3; a < argc; a++) {
        std::string kv = argv[a];
        const size_t e = kv.find('=');
        if (e == std::string::npos) {
            std::fprintf(stderr, "galhip_galaxy: bad argument %s\n", argv[a]);
            return 2;
        }
        const std::string k = kv.substr(0, e), v = kv.substr(e + 1);
        if (k == "nr") p.nr = std::atoi(v.c_str());
        else if (k == "na") p.na = std::atoi(v.c_str());
        else if (k == "hash") p.hash = std::atoi(v.c_str());
        else if (k == "soft") p.softl = std::atof(v.c_str());
        else if (k == "tsoft") p.tsoft = std::atoi(v.c_str());
        else if (k == "lscale") lscale = std::atof(v.c_str());
        else if (k == "ts") ts = std::atof(v.c_str());
        else if (k == "steps") steps = std::atoi(v.c_str());
        else if (k == "sort") sort = std::atoi(v.c_str());
        else if (k == "halo") halo = v;
        else if (k == "precision") p.precision = std::atoi(v.c_str()) == 64 ? GALHIP_FP64 : GALHIP_FP32;
        else if (k == "nstep") {
            nstep.clear();
            for (double d : list_arg(v)) nstep.push_back((int)d);
        } else if (k == "zbr") zbr = list_arg(v);
        else {
            std::fprintf(stderr, "galhip_galaxy: unknown key %s\n", k.c_str());
            return 2;
        }
    }
    if (nstep.size() != zbr.size() || nstep.empty() || nstep.size() > GALHIP_MAXZONES ||
        !(lscale > 0.0) || !(ts > 0.0) || steps < 0) {
        std::fprintf(stderr, "galhip_galaxy: need lscale, ts, steps and matching nstep/zbr\n");
        return 2;
    }
    p.nzones = (int32_t)nstep.size();
    for (size_t i = 0; i < nstep.size(); i++) {
        p.nstep[i] = nstep[i];
        p.zbr[i] = zbr[i];
    }
    Pcs in, out;
    if (!read_pcs(argv[1], in) || in.ncoor != 4) {
        std::fprintf(stderr, "galhip_galaxy: cannot read a 2-D .pcs file from %s\n", argv[1]);
        return 1;
    }
    std::vector<double> hv, hd;
    double hdr = 0.0;
    if (!halo.empty()) {
        FILE *f = std::fopen(halo.c_str(), "rb");
        int32_t hn = 0;
        bool ok = f && std::fread(&hn, 4, 1, f) == 1 && hn > 1 &&
                  std::fread(&hdr, 8, 1, f) == 1;
        if (ok) {
            hv.resize(hn);
            hd.resize(hn);
            ok = std::fread(hv.data(), 8, hn, f) == (size_t)hn &&
                 std::fread(hd.data(), 8, hn, f) == (size_t)hn;
        }
        if (f) std::fclose(f);
        if (!ok) {
            std::fprintf(stderr, "galhip_galaxy: cannot read the halo table %s\n", halo.c_str());
            return 1;
        }
    }
    const int rc = p.precision == GALHIP_FP64
                       ? run<double>(in, out, p, lscale, ts, steps, sort, nstep, zbr, hv, hd, hdr)
                       : run<float>(in, out, p, lscale, ts, steps, sort, nstep, zbr, hv, hd, hdr);
    if (rc) return rc;
    if (!write_pcs(argv[2], out)) {
        std::fprintf(stderr, "galhip_galaxy: cannot write %s\n", argv[2]);
        return 1;
    }
    return 0;
}
