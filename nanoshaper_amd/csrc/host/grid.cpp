// Host stages that stay on the CPU: the DelPhi grid, the atom jitter and the
// synthetic molecule generator used by the tests and the benchmark.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "nsgpu.h"

extern "C" int nsgpu_abi_version(void) { return NSGPU_ABI_VERSION; }

// DelPhiShared::buildGrid(scale, perfil), DelphiShared.cpp:1038-1143, with the
// atom bounds of DelPhiShared::getBounds, DelphiShared.cpp:775-797.  The grid is
// cubic: igrid = floor(scale*100/perfil*rmaxdim) forced odd, centred on the
// mid point of the atom bounding box.
extern "C" int nsgpu_build_grid(const double *xyzr, int32_t n, double scale, double perfil,
                                nsgpu_grid *out) {
    if (!xyzr || n <= 0 || !out || !(scale > 0.0) || !(perfil > 0.0) || perfil > 100.0)
        return NSGPU_E_ARG;
    double cmin[3] = {INFINITY, INFINITY, INFINITY};
    double cmax[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int32_t a = 0; a < n; ++a) {
        const double *p = xyzr + 4 * (size_t)a;
        for (int d = 0; d < 3; ++d) {
            cmin[d] = std::fmin(cmin[d], p[d] - p[3]);
            cmax[d] = std::fmax(cmax[d], p[d] + p[3]);
        }
    }
    double mid[3], rmax = -INFINITY;
    for (int d = 0; d < 3; ++d) {
        mid[d] = (cmax[d] + cmin[d]) / 2.;
        rmax = std::fmax(rmax, std::fabs(cmax[d] - mid[d]));
        rmax = std::fmax(rmax, std::fabs(cmin[d] - mid[d]));
    }
    const double rmaxdim = 2 * rmax;
    unsigned int igrid = (unsigned int)std::floor(scale * 100. / perfil * rmaxdim);
    if ((igrid % 2) == 0) igrid++;
    std::memset(out, 0, sizeof(*out));
    out->nx = out->ny = out->nz = (int32_t)igrid;
    const double half = (igrid - 1) / (2 * scale);
    out->xmin = mid[0] - half;
    out->ymin = mid[1] - half;
    out->zmin = mid[2] - half;
    out->xmax = mid[0] + half;
    out->ymax = mid[1] + half;
    out->zmax = mid[2] + half;
    out->scale = scale;
    out->side = 1. / scale;
    out->hside = 0.5 / scale;
    out->A = out->side * out->side;
    out->rmaxdim = rmaxdim;
    for (int d = 0; d < 3; ++d) out->baricenter[d] = mid[d];
    return NSGPU_OK;
}

// Park-Miller "minimal standard" generator with Schrage's factorisation, the
// reference's randnum() (tools.cpp:64-84): a = 16807, m = 2^31-1, q = 127773,
// r = 2836.  Returns state/m after advancing the state.
static double park_miller(int64_t *state) {
    const int64_t aa = 16807, mm = 2147483647, qq = 127773, rr = 2836;
    const int64_t hh = *state / qq;
    const int64_t lo = *state - hh * qq;
    const int64_t test = aa * lo - rr * hh;
    *state = (test >= 0) ? test : test + mm;
    return (double)*state / (double)mm;
}

// Jitter of ConnollySurface::buildConnollyCGAL, ConnollySurface.cpp:2484-2489:
// pos += amplitude*(u - 0.5) for x, y, z of each atom in turn.
extern "C" void nsgpu_jitter(double *xyzr, int32_t n, double amplitude, int32_t *seed) {
    if (!xyzr || !seed) return;
    int64_t s = *seed;
    for (int32_t a = 0; a < n; ++a) {
        double *p = xyzr + 4 * (size_t)a;
        p[0] += amplitude * (park_miller(&s) - 0.5);
        p[1] += amplitude * (park_miller(&s) - 0.5);
        p[2] += amplitude * (park_miller(&s) - 0.5);
    }
    *seed = (int32_t)s;
}

// ---------------------------------------------------------------------------
// Synthetic chains.  splitmix64 drives a self-avoiding random walk confined to
// a sphere; a uniform cell list (3 A cells) keeps the separation test local.
// ---------------------------------------------------------------------------
namespace {
struct Rng {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
};

struct CellList {
    double lo, inv;
    int n;
    std::vector<std::vector<int>> cells;
    CellList(double radius, double cell) {
        lo = -radius - cell;
        inv = 1.0 / cell;
        n = (int)std::ceil((2 * radius + 2 * cell) * inv) + 1;
        cells.resize((size_t)n * n * n);
    }
    int idx(double v) const {
        int i = (int)std::floor((v - lo) * inv);
        return i < 0 ? 0 : (i >= n ? n - 1 : i);
    }
    std::vector<int> &at(int i, int j, int k) { return cells[((size_t)k * n + j) * n + i]; }
};
}  // namespace

extern "C" int32_t nsgpu_generate_chain(int32_t n, double radius, uint64_t seed, double *xyzr,
                                       double *radius_used) {
    if (n <= 0 || !xyzr || !(radius > 0)) return 0;
    const double step = 1.5, sep = 3.0, sep2 = sep * sep;
    Rng rng{seed * 0x2545F4914F6CDD1Dull + 0x1234567ull};
    // the cell list covers the largest radius the confinement may grow to
    CellList cl(radius * 4.0, sep);
    double R = radius;
    std::vector<double> pos;
    pos.reserve((size_t)n * 3);
    auto ok = [&](const double p[3], int count) {
        if (p[0] * p[0] + p[1] * p[1] + p[2] * p[2] > R * R) return false;
        int ci = cl.idx(p[0]), cj = cl.idx(p[1]), ck = cl.idx(p[2]);
        for (int dk = -1; dk <= 1; ++dk)
            for (int dj = -1; dj <= 1; ++dj)
                for (int di = -1; di <= 1; ++di) {
                    int a = ci + di, b = cj + dj, c = ck + dk;
                    if (a < 0 || b < 0 || c < 0 || a >= cl.n || b >= cl.n || c >= cl.n) continue;
                    for (int q : cl.at(a, b, c)) {
                        if (q >= count - 2) continue;  // atoms within 2 bonds are exempt
                        double dx = p[0] - pos[3 * q], dy = p[1] - pos[3 * q + 1],
                               dz = p[2] - pos[3 * q + 2];
                        if (dx * dx + dy * dy + dz * dz < sep2) return false;
                    }
                }
        return true;
    };
    auto add = [&](const double p[3]) {
        int q = (int)(pos.size() / 3);
        pos.insert(pos.end(), p, p + 3);
        cl.at(cl.idx(p[0]), cl.idx(p[1]), cl.idx(p[2])).push_back(q);
    };
    auto pop = [&]() {
        int q = (int)(pos.size() / 3) - 1;
        auto &v = cl.at(cl.idx(pos[3 * q]), cl.idx(pos[3 * q + 1]), cl.idx(pos[3 * q + 2]));
        for (size_t t = 0; t < v.size(); ++t)
            if (v[t] == q) { v.erase(v.begin() + t); break; }
        pos.resize(pos.size() - 3);
    };
    const double origin[3] = {0, 0, 0};
    add(origin);
    // 'stuck' counts dead ends since the chain last grew beyond its longest
    // length so far, so back-off/re-grow cycles cannot loop forever.
    int stuck = 0, longest = 1;
    while ((int)(pos.size() / 3) < n) {
        int count = (int)(pos.size() / 3);
        const double *prev = &pos[3 * (count - 1)];
        bool placed = false;
        for (int tries = 0; tries < 48 && !placed; ++tries) {
            double z = 2 * rng.uniform() - 1, phi = 2 * M_PI * rng.uniform();
            double s = std::sqrt(std::fmax(0.0, 1 - z * z));
            double p[3] = {prev[0] + step * s * std::cos(phi), prev[1] + step * s * std::sin(phi),
                           prev[2] + step * z};
            if (ok(p, count)) { add(p); placed = true; }
        }
        if (placed) {
            if (count + 1 > longest) { longest = count + 1; stuck = 0; }
            continue;
        }
        // dead end: back off a few monomers; when the walk keeps jamming the
        // confinement is too tight for the chain, so it grows by 2 %
        if (++stuck > 64) {
            R *= 1.02;
            stuck = 0;
            if (R > radius * 3.9) break;
        }
        int back = 1 + (stuck > 8 ? (int)(rng.uniform() * 12) : 0);
        for (int b = 0; b < back && pos.size() / 3 > 1; ++b) pop();
    }
    const int32_t got = (int32_t)(pos.size() / 3);
    // element radii C/N/O/S with probabilities .62/.17/.19/.02
    for (int32_t a = 0; a < got; ++a) {
        double u = rng.uniform();
        double r = u < 0.62 ? 1.70 : (u < 0.79 ? 1.55 : (u < 0.98 ? 1.52 : 1.80));
        xyzr[4 * a] = pos[3 * a];
        xyzr[4 * a + 1] = pos[3 * a + 1];
        xyzr[4 * a + 2] = pos[3 * a + 2];
        xyzr[4 * a + 3] = r;
    }
    if (radius_used) *radius_used = R;
    return got;
}
