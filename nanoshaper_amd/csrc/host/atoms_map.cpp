// Nearest atom of points, the host stage behind the vertex-atom map of the
// mesh files and the atoms of each cavity: the acceleration grid of
// Surface::buildAtomsMap (Surface.cpp:1695-1749) and the search of
// Surface::vdwAccessible (Surface.cpp:16663-16713), which the cavity-atom
// pass repeats on the cavities' boundary nodes (Surface::getCavitiesAtoms,
// Surface.cpp:2129-2245).
//
// The map has ggrid^3 cells of side 1 / gscale centred on the grid's
// baricenter (gscale starts at 0.5 / max radius and is halved or doubled
// until ggrid = floor(1.3 gscale rmaxdim) is in [1, 100]); a point searches
// the 27 cells around its own in the reference's order and keeps the first
// atom of smallest |p - c|^2 - r^2.  Points with no atom in those cells get -1.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "nsgpu.h"

namespace {

double rintp(double x) { return std::floor(x + 0.5); }

// tools.h:35-64: dx in (0, +1, -1), then dy in (0, +1, -1), then dz in (0, -1, +1)
struct Shift {
    int d[3];
};
std::vector<Shift> shift_map() {
    std::vector<Shift> s;
    const int a[3] = {0, 1, -1}, c[3] = {0, -1, 1};
    for (int x : a)
        for (int y : a)
            for (int z : c) s.push_back({{x, y, z}});
    return s;
}

struct AtomsMap {
    double gscale = 0, gmin[3] = {0, 0, 0};
    int64_t ggrid = 0;
    std::vector<int64_t> start;  // ggrid^3 + 1 offsets into ids
    std::vector<int32_t> ids;    // atoms per cell, in atom order

    bool build(const nsgpu_grid &g, const double *xyzr, int64_t n) {
        double rmax = 0;
        for (int64_t i = 0; i < n; ++i) rmax = std::max(xyzr[4 * i + 3], rmax);
        if (!(rmax > 0)) return false;
        gscale = 0.5 / rmax;
        for (int it = 0; it < 4096; ++it) {
            ggrid = (int64_t)(unsigned int)std::floor(gscale * 1.3 * g.rmaxdim);
            if (ggrid <= 100) {
                if (ggrid <= 0)
                    gscale *= 2.0;
                else
                    break;
            } else {
                gscale *= 0.5;
            }
        }
        if (ggrid <= 0 || ggrid > 100) return false;
        const double gside = 1. / gscale;
        for (int d = 0; d < 3; ++d) gmin[d] = g.baricenter[d] - (ggrid - 1) * 0.5 * gside;
        const int64_t nc = ggrid * ggrid * ggrid;
        std::vector<int64_t> cell(n);
        std::vector<int64_t> cnt(nc + 1, 0);
        for (int64_t i = 0; i < n; ++i) {
            int64_t ix[3];
            for (int d = 0; d < 3; ++d) ix[d] = (int64_t)rintp((xyzr[4 * i + d] - gmin[d]) * gscale);
            // the reference writes past its map here; such atoms are left out
            if (ix[0] < 0 || ix[1] < 0 || ix[2] < 0 || ix[0] >= ggrid || ix[1] >= ggrid || ix[2] >= ggrid) {
                cell[i] = -1;
                continue;
            }
            cell[i] = ix[2] * ggrid * ggrid + ix[1] * ggrid + ix[0];
            ++cnt[cell[i] + 1];
        }
        start.assign(nc + 1, 0);
        for (int64_t c = 0; c < nc; ++c) start[c + 1] = start[c] + cnt[c + 1];
        ids.assign(start[nc], 0);
        std::vector<int64_t> fill(start.begin(), start.end() - 1);
        for (int64_t i = 0; i < n; ++i)
            if (cell[i] >= 0) ids[fill[cell[i]]++] = (int32_t)i;
        return true;
    }

    int32_t nearest(const double *xyzr, const double p[3], const std::vector<Shift> &sm) const {
        double best = INFINITY;
        int32_t winner = -1;
        int64_t ix[3];
        for (int d = 0; d < 3; ++d) ix[d] = (int64_t)rintp((p[d] - gmin[d]) * gscale);
        for (const Shift &s : sm) {
            const int64_t cx = ix[0] + s.d[0], cy = ix[1] + s.d[1], cz = ix[2] + s.d[2];
            if (cx >= ggrid || cy >= ggrid || cz >= ggrid || cx < 0 || cy < 0 || cz < 0) continue;
            const int64_t c = cz * ggrid * ggrid + cy * ggrid + cx;
            for (int64_t q = start[c]; q < start[c + 1]; ++q) {
                const int32_t a = ids[q];
                const double *at = xyzr + 4 * (int64_t)a;
                // DIST2 then minus r^2, in the reference's order
                double sd = (at[0] - p[0]) * (at[0] - p[0]) + (at[1] - p[1]) * (at[1] - p[1]) +
                            (at[2] - p[2]) * (at[2] - p[2]);
                sd -= at[3] * at[3];
                if (sd < best) {
                    best = sd;
                    winner = a;
                }
            }
        }
        return winner;
    }
};

}  // namespace

extern "C" int nsgpu_nearest_atoms(const nsgpu_grid *g, const double *xyzr, int64_t n, const double *points,
                                   int64_t m, int32_t *out) {
    if (!g || (n > 0 && !xyzr) || (m > 0 && (!points || !out)) || n < 0 || m < 0) return NSGPU_E_ARG;
    AtomsMap map;
    if (n == 0 || !map.build(*g, xyzr, n)) {
        for (int64_t i = 0; i < m; ++i) out[i] = -1;
        return n == 0 ? NSGPU_OK : NSGPU_E_ARG;
    }
    const std::vector<Shift> sm = shift_map();
    for (int64_t i = 0; i < m; ++i) out[i] = map.nearest(xyzr, points + 3 * i, sm);
    return NSGPU_OK;
}
