// Atom files, the host stage DelPhiShared::loadAtoms (DelphiShared.cpp:309-702):
// .xyzr lines "x y z r" and .pqr ATOM/HETATM records of 10 or 11 fields.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "nsgpu.h"

struct nsgpu_atom_set {
    std::vector<double> xyzr;    // n * 4
    std::vector<double> charge;  // n (pqr; zeros for xyzr)
    std::string err;
};

namespace {

// whole-field integer / double conversion, as the reference's pqr reader
// (DelphiShared.cpp:353-368): a trailing character fails the field
bool field_int(const std::string &s, long *v) {
    char *end = nullptr;
    *v = std::strtol(s.c_str(), &end, 10);
    return end != s.c_str() && *end == '\0';
}

bool field_double(const std::string &s, double *v) {
    char *end = nullptr;
    *v = std::strtod(s.c_str(), &end);
    return end != s.c_str() && *end == '\0';
}

// one pqr line: false when an ATOM/HETATM record does not parse; *atom tells
// whether the line is an atom record (DelphiShared.cpp:370-413)
bool pqr_line(const std::string &line, bool *atom, double p[3], double *q, double *r) {
    *atom = false;
    std::istringstream in(line);
    std::vector<std::string> f;
    std::string w;
    while (in >> w) f.push_back(w);
    if (f.empty() || (f[0] != "ATOM" && f[0] != "HETATM")) return true;
    *atom = true;
    long serial = 0, resid = 0;
    // 11 fields carry a chain name, 10 do not
    const size_t base = f.size() == 11 ? 6 : f.size() == 10 ? 5 : 0;
    if (!base) return false;
    return field_int(f[1], &serial) && field_int(f[base - 1], &resid) && field_double(f[base], &p[0]) &&
           field_double(f[base + 1], &p[1]) && field_double(f[base + 2], &p[2]) && field_double(f[base + 3], q) &&
           field_double(f[base + 4], r);
}

struct Atom {
    double p[3], r, q;
};

// every atom of the file, in file order; lines of one character or none are
// skipped (DelphiShared.cpp:511-515)
int read_all(const char *path, std::vector<Atom> *out, std::string *err) {
    std::ifstream in(path);
    if (!in) {
        *err = std::string("cannot read file ") + path;
        return NSGPU_E_ARG;
    }
    const bool pqr = std::strstr(path, ".pqr") != nullptr;
    std::string line;
    while (std::getline(in, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (line.size() <= 1) continue;
        Atom a{{0, 0, 0}, 0, 0};
        if (pqr) {
            bool atom = false;
            if (!pqr_line(line, &atom, a.p, &a.q, &a.r)) {
                *err = "cannot parse pqr line: " + line;
                return NSGPU_E_ARG;
            }
            if (!atom || a.r < 1e-20) continue;
        } else {
            // sscanf "%lf %lf %lf %lf": fields it cannot read keep their value
            std::sscanf(line.c_str(), "%lf %lf %lf %lf", &a.p[0], &a.p[1], &a.p[2], &a.r);
        }
        out->push_back(a);
    }
    return NSGPU_OK;
}

}  // namespace

extern "C" {

int nsgpu_load_atoms(const char *path, int32_t max_atoms, double domain_shrinkage, nsgpu_atom_set **out) {
    if (!path || !out) return NSGPU_E_ARG;
    *out = nullptr;
    auto *s = new nsgpu_atom_set();
    std::vector<Atom> all;
    int rc = read_all(path, &all, &s->err);
    if (rc) {
        *out = s;
        return rc;
    }
    // Domain_Shrinkage in (0, 1): keep the atoms inside the centred box
    // shrunk by that factor (DelphiShared.cpp:420-504, :622-647)
    const bool shrink = domain_shrinkage > 0. && domain_shrinkage < 1.;
    double lo[3] = {1e6, 1e6, 1e6}, hi[3] = {-1e6, -1e6, -1e6};
    if (shrink)
        for (const Atom &a : all)
            for (int d = 0; d < 3; ++d) {
                lo[d] = std::fmin(lo[d], a.p[d]);
                hi[d] = std::fmax(hi[d], a.p[d]);
            }
    double max_rad = 0;
    for (const Atom &a : all) {
        if (shrink) {
            bool keep = true;
            for (int d = 0; d < 3; ++d) {
                const double mid = 0.5 * (hi[d] + lo[d]);
                const double rlo = mid + domain_shrinkage * (lo[d] - mid), rhi = mid + domain_shrinkage * (hi[d] - mid);
                if (a.p[d] < rlo || a.p[d] > rhi) keep = false;
            }
            if (!keep) continue;
        }
        max_rad = std::fmax(a.r, max_rad);
        s->xyzr.insert(s->xyzr.end(), {a.p[0], a.p[1], a.p[2], a.r});
        s->charge.push_back(a.q);
        if (max_atoms > 0 && (int64_t)s->charge.size() >= max_atoms) break;
    }
    *out = s;
    if (s->charge.size() < 4) {
        s->err = "NanoShaper needs at least 4 atoms to work";
        return NSGPU_E_ARG;
    }
    if (max_rad == 0) {
        s->err = "all null radii";
        return NSGPU_E_ARG;
    }
    if (s->charge.size() > (size_t)INT32_MAX) {
        s->err = "too many atoms";
        return NSGPU_E_ARG;
    }
    return NSGPU_OK;
}

int32_t nsgpu_atom_set_view(const nsgpu_atom_set *s, const double **xyzr, const double **charge) {
    if (!s) return 0;
    if (xyzr) *xyzr = s->xyzr.empty() ? nullptr : s->xyzr.data();
    if (charge) *charge = s->charge.empty() ? nullptr : s->charge.data();
    return (int32_t)s->charge.size();
}

const char *nsgpu_atom_set_error(const nsgpu_atom_set *s) { return s ? s->err.c_str() : "null atom set"; }

void nsgpu_atom_set_free(nsgpu_atom_set *s) { delete s; }

}  // extern "C"
