// nsgpu_surface.hpp - header-only C++ plugin layer over the C ABI (nsgpu.h).
//
// Mirrors the reference's surface plugin interface so that code written
// against Surface / SurfaceFactory / ConfigFile / DelPhiShared keeps its shape:
//   ConfigFile            ConfigFile/ConfigFile.h:59-269 ("key = value", '#' comments)
//   DelPhiShared          DelphiShared.h:43-216 (atoms, grid, caller-visible maps)
//   Surface               Surface.h:835-979 (build, getSurf, triangulateSurface)
//   ConnollySurface "ses" ConnollySurface.h:682-692 (factory registration)
//   MeshSurface "mesh"    MeshSurface.h:211-221, MeshSurface.cpp:176-248 (OFF input)
//   BlobbySurface "blobby" BlobbySurface.h:68-79, BlobbySurface.cpp:83-229
//   SurfaceFactory        SurfaceFactory.h:15-65 (name -> instantiator, key "Surface")
// Everything on the device goes through the extern "C" entry points; this
// header adds no symbols to libnsgpu.so.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "nsgpu.h"

namespace nsgpu {

// key = value configuration, as the reference's ConfigFile
class ConfigFile {
   public:
    ConfigFile() = default;
    explicit ConfigFile(const std::string &path) {
        std::ifstream in(path);
        if (!in) throw std::runtime_error("cannot open config " + path);
        std::string line;
        while (std::getline(in, line)) parse_line(line);
    }
    void parse_line(std::string line) {
        const size_t hash = line.find('#');
        if (hash != std::string::npos) line.erase(hash);
        const size_t eq = line.find('=');
        if (eq == std::string::npos) return;
        std::string k = trim(line.substr(0, eq)), v = trim(line.substr(eq + 1));
        if (!k.empty()) kv_[k] = v;
    }
    bool keyExists(const std::string &k) const { return kv_.count(k) != 0; }
    template <class T>
    T read(const std::string &k, const T &def) const {
        auto it = kv_.find(k);
        if (it == kv_.end()) return def;
        return convert<T>(it->second);
    }
    template <class T>
    void add(const std::string &k, const T &v) {
        std::ostringstream os;
        os << v;
        kv_[k] = os.str();
    }

   private:
    static std::string trim(const std::string &s) {
        const size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
        return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
    }
    template <class T>
    static T convert(const std::string &s) {
        T v{};
        std::istringstream is(s);
        is >> v;
        return v;
    }
    std::map<std::string, std::string> kv_;
};

template <>
inline bool ConfigFile::convert<bool>(const std::string &s) {
    std::string l;
    for (char c : s) l += (char)std::tolower((unsigned char)c);
    return l == "true" || l == "1" || l == "yes";
}
template <>
inline std::string ConfigFile::convert<std::string>(const std::string &s) { return s; }

// Atoms, grid and the maps the caller reads after getSurf.
class DelPhiShared {
   public:
    DelPhiShared(bool buildEpsMap = true, bool buildStatus = true) : buildEpsMap(buildEpsMap), buildStatus(buildStatus) {}
    // xyzr: n*4 doubles
    void loadAtoms(const std::vector<double> &xyzr) {
        if (xyzr.empty() || xyzr.size() % 4) throw std::runtime_error("atoms must be x,y,z,r quadruples");
        atoms = xyzr;
    }
    // DelPhiShared::loadAtoms(fn) (DelphiShared.cpp:309-702): an .xyzr or
    // .pqr file; false (with the message in loadError) when it cannot be read
    bool loadAtoms(const std::string &fn, int maxNumAtoms = -1, double domainShrinkage = 0) {
        nsgpu_atom_set *set = nullptr;
        const int rc = nsgpu_load_atoms(fn.c_str(), maxNumAtoms, domainShrinkage, &set);
        if (rc != NSGPU_OK) {
            loadError = set ? nsgpu_atom_set_error(set) : "cannot load atoms";
            nsgpu_atom_set_free(set);
            return false;
        }
        const double *x = nullptr, *q = nullptr;
        const int32_t n = nsgpu_atom_set_view(set, &x, &q);
        atoms.assign(x, x + 4 * (size_t)n);
        charges.assign(q, q + n);
        nsgpu_atom_set_free(set);
        return true;
    }
    bool buildGrid(double scale_, double perfil_) {
        return nsgpu_build_grid(atoms.data(), (int32_t)(atoms.size() / 4), scale_, perfil_, &grid) == NSGPU_OK;
    }
    int64_t numNodes() const { return (int64_t)grid.nx * grid.ny * grid.nz; }

    bool buildEpsMap, buildStatus;
    std::vector<double> atoms;
    std::vector<double> charges;  // pqr charges (zeros for xyzr)
    std::string loadError;
    nsgpu_grid grid{};
    std::vector<int32_t> epsmap;   // [((k*ny+j)*nx+i)*3+l]
    std::vector<uint8_t> idebmap;  // 1 = out
    std::vector<int32_t> status;   // -1 in, 3 out, >=4 cavity
    std::vector<double> cavitiesSize;
    std::vector<uint8_t> cavitiesFlag;
};

// Base class of the surfaces; the ray caster is the device path.
class Surface {
   public:
    Surface(ConfigFile *cf, DelPhiShared *ds) : delphi(ds) {
        if (cf) {
            inside = cf->read<int>("Inside_code", inside);
            accurateTriangulation = cf->read<bool>("Accurate_Triangulation", accurateTriangulation);
            probe_radius = cf->read<double>("Probe_Radius", probe_radius);
            computeNormals = cf->read<bool>("Compute_Vertex_Normals", computeNormals);
            if (!cf->read<bool>("Analytical_Ray_Vs_Torus_Intersection", true)) precision = NSGPU_PREC_STURM;
            device = cf->read<int>("Device", device);
            patchBased = cf->read<bool>("Patch_Based_Algorithm", patchBased);
            auxGrid2D = cf->read<int>("Max_mesh_auxiliary_grid_2d_size", auxGrid2D);
        }
    }
    virtual ~Surface() {
        if (ctx_) nsgpu_destroy(ctx_);
    }
    Surface(const Surface &) = delete;
    Surface &operator=(const Surface &) = delete;

    virtual bool build() = 0;
    virtual int getNumPatches() = 0;
    virtual bool isPatchBasedRayTracingSupported() { return true; }
    virtual const nsgpu_patches *patches() = 0;

    // Surface::getSurf(&vol, optimizeGrids, fillCav, cavVol): colours the
    // DelPhi maps (read back into delphi) and returns true on success.
    bool getSurf(double *surf_volume, bool /*optimize_grids*/, bool fillCav, double cav_vol) {
        if (!delphi || !ensure_device()) return false;
        nsgpu_surf_opts o{};
        o.build_epsmap = delphi->buildEpsMap;
        o.build_status = delphi->buildStatus;
        o.accurate_triangulation = accurateTriangulation;
        o.fill_cavities = fillCav;
        o.cav_vol = cav_vol;
        o.inside = inside;
        o.precision = precision;
        if (nsgpu_set_vertex_normals(ctx_, computeNormals) != NSGPU_OK) return false;
        if (nsgpu_get_surf(ctx_, &o, &last_) != NSGPU_OK) return false;
        if (surf_volume) *surf_volume = last_.volume;
        // the maps are those of the grid on the device (delphi->grid, uploaded by ensure_device)
        const int64_t nv = (int64_t)grid_.nx * grid_.ny * grid_.nz;
        if (delphi->buildEpsMap) {
            delphi->epsmap.resize(nv * 3);
            delphi->idebmap.resize(nv);
            if (nsgpu_read_epsmap(ctx_, delphi->epsmap.data()) || nsgpu_read_idebmap(ctx_, delphi->idebmap.data()))
                return false;
        }
        if (delphi->buildStatus) {
            delphi->status.resize(nv);
            if (nsgpu_read_status(ctx_, delphi->status.data())) return false;
            delphi->cavitiesSize.resize(last_.num_cavities);
            delphi->cavitiesFlag.resize(last_.num_cavities);
            if (last_.num_cavities > 0 &&
                nsgpu_read_cavities(ctx_, delphi->cavitiesSize.data(), delphi->cavitiesFlag.data()) != NSGPU_OK)
                return false;
        }
        return true;
    }

    // Surface::triangulateSurface: returns the area, fills vertList/triList.
    double triangulateSurface(bool outputMesh = true, bool /*buildAtomsMapHere*/ = false, double /*iso*/ = 0.0) {
        nsgpu_mesh_info m{};
        if (!ctx_ || nsgpu_triangulate(ctx_, &m) != NSGPU_OK) return 0.0;
        if (outputMesh) {
            vertList.resize(m.num_vertices * 3);
            triList.resize(m.num_triangles * 3);
            nsgpu_read_mesh(ctx_, vertList.data(), triList.data());
            if (computeNormals) {
                normList.resize(m.num_vertices * 3);
                nsgpu_read_normals(ctx_, normList.data());
            }
        }
        return m.area;
    }

    const char *lastError() const {
        if (!err_.empty()) return err_.c_str();
        return ctx_ ? nsgpu_last_error(ctx_) : "no device context";
    }
    const nsgpu_surf_result &lastResult() const { return last_; }

    int inside = 5;
    bool accurateTriangulation = true;
    double probe_radius = 1.4;
    double randDisplacement = 1e-4;
    int precision = NSGPU_PREC_FP64;
    bool computeNormals = false;  // Compute_Vertex_Normals
    int device = 0;
    bool patchBased = true;  // Patch_Based_Algorithm (triangle meshes; ses and skin are always patch based)
    int auxGrid2D = 100;     // Max_mesh_auxiliary_grid_2d_size
    std::vector<double> vertList;
    std::vector<int32_t> triList;
    std::vector<double> normList;  // filled when computeNormals

   protected:
    // the surface's device representation (patches by default)
    virtual bool upload() {
        const nsgpu_patches *p = patches();
        return p && nsgpu_set_patches(ctx_, p, probe_radius) == NSGPU_OK;
    }
    // the device holds delphi's current grid and the surface: a grid rebuilt
    // since the last upload (another scale, other atoms) is uploaded again
    bool ensure_device() {
        if (!ctx_ && nsgpu_create(device, &ctx_) != NSGPU_OK) return false;
        if (uploaded_ && std::memcmp(&grid_, &delphi->grid, sizeof(nsgpu_grid)) != 0) uploaded_ = false;
        if (!uploaded_) {
            if (nsgpu_set_grid(ctx_, &delphi->grid) != NSGPU_OK) return false;
            grid_ = delphi->grid;
            if (!upload()) return false;
            uploaded_ = true;
        }
        return true;
    }
    DelPhiShared *delphi;
    std::string err_;
    nsgpu_ctx *ctx_ = nullptr;
    bool uploaded_ = false;
    nsgpu_grid grid_{};  // the grid on the device
    nsgpu_surf_result last_{};
};

// "ses": host patch construction (jitter + free-probe cells), device ray casting
class ConnollySurface : public Surface {
   public:
    ConnollySurface(ConfigFile *cf, DelPhiShared *ds) : Surface(cf, ds) {}
    ~ConnollySurface() override {
        if (ps_) nsgpu_patch_set_free(ps_);
    }
    bool build() override {
        if (!delphi || delphi->atoms.empty()) return false;
        jittered_ = delphi->atoms;
        int32_t seed = 1234;  // SEED, tools.h:1827
        nsgpu_jitter(jittered_.data(), (int32_t)(jittered_.size() / 4), randDisplacement, &seed);
        if (ps_) nsgpu_patch_set_free(ps_);
        ps_ = nullptr;
        if (nsgpu_ses_build(jittered_.data(), (int32_t)(jittered_.size() / 4), probe_radius, &ps_) != NSGPU_OK)
            return false;
        nsgpu_patch_set_view(ps_, &view_);
        uploaded_ = false;
        return true;
    }
    int getNumPatches() override { return ps_ ? view_.n : 0; }
    const nsgpu_patches *patches() override { return ps_ ? &view_ : nullptr; }

   private:
    std::vector<double> jittered_;
    nsgpu_patch_set *ps_ = nullptr;
    nsgpu_patches view_{};
};

// "mesh": a triangle mesh read from an OFF file (MeshSurface::loadOFF,
// MeshSurface.cpp:2350-2479: coordinates read as floats), cast on the device
// with the ray-based (Patch_Based_Algorithm = false) or patch-based rule.
class MeshSurface : public Surface {
   public:
    MeshSurface(ConfigFile *cf, DelPhiShared *ds) : Surface(cf, ds) {
        if (cf) fileName = cf->read<std::string>("Surface_File_Name", fileName);
        computeNormals = false;  // vertex normals need analytic patches on the device
    }
    bool build() override { return loadOFF(fileName); }
    bool loadOFF(const std::string &fn) {
        std::ifstream in(fn);
        if (!in) {
            err_ = "cannot read " + fn;
            return false;
        }
        std::string line;
        bool tag = false;
        long nv = -1, nt = -1;
        while (std::getline(in, line)) {
            if (line.size() <= 1) continue;
            if (!tag && line[0] == 'O') {
                tag = true;
                continue;
            }
            if (line[0] == '#') continue;
            if (tag) {
                std::istringstream h(line);
                h >> nv >> nt;
                break;
            }
        }
        if (!tag || nv <= 0 || nt <= 0) {
            err_ = "cannot read the OFF header of " + fn;
            return false;
        }
        verts.clear();
        tris.clear();
        while ((long)verts.size() < 3 * nv && std::getline(in, line)) {
            if (line.size() <= 1) continue;
            float a = 0, b = 0, c = 0;
            std::sscanf(line.c_str(), "%f %f %f", &a, &b, &c);
            verts.insert(verts.end(), {(double)a, (double)b, (double)c});
        }
        while ((long)tris.size() < 3 * nt && std::getline(in, line)) {
            if (line.size() <= 1) continue;
            int k = 0, a = 0, b = 0, c = 0;
            std::sscanf(line.c_str(), "%d %d %d %d", &k, &a, &b, &c);
            if (k != 3) {
                err_ = "non triangular mesh";
                return false;
            }
            tris.insert(tris.end(), {a, b, c});
        }
        uploaded_ = false;
        return (long)verts.size() == 3 * nv && (long)tris.size() == 3 * nt;
    }
    int getNumPatches() override { return (int)(tris.size() / 3); }
    const nsgpu_patches *patches() override { return nullptr; }
    std::string fileName = "mesh.off";
    std::vector<double> verts;
    std::vector<int32_t> tris;

   protected:
    bool upload() override {
        return nsgpu_set_mesh(ctx_, verts.data(), (int32_t)(verts.size() / 3), tris.data(), (int32_t)(tris.size() / 3),
                              patchBased ? NSGPU_MESH_PATCH_BASED : NSGPU_MESH_RAY_BASED, auxGrid2D) == NSGPU_OK;
    }
};

// "blobby": the blobby mesh built on the device from the atoms (Blobbyness,
// cutoff 6 A), then cast as a ray-based mesh.  As in the reference, it needs
// Patch_Based_Algorithm = false (main.cpp:1154-1161).
class BlobbySurface : public Surface {
   public:
    BlobbySurface(ConfigFile *cf, DelPhiShared *ds) : Surface(cf, ds) {
        if (cf) B = cf->read<double>("Blobbyness", B);
        if (B >= 0) B = -2.5;  // setBlobbyness, BlobbySurface.cpp:46-54
        computeNormals = false;
    }
    bool build() override {
        if (patchBased) {
            err_ = "This surface type does not support patch based ray-tracing: set Patch_Based_Algorithm = false";
            return false;
        }
        if (!delphi || delphi->atoms.empty()) return false;
        uploaded_ = false;
        return ensure_device();
    }
    int getNumPatches() override { return 0; }  // BlobbySurface.cpp:71-74
    const nsgpu_patches *patches() override { return nullptr; }
    double B = -2.5;
    double cutoff = 6.0;  // DEFAULT_CUTOFF, BlobbySurface.h:19
    nsgpu_mesh_info info{};

   protected:
    // the blobby mesh lives on the grid: built again whenever the grid is
    bool upload() override {
        return nsgpu_blobby_build(ctx_, delphi->atoms.data(), (int32_t)(delphi->atoms.size() / 4), B, cutoff,
                                  auxGrid2D, &info) == NSGPU_OK;
    }
};

typedef Surface *(*surface_instantiator)(ConfigFile *, DelPhiShared *);

class SurfaceFactory {
   public:
    void add(const std::string &name, surface_instantiator fn) { reg_[name] = fn; }
    // reads key "Surface"; NULL when the name is unknown (SurfaceFactory.h:29-38)
    Surface *create(ConfigFile *conf, DelPhiShared *ds) {
        const std::string name = conf ? conf->read<std::string>("Surface", std::string("ses")) : "ses";
        auto it = reg_.find(name);
        return it == reg_.end() ? nullptr : it->second(conf, ds);
    }
    bool has(const std::string &name) const { return reg_.count(name) != 0; }

   private:
    std::map<std::string, surface_instantiator> reg_;
};

inline SurfaceFactory &surfaceFactory() {
    static SurfaceFactory f;
    static bool init = [] {
        f.add("ses", [](ConfigFile *c, DelPhiShared *d) -> Surface * { return new ConnollySurface(c, d); });
        f.add("mesh", [](ConfigFile *c, DelPhiShared *d) -> Surface * { return new MeshSurface(c, d); });
        f.add("blobby", [](ConfigFile *c, DelPhiShared *d) -> Surface * { return new BlobbySurface(c, d); });
        return true;
    }();
    (void)init;
    return f;
}

}  // namespace nsgpu
