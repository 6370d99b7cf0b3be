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
//   SkinSurface "skin"    SkinSurface.h:516-526 (mixed cells from a regular
//                         triangulation the caller plugs in)
//   SurfaceFactory        SurfaceFactory.h:15-65 (name -> instantiator, key "Surface")
//   normalMode            main.cpp:1032-1204 (init's keys and checks), :1253-1320
//   NanoShaper            nanoshaper.h:294-424, nanoshaper.cpp (the C++ API)
// Everything on the device goes through the extern "C" entry points; this
// header adds no symbols to libnsgpu.so.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <array>
#include <cstring>
#include <fstream>
#include <functional>
#include <map>
#include <memory>
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
    // boundary grid points of getSurf with projBGP (Surface.cpp:1544-1575):
    // ibgp [nbgp*3] 0-based node indices, scspos/scsnor [nbgp*3]
    int nbgp = 0;
    std::vector<int32_t> ibgp;
    std::vector<double> scspos, scsnor;

    // DelPhiShared::saveBGP (DelphiShared.cpp:1421-1455): rootFile + fname +
    // ".projections.txt" (count, then 1-based indices, projection and normal
    // per line) and rootFile + "surf.txt" (projection and normal per line)
    bool saveBGP(const std::string &rootFile, const std::string &fname) const {
        if (scspos.size() < (size_t)nbgp * 3 || scsnor.size() < (size_t)nbgp * 3) return false;
        FILE *fp = std::fopen((rootFile + fname + ".projections.txt").c_str(), "w");
        FILE *fp2 = std::fopen((rootFile + "surf.txt").c_str(), "w");
        if (!fp || !fp2) {
            if (fp) std::fclose(fp);
            if (fp2) std::fclose(fp2);
            return false;
        }
        std::fprintf(fp, "%d\n", nbgp);
        for (int i = 0; i < nbgp; ++i) {
            const double *p = &scspos[3 * (size_t)i], *q = &scsnor[3 * (size_t)i];
            std::fprintf(fp, "%d %d %d %lf %lf %lf %lf %lf %lf", ibgp[3 * (size_t)i] + 1, ibgp[3 * (size_t)i + 1] + 1,
                         ibgp[3 * (size_t)i + 2] + 1, p[0], p[1], p[2], q[0], q[1], q[2]);
            std::fprintf(fp2, "%lf %lf %lf %lf %lf %lf", p[0], p[1], p[2], q[0], q[1], q[2]);
            if (i < nbgp - 1) {
                std::fprintf(fp, "\n");
                std::fprintf(fp2, "\n");
            }
        }
        std::fclose(fp);
        std::fclose(fp2);
        return true;
    }
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
            projBGP = cf->read<bool>("Project_boundary_grid_points", projBGP);
            auxGrid3D = cf->read<int>("Max_ses_patches_auxiliary_grid_size", auxGrid3D);
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
        // phase 2: boundary grid points projected on the surface (SES)
        delphi->nbgp = 0;
        if (projBGP && delphi->buildEpsMap) {
            int64_t counts[4] = {0, 0, 0, 0};
            if (nsgpu_project_bgp(ctx_, auxGrid3D, counts) != NSGPU_OK) return false;
            delphi->nbgp = (int)counts[0];
            delphi->ibgp.resize((size_t)counts[0] * 3);
            delphi->scspos.resize((size_t)counts[0] * 3);
            delphi->scsnor.resize((size_t)counts[0] * 3);
            if (nsgpu_read_bgp(ctx_, delphi->ibgp.data(), delphi->scspos.data(), delphi->scsnor.data()) != NSGPU_OK)
                return false;
        }
        return true;
    }

    // Surface::triangulateSurface: returns the area, fills vertList/triList.
    double triangulateSurface(bool outputMesh = true, bool /*buildAtomsMapHere*/ = false, double /*iso*/ = 0.0) {
        nsgpu_mesh_info m{};
        if (!ctx_ || nsgpu_triangulate(ctx_, &m) != NSGPU_OK) return 0.0;
        mesh_ = m;
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

    // Surface::smoothSurface (Surface.cpp:13089-13240) on the last mesh, and
    // on its normals when computeNormals is on
    bool smoothSurface() {
        nsgpu_mesh_info m{};
        if (!ctx_ || nsgpu_smooth_surface(ctx_) != NSGPU_OK) return false;
        if (!last_mesh(m)) return false;
        vertList.resize(m.num_vertices * 3);
        triList.resize(m.num_triangles * 3);
        if (nsgpu_read_mesh(ctx_, vertList.data(), triList.data()) != NSGPU_OK) return false;
        if (computeNormals) {
            normList.resize(m.num_vertices * 3);
            if (nsgpu_read_normals(ctx_, normList.data()) != NSGPU_OK) return false;
        }
        return true;
    }
    void setComputeNormals(bool on) { computeNormals = on; }
    // Surface::setProjBGP / getProjBGP (Surface.h:1054-1062)
    void setProjBGP(bool on) { projBGP = on; }
    bool getProjBGP() const { return projBGP; }
    bool getComputeNormals() const { return computeNormals; }
    void setRandDisplacement(double r) { randDisplacement = r; }
    int64_t getNumVertices() const { return (int64_t)(vertList.size() / 3); }
    int64_t getNumTriangles() const { return (int64_t)(triList.size() / 3); }
    // the device context holding the surface (created and filled on first use)
    nsgpu_ctx *context() { return ensure_device() ? ctx_ : nullptr; }
    const nsgpu_grid &deviceGrid() const { return grid_; }

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
    bool projBGP = false;    // Project_boundary_grid_points (SES patch sets)
    int auxGrid3D = 50;      // Max_ses_patches_auxiliary_grid_size
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
    nsgpu_mesh_info mesh_{};  // the last triangulation
    bool last_mesh(nsgpu_mesh_info &m) const {
        m = mesh_;
        return mesh_.num_vertices || mesh_.num_triangles;
    }
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

// "skin": the skin of the jittered atoms (Skin_Surface_Parameter s) from
// the mixed cells the host builds (nsgpu_skin_build) out of the regular
// triangulation of the weighted points nsgpu_skin_points gives.  The
// reference takes that triangulation from CGAL's Regular_triangulation_3
// (SkinSurface.cpp:1271-1811); here the library's own triangulator does it
// (nsgpu_skin_patches), unless the caller plugs in another: it receives the
// (n + 6) weighted points (x, y, z, w) and returns the finite tetrahedra and
// their neighbours (-1 on the hull), as nsgpu_skin_build takes them.
class SkinSurface : public Surface {
   public:
    using Triangulator = std::function<bool(const std::vector<double> &wpts, std::vector<int32_t> &tets,
                                            std::vector<int32_t> &nbrs)>;
    static Triangulator &triangulator() {
        static Triangulator t;
        return t;
    }
    SkinSurface(ConfigFile *cf, DelPhiShared *ds) : Surface(cf, ds) {
        if (cf) s = cf->read<double>("Skin_Surface_Parameter", s);
    }
    ~SkinSurface() override {
        if (ps_) nsgpu_patch_set_free(ps_);
    }
    bool build() override {
        if (!delphi || delphi->atoms.empty()) return false;
        jittered_ = delphi->atoms;
        int32_t seed = 1234;  // SEED, tools.h:1827
        const int32_t n = (int32_t)(jittered_.size() / 4);
        nsgpu_jitter(jittered_.data(), n, randDisplacement, &seed);
        if (ps_) nsgpu_patch_set_free(ps_);
        ps_ = nullptr;
        if (!triangulator()) {
            if (nsgpu_skin_patches(jittered_.data(), n, s, &ps_) != NSGPU_OK) {
                err_ = "the skin's regular triangulation or mixed cells failed";
                return false;
            }
        } else {
            std::vector<double> w((size_t)(n + 6) * 4);
            if (nsgpu_skin_points(jittered_.data(), n, s, w.data()) != NSGPU_OK) return false;
            std::vector<int32_t> tets, nbrs;
            if (!triangulator()(w, tets, nbrs) || tets.size() % 4 || nbrs.size() != tets.size()) {
                err_ = "the regular triangulation failed";
                return false;
            }
            if (nsgpu_skin_build(jittered_.data(), n, s, tets.data(), nbrs.data(), (int32_t)(tets.size() / 4), &ps_) !=
                NSGPU_OK)
                return false;
        }
        nsgpu_patch_set_view(ps_, &view_);
        uploaded_ = false;
        return true;
    }
    int getNumPatches() override { return ps_ ? view_.n : 0; }
    const nsgpu_patches *patches() override { return ps_ ? &view_ : nullptr; }
    double s = 0.45;

   private:
    std::vector<double> jittered_;
    nsgpu_patch_set *ps_ = nullptr;
    nsgpu_patches view_{};
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
        f.add("skin", [](ConfigFile *c, DelPhiShared *d) -> Surface * { return new SkinSurface(c, d); });
        return true;
    }();
    (void)init;
    return f;
}

// The normal-mode keys of the CLI (init, main.cpp:1032-1204) with the
// reference's defaults; error is set for the flag combinations it refuses.
struct Configuration {
    bool buildEpsmaps, buildStatus, fillCavities, accTri, tri, optimizeGrids, smoothing, patchBased, projBGP;
    double cavVol, scale, perfill, domainShrinkage;
    int maxNumAtoms;
    std::string molFile, surface, operativeMode, error;
    explicit Configuration(const ConfigFile &cf) {
        buildEpsmaps = cf.read<bool>("Build_epsilon_maps", false);
        const bool saveEps = cf.read<bool>("Save_eps_maps", false);
        buildStatus = cf.read<bool>("Build_status_map", false);
        const bool saveIdeb = cf.read<bool>("Save_ideb_map", false);
        fillCavities = cf.read<bool>("Cavity_Detection_Filling", false);
        projBGP = cf.read<bool>("Project_boundary_grid_points", false);
        accTri = cf.read<bool>("Accurate_Triangulation", false);
        tri = cf.read<bool>("Triangulation", false);
        operativeMode = cf.read<std::string>("Operative_Mode", "normal");
        patchBased = cf.read<bool>("Patch_Based_Algorithm", true);
        maxNumAtoms = cf.read<int>("Max_Num_Atoms", -1);
        domainShrinkage = cf.read<double>("Domain_Shrinkage", 0.0);
        optimizeGrids = cf.read<bool>("Optimize_Grids", true);
        surface = cf.read<std::string>("Surface", "ses");
        cavVol = cf.read<double>("Conditional_Volume_Filling_Value", 11.4);
        scale = cf.read<double>("Grid_scale", 2.0);
        perfill = cf.read<double>("Grid_perfil", 80.0);
        molFile = cf.read<std::string>("XYZR_FileName", "temp.xyzr");
        smoothing = cf.read<bool>("Smooth_Mesh", false);
        if (!buildEpsmaps && saveEps) error = "Asked to save epsmap without builiding it";
        else if (!buildEpsmaps && projBGP) error = "Cannot project boundary grid points without an epsilon map";
        else if (!accTri && !buildStatus && tri) error = "If non analytical triangulation is enabled status map is needed";
        else if (fillCavities && !buildStatus) error = "If cavity detection is enabled status map is needed";
        else if (saveIdeb && !buildEpsmaps) error = "Idebmap is computed only if epsilon map is enabled";
        else if (patchBased && surface != "ses" && surface != "skin" && surface != "mesh")
            error = "This surface type does not support patch based ray-tracing";
        else if (operativeMode != "normal") error = "only the normal operative mode is built";
    }
};

// normalMode (main.cpp:1253-1320): build, getSurf, and triangulateSurface
// (with the Smooth_Mesh pass) when Triangulation is on.
struct NormalModeResult {
    bool ok = false;
    double volume = 0, area = 0;
    std::string error;
};
inline NormalModeResult normalMode(Surface *surf, DelPhiShared *dg, const Configuration &conf) {
    NormalModeResult r;
    if (!surf || !dg) {
        r.error = "no surface";
        return r;
    }
    if (!surf->build()) {
        r.error = std::string("Surface construction failed: ") + surf->lastError();
        return r;
    }
    if (!surf->getSurf(&r.volume, conf.optimizeGrids, conf.fillCavities, conf.cavVol)) {
        r.error = std::string("getSurf failed: ") + surf->lastError();
        return r;
    }
    if (conf.tri) {
        r.area = surf->triangulateSurface(!conf.smoothing, !conf.smoothing);
        if (conf.smoothing && !surf->smoothSurface()) {
            r.error = std::string("smoothSurface failed: ") + surf->lastError();
            return r;
        }
    }
    r.ok = true;
    return r;
}

// A configuration file run as the CLI runs it: the atom file relative to
// the configuration's directory.
inline NormalModeResult runConfiguration(const std::string &prm, DelPhiShared &dg, std::unique_ptr<Surface> &surf) {
    NormalModeResult r;
    ConfigFile cf(prm);
    Configuration conf(cf);
    if (!conf.error.empty()) {
        r.error = conf.error;
        return r;
    }
    std::string mol = conf.molFile;
    const size_t slash = prm.find_last_of('/');
    if (!mol.empty() && mol[0] != '/' && slash != std::string::npos) mol = prm.substr(0, slash + 1) + mol;
    dg.buildEpsMap = conf.buildEpsmaps;
    dg.buildStatus = conf.buildStatus;
    if (!dg.loadAtoms(mol, conf.maxNumAtoms, conf.domainShrinkage)) {
        r.error = dg.loadError;
        return r;
    }
    if (!dg.buildGrid(conf.scale, conf.perfill)) {
        r.error = "buildGrid failed";
        return r;
    }
    surf.reset(surfaceFactory().create(&cf, &dg));
    if (!surf) {
        r.error = "unknown surface " + conf.surface;
        return r;
    }
    return normalMode(surf.get(), &dg, conf);
}

// ------------------------------------------------------------------ NanoShaper
// Atom of the C++ API (nanoshaper.h:227-292): centre, radius, charge.
struct Atom {
    double pos[3] = {0, 0, 0};
    double radius = 0, charge = 0;
    int dielectric = 0;
    Atom() = default;
    Atom(double x, double y, double z, double r, double q = 0) : pos{x, y, z}, radius(r), charge(q) {}
};

enum surface_type { ses, skin, blobby };

// Grid-ray crossings of colourGrid with collectGridRays (raytracer_datatype.h:9-43):
// per direction (0 x, 1 y, 2 z) and ray (its two other coordinates), the
// kept segment ends along the ray and their normals.
struct crossings_t {
    double point[2] = {0, 0};
    unsigned dir = 0;
    bool init = false;
    std::vector<double> inters, normals;
};
struct map_compare {
    bool operator()(const std::array<double, 2> &a, const std::array<double, 2> &b) const {
        const double tol = 1.e-3;
        if (a[0] - b[0] < -tol) return true;
        if (std::fabs(a[0] - b[0]) <= tol) return a[1] - b[1] < -tol;
        return false;
    }
};
using rays_t = std::array<std::map<std::array<double, 2>, crossings_t, map_compare>, 3>;

// NS::NanoShaper (nanoshaper.h:294-424): atoms in, surface built on the host
// and cast on the device, grid queries and single axis-aligned rays.
class NanoShaper {
   public:
    NanoShaper(const std::vector<Atom> &ats, surface_type flag, double parameter, double stern_layer_thickness = 0.,
               unsigned nt = 1, const std::string *configFile = nullptr) {
        initConstructor(ats, flag, parameter, stern_layer_thickness, nt, configFile);
    }
    NanoShaper() = default;

    // nanoshaper.cpp:30-91: defaults, then the file, then the arguments
    void initConstructor(const std::vector<Atom> &ats, surface_type flag, double parameter, double /*stern*/,
                         unsigned num_cores = 1, const std::string *configFile = nullptr) {
        initConfig(configFile);
        atoms_ = ats;
        currentPanel_ = -1;
        cf_.add<int>("Number_thread", (int)num_cores);
        if (flag == skin) {
            cf_.add<std::string>("Surface", "skin");
            cf_.add<double>("Skin_Surface_Parameter", parameter);
        } else if (flag == ses) {
            cf_.add<std::string>("Surface", "ses");
            cf_.add<double>("Probe_Radius", parameter);
        } else {
            cf_.add<std::string>("Surface", "blobby");
            cf_.add<double>("Blobbyness", parameter);
            cf_.add<bool>("Patch_Based_Algorithm", false);
        }
    }

    // initConfig (nanoshaper.cpp:707-805): the defaults of the API
    void initConfig(const std::string *filePath = nullptr) {
        cf_ = filePath ? ConfigFile(*filePath) : ConfigFile();
        auto def = [&](const char *k, const std::string &v) {
            if (!cf_.keyExists(k)) cf_.parse_line(std::string(k) + " = " + v);
        };
        def("Build_epsilon_maps", "false");
        def("Build_status_map", "true");
        def("Cavity_Detection_Filling", "false");
        def("Conditional_Volume_Filling_Value", "11.4");
        def("Accurate_Triangulation", "false");
        def("Triangulation", "false");
        def("Smooth_Mesh", "true");
        def("Optimize_Grids", "true");
        def("Grid_scale", "1.0");
        def("Grid_perfil", "80.0");
        def("Probe_Radius", "1.4");
        def("Skin_Surface_Parameter", "0.45");
        def("Blobbyness", "-2.5");
        def("Compute_Vertex_Normals", "false");
        def("Patch_Based_Algorithm", "true");
        def("Analytical_Ray_Vs_Torus_Intersection", "true");
        def("Max_Num_Atoms", "-1");
        def("Domain_Shrinkage", "0");
    }

    template <class T>
    T readConfig(const std::string &k) { return cf_.read<T>(k, T()); }
    template <class T>
    void setConfig(const std::string &k, const T &v) { cf_.add<T>(k, v); }

    // buildAnalyticalSurface (nanoshaper.cpp:300-368): grid by perfil, the
    // surface of key "Surface", its host build
    bool buildAnalyticalSurface(double randDisplacement = -1) {
        clean();
        if (cf_.read<bool>("PB_grid_mode", false)) {
            err_ = "PB_grid_mode (explicit grid bounds) is not built";
            return false;
        }
        ds_.reset(new DelPhiShared(cf_.read<bool>("Build_epsilon_maps", false), cf_.read<bool>("Build_status_map", true)));
        std::vector<double> xyzr;
        for (const Atom &a : atoms_) xyzr.insert(xyzr.end(), {a.pos[0], a.pos[1], a.pos[2], a.radius});
        ds_->loadAtoms(xyzr);
        if (!ds_->buildGrid(cf_.read<double>("Grid_scale", 1.0), cf_.read<double>("Grid_perfil", 80.0))) return false;
        surf_.reset(surfaceFactory().create(&cf_, ds_.get()));
        if (!surf_) return false;
        if (randDisplacement >= 0) surf_->setRandDisplacement(randDisplacement);
        return surf_->build();
    }

    bool colourGrid() {
        double v = 0;
        return colourGrid(&v);
    }
    // colourGrid (nanoshaper.cpp:498-690), with the grid rays' crossings
    // collected when setCollectGridRays(true)
    bool colourGrid(double *surf_volume) {
        if (!surf_) return false;
        if (!surf_->getSurf(surf_volume, cf_.read<bool>("Optimize_Grids", true),
                            cf_.read<bool>("Cavity_Detection_Filling", false),
                            cf_.read<double>("Conditional_Volume_Filling_Value", 11.4)))
            return false;
        return collectGridRays_ ? collectRays() : true;
    }

    // triangulate (nanoshaper.cpp:371-388)
    bool triangulate(double *surf_area) {
        if (!surf_ || !ds_) return false;
        const bool smooth = cf_.read<bool>("Smooth_Mesh", true);
        *surf_area = surf_->triangulateSurface(!smooth, !smooth);
        if (*surf_area == 0) return false;
        if (smooth) return surf_->smoothSurface();
        return true;
    }

    void clean() {
        surf_.reset();
        ds_.reset();
        currentPanel_ = -1;
        for (auto &m : rays_) m.clear();
    }

    // panelResolver (nanoshaper.cpp:190-206): x -> panel 0, z -> panel 1, y -> panel 2
    static int panelResolver(unsigned direction) { return direction == 0 ? 0 : direction == 2 ? 1 : direction == 1 ? 2 : -1; }

    bool setDirection(unsigned direction) {
        if (direction > 2 || !surf_) return false;
        currentPanel_ = panelResolver(direction);
        return true;
    }

    // castAxisOrientedRay (nanoshaper.cpp:220-297): the ray from pa to the
    // coordinate coord along the current direction; hits as (coordinate,
    // normal or nullptr), near-coincident ones dropped.  -1 on a wrong
    // direction, a non-progressive ray, no surface or no intersection.
    short castAxisOrientedRay(double *pa, double coord, std::vector<std::pair<double, double *>> &intersections,
                              unsigned direction, bool computeNormals = false) {
        const int desired = panelResolver(direction);
        if (desired < 0 || currentPanel_ != desired || pa[direction] >= coord || !surf_) return -1;
        nsgpu_ctx *ctx = surf_->context();
        if (!ctx) return -1;
        int32_t maxh = 64;
        for (;;) {
            int32_t count = 0;
            hitbuf_.assign(maxh, 0.0);
            normbuf_.assign(computeNormals ? 3 * maxh : 0, 0.0);
            if (nsgpu_cast_rays(ctx, (int32_t)direction, 1, pa, &coord, maxh, &count, hitbuf_.data(),
                                computeNormals ? normbuf_.data() : nullptr) != NSGPU_OK)
                return -1;
            if (count < 0) return -1;
            if (count > maxh) {
                maxh = count;
                continue;
            }
            if (count == 0) return -1;
            for (int32_t q = 0; q < count; ++q)
                intersections.push_back({hitbuf_[q], computeNormals ? &normbuf_[3 * q] : nullptr});
            return 1;
        }
    }

    // getGridPointStatus: +1 outside, -1 inside, 0 unknown
    int getGridPointStatus(unsigned ix, unsigned iy, unsigned iz) const {
        if (!ds_ || ds_->status.empty()) return 0;
        const nsgpu_grid &g = ds_->grid;
        if (ix >= (unsigned)g.nx || iy >= (unsigned)g.ny || iz >= (unsigned)g.nz) return 0;
        return ds_->status[((size_t)iz * g.ny + iy) * g.nx + ix] < 0 ? -1 : +1;
    }
    bool getGridSize(std::vector<unsigned> &p) const {
        if (!ds_ || ds_->status.empty()) return false;
        p = {(unsigned)ds_->grid.nx, (unsigned)ds_->grid.ny, (unsigned)ds_->grid.nz};
        return true;
    }
    bool getGridPointCoordinates(unsigned ix, unsigned iy, unsigned iz, std::vector<double> &p) const {
        if (!ds_ || ds_->status.empty()) return false;
        const nsgpu_grid &g = ds_->grid;
        if (ix >= (unsigned)g.nx || iy >= (unsigned)g.ny || iz >= (unsigned)g.nz) return false;
        p = {g.xmin + ix * g.side, g.ymin + iy * g.side, g.zmin + iz * g.side};
        return true;
    }
    // getNearestGridPoint (nanoshaper.cpp:391-449): indices saturated to the
    // grid; true when p lies outside it (saturation), as the reference
    bool getNearestGridPoint(double p[3], std::vector<unsigned> &indices) const {
        if (!ds_ || ds_->status.empty()) return false;
        const nsgpu_grid &g = ds_->grid;
        const double lo[3] = {g.xmin, g.ymin, g.zmin};
        const int n[3] = {g.nx, g.ny, g.nz};
        bool outside = false;
        indices.clear();
        for (int d = 0; d < 3; ++d) {
            int i = (int)std::floor((p[d] - lo[d]) / g.side + 0.5);
            if (i < 0) {
                i = 0;
                outside = true;
            }
            if (i >= n[d]) {
                i = n[d] - 1;
                outside = true;
            }
            indices.push_back((unsigned)i);
        }
        return outside;
    }

    void setCollectGridRays(bool c) { collectGridRays_ = c; }
    void setGridTol(double t) { gridTol_ = t; }
    void setOptimizeGrids(bool o) { cf_.add<bool>("Optimize_Grids", o); }
    rays_t &getRaysMap() { return rays_; }

    Surface *surface() { return surf_.get(); }
    DelPhiShared *delphi() { return ds_.get(); }
    const char *lastError() const { return !err_.empty() ? err_.c_str() : surf_ ? surf_->lastError() : "no surface"; }

   private:
    // the kept segment ends of every traced grid ray (the pairing rule of the
    // fill) and their analytic normals, cast again along the same grid lines
    bool collectRays() {
        nsgpu_ctx *ctx = surf_->context();
        if (!ctx) return false;
        for (auto &m : rays_) m.clear();
        const nsgpu_grid &g = surf_->deviceGrid();
        const int N = std::max(g.nx, std::max(g.ny, g.nz));
        const int panels = ds_->buildEpsMap ? 3 : (ds_->buildStatus ? 1 : 0);
        const double lo[3] = {g.xmin, g.ymin, g.zmin};
        const int n[3] = {g.nx, g.ny, g.nz};
        for (int slot = 0; slot < 3; ++slot) {
            // slot 0: x rays over (y, z); slot 1: z rays over (x, y); slot 2: y rays over (x, z)
            const int dir = slot == 0 ? 0 : (slot == 1 ? 2 : 1);
            const int a0 = slot == 0 ? 1 : 0, a1 = slot == 1 ? 1 : 2;
            for (int j = 0; j < n[a1]; ++j)
                for (int i = 0; i < n[a0]; ++i) {
                    std::array<double, 2> pt = {lo[a0] + i * g.side, lo[a1] + j * g.side};
                    crossings_t &c = rays_[dir][pt];
                    c.point[0] = pt[0];
                    c.point[1] = pt[1];
                    c.dir = (unsigned)dir;
                    c.init = true;
                }
            if (slot >= panels) continue;
            const int maxh = std::max(1, surf_->lastResult().max_hits_per_ray);
            std::vector<int32_t> cnt((size_t)N * N);
            std::vector<double> t((size_t)N * N * maxh);
            if (nsgpu_read_ray_hits(ctx, slot, maxh, cnt.data(), t.data()) != NSGPU_OK) return false;
            const double start = lo[dir], len = (n[dir] - 1) * g.side;
            std::vector<double> orig, coord;
            std::vector<int> ray;
            for (int r = 0; r < N * N; ++r) {
                if (cnt[r] == 0 || (cnt[r] & 1)) continue;
                const int m = r % N, nn = r / N;
                double o[3];
                o[dir] = start;
                o[a0] = lo[a0] + m * g.side;
                o[a1] = lo[a1] + nn * g.side;
                orig.insert(orig.end(), {o[0], o[1], o[2]});
                coord.push_back(start + len);
                ray.push_back(r);
            }
            const int nr = (int)ray.size();
            std::vector<int32_t> hc(nr);
            std::vector<double> hh((size_t)nr * maxh), hn((size_t)nr * maxh * 3);
            if (nr && nsgpu_cast_rays(ctx, dir, nr, orig.data(), coord.data(), maxh, hc.data(), hh.data(),
                                      hn.data()) != NSGPU_OK)
                return false;
            for (int q = 0; q < nr; ++q) {
                const int r = ray[q];
                crossings_t &c = rays_[dir][{orig[3 * q + a0], orig[3 * q + a1]}];
                const double *T = &t[(size_t)r * maxh];
                for (int h = 0; h + 1 < cnt[r]; h += 2) {
                    if (std::fabs(T[h + 1] - T[h]) < 1e-8) continue;  // EPS_INT pairing
                    for (int e = 0; e < 2; ++e) {
                        const double x = start + len * T[h + e];
                        c.inters.push_back(x);
                        // the normal of the cast depth nearest this one
                        int best = 0;
                        for (int k = 1; k < std::min(hc[q], maxh); ++k)
                            if (std::fabs(hh[(size_t)q * maxh + k] - x) < std::fabs(hh[(size_t)q * maxh + best] - x))
                                best = k;
                        const double *nv = &hn[((size_t)q * maxh + best) * 3];
                        c.normals.insert(c.normals.end(), {nv[0], nv[1], nv[2]});
                    }
                }
            }
        }
        return true;
    }

    ConfigFile cf_;
    std::vector<Atom> atoms_;
    std::unique_ptr<DelPhiShared> ds_;
    std::unique_ptr<Surface> surf_;
    int currentPanel_ = -1;
    bool collectGridRays_ = false;
    double gridTol_ = 1e-2;
    rays_t rays_;
    std::vector<double> hitbuf_, normbuf_;
    std::string err_;
};

}  // namespace nsgpu
