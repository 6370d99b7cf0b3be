// Client of the C++ plugin layer, written the way reference code drives a
// surface (main.cpp:748-790, normalMode :1253-1320): config -> DelPhiShared ->
// surfaceFactory().create -> build -> getSurf -> triangulateSurface.
// argv[1] = "host" runs the host stages only, "gpu" the whole path.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>

#include "nsgpu_surface.hpp"

// "blobby": the blobby surface of the same atoms; "mesh FILE": an OFF mesh
static int mesh_modes(int argc, char **argv) {
    nsgpu::ConfigFile cf;
    std::vector<double> xyzr(330 * 4);
    double radius_used = 0;
    const int n = nsgpu_generate_chain(330, 12.0, 0, xyzr.data(), &radius_used);
    xyzr.resize(n * 4);
    nsgpu::DelPhiShared ds(true, true);
    ds.loadAtoms(xyzr);
    if (!ds.buildGrid(2.0, 80.0)) return 2;
    if (std::strcmp(argv[1], "blobby") == 0) {
        cf.parse_line("Surface = blobby");
        cf.parse_line("Blobbyness = -2.5");
        std::unique_ptr<nsgpu::Surface> pb(nsgpu::surfaceFactory().create(&cf, &ds));
        if (!pb || pb->build()) return 10;  // patch-based ray casting is refused, as in the reference
        cf.parse_line("Patch_Based_Algorithm = false");
    } else {
        if (argc < 3) return 11;
        cf.parse_line("Surface = mesh");
        cf.parse_line(std::string("Surface_File_Name = ") + argv[2]);
        cf.parse_line("Patch_Based_Algorithm = false");
    }
    std::unique_ptr<nsgpu::Surface> surf(nsgpu::surfaceFactory().create(&cf, &ds));
    if (!surf || !surf->build()) {
        std::printf("build failed: %s\n", surf ? surf->lastError() : "no surface");
        return 12;
    }
    double vol = 0;
    if (!surf->getSurf(&vol, true, true, 11.4)) {
        std::printf("getSurf failed: %s\n", surf->lastError());
        return 13;
    }
    const double area = surf->triangulateSurface(true);
    long inside = 0;
    for (int v : ds.status) inside += (v == -1);
    std::printf("volume %.9f area %.9f triangles %zu inside %ld\n", vol, area, surf->triList.size() / 3, inside);
    return 0;
}

// "facade": the NS::NanoShaper API (nanoshaper.h) over the device path
static int facade_mode() {
    std::vector<double> xyzr(330 * 4);
    double radius_used = 0;
    const int n = nsgpu_generate_chain(330, 12.0, 0, xyzr.data(), &radius_used);
    std::vector<nsgpu::Atom> atoms;
    for (int i = 0; i < n; ++i) atoms.emplace_back(xyzr[4 * i], xyzr[4 * i + 1], xyzr[4 * i + 2], xyzr[4 * i + 3]);
    nsgpu::NanoShaper ns(atoms, nsgpu::ses, 1.4, 0.0, 1);
    ns.setConfig<double>("Grid_scale", 2.0);
    ns.setConfig<double>("Grid_perfil", 90.0);
    ns.setConfig<bool>("Build_epsilon_maps", true);
    ns.setConfig<bool>("Accurate_Triangulation", true);
    ns.setConfig<bool>("Compute_Vertex_Normals", true);
    ns.setConfig<bool>("Smooth_Mesh", false);
    if (ns.readConfig<double>("Probe_Radius") != 1.4) return 20;
    if (!ns.buildAnalyticalSurface()) return 21;
    double vol = 0, area = 0;
    ns.setCollectGridRays(true);
    if (!ns.colourGrid(&vol)) return 22;
    if (!ns.triangulate(&area)) return 23;
    std::vector<unsigned> size, idx;
    if (!ns.getGridSize(size) || size.size() != 3) return 24;
    std::vector<double> p0;
    if (!ns.getGridPointCoordinates(size[0] / 2, size[1] / 2, size[2] / 2, p0)) return 25;
    double q[3] = {p0[0] + 0.1, p0[1] - 0.1, p0[2]};
    if (ns.getNearestGridPoint(q, idx) || idx[0] != size[0] / 2 || idx[1] != size[1] / 2) return 26;
    double far[3] = {-1e4, p0[1], p0[2]};
    if (!ns.getNearestGridPoint(far, idx) || idx[0] != 0) return 27;  // saturated
    std::vector<double> c0;
    ns.getGridPointCoordinates(0, size[1] / 2, size[2] / 2, c0);
    std::vector<double> cN;
    ns.getGridPointCoordinates(size[0] - 1, 0, 0, cN);
    // a ray along x through the middle of the grid, before and after setDirection
    double pa[3] = {c0[0], c0[1], c0[2]};
    std::vector<std::pair<double, double *>> hits;
    if (ns.castAxisOrientedRay(pa, cN[0], hits, 0, true) != -1) return 28;  // no direction set yet
    if (!ns.setDirection(0)) return 29;
    if (ns.castAxisOrientedRay(pa, pa[0] - 1.0, hits, 0, true) != -1) return 30;  // not progressive
    if (ns.castAxisOrientedRay(pa, cN[0], hits, 0, true) != 1 || hits.empty()) return 31;
    for (auto &h : hits) {
        const double *nv = h.second;
        if (!nv || std::fabs(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2] - 1.0) > 1e-9) return 32;
    }
    const int status = ns.getGridPointStatus(size[0] / 2, size[1] / 2, size[2] / 2);
    size_t crossings = 0;
    for (auto &kv : ns.getRaysMap()[0]) crossings += kv.second.inters.size();
    for (auto &kv : ns.getRaysMap()[1]) crossings += kv.second.inters.size();
    for (auto &kv : ns.getRaysMap()[2]) crossings += kv.second.inters.size();
    std::printf("facade volume %.9f area %.9f hits %zu first %.12f status %d rays %zu crossings %zu\n", vol, area,
                hits.size(), hits[0].first, status,
                ns.getRaysMap()[0].size() + ns.getRaysMap()[1].size() + ns.getRaysMap()[2].size(), crossings);
    return 0;
}

// "prm FILE": a configuration file as the CLI's normal mode runs it
static int prm_mode(const char *prm) {
    nsgpu::DelPhiShared dg;
    std::unique_ptr<nsgpu::Surface> surf;
    const nsgpu::NormalModeResult r = nsgpu::runConfiguration(prm, dg, surf);
    if (!r.ok) {
        std::printf("error %s\n", r.error.c_str());
        return 40;
    }
    long inside = 0;
    for (int v : dg.status) inside += (v == -1);
    std::printf("prm volume %.9f area %.9f triangles %lld inside %ld atoms %zu\n", r.volume, r.area,
                (long long)surf->getNumTriangles(), inside, dg.atoms.size() / 4);
    return 0;
}

// "rescale": getSurf, then a grid built at another scale and getSurf again
// on the same surface (the maps follow the new grid)
static int rescale_mode() {
    std::vector<double> xyzr(200 * 4);
    double radius_used = 0;
    const int n = nsgpu_generate_chain(200, 10.0, 1, xyzr.data(), &radius_used);
    xyzr.resize(n * 4);
    nsgpu::ConfigFile cf;
    cf.parse_line("Surface = ses");
    cf.parse_line("Accurate_Triangulation = false");
    nsgpu::DelPhiShared ds(true, true);
    ds.loadAtoms(xyzr);
    std::unique_ptr<nsgpu::Surface> surf(nsgpu::surfaceFactory().create(&cf, &ds));
    if (!surf->build()) return 50;
    double v1 = 0, v2 = 0, v3 = 0;
    if (!ds.buildGrid(2.0, 80.0) || !surf->getSurf(&v1, true, true, 11.4)) return 51;
    const size_t n1 = ds.status.size();
    if (!ds.buildGrid(3.0, 80.0) || !surf->getSurf(&v2, true, true, 11.4)) return 52;
    if (ds.status.size() != (size_t)ds.numNodes() || ds.status.size() <= n1) return 53;
    // a fresh surface on the finer grid gives the same volume
    std::unique_ptr<nsgpu::Surface> fresh(nsgpu::surfaceFactory().create(&cf, &ds));
    if (!fresh->build() || !fresh->getSurf(&v3, true, true, 11.4)) return 54;
    std::printf("rescale %.9f %.9f %.9f nodes %zu %zu\n", v1, v2, v3, n1, ds.status.size());
    return v2 == v3 ? 0 : 55;
}

// "bgp DIR": getSurf with Project_boundary_grid_points on the 330-atom chain,
// the projections saved with saveBGP into DIR
static int bgp_mode(const char *dir) {
    std::vector<double> xyzr(330 * 4);
    double radius_used = 0;
    const int n = nsgpu_generate_chain(330, 12.0, 0, xyzr.data(), &radius_used);
    xyzr.resize(n * 4);
    nsgpu::ConfigFile cf;
    cf.parse_line("Surface = ses");
    cf.parse_line("Accurate_Triangulation = false");
    cf.parse_line("Project_boundary_grid_points = true");
    nsgpu::DelPhiShared ds(true, true);
    ds.loadAtoms(xyzr);
    if (!ds.buildGrid(2.0, 90.0)) return 60;
    std::unique_ptr<nsgpu::Surface> surf(nsgpu::surfaceFactory().create(&cf, &ds));
    double vol = 0;
    if (!surf->build() || !surf->getSurf(&vol, true, true, 11.4)) return 61;
    if (!ds.saveBGP(std::string(dir) + "/", "mol")) return 62;
    double sum = 0;
    for (double v : ds.scspos) sum += v;
    std::printf("bgp %d sum %.9f\n", ds.nbgp, sum);
    return ds.nbgp > 0 ? 0 : 63;
}

// skin through the factory with the library's own regular triangulation
static int skin_host_mode() {
    nsgpu::ConfigFile cf;
    cf.parse_line("Surface = skin");
    cf.parse_line("Skin_Surface_Parameter = 0.45");
    std::vector<double> xyzr(330 * 4);
    double radius_used = 0;
    const int n = nsgpu_generate_chain(330, 12.0, 0, xyzr.data(), &radius_used);
    xyzr.resize(n * 4);
    nsgpu::DelPhiShared ds(true, true);
    ds.loadAtoms(xyzr);
    if (!ds.buildGrid(2.0, 80.0)) return 2;
    std::unique_ptr<nsgpu::Surface> surf(nsgpu::surfaceFactory().create(&cf, &ds));
    if (!surf) return 3;
    if (!surf->build()) {
        std::printf("build failed: %s\n", surf->lastError());
        return 5;
    }
    std::printf("skin patches %d\n", surf->getNumPatches());
    return 0;
}

int main(int argc, char **argv) {
    if (argc > 1 && (std::strcmp(argv[1], "blobby") == 0 || std::strcmp(argv[1], "mesh") == 0))
        return mesh_modes(argc, argv);
    if (argc > 1 && std::strcmp(argv[1], "facade") == 0) return facade_mode();
    if (argc > 2 && std::strcmp(argv[1], "prm") == 0) return prm_mode(argv[2]);
    if (argc > 2 && std::strcmp(argv[1], "bgp") == 0) return bgp_mode(argv[2]);
    if (argc > 1 && std::strcmp(argv[1], "rescale") == 0) return rescale_mode();
    if (argc > 1 && std::strcmp(argv[1], "skinhost") == 0) return skin_host_mode();
    const bool gpu = argc > 1 && std::strcmp(argv[1], "gpu") == 0;
    nsgpu::ConfigFile cf;
    cf.parse_line("Surface = ses   # surface type");
    cf.parse_line("Grid_scale = 2.0");
    cf.parse_line("Grid_perfil = 90.0");
    cf.parse_line("Probe_Radius = 1.4");
    cf.parse_line("Accurate_Triangulation = true");
    cf.parse_line("Compute_Vertex_Normals = true");
    std::vector<double> xyzr(330 * 4);
    double radius_used = 0;
    const int n = nsgpu_generate_chain(330, 12.0, 0, xyzr.data(), &radius_used);
    xyzr.resize(n * 4);
    nsgpu::DelPhiShared ds(true, true);
    ds.loadAtoms(xyzr);
    if (!ds.buildGrid(cf.read<double>("Grid_scale", 2.0), cf.read<double>("Grid_perfil", 80.0))) return 2;
    std::unique_ptr<nsgpu::Surface> surf(nsgpu::surfaceFactory().create(&cf, &ds));
    if (!surf) return 3;
    nsgpu::ConfigFile bad;
    bad.parse_line("Surface = nonexistent");
    if (nsgpu::surfaceFactory().create(&bad, &ds) != nullptr) return 4;
    if (!surf->build()) return 5;
    std::printf("atoms %d grid %d patches %d\n", n, ds.grid.nx, surf->getNumPatches());
    if (!gpu) return 0;
    double vol = 0;
    if (!surf->getSurf(&vol, true, true, 11.4)) {
        std::printf("getSurf failed: %s\n", surf->lastError());
        return 6;
    }
    const double area = surf->triangulateSurface(true);
    if (surf->normList.size() != surf->vertList.size()) return 7;
    for (size_t q = 0; q + 2 < surf->normList.size(); q += 3) {
        const double *nv = &surf->normList[q];
        if (std::fabs(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2] - 1.0) > 1e-9) return 8;
    }
    long inside = 0;
    for (int v : ds.status) inside += (v == -1);
    std::printf("volume %.9f area %.9f triangles %zu inside %ld cavities %zu\n", vol, area, surf->triList.size() / 3,
                inside, ds.cavitiesSize.size());
    return 0;
}
