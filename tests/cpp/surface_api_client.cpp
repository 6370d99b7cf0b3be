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

int main(int argc, char **argv) {
    if (argc > 1 && (std::strcmp(argv[1], "blobby") == 0 || std::strcmp(argv[1], "mesh") == 0))
        return mesh_modes(argc, argv);
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
