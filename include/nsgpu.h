/*
 * nsgpu.h - C ABI of the MI355X (gfx950) grid ray-casting surface builder.
 *
 * This is the drop-in boundary for NanoShaper's grid ray caster.  Every entry
 * point takes plain pointers and sizes; no C++ or torch types cross it.  The
 * host side of the reference (config parsing, atom loading, the SES patch
 * construction) runs on the CPU and hands the patch arrays over with
 * nsgpu_set_patches(); everything from there to the maps, the cavity labels and
 * the marching-cubes mesh runs in HIP kernels on the device.
 *
 * Citations name the reference interface each entry point replaces
 * (paths relative to the reference's src/ directory).
 *
 * Return convention: 0 on success, a negative NSGPU_E* code on failure; the
 * message of the last failure is available from nsgpu_last_error().
 */
#ifndef NSGPU_H
#define NSGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NSGPU_ABI_VERSION 1

/* error codes */
#define NSGPU_OK 0
#define NSGPU_E_ARG -1      /* invalid argument (null pointer, bad size)        */
#define NSGPU_E_DEVICE -2   /* HIP runtime error                                 */
#define NSGPU_E_STATE -3    /* call order violated (e.g. get_surf before grid)   */
#define NSGPU_E_NOMEM -4    /* device or host allocation failed                  */
#define NSGPU_E_BUILD -5    /* host patch construction failed                    */

/* DelPhi map codes (DelphiShared.h:21-33, ConnollySurface.cpp:214). */
#define NSGPU_STATUS_INSIDE -1
#define NSGPU_STATUS_TEMP_OUT 2
#define NSGPU_STATUS_OUT 3
#define NSGPU_STATUS_FIRST_CAV 4

/* ------------------------------------------------------------------------ */
/* Grid descriptor: the result of DelPhiShared::buildGrid                    */
/* (DelphiShared.cpp:1038-1201).  Node i sits at x[i] = xmin + i*side.       */
/* ------------------------------------------------------------------------ */
typedef struct {
    int32_t nx, ny, nz;
    int32_t pad_;
    double xmin, ymin, zmin;
    double xmax, ymax, zmax;
    double scale;  /* grid points per Angstrom                  */
    double side;   /* 1/scale                                   */
    double hside;  /* side/2                                    */
    double A;      /* side*side, pixel area of one ray           */
    double rmaxdim;
    double baricenter[3];
} nsgpu_grid;

/* ------------------------------------------------------------------------ */
/* Patch set (SoA).  One entry per SES cell of ConnollySurface              */
/* (ConnollySurface.h:31-167).  Each patch is a primitive (sphere or torus)  */
/* plus a feasibility region; a ray/primitive hit is kept when the hit point */
/* is feasible (ConnollySurface::isFeasible, ConnollySurface.cpp:4769-5021). */
/*                                                                          */
/* All clipping tests are normalised to "value > 0 (or < 0) rejects":       */
/*   POINT        data[off + 4c .. +3] = clipping sphere (cx,cy,cz,r^2);    */
/*                hit rejected when |p-c|^2 - r^2 < 0.                     */
/*   FACET_REG/   data[off + 4c .. +3] = plane (a,b,c,d);                  */
/*   FACET_SING   hit rejected when a*x+b*y+c*z+d > 0.  Singular facets     */
/*                carry the 4 trimming planes, regular facets planes 1..3,  */
/*                both followed by their self-intersection planes with the  */
/*                sign already applied.                                    */
/*   TORUS_REG/   data[off .. off+13] = torus centre (3), inverse frame     */
/*   TORUS_SING   rotation (9, row major), major radius R, self-            */
/*                intersection radius (<0 = none); then cnt pairs of planes */
/*                (8 doubles per pair) with the sign already applied;       */
/*                flags bit c = pair c is "acute" (AND) rather than OR.     */
/*                Singular tori have cnt = 0.                               */
/* bsph[4*p..] is the bounding sphere intersected first: the probe sphere   */
/* (facets), the torus clipping sphere (tori) or the atom (points).          */
/* The torus minor radius and the facet sphere radius are the probe radius. */
/* ------------------------------------------------------------------------ */
#define NSGPU_PATCH_POINT 0
#define NSGPU_PATCH_FACET_REG 1
#define NSGPU_PATCH_FACET_SING 2
#define NSGPU_PATCH_TORUS_REG 3
#define NSGPU_PATCH_TORUS_SING 4
#define NSGPU_TORUS_HDR 14
/* Skin mixed cells (SkinSurface.h:18-101), quadric patches:              */
/*   data[off .. off+6)   bounding box (down x, y, z, up x, y, z) of the    */
/*                        cell's points (the atom's sphere for SKIN_POINT); */
/*   data[off+6 .. +16)   quadric Q in the reference's order q11 q22 q33    */
/*                        q44 q12 q23 q34 q13 q24 q14 (SkinSurface.cpp:4508);*/
/*   then cnt planes (a,b,c,d): a hit is rejected when a x+b y+c z+d > 0    */
/*   (the point cell's planes stored negated, so one rule serves all).      */
/* bsph: the atom (centre, radius) of a SKIN_POINT cell, zeros otherwise.   */
#define NSGPU_PATCH_SKIN_POINT 5 /* Del0 x Vor3: sphere of the atom        */
#define NSGPU_PATCH_SKIN_EDGE 6  /* Del1 x Vor2: hyperboloid               */
#define NSGPU_PATCH_SKIN_FACET 7 /* Del2 x Vor1: hyperboloid               */
#define NSGPU_PATCH_SKIN_TETRA 8 /* Del3 x Vor0: sphere                    */
#define NSGPU_SKIN_HDR 16

typedef struct {
    int32_t n;           /* number of patches                           */
    int32_t n_data;      /* doubles in data[]                            */
    const int32_t *kind; /* [n]   NSGPU_PATCH_*                          */
    const int32_t *off;  /* [n]   offset into data[]                     */
    const int32_t *cnt;  /* [n]   clip items (spheres/planes/plane pairs) */
    const int32_t *flags;/* [n]   torus: acute bit per plane pair         */
    const int32_t *atom; /* [n*3] atom ids (point: id,-1,-1; torus: i,j,-1; facet: i,j,k) */
    const double *bsph;  /* [n*4] bounding sphere                        */
    const double *data;  /* [n_data] pool                                */
} nsgpu_patches;

/* ------------------------------------------------------------------------ */
/* Surface::getSurf options (Surface.h:890) and results.                    */
/* ------------------------------------------------------------------------ */
#define NSGPU_PREC_FP64 0  /* reference arithmetic throughout (closed-form quartic)      */
#define NSGPU_PREC_MIXED 1 /* torus roots: FP32-seeded resolvent, FP64 Newton polish   */
#define NSGPU_PREC_STURM 2 /* torus roots by Sturm bisection, the reference's path for  */
                           /* Analytical_Ray_Vs_Torus_Intersection = false              */

typedef struct {
    int32_t build_epsmap;           /* DelPhiShared::buildEpsMap (epsmap + idebmap)      */
    int32_t build_status;           /* DelPhiShared::buildStatus (status map)            */
    int32_t accurate_triangulation; /* Surface::accurateTriangulation (edge-ray panels)  */
    int32_t fill_cavities;          /* getSurf(..., fillCav, cav_vol)                    */
    double cav_vol;                 /* Conditional_Volume_Filling_Value                  */
    int32_t inside;                 /* epsmap inside code (5 for ses)                    */
    int32_t precision;              /* NSGPU_PREC_*                                      */
} nsgpu_surf_opts;

typedef struct {
    double volume;         /* Surface::getSurf *surf_volume (Surface.cpp:1677-1684)   */
    int64_t total_rays;    /* rays traced over all panels and phases                  */
    int64_t odd_rays;      /* rays with an odd hit count (dropped, as the reference)  */
    int64_t total_hits;    /* feasible ray/patch hits                                 */
    int64_t candidates;    /* ray/patch pairs tested after bounding-box culling       */
    int32_t num_cavities;  /* getCavitiesWithBilevelStatusMap() return value          */
    int32_t max_hits_per_ray;
} nsgpu_surf_result;

typedef struct {
    int64_t num_vertices;
    int64_t num_triangles;
    int64_t num_analytic_vertices; /* vertices from edge-ray hits (assembleVerticesList) */
    double area;                   /* triangulateSurface return value                  */
} nsgpu_mesh_info;

typedef struct nsgpu_ctx nsgpu_ctx;

/* -------------------- context --------------------------------------------- */
/* One context = one DelPhiShared grid + one Surface on one device.           */
/* Calls on one context are not thread safe; distinct contexts are.           */
/* HIP devices visible to the library (0 without one). */
int nsgpu_device_count(void);
int nsgpu_create(int device, nsgpu_ctx **out);
void nsgpu_destroy(nsgpu_ctx *ctx);
const char *nsgpu_last_error(const nsgpu_ctx *ctx);
int nsgpu_abi_version(void);

/* -------------------- host stages that stay on the CPU --------------------- */
/* DelPhiShared::loadAtoms (DelphiShared.cpp:309-702): atoms of an .xyzr file
 * ("x y z r" per line) or, when the path contains ".pqr", of the ATOM/HETATM
 * records of a PQR file (10 or 11 fields, radius below 1e-20 skipped).
 * max_atoms > 0 stops after that many (Max_Num_Atoms); domain_shrinkage in
 * (0, 1) keeps the atoms inside the centred bounding box shrunk by that factor
 * (Domain_Shrinkage).  Fewer than 4 atoms or all radii null fail, as in the
 * reference.  The set is library-owned; on failure *out still holds the
 * message (nsgpu_atom_set_error) and must be freed. */
typedef struct nsgpu_atom_set nsgpu_atom_set;
int nsgpu_load_atoms(const char *path, int32_t max_atoms, double domain_shrinkage, nsgpu_atom_set **out);
/* number of atoms; *xyzr (n*4 doubles) and *charge (n, zeros for xyzr files) */
int32_t nsgpu_atom_set_view(const nsgpu_atom_set *s, const double **xyzr, const double **charge);
const char *nsgpu_atom_set_error(const nsgpu_atom_set *s);
void nsgpu_atom_set_free(nsgpu_atom_set *s);

/* Nearest atom of each point (m points, 3 doubles each): the acceleration
 * grid of Surface::buildAtomsMap (Surface.cpp:1695-1749) over the grid's
 * baricenter and rmaxdim, and the 27-cell search of Surface::vdwAccessible
 * (Surface.cpp:16663-16713): the first atom of smallest |p - c|^2 - r^2 in
 * the reference's cell order, or -1.  The vertex-atom map of the mesh files
 * and, over the cavities' boundary nodes, Surface::getCavitiesAtoms
 * (Surface.cpp:2129-2245). */
int nsgpu_nearest_atoms(const nsgpu_grid *grid, const double *xyzr, int64_t n, const double *points, int64_t m,
                        int32_t *out);

/* DelPhiShared::buildGrid(scale, perfil) (DelphiShared.cpp:1038-1201),
 * bounds from DelPhiShared::getBounds (DelphiShared.cpp:775-797).
 * xyzr: n*4 doubles (x, y, z, radius). */
int nsgpu_build_grid(const double *xyzr, int32_t n, double scale, double perfil,
                     nsgpu_grid *out);

/* Atom jitter of ConnollySurface::buildConnollyCGAL (ConnollySurface.cpp:2484-2489)
 * with the Park-Miller generator randnum() (tools.cpp:64-84).  *seed is the
 * generator state (the reference starts at SEED = 1234, tools.h:1827) and is
 * advanced in place. */
void nsgpu_jitter(double *xyzr, int32_t n, double amplitude, int32_t *seed);

/* SES patch construction, the host stage ConnollySurface::build
 * (ConnollySurface.cpp:2326-3316) emits: point, facet and edge cells with
 * their clipping geometry.  Built here without CGAL by free-probe enumeration
 * (every probe position tangent to three atoms that no atom overlaps).  The
 * patch set is owned by the library. */
typedef struct nsgpu_patch_set nsgpu_patch_set;
int nsgpu_ses_build(const double *xyzr, int32_t n, double probe_radius,
                    nsgpu_patch_set **out);
void nsgpu_patch_set_view(const nsgpu_patch_set *ps, nsgpu_patches *view);
void nsgpu_patch_set_free(nsgpu_patch_set *ps);
/* counts[5] = number of patches of each NSGPU_PATCH_* kind */
void nsgpu_patch_set_counts(const nsgpu_patch_set *ps, int64_t counts[5]);

/* Skin mixed complex, the host stage SkinSurface::buildSkinCGAL
 * (SkinSurface.cpp:1271-1811) minus the triangulation itself: CGAL's
 * Regular_triangulation_3 in the reference, any regular triangulator here.
 * nsgpu_skin_points writes the (n + 6) weighted points to triangulate: the
 * atoms (x, y, z, r^2/s) and six dummy points of weight -1 (:1290-1345).
 * nsgpu_skin_build takes the finite tetrahedra of their regular
 * triangulation (tets [ntets*4] point indices) with their neighbours (nbrs
 * [ntets*4], the tetrahedron opposite vertex k, -1 on the hull) and emits the
 * mixed cells (BuildDelaunayTetraCells/PointCells/EdgeCells/FacetCells,
 * :348-1210) as a patch set of NSGPU_PATCH_SKIN_* kinds.  xyzr are the
 * atoms after jitter; s is Skin_Surface_Parameter (0 < s < 1). */
int nsgpu_skin_points(const double *xyzr, int32_t n, double s, double *wpts);
/* The regular triangulation of those weighted points, in place of CGAL's
 * Regular_triangulation_3 (SkinSurface.cpp:1271-1345): the six dummies as an
 * octahedron, then the atoms inserted in Morton order with exact predicates;
 * hidden atoms are left out, as CGAL leaves them.  The handle holds the
 * finite tetrahedra and their neighbours in nsgpu_skin_build's layout
 * (count, then read into tets[count*4] and nbrs[count*4]).  NSGPU_E_BUILD if
 * an atom lies outside the dummies' hull or meets an exact degeneracy the
 * insertion cannot star. */
typedef struct nsgpu_regular_tri nsgpu_regular_tri;
int nsgpu_skin_triangulate(const double *xyzr, int32_t n, double s, nsgpu_regular_tri **out);
/* The same for any weighted points wpts[n*4] (x, y, z, w): six dummies of
 * weight -1 around them (indices n .. n+5 in the tetrahedra) close the hull.
 * With weights (r + probe)^2 the tetrahedra carry the weighted alpha complex
 * of ConnollySurface (CGAL's Fixed_alpha_shape_3 at alpha 0). */
int nsgpu_regular_triangulate(const double *wpts, int32_t n, nsgpu_regular_tri **out);
int32_t nsgpu_regular_tri_count(const nsgpu_regular_tri *tri);
void nsgpu_regular_tri_read(const nsgpu_regular_tri *tri, int32_t *tets, int32_t *nbrs);
void nsgpu_regular_tri_free(nsgpu_regular_tri *tri);
/* The whole host stage of SkinSurface::buildSkinCGAL: nsgpu_skin_points,
 * nsgpu_skin_triangulate and nsgpu_skin_build in one call. */
int nsgpu_skin_patches(const double *xyzr, int32_t n, double s, nsgpu_patch_set **out);
/* The triangulation's exact predicates (exposed for their tests):
 * orient3d = sign of det[b - a; c - a; d - a]; power_test on weighted points
 * (x, y, z, w) = the sign that is negative when p is in conflict with the
 * positively oriented cell (a, b, c, d). */
int nsgpu_orient3d(const double *a, const double *b, const double *c, const double *d);
int nsgpu_power_test(const double *a, const double *b, const double *c, const double *d, const double *p);
int nsgpu_skin_build(const double *xyzr, int32_t n, double s, const int32_t *tets, const int32_t *nbrs,
                     int32_t ntets, nsgpu_patch_set **out);

/* -------------------- device hot path ------------------------------------- */
/* Grid upload: replaces the DelPhiShared arrays the reference colours in place. */
int nsgpu_set_grid(nsgpu_ctx *ctx, const nsgpu_grid *grid);

/* Patch upload (device copy of the SoA); probe_radius is Surface::probe_radius. */
int nsgpu_set_patches(nsgpu_ctx *ctx, const nsgpu_patches *p, double probe_radius);

/* Surface::getSurf(&vol, optimizeGrids, fillCav, cavVol) (Surface.cpp:398-1689):
 * ray-patch intersection over the 3 grid panels (+3 edge panels when
 * accurate_triangulation), per-ray hit sort, epsmap/idebmap/status fill,
 * corner bitmap and vertex records, cavity labelling and conditional filling
 * (getCavitiesWithBilevelStatusMap Surface.cpp:1994-2126, fillCavities
 * :2355-2487), analytic vertex assembly (assembleVerticesList :7415-7660).
 * Maps stay on the device until read back.  Returns once *res is complete;
 * the vertex assembly may still be running on the context's stream, and every
 * later call on the context is ordered after it. */
int nsgpu_get_surf(nsgpu_ctx *ctx, const nsgpu_surf_opts *opts, nsgpu_surf_result *res);

/* nsgpu_get_surf without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The maps
 * are written to the caller's device arrays in the layouts of the read calls
 * below (epsmap and idebmap when opts->build_epsmap, status when
 * opts->build_status; the others may be NULL).  Every scratch buffer of the
 * call comes from one device array of NSGPU_NOALLOC_SCRATCH_BYTES per GPU,
 * held by the library until the process ends; a call whose buffers do not fit
 * fails with NSGPU_E_NOMEM and says how many bytes it needed (grids up to
 * about 150^3 nodes fit).  Calls on one GPU share that array: calls that
 * overlap on one GPU race.  Edge-ray panels are refused
 * (opts->accurate_triangulation must be 0), as are more cavity components
 * than the device numbering holds (4,096).  Returns when the maps are
 * written; *res as nsgpu_get_surf.  The context's own map and ray buffers do
 * not receive this call's results (the read calls refuse until the next
 * nsgpu_get_surf), and no cavity sizes are kept. */
#define NSGPU_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int nsgpu_get_surf_noalloc(nsgpu_ctx *ctx, const nsgpu_surf_opts *opts, nsgpu_surf_result *res, int32_t *epsmap,
                           uint8_t *idebmap, int32_t *status);

/* Surface::triangulateSurface(outputMesh, buildAtomsMap, iso=0) (Surface.cpp:9909-10520):
 * classify cubes from the corner bitmap, add the missing edge midpoints,
 * emit triangles (reversed winding, as the reference), sum the area. */
int nsgpu_triangulate(nsgpu_ctx *ctx, nsgpu_mesh_info *info);

/* The ray/patch intersection of get_surf and surface_gradient.
 * NSGPU_CAST_PATCHES (default): one workgroup per (patch, panel slot) walks
 * the pixel rectangle of the patch, as getPatchIntersectionData does
 * (ConnollySurface.cpp:4319-4591).  NSGPU_CAST_TILES: a uniform grid of 8x8
 * pixel cells lists the patches meeting each cell (the ray-centric auxiliary
 * grid of ConnollySurface::preProcessPanel, ConnollySurface.cpp:3318-3560, on
 * the grid's pixels, built on the device at the first call after set_grid or
 * set_patches); one wave per cell and slot casts the cell's 64 rays against
 * its list.  Both give the same depths, maps and meshes. */
#define NSGPU_CAST_PATCHES 0
#define NSGPU_CAST_TILES 1
int nsgpu_set_cast_mode(nsgpu_ctx *ctx, int mode);

/* NS::NanoShaper::castAxisOrientedRay (nanoshaper.cpp:220-297) for a batch of
 * rays, on the device: ray r starts at origins[3r..3r+2] and runs along
 * direction (0 = x, 1 = y, 2 = z) to the coordinate coords[r] on that axis.
 * Its intersections with the patches of its cell of the auxiliary grid
 * (ConnollySurface::getRayIntersection, ConnollySurface.cpp:5494-5608, the
 * roots of the whole line as the reference keeps them) are returned as
 * coordinates along the axis in ascending order, hits[r*max_hits ...], with a
 * depth within EPS_INT (1e-8) of the last kept one dropped; normals (may be
 * NULL) receives their analytic normals [r*max_hits*3 ...].  counts[r]: the
 * hits kept, more than max_hits when the list was cut (call again with a
 * larger max_hits), -1 when coords[r] <= the origin's coordinate, -2 when the
 * ray lies outside the grid.  Host arrays; returns when they are written. */
int nsgpu_cast_rays(nsgpu_ctx *ctx, int32_t direction, int32_t n, const double *origins, const double *coords,
                    int32_t max_hits, int32_t *counts, double *hits, double *normals);

/* Wait for all work queued on the context's stream. */
int nsgpu_synchronize(nsgpu_ctx *ctx);

/* Stage timing with HIP events on the context's stream (no counterpart in the
 * reference, whose stages print wall-clock times, e.g. Surface.cpp:1093-1096).
 * When on, get_surf/triangulate record one event pair per stage; stage_times
 * writes up to n durations in ms in the order intersect, sort, fill, corners,
 * cavities, vertices, marching cubes, adjoint, and returns the number of
 * stages. */
int nsgpu_set_profiling(nsgpu_ctx *ctx, int on);
int nsgpu_stage_times(nsgpu_ctx *ctx, double *ms, int n);

/* Path counts of the ray/patch candidate set for the roofline (no
 * counterpart in the reference).  Runs a counting copy of the intersection
 * kernel that writes no hits; counts[8] receives: bounding-box candidates,
 * disc-cull passes, ray/sphere hits, torus quartics, roots tested,
 * feasible hits, feasibility items tested, feasibility calls. */
int nsgpu_intersect_path_counts(nsgpu_ctx *ctx, const nsgpu_surf_opts *opts, int64_t *counts, int n);

/* -------------------- triangle-mesh surfaces ------------------------------- */
/* MeshSurface::load (MeshSurface.cpp:2177-2207; loadOFF :2350-2479 reads the
 * coordinates as floats) as the context's surface: get_surf then casts the
 * rays against the triangles (intersect_triangle, :1487-1553, |det| < 1e-5
 * rejected) instead of SES patches; everything after the ray lists is the
 * same.  ray_based selects Patch_Based_Algorithm = false (the mode blobby
 * surfaces require, main.cpp:1154-1161): rays test the triangles of their
 * cell of the auxiliary 2D grid of at most aux_grid_dim cells per side
 * (Max_mesh_auxiliary_grid_2d_size, default 100; preProcessPanel :338-575,
 * getRayIntersection :1756-1855), edge rays with direction pb - pa;
 * otherwise getPatchIntersectionData over the triangles (:1264-1484).
 * verts [nv*3], tris [nt*3] (vertex indices).  Needs the grid first. */
#define NSGPU_MESH_PATCH_BASED 0
#define NSGPU_MESH_RAY_BASED 1
int nsgpu_set_mesh(nsgpu_ctx *ctx, const double *verts, int32_t nv, const int32_t *tris, int32_t nt,
                   int32_t ray_based, int32_t aux_grid_dim);

/* BlobbySurface::build (BlobbySurface.cpp:83-229): the field
 * sum_atoms exp(B (|p - c|^2 / r^2 - 1)) at the nodes p = x - hside of each
 * atom's box of +-(int)(cutoff/side + 0.5) nodes, marching cubes at iso 1.0
 * with interpolated vertices (getVertices, Surface.cpp:8143-8240), one
 * smoothing pass (smoothSurface :13089-13240), the OFF text round trip
 * ("%.3f" written, read back as floats), installed as a ray-based mesh
 * surface.  B >= 0 is replaced by -2.5 (setBlobbyness); the reference's cutoff
 * is 6 A (BlobbySurface.h:19).  info (may be NULL) receives the size of the
 * blobby mesh and the area of its marching cubes (before smoothing). */
int nsgpu_blobby_build(nsgpu_ctx *ctx, const double *xyzr, int32_t n, double B, double cutoff,
                       int32_t aux_grid_dim, nsgpu_mesh_info *info);
/* Size and contents of the mesh surface; raw = 1 reads the blobby build's
 * smoothed vertices before the text round trip.  The field of the last
 * blobby build: one double per node, flat index (k*ny+j)*nx+i. */
int nsgpu_surface_mesh_size(nsgpu_ctx *ctx, int64_t *nv, int64_t *nt);
int nsgpu_read_surface_mesh(nsgpu_ctx *ctx, int32_t raw, double *verts, int32_t *tris);
int nsgpu_read_blobby_field(nsgpu_ctx *ctx, double *field);

/* -------------------- differentiation (no counterpart in the reference) ---- */
/* Atoms (x, y, z, r per atom, after jitter) the current patch set was built
 * from; the patch set must carry atom ids (nsgpu_patches.atom). */
int nsgpu_set_atoms(nsgpu_ctx *ctx, const double *xyzr, int32_t n);

/* Enclosed volume and area estimate with their gradients with respect to the
 * atoms: traces the three grid-ray panels (the volume is getSurf's with the
 * epsmap and no edge rays, Surface.cpp:1677-1684), then differentiates every
 * kept hit through the implicit equation of its patch's sphere or torus.
 * area = A * sum over the hits of 1/|n|_1 (n the unit normal).  dvolume and
 * darea (either may be NULL) receive n_atoms*4 derivatives d/d(x, y, z, r);
 * res (may be NULL) the ray statistics of the trace.  Leaves no maps (call
 * get_surf again for them); nsgpu_read_ray_hits reads its three panels. */
int nsgpu_surface_gradient(nsgpu_ctx *ctx, double *volume, double *area, double *dvolume, double *darea,
                           nsgpu_surf_result *res);

/* -------------------- read back (caller-owned buffers, DelPhi layouts) ----- */
/* epsmap int32 [((k*ny+j)*nx+i)*3+l]; idebmap uint8 [(k*ny+j)*nx+i], 1 = out;
 * status int32 [(k*ny+j)*nx+i]; corner bits uint32 [ceil(nx*ny*nz/32)],
 * bit (flat & 31) of word (flat >> 5), 1 = outside (tools.h:1324-1413). */
int nsgpu_read_epsmap(nsgpu_ctx *ctx, int32_t *out);
int nsgpu_read_idebmap(nsgpu_ctx *ctx, uint8_t *out);
int nsgpu_read_status(nsgpu_ctx *ctx, int32_t *out);
int nsgpu_read_corner_bits(nsgpu_ctx *ctx, uint32_t *out);
/* cavity sizes in A^3 (DelPhiShared::cavitiesSize) and filled flags. */
int nsgpu_read_cavities(nsgpu_ctx *ctx, double *sizes, uint8_t *filled);
/* mesh: verts [num_vertices*3], tris [num_triangles*3] */
int nsgpu_read_mesh(nsgpu_ctx *ctx, double *verts, int32_t *tris);
/* Vertex normals of the last mesh (Compute_Vertex_Normals), normals
 * [num_vertices*3]: analytic vertices get the analytic normal of the patch
 * at their intersection (ConnollySurface::getNormal, ConnollySurface.cpp:
 * 5736-5819), mid-point vertices the normalised sum of their triangles'
 * plane3points normals in triangle order (approximateNormals, Surface.cpp:
 * 10588-10737). */
int nsgpu_read_normals(nsgpu_ctx *ctx, double *normals);
/* Surface::smoothSurface (Surface.cpp:13089-13240), the Smooth_Mesh pass of
 * triangulate: every vertex of the last mesh moves to the mean of itself and
 * the average of its triangles' other vertices; with vertex normals on, each
 * normal becomes the normalised sum of itself and its neighbours' average
 * (nsgpu_read_normals then returns the smoothed normals). */
int nsgpu_smooth_surface(nsgpu_ctx *ctx);
/* Boundary grid points, phase 2 of Surface::getSurf with
 * Project_boundary_grid_points (Surface.cpp:1476-1588, projector :4657-4680),
 * from the epsmap of the last get_surf: every node with 1 <= i, j, k < n
 * whose six epsmap faces (its own three and the +x/+y/+z faces of its -x, -y,
 * -z neighbours) are not all equal, in z, y, x loop order.  A point with a
 * face outside is projected on the nearest feasible SES patch of its cell of
 * the auxiliary grid of at most aux_grid_dim cells per side
 * (Max_ses_patches_auxiliary_grid_size, default 50; ConnollySurface::
 * buildAuxiliaryGrid ConnollySurface.cpp:3666-3909, getProjection
 * :5822-6106): atom and probe spheres, tori (tested against their planes and
 * self-intersection sphere, not their clipping sphere, as the reference
 * does); with no feasible patch, the probe positions sampled along the circle
 * of each self-intersecting regular torus of the cell (500 samples, kept when
 * a facet on that edge accepts the point, needs the patches' atom ids); else
 * the node itself.  Normals are zero, as the reference leaves them for the
 * SES.  A point with no face outside (only with several dielectrics) keeps
 * its node.  counts (may be NULL) [4]: points, points with no face outside,
 * points placed by the sampling, points left at their node.  SES patch sets
 * only (no skin cells, no mesh surface). */
int nsgpu_project_bgp(nsgpu_ctx *ctx, int32_t aux_grid_dim, int64_t *counts);
/* The last nsgpu_project_bgp: ibgp [n*3] node indices (0-based; the
 * reference's DelPhi binding and saveBGP add 1, DelphiShared.cpp:286-290,
 * :1437-1441), scspos [n*3] projections, scsnor [n*3] normals; any may be
 * NULL. */
int nsgpu_read_bgp(nsgpu_ctx *ctx, int32_t *ibgp, double *scspos, double *scsnor);
/* Pocket detection: the grid part of the CLI's pocket mode (pocketMode,
 * main.cpp:1411-1660) on two surfaces of the same atoms on one grid, both
 * after get_surf with the status map:
 *   fat    probe Pocket_Radius_Big (3.0), cavities detected and all filled
 *          (fill_cavities = 1, cav_vol = INFINITY);
 *   small  probe Pocket_Radius_Small (1.4), cavities detected (fill_cavities
 *          = 1; the reference detects them after the difference when
 *          cavities and pockets are kept together, main.cpp:1494, :1629-1646).
 * The difference map (not out in fat, out in small; small's cavities count as
 * out when cav_and_pockets, Surface::difference Surface.cpp:11471-11718) is
 * filtered by the filter_probe ball (operator-= sets 1.4: a node stays when it
 * lies in a ball of that radius holding only difference nodes,
 * filterCavities :2751-3010, cav2out :2649-2688), split into 6-connected
 * components numbered from 4 in getCavities' alternating scan order (:1877-
 * 1991), components of at most 11.4 * num_mol A^3 marked filled
 * (fillCavities), the filter applied again to the rest (main.cpp:1616; a
 * component it empties is filled too), and a kept component is a pocket when
 * one of its nodes lies in small's outside component (markPockets,
 * DelphiShared.cpp:1530-1584).  fat's status map (nsgpu_read_status) then
 * holds the reference's labels: -id at the support points (the centres of the
 * balls that fit), id at the other kept nodes and in filled components, -1
 * elsewhere.  Both contexts on one device; fat's device buffers hold the work
 * arrays. */
typedef struct {
    double filter_probe;     /* probe of the shape filter, 1.4 in the reference     */
    double num_mol;          /* Num_Wat_Pocket (2): volume threshold 11.4 * num_mol */
    int32_t cav_and_pockets; /* Pockets_And_Cavities (1)                            */
    int32_t pad_;
} nsgpu_pocket_opts;
typedef struct {
    int32_t num_components;  /* components of the filtered difference map          */
    int32_t num_kept;        /* not filled: the cav%d.txt files of cavitiesToAtoms  */
    int64_t support_points;  /* support points of the kept components               */
} nsgpu_pocket_result;
int nsgpu_pockets(nsgpu_ctx *fat, nsgpu_ctx *small, const nsgpu_pocket_opts *opts, nsgpu_pocket_result *res);
/* The last nsgpu_pockets, per component in id order: volumes (node count *
 * side^3, DelPhiShared::cavitiesSize), filled flags (cavitiesFlag), pocket
 * flags (meaningful for kept components); offsets[num_kept + 1] and support
 * [support_points]: the flat node indices of each kept component's support
 * points (the points cavitiesToAtoms writes, DelphiShared.cpp:1669-1745), in
 * flat order.  Any pointer may be NULL. */
int nsgpu_read_pockets(nsgpu_ctx *ctx, double *volumes, uint8_t *filled, uint8_t *is_pocket, int64_t *offsets,
                       int64_t *support);
/* The components' ids of the last nsgpu_pockets before the second filter
 * (the status map DelPhiShared::saveCavities and getCavitiesAtoms read,
 * main.cpp:1590-1610): labels [nx*ny*nz], flat index (k*ny+j)*nx+i, -1 off
 * the components. */
int nsgpu_read_pocket_labels(nsgpu_ctx *ctx, int32_t *labels);
/* Compute_Vertex_Normals: when on, get_surf keeps the patch of every depth so
 * that nsgpu_read_normals can follow the mesh (off by default, as in the
 * reference; it costs the sort a second permuted array). */
int nsgpu_set_vertex_normals(nsgpu_ctx *ctx, int on);
/* Sorted hits of every ray of one panel slot (0..2 grid rays, 3..5 edge rays):
 * counts[N*N] and t[N*N*max_hits], ray (m,n) at index n*N+m, N = max(nx,ny,nz).
 * max_hits must be >= nsgpu_surf_result.max_hits_per_ray.  Reads the rays of
 * the last get_surf or surface_gradient. */
int nsgpu_read_ray_hits(nsgpu_ctx *ctx, int32_t slot, int32_t max_hits, int32_t *counts,
                        double *t);

/* -------------------- batches (no counterpart in the reference) ----------- */
/* Many frames (molecules or conformations, uneven sizes), one grid and patch
 * set each, computed by one graph launch.  A batch owns one context per frame;
 * nsgpu_batch_frame gives it to the nsgpu_read_* calls. */
typedef struct nsgpu_batch nsgpu_batch;

int nsgpu_batch_create(int device, int32_t n_frames, nsgpu_batch **out);
void nsgpu_batch_destroy(nsgpu_batch *batch);
const char *nsgpu_batch_last_error(const nsgpu_batch *batch);
nsgpu_ctx *nsgpu_batch_frame(nsgpu_batch *batch, int32_t frame);
/* nsgpu_set_grid + nsgpu_set_patches of one frame (the graph is captured
 * again at the next nsgpu_batch_run). */
int nsgpu_batch_set_frame(nsgpu_batch *batch, int32_t frame, const nsgpu_grid *grid,
                          const nsgpu_patches *patches, double probe_radius);
/* Runs every frame once eagerly (sizing its buffers and ray lists), then
 * captures all frames' getSurf work into one graph: frames in decreasing
 * grid size, dealt round-robin onto `chains` concurrent chains, each frame's
 * intersection its own launches.  chains <= 0 autotunes the launch shape
 * (nsgpu_batch_prepare_shape).
 * Maps and cavities only: opts->accurate_triangulation must be 0. */
int nsgpu_batch_prepare(nsgpu_batch *batch, const nsgpu_surf_opts *opts, int32_t chains);
int32_t nsgpu_batch_chains(const nsgpu_batch *batch);
/* Launch shape of a batch graph.  cast:
 *   NSGPU_BATCH_CAST_FRAMES      each frame's intersection is its own launches
 *                                inside its chain;
 *   NSGPU_BATCH_CAST_JOINT       one launch per patch class intersects every
 *                                frame's patches (a workgroup per patch item),
 *                                between the chains' clears and their
 *                                sort/fill/cavity work;
 *   NSGPU_BATCH_CAST_PERSISTENT  as JOINT, with `workgroups` persistent
 *                                workgroups per class taking `chunk` items
 *                                at a time from a device queue (0: the
 *                                resident limit; 16 items).
 * The joint casts need SES frames traced patch by patch (no skin cells, mesh
 * surface or cell-grid cast).  chains: concurrent chains of per-frame work. */
enum { NSGPU_BATCH_CAST_FRAMES = 0, NSGPU_BATCH_CAST_JOINT = 1, NSGPU_BATCH_CAST_PERSISTENT = 2 };
typedef struct nsgpu_batch_shape {
    int32_t cast;
    int32_t chains;
    int32_t workgroups;
    int32_t chunk;
} nsgpu_batch_shape;
/* nsgpu_batch_prepare with a given shape, or (shape == NULL) the autotuned
 * one: every candidate (per-frame casts on 1, 2, 4, ... 64 chains; with at
 * least 8 SES frames also joint and persistent casts on 16, 32, 64 chains,
 * persistent at the resident limit with 16-item chunks) is captured and
 * launched twice; the fastest is kept.
 * nsgpu_batch_prepare(batch, opts, 0) autotunes the same way. */
int nsgpu_batch_prepare_shape(nsgpu_batch *batch, const nsgpu_surf_opts *opts, const nsgpu_batch_shape *shape);
/* The shape of the captured graph (workgroups and chunk as launched). */
int nsgpu_batch_get_shape(const nsgpu_batch *batch, nsgpu_batch_shape *shape);
/* One launch of the graph.  latency_ms[n_frames] (may be NULL) receives each
 * frame's completion time after the launch (device clock); *redone (may be
 * NULL) the frames that were finished on the host path (ray lists outgrown,
 * or more cavity components than the device numbering holds).  A frame whose
 * context changed since the capture (new inputs or options through
 * nsgpu_batch_frame, a device buffer that grew, a redone frame) is first run
 * eagerly and the graph captured again, so the launch never reads a freed
 * buffer. */
int nsgpu_batch_run(nsgpu_batch *batch, double *latency_ms, int32_t *redone);
/* The per-frame launch loop over the same frames: nsgpu_get_surf on each in
 * turn; latency_ms[n_frames] (may be NULL) receives each call's wall time. */
int nsgpu_batch_run_loop(nsgpu_batch *batch, double *latency_ms);
int nsgpu_batch_result(nsgpu_batch *batch, int32_t frame, nsgpu_surf_result *res);

/* -------------------- synthetic inputs ------------------------------------ */
/* Seeded self-avoiding chain: 1.5 A steps, >= 3.0 A separation between atoms
 * at least 3 bonds apart, centres confined to a sphere of radius R, radii
 * C/N/O/S = 1.70/1.55/1.52/1.80 A with probabilities .62/.17/.19/.02.
 * The confinement radius grows by 2 % whenever the walk jams; the radius
 * finally used is written to *radius_used.  Writes n*4 doubles (x,y,z,r).
 * Returns the number of atoms placed. */
int32_t nsgpu_generate_chain(int32_t n, double radius, uint64_t seed, double *xyzr,
                             double *radius_used);

#ifdef __cplusplus
}
#endif

#endif /* NSGPU_H */
