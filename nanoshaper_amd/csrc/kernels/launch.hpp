// Host-side launch interface of the kernels (defined in kernels/*.hip).
#pragma once
#include "common.hpp"

namespace nsg {

struct FillOpts {
    int build_eps, build_status, inside;
    int slot_x, slot_z, slot_y;     // grid-ray slots of panels 0 (x), 1 (z), 2 (y); -1 = not traced
    int eslot_x, eslot_z, eslot_y;  // edge-ray slots
    // status of a node outside the surface: NSGPU_STATUS_OUT when the cavity
    // pass follows (its first component, the one with the smallest flat
    // index, is numbered NSGPU_STATUS_OUT, and only the other components'
    // nodes are relabelled), else NSGPU_STATUS_TEMP_OUT
    int out_code;
    // when set, the fill also sets every word of this corner bitmap to all
    // ones (outside), ahead of k_corners_bits
    uint32_t *init_bits;
};

// Uniform cell grid over each panel's plane: kCell x kCell pixel cells, per
// (panel, cell) a count, an offset into list and the patch ids in ascending
// order; cells of panel p at [p * nc * nc, (p + 1) * nc * nc).
constexpr int kCell = 8;
struct CellGrid {
    int nc;  // cells per side
    int *count, *off, *cursor, *list;
};
int cell_grid_side(const DevRays &);
// count the (panel, cell) entries; scan count into off (host); then fill the lists
void launch_cells_count(hipStream_t, const DevGrid &, const DevPatches &, const CellGrid &);
void launch_cells_fill(hipStream_t, const DevGrid &, const DevPatches &, const CellGrid &);
// ray tiles: one wave per (cell, slot); candidate partials part[slot * nc * nc + cell]
int64_t tile_workgroups(const DevRays &, int nc);
void launch_cast_tiles(hipStream_t, const DevGrid &, const DevPatches &, const DevRays &, const CellGrid &, double probe,
                       unsigned long long *part, int roots);
// arbitrary axis-aligned rays of one panel (device arrays)
void launch_cast_rays(hipStream_t, const DevGrid &, const DevPatches &, const CellGrid &, double probe, int panel,
                      int n, const double *orig, const double *coord, int maxh, int *count, double *hits,
                      double *normals);

constexpr int kPathCounters = 8;
int64_t intersect_workgroups(const DevPatches &, const DevRays &);
int64_t sort_blocks(const DevRays &);
// torus root solvers of the intersection kernel
enum { kRootsClosed = 0, kRootsMixed = 1, kRootsSturm = 2 };
// One frame of a multi-frame cast: its grid, patches, rays, probe radius,
// candidate partials (the rows of launch_intersect), and the first item of
// each class (0 plain, 1 tori) in the batch's item lists.
struct FrameCast {
    DevGrid g;
    DevPatches P;
    DevRays R;
    double probe;
    unsigned long long *part;
    int64_t first[2];
};
// The frames of a multi-frame cast in item order.  Items of class c of frame
// f are [first[c], first[c] + count) of that class's list; a frame's plain
// items are its plain patches, its torus items (torus rank, slot) pairs.
struct BatchCast {
    const FrameCast *f;
    int nf;
    int64_t items[2];     // items per class
    unsigned int *queue;  // [2] next item of each class, zero before the launch
    int persistent;       // 1: workgroups take items from the queue; 0: one item per workgroup
    int chunk;            // items per queue take (persistent)
};
// One launch over every frame's items of class cls; grid = workgroups of a
// persistent launch (ignored when B.persistent == 0).
void launch_intersect_frames(hipStream_t, const BatchCast &B, int cls, int roots, int grid);
// One frame's intersection, a workgroup per item of each class (part:
// intersect_workgroups() rows, zeroed by the caller).
void launch_intersect(hipStream_t, const DevGrid &, const DevPatches &, const DevRays &, double,
                      unsigned long long *part, int roots);
// counting variant: PC_N path counters per row of part
void launch_intersect_count(hipStream_t, const DevGrid &, const DevPatches &, const DevRays &, double,
                            unsigned long long *part);
// resident workgroups per CU of the class's multi-frame kernel
int intersect_frames_occupancy(int cls, int roots);
// skin cells: order[n - n_skin, n), same per-workgroup candidate rows
void launch_intersect_skin(hipStream_t, const DevGrid &, const DevPatches &, const DevRays &, unsigned long long *part);
void launch_sort(hipStream_t, const DevGrid &, const DevRays &, unsigned long long *part);
// zero_out = false: the caller cleared out[0 .. ncols) already
void launch_reduce_cols(hipStream_t, const unsigned long long *in, int64_t nrows, int ncols, unsigned int max_mask,
                        unsigned long long *out, bool zero_out = true);
void launch_fill_bits(hipStream_t, const DevGrid &, const DevRays &, const FillOpts &, int *, uint8_t *, int *);
// bits must start at all ones
void launch_corners_bits(hipStream_t, const DevGrid &, const DevRays &, const FillOpts &, uint32_t *);
int64_t sum_partials_count(int64_t n);
void launch_sum_blocks(hipStream_t, const double *, int64_t, double *);
// k_sum_blocks over n = *cnt + *off <= ncap elements (partials past n are 0)
void launch_sum_blocks_scan(hipStream_t, const double *v, int64_t ncap, const int *cnt, const int *off,
                            double *partial);
// x-runs of TEMP_OUT nodes: rows * ccl_row_capacity() slots each of rx0, rx1
// and R (R is indexed by compact run id), one count and one offset per row
struct CclRuns {
    int *rx0, *rx1, *R, *rowcnt, *roff;
    // the runs outside the component of run 0 (run id and row), *nlist of them
    int *list_c, *list_row, *nlist;
};
int ccl_row_capacity(const DevGrid &);
// runs from the x grid rays' status plane (R.cov_st); scan_scratch: scan_scratch_ints(rows) ints;
// nroots is zeroed here; the keys of root slots below keycap start cleared
void launch_ccl(hipStream_t, const DevGrid &, const DevRays &R, const CclRuns &, int *scan_scratch, int *nroots,
                int *rootlist, unsigned int *bkey, int *count, int keycap);
// components with slot >= capacity are skipped
void launch_ccl_keys(hipStream_t, const DevGrid &, const CclRuns &, unsigned int *bkey, int *count, int capacity);
// done: the replay's info[0]; the pass does nothing when it is 0
void launch_ccl_label(hipStream_t, const DevGrid &, const CclRuns &, const int *label, const uint8_t *filled,
                      const int *done, int, int, int, int *, int *, uint8_t *, uint32_t *);
// adjoint of the three-panel volume and of the area estimate
struct AdjointArgs {
    const double *atoms;  // x, y, z, r per atom
    double probe;
    double area_px;  // A = side^2
    double flag;     // panels in the volume sum (3)
    double *grad_v, *grad_a;  // 4 per atom, accumulated (zeroed by the caller)
    double *sums;             // 2 per block: volume and area partials
};
int64_t adjoint_blocks(const DevRays &);
void launch_adjoint(hipStream_t, const DevGrid &, const DevPatches &, const DevRays &, const AdjointArgs &);
// fills (graph-safe clears) and the volume reduction
void launch_fill32(hipStream_t, void *p, uint32_t v, int64_t n_words);
// up to four fills (32-bit words) in one launch
struct FillSeg {
    uint32_t *p;
    uint32_t v;
    int64_t n;
};
struct FillList {
    FillSeg s[4];
    int k;
};
void launch_fill_multi(hipStream_t, const FillList &);
struct VolumeSlots {
    int n, flag_sum;
    int slot[6];
    double area_px;
};
void launch_volume(hipStream_t, const double *partial, int64_t per, const VolumeSlots &, double *out);
// k_sum_blocks over each slot vs.slot[k] of v (slot s at v + s * n, partials at partial + s * per)
void launch_sum_slots(hipStream_t, const double *v, int64_t n, const VolumeSlots &vs, double *partial);
// writes the device wall clock
void launch_stamp(hipStream_t, unsigned long long *out);
// on-device replay of the cavity numbering (up to kReplayMax components)
constexpr int kReplayMax = 4096;
struct CavityReplay {
    const int *nroots, *rootlist;
    const unsigned int *bkey;
    const int *count;
    double cube, vol;  // node volume side^3, filling threshold
    int *label;        // per component
    uint8_t *filled;   // per component
    double *cav_size;  // per cavity label - 4
    uint8_t *cav_filled;
    int *info;  // [0] replay done (0: too many components), [1] number of cavities
};
void launch_ccl_clear_keys(hipStream_t, const int *nroots, unsigned int *bkey, int *count, int64_t capacity);
void launch_ccl_replay(hipStream_t, const CavityReplay &);
int64_t scan_scratch_ints(int64_t n);
// total (optional): total[0] = sum of the n inputs, total[1] = 0
void launch_scan(hipStream_t, const int *, int64_t, int *, int *, int *total = nullptr);
// up to 8 device ints and nd doubles gathered into one device buffer
// (ints at dst[0..8), doubles from dst + 8), for a single copy to the host
struct GatherSrc {
    const int *src[8];
    int n;
    const double *dsrc;
    int64_t nd;
};
void launch_gather(hipStream_t, const GatherSrc &, int *dst);
void mc_upload_tables(hipStream_t);
void launch_vtx_count(hipStream_t, const DevGrid &, const DevRays &, const int[3], const uint32_t *, int *);
// The analytic vertices of edge ray r (panel * N^2 + n * N + m) are ids
// offs[r] .. offs[r] + counts[r] - 1 (counts from launch_vtx_count, offs its
// exclusive scan), in increasing node order along the ray.  vbits: for word
// w = ia / 64 of the ray (W = (N + 63) / 64 words), vbits[w * 3N^2 + r] has
// bit ia % 64 set when node ia along the ray has a vertex, and vpre[same]
// counts the ray's vertices in the words before w.  vslot[v] receives the
// vertex's edge (3 * node + direction), vhit[v] the index of its depth in
// R.t / R.pid; vertices v >= vcap are dropped (vcap bounds the total; the
// host checks it)
void launch_vtx_write(hipStream_t, const DevGrid &, const DevRays &, const int[3], const uint32_t *, const int *offs,
                      double *, uint64_t *vbits, int *vpre, int64_t *vslot, int64_t *vhit, int64_t vcap);
int64_t mc_num_segments(const DevGrid &);
void launch_mc_classify(hipStream_t, const DevGrid &, const uint32_t *bits, uint32_t *segmask, int *segcnt);
void launch_mc_compact(hipStream_t, const DevGrid &, const uint32_t *bits, const uint32_t *segmask, const int *segoff,
                       int *anode, uint8_t *acase);
// Where the marching cubes find the vertex of a cube edge: the analytic
// vertices by edge ray (launch_vtx_write's offs, vbits and vpre; vbits ==
// nullptr for a scalar field, which has none), and the mid points through
// the active index of the cube that adds them (the segment masks of
// launch_mc_classify and their exclusive scan), its mid-point mask and offset.
struct McVertexMap {
    const int *offs;
    const uint64_t *vbits;
    const int *vpre;
    int N;
    const uint32_t *segmask;
    const int *segoff;
    const uint8_t *amask;  // bit 0, 1, 2: the cube adds a mid point on edge 5, 6, 10
    const int *avoff;      // exclusive scan of the mid points per active cube
    int base;              // id of the first mid point
};
// per active cube: amask (its mid points), nv (their count) and nt (its triangles)
void launch_mc_counts(hipStream_t, const DevGrid &, int nact, const int *anode, const uint8_t *acase,
                      const McVertexMap &, uint8_t *amask, int *nv, int *nt);
// vorig[v - base] = 4 * (active cube) + (0, 1, 2 for edge 5, 6, 10) of each mid point
// field == nullptr: edge mid points; else vertexInterp at iso on the scalar
// field (one double per node)
void launch_mc_vwrite(hipStream_t, const DevGrid &, int nact, const int *anode, const uint8_t *acase,
                      const McVertexMap &, double *verts, int64_t *vslot, int *vorig, const double *field = nullptr,
                      double iso = 0.);
// vertex normals: analytic at the depths the vertices come from (the ray
// lists must carry patch ids), mid points from their adjacent triangles
struct NormalArgs {
    const double *verts;
    const int64_t *vhit;
    const int *vorig;
    const int *anode, *ant, *atoff, *tris;
    int64_t n_analytic, nverts;
    int nact;
    int eslot0;  // first edge-ray slot
    double probe;
    double *normals;
};
void launch_normals(hipStream_t, const DevGrid &, const DevPatches &, const DevRays &, const NormalArgs &);
// reverse: triangles written (t2, t1, t0) as for the analytic corner bitmap;
// the scalar-field marching cubes keeps (t0, t1, t2)
void launch_mc_twrite(hipStream_t, const DevGrid &, int nact, const int *anode, const uint8_t *acase,
                      const McVertexMap &, const int *toff, const double *verts, int *tris, double *tarea,
                      int reverse = 1);

// blobby scalar field (BlobbySurface::build)
struct BlobbyArgs {
    const double *xyzr;  // 4 per atom
    int n;
    double B;
    int ng;                  // box half-width in nodes, (int)(cutoff / side + 0.5)
    int nbx, nby, nbz;       // bricks of 8^3 nodes
    int *box;                // 8 per atom: box [s, e) per axis, then r^2 (a double)
    int *bcount, *boff, *bcur, *blist;
    double *field;           // one per node
    int *overflow;           // largest brick list when one exceeds blobby_brick_cap()
};
void blobby_bricks(const DevGrid &, int *nbx, int *nby, int *nbz);
int blobby_brick_cap();
void launch_blobby_pairs(hipStream_t, const DevGrid &, const BlobbyArgs &, int scatter);
void launch_blobby_field(hipStream_t, const DevGrid &, const BlobbyArgs &);
void launch_field_bits(hipStream_t, const double *field, int64_t nvox, double iso, uint32_t *bits);
// smoothing pass and the OFF text round trip of the blobby mesh
struct SmoothArgs {
    const double *vin;  // marching-cubes vertices
    const int64_t *vslot;  // edge slot 3 * node + direction of each vertex
    int64_t nv;
    const int *anode, *ant, *atoff, *tris;  // active cubes, their triangle counts and offsets, triangles
    int nact;
    double *vout;
    // vertex normals smoothed with the vertices (null: none)
    const double *nin;
    double *nout;
    // grid-sized cube -> active index map (launch_smooth fills it; an entry is
    // trusted only where anode confirms it, so the map is never cleared)
    int *cube_index;
};
void launch_smooth(hipStream_t, const DevGrid &, const SmoothArgs &);
void launch_quantize(hipStream_t, const double *in, int64_t n, double *out);
// triangle-mesh ray casting
struct MeshArgs {
    const double *v;  // 3 per vertex
    const int *tri;   // 3 per triangle
    int nt;
    int ray_based;    // 1: auxiliary-grid candidates (Patch_Based_Algorithm = false), 0: bounding boxes
    double min2d[3];  // auxiliary grid (preProcessPanel)
    double scale2d;
    int n2d;
    // ray ranges of the auxiliary cells: [((d*2 + phase)*2 + 0)*n2d + c] first ray
    // along axis d with cell >= c, [...+ 1)*n2d + c] last ray with cell <= c
    const int *rr;
};
int64_t mesh_workgroups(int64_t nt, const DevRays &);
void launch_intersect_mesh(hipStream_t, const DevGrid &, const DevRays &, const MeshArgs &, unsigned long long *part);


// probe positions sampled along a torus circle (coiNum, ConnollySurface.cpp:5927)
constexpr int kCoiSamples = 500;
// Boundary grid points (Surface.cpp:1476-1588) and their SES projection
// (ConnollySurface::getProjection, ConnollySurface.cpp:5822-6106).
struct BgpArgs {
    const int *eps;                  // epsmap
    int W;                           // bitmask words per grid row (ceil(nx / 64))
    const unsigned long long *bits;  // [nz * ny * W] boundary-point bitmask
    const int *woff;                 // exclusive scan of the words' popcounts
    // auxiliary grid of na^3 cells (buildAuxiliaryGrid): node ranges of each
    // cell per axis at range[2 * (axis * na + a)], [lo, hi]; patch lists
    int na;
    double amin[3], ascale;
    const int *range;
    int *ccount, *coff, *ccursor, *clist;
    // facets by atom pair (ascending pair key), and per patch the range
    // [efac[2p], efac[2p] + efac[2p+1]) of fval holding the facets on a torus
    // patch's atom pair, for the sampled projection near self-intersecting tori
    const int *fval;
    const int *efac;
    const double *coi;  // cos, sin of the kCoiSamples sample angles (getCoi)
    // outputs: node indices, projections, normals (zero, as the reference's)
    int *ibgp;
    double *pos, *nor;
    // [0] sampled projections, [1] points left at their node, [2] points
    // with no face outside, [3] points queued for the sampled projection
    int *counters;
    // node -> auxiliary cell index per axis (x at [0, N), y at [N, 2N), z at
    // [2N, 3N), N = the grid's longest side); cells holding boundary points
    // (flag, exclusive scan, list) and the queue of (point, cell) pairs
    const int *node_aux;
    int *cflag, *coff_active, *active, *fq;
};
void launch_bgp_flags(hipStream_t, const DevGrid &, const BgpArgs &, unsigned long long *bits, int *wcount,
                      int *n_internal);
void launch_bgp_cells(hipStream_t, const BgpArgs &);
void launch_bgp_bin(hipStream_t, const DevPatches &, const BgpArgs &, bool fill);
void launch_bgp_project(hipStream_t, const DevGrid &, const DevPatches &, const BgpArgs &, double probe,
                        int64_t n_active);


// Pocket detection on bit planes (rows of W words, word (row, w) = nodes
// x = 64 w .. 64 w + 63 of row k * ny + j); see kernels/pocket.hip.
struct PocketArgs {
    int W;
    int64_t nwords;
    int *st1;         // fat-probe status, rewritten to the pocket labels
    const int *st2;   // small-probe status (outside component = NSGPU_STATUS_OUT)
    int cav_and_pockets;
    const int *ball;  // (dy, dz, half width along x) per row of the probe's digital ball
    int nball;
    unsigned long long *D, *U, *S2, *K2, *roots;
    int *parent, *rcount, *roff;
    // per component (index = rank of its root, the smallest flat index)
    int *ccount, *croot, *cfilled, *cid, *ckept, *cpocket;
    long long *cbkey;
    // support points in flat order
    int *scount, *soff, *sup_comp;
    long long *sup_node;
};
void launch_pk_diff(hipStream_t, const DevGrid &, const PocketArgs &);
void launch_pk_morph(hipStream_t, const DevGrid &, const PocketArgs &, const unsigned long long *in,
                     const unsigned long long *and_mask, unsigned long long *out, bool dilate);
void launch_pk_components(hipStream_t, const DevGrid &, const PocketArgs &, const unsigned long long *m);
void launch_pk_stats(hipStream_t, const DevGrid &, const PocketArgs &);
void launch_pk_union_kept(hipStream_t, const DevGrid &, const PocketArgs &);
void launch_pk_label(hipStream_t, const DevGrid &, const PocketArgs &);
void launch_pk_support_list(hipStream_t, const DevGrid &, const PocketArgs &);
void launch_pk_comp_labels(hipStream_t, const DevGrid &, const PocketArgs &, int *out);

}  // namespace nsg
