// C ABI of the device hot path: one context = one grid + one patch set on one
// device, driving the kernels of kernels/*.hip on a private HIP stream.
// Device buffers are grown on demand and reused across calls.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../kernels/launch.hpp"
#include "internal.hpp"
#include "nsgpu.h"


using namespace nsg;

namespace {
// set while a graph is being captured: buffers must already be large enough
thread_local bool g_no_alloc = false;

// set while nsgpu_get_surf_noalloc runs: buffers are carved from this
// caller-provided device region (bump allocation) instead of hipMalloc
struct Arena {
    char *base;
    size_t cap, used, need;
};
thread_local Arena *g_arena = nullptr;

// grow-only device buffer; serial counts its (re)allocations, so that a
// captured graph holding its address can tell that it moved
struct DBuf {
    void *p = nullptr;
    size_t cap = 0;
    uint32_t serial = 0;
    hipError_t ensure(size_t bytes) {
        if (bytes <= cap && p) return hipSuccess;
        if (g_arena) return carve(bytes);
        if (g_no_alloc) return hipErrorOutOfMemory;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        ++serial;
        size_t b = std::max<size_t>(bytes, 256);
        hipError_t e = hipMalloc(&p, b);
        if (e == hipSuccess) cap = b;
        return e;
    }
    // the next 256-byte aligned bytes of the arena; the region it had is
    // left behind (the arena restarts at every attempt of the call)
    hipError_t carve(size_t bytes) {
        const size_t off = (g_arena->used + 255) & ~(size_t)255;
        if (off > g_arena->cap || bytes > g_arena->cap - off) {
            g_arena->need = std::max(g_arena->need, off + bytes);
            return hipErrorOutOfMemory;
        }
        p = g_arena->base + off;
        cap = bytes;
        g_arena->used = off + bytes;
        ++serial;
        return hipSuccess;
    }
    void release() {
        const bool in_arena = g_arena && (char *)p >= g_arena->base && (char *)p < g_arena->base + g_arena->cap;
        if (p && !in_arena) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        ++serial;
    }
    template <class T>
    T *as() const { return (T *)p; }
};

enum Stage { ST_INTERSECT, ST_SORT, ST_FILL, ST_CORNERS, ST_CAV, ST_VERTS, ST_MC, ST_ADJOINT, ST_COUNT };

// device result block of one getSurf, 64-bit words: hits, odd rays, longest
// ray, candidates, volume (a double), cavity replay (int done, int count)
enum { RES_HITS, RES_ODD, RES_MAX, RES_CAND, RES_VOL, RES_CAV, RES_WORDS = 8 };
struct HostResults {
    unsigned long long u[RES_WORDS];
    double volume() const {
        double v;
        std::memcpy(&v, &u[RES_VOL], 8);
        return v;
    }
    const int *cav_info() const { return reinterpret_cast<const int *>(&u[RES_CAV]); }
};
}  // namespace

struct nsgpu_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    // grid
    bool have_grid = false;
    nsgpu_grid grid{};
    DBuf gx, gy, gz;
    // patches
    bool have_patches = false;
    int npatch = 0, npatch_plain = 0, npatch_skin = 0;
    double probe = 1.4;
    DBuf pkind, poff, pcnt, pflags, pbsph, pdata, porder, patom;
    int patch_max_atom = -1;  // largest atom id the patches name (-1: no atom ids given)
    // atoms of the patch set (adjoint)
    DBuf atoms, rpid, gradv, grada, asums;
    int natoms = 0;
    bool record_pid = false;
    // rays
    int K = 16;
    DBuf rcnt, rt, rvol, partial, stats, cand, cov, covst;
    int nslots = 0, panels[2] = {0, 0};
    // maps
    DBuf eps, ideb, status, bits;
    // cavities
    DBuf crx0, crx1, cpar, crowcnt, croff, nroots, rootlist, bkey, ccount, clabel, cfilled, dcavsize, dcavfilled,
        hostdone, clist, clrow;
    // result block (device) and its pinned host mirror
    DBuf dres;
    HostResults *hres = nullptr;
    // pinned staging for the small reads between stages (totals, partial
    // sums): copies into pageable memory are staged and serialised
    void *hpin = nullptr;
    size_t hpin_cap = 0;
    DBuf dgather;  // device side of those reads (launch_gather)
    DBuf amask;    // per active cube: the mid points it adds (bits 0, 1, 2: edges 5, 6, 10)
    DBuf cubeidx;  // smoothing: node -> active cube index (validated, never cleared)
    bool capturing = false;  // enqueueing into a graph: no stage events
    std::vector<double> cav_size;
    std::vector<uint8_t> cav_filled;
    bool cav_on_device = false;  // cav_size/cav_filled not yet copied from the device
    // vertices and mesh
    // vcount, voffs: analytic vertices per edge ray and their exclusive scan;
    // vbits, vpre: each edge ray's node bitmap of vertices and its per-word
    // prefix counts (the marching cubes find an edge's vertex from voffs and these)
    DBuf vcount, voffs, vbits, vpre, scan_scratch, verts, tris, tarea;
    DBuf vslot;
    // the analytic vertex count of the last assemble is still on the device
    // (dnv[0]); n_analytic and nverts wait for resolve_nv
    DBuf dnv;
    bool nv_pending = false;
    int64_t nv_cap = 0;
    // getSurf already classified the cubes for triangulate (segmask, segcnt,
    // segoff hold the result) and read the active-cube count
    bool pre_mc = false;
    int64_t pre_nact = 0;
    // vertex provenance for the normals: depth of each analytic vertex, cube
    // and edge of each mid point
    DBuf vhit, vorig, normals;
    bool hits_have_patches = false;  // the ray lists of the last trace carry patch ids
    bool normals_smoothed = false;   // normals holds the smoothed normals of the current mesh
    bool want_normals = false;       // keep patch ids for the vertex normals (Compute_Vertex_Normals)
    DBuf segmask, segcnt, segoff, anode, acase, anv, ant, avoff, atoff;
    // triangle-mesh surface (MeshSurface, BlobbySurface): vertices as loaded,
    // triangles, ray ranges of the auxiliary 2D grid; the blobby build also
    // keeps its field, its smoothed vertices before the OFF text round trip and
    // the scratch of the field and smoothing passes
    bool surface_is_mesh = false;
    int mesh_ray_based = 1;
    int64_t mnv = 0, mnt = 0;
    DBuf mverts, mtris, mrr, mraw, field, fbits, batoms, bbox, bcount, boff, bcur, blist, flag;
    bool have_field = false, have_raw = false;
    double m_min2d[3] = {0, 0, 0}, m_scale2d = 1;
    int m_n2d = 1, m_aux_dim = 100;
    int64_t n_analytic = 0, nverts = 0, ntris = 0, nact = 0;
    double area = 0;
    bool have_trace = false;  // sorted ray lists of the last get_surf / surface_gradient
    bool have_surf = false, have_vertices = false, have_mesh = false;
    nsgpu_surf_opts opts{};
    nsgpu_surf_result res{};
    // ray/patch intersection: patch-centric workgroups or ray tiles over a
    // cell grid of the patches (built for the current grid and patch set)
    int cast_mode = NSGPU_CAST_PATCHES;
    DBuf cgcount, cgoff, cgcursor, cglist, crorig, crcoord, crcount, crhits, crnorm;
    int cg_nc = 0;
    bool cg_valid = false;
    // bumped by every call that replaces the context's inputs (grid,
    // patches, mesh, atoms); with the buffers' serials it tells a batch
    // whether its captured graph still matches this context
    uint64_t setgen = 0;
    // boundary grid points of the last nsgpu_project_bgp: bitmask and word
    // offsets, the auxiliary grid (node ranges, patch lists), facets by atom
    // pair, outputs and counters
    DBuf bgbits, bgwcount, bgwoff, bgrange, bgccount, bgcoff, bgccursor, bgclist, bgfkey, bgfval, bgib, bgpos, bgnor,
        bgcnt, bgcoi, bgnaux, bgcflag, bgcoffa, bgactive, bgfq;
    int64_t nbgp = 0;
    bool have_bgp = false;
    // pockets (nsgpu_pockets, on the fat-probe context): bit planes, union-find,
    // per-component arrays, support lists; host results in id order
    DBuf pkD, pkS1, pkK, pkU, pkS2, pkK2, pkroots, pkparent, pkrcount, pkroff, pkball, pkccount, pkcroot, pkcfilled,
        pkcid, pkckept, pkcpocket, pkcbkey, pkscount, pksoff, pksupn, pksupc, pklab;
    std::vector<double> pk_vol;
    std::vector<uint8_t> pk_filled, pk_pocket;
    std::vector<int64_t> pk_off, pk_sup;
    bool have_pockets = false;
    PocketArgs pkA{};  // the device arrays of the last nsgpu_pockets
    // the cached tables of bgp_tables and what they were built for
    BgpArgs bgA{};
    uint64_t bg_setgen = 0;
    int bg_aux = -1;
    bool bg_valid = false;
    // stage timing
    bool profile = false;
    hipEvent_t ev[2 * ST_COUNT] = {};
    bool recorded[ST_COUNT] = {};  // both events of the stage recorded since the last query
};

#define HIP_TRY(ctx, expr)                                                                           \
    do {                                                                                             \
        hipError_t e_ = (expr);                                                                      \
        if (e_ != hipSuccess) {                                                                      \
            (ctx)->err = std::string(#expr) + ": " + hipGetErrorString(e_);                          \
            return NSGPU_E_DEVICE;                                                                   \
        }                                                                                            \
    } while (0)

static int fail(nsgpu_ctx *c, int code, const char *msg) {
    c->err = msg;
    return code;
}

static DevGrid dev_grid(const nsgpu_ctx *c) {
    DevGrid g;
    g.nx = c->grid.nx;
    g.ny = c->grid.ny;
    g.nz = c->grid.nz;
    g.N = std::max(g.nx, std::max(g.ny, g.nz));
    g.xmin = c->grid.xmin;
    g.ymin = c->grid.ymin;
    g.zmin = c->grid.zmin;
    g.xmax = c->grid.xmax;
    g.ymax = c->grid.ymax;
    g.zmax = c->grid.zmax;
    g.scale = c->grid.scale;
    g.side = c->grid.side;
    g.hside = c->grid.hside;
    g.x = c->gx.as<double>();
    g.y = c->gy.as<double>();
    g.z = c->gz.as<double>();
    return g;
}

static DevRays dev_rays(const nsgpu_ctx *c) {
    DevRays r;
    r.N = std::max(c->grid.nx, std::max(c->grid.ny, c->grid.nz));
    r.K = c->K;
    r.nslots = c->nslots;
    r.panels0 = c->panels[0];
    r.cnt = c->rcnt.as<int>();
    r.t = c->rt.as<double>();
    r.rvol = c->rvol.as<double>();
    r.W = (r.N + 63) / 64;
    r.cov = c->cov.as<unsigned long long>();
    r.cov_st = c->covst.as<unsigned long long>();
    r.pid = c->record_pid ? c->rpid.as<int>() : nullptr;
    return r;
}

static void mark(nsgpu_ctx *c, int stage, int end, hipStream_t st) {
    if (!c->profile || c->capturing) return;
    if (end == 0) c->recorded[stage] = false;
    if (hipEventRecord(c->ev[2 * stage + end], st) == hipSuccess && end == 1) c->recorded[stage] = true;
}

extern "C" {

int nsgpu_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

int nsgpu_create(int device, nsgpu_ctx **out) {
    if (!out) return NSGPU_E_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return NSGPU_E_DEVICE;
    if (device < 0 || device >= ndev) return NSGPU_E_ARG;
    if (hipSetDevice(device) != hipSuccess) return NSGPU_E_DEVICE;
    auto *c = new nsgpu_ctx();
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return NSGPU_E_DEVICE;
    }
    for (auto &e : c->ev) (void)hipEventCreate(&e);
    if (hipHostMalloc((void **)&c->hres, sizeof(HostResults), hipHostMallocDefault) != hipSuccess ||
        c->dres.ensure(sizeof(HostResults)) != hipSuccess) {
        if (c->hres) (void)hipHostFree(c->hres);
        (void)hipStreamDestroy(c->stream);
        delete c;
        return NSGPU_E_DEVICE;
    }
    std::memset(c->hres, 0, sizeof(HostResults));
    mc_upload_tables(c->stream);
    nsg_scratch_setup(device);
    *out = c;
    return NSGPU_OK;
}

}  // extern "C"

// every device buffer of a context
template <class F>
static void for_each_buf(nsgpu_ctx *c, F fn) {
    DBuf *all[] = {&c->gx, &c->gy, &c->gz, &c->pkind, &c->poff, &c->pcnt, &c->pflags, &c->pbsph, &c->pdata, &c->porder, &c->patom, &c->atoms, &c->rpid, &c->gradv, &c->grada, &c->asums,
                   &c->rcnt, &c->rt, &c->rvol, &c->partial, &c->stats, &c->cand, &c->cov, &c->covst, &c->eps, &c->ideb, &c->status,
                   &c->bits, &c->crx0, &c->crx1, &c->cpar, &c->crowcnt, &c->croff, &c->nroots, &c->rootlist, &c->bkey, &c->ccount, &c->clabel, &c->cfilled, &c->dcavsize, &c->dcavfilled, &c->hostdone, &c->dres, &c->dnv, &c->dgather, &c->amask, &c->clist, &c->clrow, &c->cubeidx,
                   &c->vslot, &c->vhit, &c->vorig, &c->normals, &c->vcount, &c->voffs, &c->vbits, &c->vpre, &c->scan_scratch, &c->verts, &c->tris, &c->tarea,
                   &c->segmask, &c->segcnt, &c->segoff, &c->anode, &c->acase, &c->anv, &c->ant, &c->avoff, &c->atoff, &c->mverts, &c->mtris,
                   &c->mrr, &c->mraw, &c->field, &c->fbits, &c->batoms, &c->bbox, &c->bcount, &c->boff, &c->bcur,
                   &c->blist, &c->flag, &c->cgcount, &c->cgoff, &c->cgcursor, &c->cglist, &c->crorig,
                   &c->crcoord, &c->crcount, &c->crhits, &c->crnorm, &c->bgbits, &c->bgwcount, &c->bgwoff,
                   &c->bgrange, &c->bgccount, &c->bgcoff, &c->bgccursor, &c->bgclist, &c->bgfkey, &c->bgfval,
                   &c->bgib, &c->bgpos, &c->bgnor, &c->bgcnt, &c->bgcoi, &c->bgnaux, &c->bgcflag,
                   &c->bgcoffa, &c->bgactive, &c->bgfq, &c->pkD, &c->pkS1, &c->pkK, &c->pkU, &c->pkS2, &c->pkK2,
                   &c->pkroots, &c->pkparent, &c->pkrcount, &c->pkroff, &c->pkball, &c->pkccount, &c->pkcroot,
                   &c->pkcfilled, &c->pkcid, &c->pkckept, &c->pkcpocket, &c->pkcbkey, &c->pkscount, &c->pksoff,
                   &c->pksupn, &c->pksupc, &c->pklab};
    for (DBuf *b : all) fn(b);
}

// state a captured graph depends on: the inputs, every buffer address (its
// allocation serial) and the ray-list capacity
static uint64_t ctx_generation(nsgpu_ctx *c) {
    uint64_t h = c->setgen * 0x9e3779b97f4a7c15ull + (uint64_t)c->K;
    for_each_buf(c, [&](DBuf *b) { h = h * 1000003ull + b->serial; });
    return h;
}

extern "C" {

void nsgpu_destroy(nsgpu_ctx *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for_each_buf(c, [](DBuf *b) { b->release(); });
    for (auto &e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->hres) (void)hipHostFree(c->hres);
    if (c->hpin) (void)hipHostFree(c->hpin);
    (void)hipStreamDestroy(c->stream);
    delete c;
}

const char *nsgpu_last_error(const nsgpu_ctx *c) { return c ? c->err.c_str() : "null context"; }

int nsgpu_set_profiling(nsgpu_ctx *c, int on) {
    if (!c) return NSGPU_E_ARG;
    c->profile = on != 0;
    return NSGPU_OK;
}

int nsgpu_stage_times(nsgpu_ctx *c, double *ms, int n) {
    if (!c || !ms) return NSGPU_E_ARG;
    (void)hipSetDevice(c->device);
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    for (int s = 0; s < ST_COUNT && s < n; ++s) {
        float t = 0;
        // stages that did not run in the last call report 0
        if (!c->recorded[s] || hipEventElapsedTime(&t, c->ev[2 * s], c->ev[2 * s + 1]) != hipSuccess) t = 0;
        ms[s] = t;
        c->recorded[s] = false;
    }
    (void)hipGetLastError();  // a failed query must not surface in a later call
    return ST_COUNT;
}

static int setup_aux_grid(nsgpu_ctx *c, int aux_dim, int ray_based);

int nsgpu_set_grid(nsgpu_ctx *c, const nsgpu_grid *g) {
    if (!c || !g || g->nx <= 1 || g->ny <= 1 || g->nz <= 1) return NSGPU_E_ARG;
    if ((int64_t)g->nx * g->ny * g->nz >= (int64_t)1 << 31)
        return fail(c, NSGPU_E_ARG, "grid exceeds 2^31 nodes");
    ++c->setgen;
    c->cg_valid = false;
    HIP_TRY(c, hipSetDevice(c->device));
    c->grid = *g;
    std::vector<double> x(g->nx), y(g->ny), z(g->nz);
    // DelphiShared.cpp:1135-1142
    for (int i = 0; i < g->nx; i++) x[i] = g->xmin + i * g->side;
    for (int i = 0; i < g->ny; i++) y[i] = g->ymin + i * g->side;
    for (int i = 0; i < g->nz; i++) z[i] = g->zmin + i * g->side;
    HIP_TRY(c, c->gx.ensure(x.size() * 8));
    HIP_TRY(c, c->gy.ensure(y.size() * 8));
    HIP_TRY(c, c->gz.ensure(z.size() * 8));
    HIP_TRY(c, hipMemcpyAsync(c->gx.p, x.data(), x.size() * 8, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->gy.p, y.data(), y.size() * 8, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->gz.p, z.data(), z.size() * 8, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->have_grid = true;
    c->have_trace = c->have_surf = c->have_vertices = c->have_mesh = false;
    c->have_field = false;
    if (c->surface_is_mesh) return setup_aux_grid(c, c->m_aux_dim, c->mesh_ray_based);
    return NSGPU_OK;
}

int nsgpu_set_patches(nsgpu_ctx *c, const nsgpu_patches *p, double probe_radius) {
    if (!c || !p || p->n < 0 || p->n_data < 0 || !(probe_radius > 0)) return NSGPU_E_ARG;
    if (p->n > 0 && (!p->kind || !p->off || !p->cnt || !p->flags || !p->bsph || (!p->data && p->n_data)))
        return NSGPU_E_ARG;
    ++c->setgen;
    c->cg_valid = false;
    for (int i = 0; i < p->n; ++i) {
        const int k = p->kind[i];
        if (k < NSGPU_PATCH_POINT || k > NSGPU_PATCH_SKIN_TETRA) return fail(c, NSGPU_E_ARG, "bad patch kind");
        const int64_t need = (k >= NSGPU_PATCH_SKIN_POINT)  ? NSGPU_SKIN_HDR + 4LL * p->cnt[i]
                             : (k >= NSGPU_PATCH_TORUS_REG) ? NSGPU_TORUS_HDR + 8LL * p->cnt[i]
                                                            : 4LL * p->cnt[i];
        if (p->off[i] < 0 || p->cnt[i] < 0 || (int64_t)p->off[i] + need > p->n_data)
            return fail(c, NSGPU_E_ARG, "patch data out of range");
        if (k == NSGPU_PATCH_TORUS_REG && (p->cnt[i] < 1 || p->cnt[i] > 31))
            return fail(c, NSGPU_E_ARG, "regular torus needs 1..31 plane pairs");
    }
    HIP_TRY(c, hipSetDevice(c->device));
    const size_t n = (size_t)p->n;
    HIP_TRY(c, c->pkind.ensure(n * 4));
    HIP_TRY(c, c->poff.ensure(n * 4));
    HIP_TRY(c, c->pcnt.ensure(n * 4));
    HIP_TRY(c, c->pflags.ensure(n * 4));
    HIP_TRY(c, c->pbsph.ensure(n * 32));
    HIP_TRY(c, c->pdata.ensure((size_t)p->n_data * 8));
    if (n) {
        HIP_TRY(c, hipMemcpyAsync(c->pkind.p, p->kind, n * 4, hipMemcpyHostToDevice, c->stream));
        HIP_TRY(c, hipMemcpyAsync(c->poff.p, p->off, n * 4, hipMemcpyHostToDevice, c->stream));
        HIP_TRY(c, hipMemcpyAsync(c->pcnt.p, p->cnt, n * 4, hipMemcpyHostToDevice, c->stream));
        HIP_TRY(c, hipMemcpyAsync(c->pflags.p, p->flags, n * 4, hipMemcpyHostToDevice, c->stream));
        HIP_TRY(c, hipMemcpyAsync(c->pbsph.p, p->bsph, n * 32, hipMemcpyHostToDevice, c->stream));
    }
    if (p->n_data)
        HIP_TRY(c, hipMemcpyAsync(c->pdata.p, p->data, (size_t)p->n_data * 8, hipMemcpyHostToDevice, c->stream));
    // patch ids by class: spheres and facets first, then tori, then skin
    // cells; within the SES classes largest bounding sphere first, so that the
    // waves of the largest rectangles start first
    std::vector<int> order;
    order.reserve(n);
    auto by_radius = [&](size_t from) {
        std::stable_sort(order.begin() + (std::ptrdiff_t)from, order.end(),
                         [&](int u, int v) { return p->bsph[4 * u + 3] > p->bsph[4 * v + 3]; });
    };
    for (int i = 0; i < p->n; ++i)
        if (p->kind[i] < NSGPU_PATCH_TORUS_REG) order.push_back(i);
    by_radius(0);
    const int n_plain = (int)order.size();
    for (int i = 0; i < p->n; ++i)
        if (p->kind[i] == NSGPU_PATCH_TORUS_REG || p->kind[i] == NSGPU_PATCH_TORUS_SING) order.push_back(i);
    by_radius((size_t)n_plain);
    const int n_skin = p->n - (int)order.size();
    // skin cells largest first (by the summed faces of their bounding box, the
    // pixels their rectangles cover over the three panels), so that the waves
    // with the largest cells start first and the small ones fill in after
    const size_t skin0 = order.size();
    for (int i = 0; i < p->n; ++i)
        if (p->kind[i] >= NSGPU_PATCH_SKIN_POINT) order.push_back(i);
    {
        std::vector<std::pair<double, int>> key;
        key.reserve(order.size() - skin0);
        for (size_t q = skin0; q < order.size(); ++q) {
            const double *b = p->data + p->off[order[q]];
            const double dx = b[3] - b[0], dy = b[4] - b[1], dz = b[5] - b[2];
            key.emplace_back(-(dx * dy + dy * dz + dz * dx), order[q]);
        }
        std::stable_sort(key.begin(), key.end(),
                         [](const std::pair<double, int> &u, const std::pair<double, int> &v) { return u.first < v.first; });
        for (size_t q = 0; q < key.size(); ++q) order[skin0 + q] = key[q].second;
    }
    HIP_TRY(c, c->porder.ensure(std::max<size_t>(n, 1) * 4));
    if (n) HIP_TRY(c, hipMemcpyAsync(c->porder.p, order.data(), n * 4, hipMemcpyHostToDevice, c->stream));
    // atom ids (optional: only the adjoint needs them)
    c->patch_max_atom = -1;
    HIP_TRY(c, c->patom.ensure(std::max<size_t>(n, 1) * 12));
    if (n && p->atom) {
        int mx = 0;
        for (size_t i = 0; i < 3 * n; ++i) mx = std::max(mx, p->atom[i]);
        c->patch_max_atom = mx;
        HIP_TRY(c, hipMemcpyAsync(c->patom.p, p->atom, n * 12, hipMemcpyHostToDevice, c->stream));
    }
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->npatch = p->n;
    c->npatch_plain = n_plain;
    c->npatch_skin = n_skin;
    c->probe = probe_radius;
    c->have_patches = true;
    c->surface_is_mesh = false;
    c->have_trace = c->have_surf = c->have_vertices = c->have_mesh = false;
    return NSGPU_OK;
}

int nsgpu_intersect_path_counts(nsgpu_ctx *c, const nsgpu_surf_opts *o, int64_t *counts, int n) {
    if (!c || !o || !counts || n < 8) return NSGPU_E_ARG;
    if (!c->have_grid || !c->have_patches) return fail(c, NSGPU_E_STATE, "set_grid and set_patches first");
    if (c->surface_is_mesh || c->npatch_skin) return fail(c, NSGPU_E_STATE, "path counts need an SES patch set");
    HIP_TRY(c, hipSetDevice(c->device));
    int panels[2] = {3, 3};
    if (!o->build_status && !o->build_epsmap) panels[0] = 0;
    else if (o->build_status && !o->build_epsmap) panels[0] = 1;
    if (!o->accurate_triangulation) panels[1] = 0;
    const DevGrid g = dev_grid(c);
    DevRays R = dev_rays(c);
    R.nslots = panels[0] + panels[1];
    R.panels0 = panels[0];
    DevPatches P;
    P.n = c->npatch;
    P.kind = c->pkind.as<int>();
    P.off = c->poff.as<int>();
    P.cnt = c->pcnt.as<int>();
    P.flags = c->pflags.as<int>();
    P.bsph = c->pbsph.as<double>();
    P.data = c->pdata.as<double>();
    P.order = c->porder.as<int>();
    P.n_plain = c->npatch_plain;
    P.n_skin = c->npatch_skin;
    P.atom = c->patom.as<int>();
    const int64_t rows = intersect_workgroups(P, R);
    DBuf part, out;
    HIP_TRY(c, part.ensure((size_t)std::max<int64_t>(rows, 1) * kPathCounters * 8));
    HIP_TRY(c, out.ensure(kPathCounters * 8));
    HIP_TRY(c, hipMemsetAsync(part.p, 0, (size_t)std::max<int64_t>(rows, 1) * kPathCounters * 8, c->stream));
    HIP_TRY(c, hipMemsetAsync(out.p, 0, kPathCounters * 8, c->stream));
    launch_intersect_count(c->stream, g, P, R, c->probe, part.as<unsigned long long>());
    if (rows) launch_reduce_cols(c->stream, part.as<unsigned long long>(), rows, kPathCounters, 0u,
                                 out.as<unsigned long long>());
    HIP_TRY(c, hipGetLastError());
    unsigned long long h[kPathCounters];
    HIP_TRY(c, hipMemcpyAsync(h, out.p, sizeof h, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    part.release();
    out.release();
    for (int q = 0; q < kPathCounters; ++q) counts[q] = (int64_t)h[q];
    return NSGPU_OK;
}

int nsgpu_synchronize(nsgpu_ctx *c) {
    if (!c) return NSGPU_E_ARG;
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return NSGPU_OK;
}

// ray tracing and sort; grows the per-ray capacity and repeats when a ray overflows
// ---------------------------------------------------------------------------
// getSurf as a host-synchronisation-free sequence on one stream: every size
// that the device decides (hits per ray, cavity count, volume) stays on the
// device until the single copy of the result block at the end, so the same
// sequence runs eagerly (get_surf) and inside a captured graph (batches).
// Clears are fill kernels, not memset nodes.

// ray lists: intersection and per-ray sort; statistics into c->dres[0..3]
static MeshArgs mesh_args(const nsgpu_ctx *c) {
    MeshArgs a;
    a.v = c->mverts.as<double>();
    a.tri = c->mtris.as<int>();
    a.nt = (int)c->mnt;
    a.ray_based = c->mesh_ray_based;
    for (int d = 0; d < 3; ++d) a.min2d[d] = c->m_min2d[d];
    a.scale2d = c->m_scale2d;
    a.n2d = c->m_n2d;
    a.rr = c->mrr.as<int>();
    return a;
}

static DevPatches dev_patches(const nsgpu_ctx *c) {
    DevPatches P;
    P.n = c->npatch;
    P.kind = c->pkind.as<int>();
    P.off = c->poff.as<int>();
    P.cnt = c->pcnt.as<int>();
    P.flags = c->pflags.as<int>();
    P.bsph = c->pbsph.as<double>();
    P.data = c->pdata.as<double>();
    P.order = c->porder.as<int>();
    P.n_plain = c->npatch_plain;
    P.n_skin = c->npatch_skin;
    P.atom = c->patom.as<int>();
    return P;
}

static CellGrid cell_grid(const nsgpu_ctx *c) {
    return CellGrid{c->cg_nc, c->cgcount.as<int>(), c->cgoff.as<int>(), c->cgcursor.as<int>(), c->cglist.as<int>()};
}

static int read_total(nsgpu_ctx *c, const int *counts, const int *offs, int64_t n, int64_t *total);

// The cell grid of the current grid and patch set: count, scan, one read of
// the total (the lists' size), fill.  Built eagerly, never inside a capture.
static int ensure_cells(nsgpu_ctx *c) {
    if (c->cg_valid) return NSGPU_OK;
    if (c->capturing) return fail(c, NSGPU_E_STATE, "the cell grid is built before a capture");
    const DevGrid g = dev_grid(c);
    const DevPatches P = dev_patches(c);
    const int nc = cell_grid_side(dev_rays(c));
    const int64_t cells = 3 * (int64_t)nc * nc;
    HIP_TRY(c, c->cgcount.ensure((size_t)cells * 4));
    HIP_TRY(c, c->cgoff.ensure((size_t)cells * 4));
    HIP_TRY(c, c->cgcursor.ensure((size_t)cells * 4));
    HIP_TRY(c, c->scan_scratch.ensure((size_t)scan_scratch_ints(cells) * 4));
    c->cg_nc = nc;
    CellGrid C = cell_grid(c);
    launch_cells_count(c->stream, g, P, C);
    launch_scan(c->stream, C.count, cells, C.off, c->scan_scratch.as<int>());
    HIP_TRY(c, hipGetLastError());
    int64_t total = 0;
    int rc = read_total(c, C.count, C.off, cells, &total);
    if (rc) return rc;
    if (total > INT32_MAX) return fail(c, NSGPU_E_NOMEM, "cell grid lists exceed 2^31 entries");
    HIP_TRY(c, c->cglist.ensure((size_t)std::max<int64_t>(total, 1) * 4));
    C = cell_grid(c);
    launch_cells_fill(c->stream, g, P, C);
    HIP_TRY(c, hipGetLastError());
    c->cg_valid = true;
    return NSGPU_OK;
}

// Rows of the per-workgroup candidate partials of the intersection.
static int64_t intersect_rows(const nsgpu_ctx *c) {
    const DevPatches P = dev_patches(c);
    const DevRays R = dev_rays(c);
    const bool tiles = c->cast_mode == NSGPU_CAST_TILES && !c->surface_is_mesh;
    return std::max<int64_t>(c->surface_is_mesh ? mesh_workgroups(c->mnt, R)
                             : tiles             ? tile_workgroups(R, c->cg_nc)
                                                 : intersect_workgroups(P, R),
                             1);
}

// The trace in three parts (buffers and clears; the intersection; sort and
// statistics), so that a batch can run every frame's intersection in one
// launch between the first part and the last.
static int trace_head(nsgpu_ctx *c, hipStream_t st) {
    const DevGrid g = dev_grid(c);
    const int64_t N = g.N;
    const int64_t nrays = N * N * c->nslots;
    const bool tiles = c->cast_mode == NSGPU_CAST_TILES && !c->surface_is_mesh;
    if (tiles && !c->cg_valid) return fail(c, NSGPU_E_STATE, "the cell grid is not built");
    HIP_TRY(c, c->rcnt.ensure(nrays * 4));
    HIP_TRY(c, c->rvol.ensure(nrays * 8));
    const int64_t W = (N + 63) / 64;
    HIP_TRY(c, c->cov.ensure((size_t)nrays * W * 8));
    HIP_TRY(c, c->covst.ensure((size_t)N * N * W * 8));
    HIP_TRY(c, c->rt.ensure((size_t)nrays * c->K * 8));
    if (c->record_pid) HIP_TRY(c, c->rpid.ensure((size_t)nrays * c->K * 4));
    const DevRays R = dev_rays(c);
    // per-workgroup partials: candidates of the intersection, (hits, odd, longest) of the sort
    const int64_t irows = intersect_rows(c), srows = sort_blocks(R);
    HIP_TRY(c, c->cand.ensure((size_t)irows * 8));
    HIP_TRY(c, c->stats.ensure((size_t)srows * 3 * 8));
    unsigned long long *sout = c->dres.as<unsigned long long>();
    // one launch clears the ray counts, the candidate partials and the result
    // words the column reductions add into (hits, odd rays, longest, candidates)
    FillList fl{};
    fl.s[0] = FillSeg{c->rcnt.as<uint32_t>(), 0u, nrays};
    fl.s[1] = FillSeg{c->cand.as<uint32_t>(), 0u, irows * 2};
    fl.s[2] = FillSeg{reinterpret_cast<uint32_t *>(sout + RES_HITS), 0u, 2 * (RES_CAND + 1 - RES_HITS)};
    fl.k = 3;
    launch_fill_multi(st, fl);
    return NSGPU_OK;
}

static int torus_roots(const nsgpu_ctx *c) {
    return c->opts.precision == NSGPU_PREC_MIXED   ? kRootsMixed
           : c->opts.precision == NSGPU_PREC_STURM ? kRootsSturm
                                                   : kRootsClosed;
}

// the frame as the intersection kernels see it (its items first in the lists)
static FrameCast frame_cast(const nsgpu_ctx *c, const DevRays &R, unsigned long long *part) {
    FrameCast F{};
    F.g = dev_grid(c);
    F.P = dev_patches(c);
    F.R = R;
    F.probe = c->probe;
    F.part = part;
    return F;
}

static int trace_cast(nsgpu_ctx *c, hipStream_t st) {
    const DevGrid g = dev_grid(c);
    const DevPatches P = dev_patches(c);
    const DevRays R = dev_rays(c);
    const bool tiles = c->cast_mode == NSGPU_CAST_TILES && !c->surface_is_mesh;
    mark(c, ST_INTERSECT, 0, st);
    const int roots = torus_roots(c);
    if (c->surface_is_mesh) {
        launch_intersect_mesh(st, g, R, mesh_args(c), c->cand.as<unsigned long long>());
    } else if (tiles) {
        launch_cast_tiles(st, g, P, R, cell_grid(c), c->probe, c->cand.as<unsigned long long>(), roots);
    } else {
        launch_intersect(st, g, P, R, c->probe, c->cand.as<unsigned long long>(), roots);
    }
    HIP_TRY(c, hipGetLastError());
    mark(c, ST_INTERSECT, 1, st);
    return NSGPU_OK;
}

static int trace_tail(nsgpu_ctx *c, hipStream_t st) {
    const DevGrid g = dev_grid(c);
    const DevRays R = dev_rays(c);
    const int64_t irows = intersect_rows(c), srows = sort_blocks(R);
    unsigned long long *spart = c->stats.as<unsigned long long>(), *sout = c->dres.as<unsigned long long>();
    mark(c, ST_SORT, 0, st);
    launch_sort(st, g, R, spart);
    launch_reduce_cols(st, spart, srows, 3, 1u << 2, sout + RES_HITS, false);
    launch_reduce_cols(st, c->cand.as<unsigned long long>(), irows, 1, 0u, sout + RES_CAND, false);
    HIP_TRY(c, hipGetLastError());
    mark(c, ST_SORT, 1, st);
    return NSGPU_OK;
}

static int enqueue_trace(nsgpu_ctx *c, hipStream_t st) {
    int rc = trace_head(c, st);
    if (!rc) rc = trace_cast(c, st);
    if (!rc) rc = trace_tail(c, st);
    return rc;
}

// copy the result block to the pinned host mirror
static int enqueue_results(nsgpu_ctx *c, hipStream_t st) {
    HIP_TRY(c, hipMemcpyAsync(c->hres, c->dres.p, sizeof(HostResults), hipMemcpyDeviceToHost, st));
    return NSGPU_OK;
}

static void take_trace_stats(nsgpu_ctx *c) {
    c->res.total_hits = (int64_t)c->hres->u[RES_HITS];
    c->res.odd_rays = (int64_t)c->hres->u[RES_ODD];
    c->res.max_hits_per_ray = (int)c->hres->u[RES_MAX];
    c->res.candidates = (int64_t)c->hres->u[RES_CAND];
}

// grows the per-ray capacity when the last trace overflowed it; true if it did
static bool regrow(nsgpu_ctx *c) {
    const int longest = (int)c->hres->u[RES_MAX];
    if (longest <= c->K) return false;
    int k = c->K;
    while (k < longest) k *= 2;
    c->K = k;
    return true;
}

// eager trace (adjoint path): enqueue, wait, grow the ray lists until they hold every hit
static int trace(nsgpu_ctx *c) {
    if (c->cast_mode == NSGPU_CAST_TILES && !c->surface_is_mesh) {
        const int rc = ensure_cells(c);
        if (rc) return rc;
    }
    for (int attempt = 0; attempt < 4; ++attempt) {
        int rc = enqueue_trace(c, c->stream);
        if (rc) return rc;
        rc = enqueue_results(c, c->stream);
        if (rc) return rc;
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        if (regrow(c)) continue;
        take_trace_stats(c);
        c->have_trace = true;
        return NSGPU_OK;
    }
    return fail(c, NSGPU_E_STATE, "ray hit capacity did not converge");
}

static CclRuns ccl_runs(nsgpu_ctx *c) {
    return CclRuns{c->crx0.as<int>(), c->crx1.as<int>(), c->cpar.as<int>(), c->crowcnt.as<int>(),
                   c->croff.as<int>(), c->clist.as<int>(), c->clrow.as<int>(), c->nroots.as<int>() + 1};
}

// cavity detection, numbering and filling (getCavitiesWithBilevelStatusMap +
// fillCavities) for up to kReplayMax components on the device
static int enqueue_cavities(nsgpu_ctx *c, hipStream_t st) {
    const DevGrid g = dev_grid(c);
    const int64_t rows = (int64_t)g.ny * g.nz;
    const int64_t slots = rows * ccl_row_capacity(g);
    if (slots > INT32_MAX) return fail(c, NSGPU_E_ARG, "grid too large for 32-bit run ids");
    HIP_TRY(c, c->crx0.ensure(slots * 4));
    HIP_TRY(c, c->crx1.ensure(slots * 4));
    HIP_TRY(c, c->cpar.ensure(slots * 4));
    HIP_TRY(c, c->crowcnt.ensure(rows * 4));
    HIP_TRY(c, c->croff.ensure(rows * 4));
    HIP_TRY(c, c->scan_scratch.ensure((size_t)scan_scratch_ints(rows) * 4));
    HIP_TRY(c, c->nroots.ensure(8));  // root count, run list length
    HIP_TRY(c, c->rootlist.ensure(slots * 4));
    HIP_TRY(c, c->clist.ensure(slots * 4));
    HIP_TRY(c, c->clrow.ensure(slots * 4));
    HIP_TRY(c, c->bkey.ensure((size_t)kReplayMax * 4));
    HIP_TRY(c, c->ccount.ensure((size_t)kReplayMax * 4));
    HIP_TRY(c, c->clabel.ensure((size_t)kReplayMax * 4));
    HIP_TRY(c, c->cfilled.ensure((size_t)kReplayMax));
    HIP_TRY(c, c->dcavsize.ensure((size_t)kReplayMax * 8));
    HIP_TRY(c, c->dcavfilled.ensure((size_t)kReplayMax));
    const CclRuns runs = ccl_runs(c);
    launch_ccl(st, g, dev_rays(c), runs, c->scan_scratch.as<int>(), c->nroots.as<int>(), c->rootlist.as<int>(),
               c->bkey.as<unsigned int>(), c->ccount.as<int>(), kReplayMax);
    launch_ccl_keys(st, g, runs, c->bkey.as<unsigned int>(), c->ccount.as<int>(), kReplayMax);
    CavityReplay r;
    r.nroots = c->nroots.as<int>();
    r.rootlist = c->rootlist.as<int>();
    r.bkey = c->bkey.as<unsigned int>();
    r.count = c->ccount.as<int>();
    r.cube = c->grid.side * c->grid.side * c->grid.side;
    r.vol = c->opts.cav_vol < 0 ? 11.4 : c->opts.cav_vol;
    r.label = c->clabel.as<int>();
    r.filled = c->cfilled.as<uint8_t>();
    r.cav_size = c->dcavsize.as<double>();
    r.cav_filled = c->dcavfilled.as<uint8_t>();
    r.info = (int *)(c->dres.as<unsigned long long>() + RES_CAV);
    launch_ccl_replay(st, r);
    launch_ccl_label(st, g, runs, c->clabel.as<int>(), c->cfilled.as<uint8_t>(), r.info, c->opts.build_epsmap,
                     c->opts.accurate_triangulation, c->opts.inside, c->status.as<int>(), c->eps.as<int>(),
                     c->ideb.as<uint8_t>(), c->bits.as<uint32_t>());
    HIP_TRY(c, hipGetLastError());
    return NSGPU_OK;
}

// host numbering when the component count exceeds the device replay
static int cavities_on_host(nsgpu_ctx *c) {
    const DevGrid g = dev_grid(c);
    int nroots = 0;
    HIP_TRY(c, hipMemcpyAsync(&nroots, c->nroots.p, 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    HIP_TRY(c, c->bkey.ensure((size_t)nroots * 4));
    HIP_TRY(c, c->ccount.ensure((size_t)nroots * 4));
    launch_ccl_clear_keys(c->stream, c->nroots.as<int>(), c->bkey.as<unsigned int>(), c->ccount.as<int>(), nroots);
    const CclRuns runs = ccl_runs(c);
    launch_ccl_keys(c->stream, g, runs, c->bkey.as<unsigned int>(), c->ccount.as<int>(), nroots);
    HIP_TRY(c, hipGetLastError());
    std::vector<int> roots(nroots), count(nroots);
    std::vector<unsigned int> bkey(nroots);
    HIP_TRY(c, hipMemcpyAsync(roots.data(), c->rootlist.p, (size_t)nroots * 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipMemcpyAsync(bkey.data(), c->bkey.p, (size_t)nroots * 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipMemcpyAsync(count.data(), c->ccount.p, (size_t)nroots * 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    // replay the alternating scan of getCavitiesWithBilevelStatusMap: even
    // passes take the component met first scanning k upwards (smallest flat
    // index = the root), odd passes the first one scanning k downwards
    std::vector<int> byF(nroots), byB(nroots);
    for (int s = 0; s < nroots; ++s) byF[s] = byB[s] = s;
    std::sort(byF.begin(), byF.end(), [&](int a, int b) { return roots[a] < roots[b]; });
    std::sort(byB.begin(), byB.end(), [&](int a, int b) { return bkey[a] < bkey[b]; });
    std::vector<int> label(nroots, 0);
    std::vector<char> done(nroots, 0);
    size_t pf = 0, pb = 0;
    int id = NSGPU_STATUS_TEMP_OUT;
    for (int t = 0; t < nroots; ++t) {
        int pick;
        if (t % 2 == 0) {
            while (done[byF[pf]]) ++pf;
            pick = byF[pf];
        } else {
            while (done[byB[pb]]) ++pb;
            pick = byB[pb];
        }
        done[pick] = 1;
        label[pick] = ++id;
    }
    const double vol = c->opts.cav_vol < 0 ? 11.4 : c->opts.cav_vol;
    const double cube = c->grid.side * c->grid.side * c->grid.side;
    const int ncav = std::max(id - NSGPU_STATUS_OUT, 0);
    c->cav_size.assign(ncav, 0.0);
    c->cav_filled.assign(ncav, 0);
    std::vector<uint8_t> filled(nroots, 0);
    for (int s = 0; s < nroots; ++s) {
        if (label[s] < NSGPU_STATUS_FIRST_CAV) continue;
        const double v = count[s] * cube;
        c->cav_size[label[s] - NSGPU_STATUS_FIRST_CAV] = v;
        filled[s] = v <= vol;
        c->cav_filled[label[s] - NSGPU_STATUS_FIRST_CAV] = filled[s];
    }
    c->res.num_cavities = ncav;
    const int one = 1;
    HIP_TRY(c, c->clabel.ensure((size_t)nroots * 4));
    HIP_TRY(c, c->cfilled.ensure((size_t)nroots));
    HIP_TRY(c, c->hostdone.ensure(4));
    HIP_TRY(c, hipMemcpyAsync(c->clabel.p, label.data(), (size_t)nroots * 4, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->cfilled.p, filled.data(), (size_t)nroots, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->hostdone.p, &one, 4, hipMemcpyHostToDevice, c->stream));
    launch_ccl_label(c->stream, g, runs, c->clabel.as<int>(), c->cfilled.as<uint8_t>(), c->hostdone.as<int>(),
                     c->opts.build_epsmap, c->opts.accurate_triangulation, c->opts.inside, c->status.as<int>(),
                     c->eps.as<int>(), c->ideb.as<uint8_t>(), c->bits.as<uint32_t>());
    HIP_TRY(c, hipGetLastError());
    // the host vectors above must outlive the copies
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return NSGPU_OK;
}

// panel rule of getSurf (Surface.cpp:652-659) and the per-call state
static int begin_surf(nsgpu_ctx *c, const nsgpu_surf_opts *o) {
    if (!c->have_grid || !c->have_patches) return fail(c, NSGPU_E_STATE, "set_grid and set_patches first");
    if (c->surface_is_mesh && c->want_normals && o->accurate_triangulation)
        return fail(c, NSGPU_E_STATE, "vertex normals need analytic patches (ses or skin)");
    if (o->precision != NSGPU_PREC_FP64 && o->precision != NSGPU_PREC_MIXED && o->precision != NSGPU_PREC_STURM)
        return fail(c, NSGPU_E_ARG, "unknown precision mode");
    c->opts = *o;
    std::memset(&c->res, 0, sizeof c->res);
    c->have_trace = c->have_surf = c->have_vertices = c->have_mesh = false;
    c->have_bgp = false;
    c->have_pockets = false;
    c->panels[0] = 3;
    c->panels[1] = 3;
    if (!o->build_status && !o->build_epsmap) c->panels[0] = 0;
    else if (o->build_status && !o->build_epsmap) c->panels[0] = 1;
    if (!o->accurate_triangulation) c->panels[1] = 0;
    c->nslots = c->panels[0] + c->panels[1];
    if (c->nslots == 0) return fail(c, NSGPU_E_ARG, "nothing to compute (no epsmap, status or triangulation)");
    return NSGPU_OK;
}

static FillOpts fill_opts(const nsgpu_ctx *c) {
    const nsgpu_surf_opts *o = &c->opts;
    FillOpts fo;
    fo.build_eps = o->build_epsmap;
    fo.build_status = o->build_status;
    fo.inside = o->inside;
    fo.slot_x = c->panels[0] >= 1 ? 0 : -1;
    fo.slot_z = (c->panels[0] >= 3 && o->build_epsmap) ? 1 : -1;
    fo.slot_y = (c->panels[0] >= 3 && o->build_epsmap) ? 2 : -1;
    fo.eslot_x = c->panels[0] + 0;
    fo.eslot_z = c->panels[0] + 1;
    fo.eslot_y = c->panels[0] + 2;
    fo.out_code = (o->fill_cavities && o->build_status) ? NSGPU_STATUS_OUT : NSGPU_STATUS_TEMP_OUT;
    fo.init_bits = nullptr;
    return fo;
}

// slots in the volume sum and their count (Surface.cpp:1677-1684), rays traced
static VolumeSlots volume_slots(const nsgpu_ctx *c, int64_t *traced) {
    const nsgpu_surf_opts *o = &c->opts;
    const int64_t N = std::max(c->grid.nx, std::max(c->grid.ny, c->grid.nz));
    int fill_panels = 3;
    if (!o->build_status && !o->build_epsmap) fill_panels = 0;
    else if (o->build_status && !o->build_epsmap) fill_panels = 1;
    if (o->accurate_triangulation) fill_panels = 3;
    VolumeSlots vs{};
    vs.area_px = c->grid.A;
    *traced = 0;
    for (int p = 0; p < fill_panels; ++p) {
        if ((o->build_status && p == 0) || o->build_epsmap) {
            ++vs.flag_sum;
            *traced += N * N;
            if (p < c->panels[0]) vs.slot[vs.n++] = p;
        }
        if (o->accurate_triangulation) {
            ++vs.flag_sum;
            *traced += N * N;
            vs.slot[vs.n++] = c->panels[0] + p;
        }
    }
    return vs;
}

// everything of getSurf up to the cavities, then the result block copy
// getSurf after the trace: maps, corner bitmap, volume, cavities, results
static int enqueue_maps(nsgpu_ctx *c, hipStream_t st) {
    int rc;
    const DevGrid g = dev_grid(c);
    const DevRays R = dev_rays(c);
    const nsgpu_surf_opts *o = &c->opts;
    const int64_t nvox = (int64_t)g.nx * g.ny * g.nz;
    const int64_t nwords = (nvox + 31) / 32;
    FillOpts fo = fill_opts(c);
    // maps (buffers always exist so that later stages can index them)
    HIP_TRY(c, c->eps.ensure((size_t)nvox * 12));
    HIP_TRY(c, c->ideb.ensure((size_t)nvox));
    HIP_TRY(c, c->status.ensure((size_t)nvox * 4));
    HIP_TRY(c, c->bits.ensure((size_t)nwords * 4));
    // the fill's lanes set the corner bitmap to all outside on their way
    if (o->accurate_triangulation && c->panels[0] > 0) fo.init_bits = c->bits.as<uint32_t>();
    mark(c, ST_FILL, 0, st);
    if (c->panels[0] > 0)
        launch_fill_bits(st, g, R, fo, c->eps.as<int>(), c->ideb.as<uint8_t>(), c->status.as<int>());
    HIP_TRY(c, hipGetLastError());
    mark(c, ST_FILL, 1, st);
    mark(c, ST_CORNERS, 0, st);
    if (o->accurate_triangulation) {
        if (!fo.init_bits) launch_fill32(st, c->bits.p, 0xffffffffu, nwords);
        launch_corners_bits(st, g, R, fo, c->bits.as<uint32_t>());
    }
    HIP_TRY(c, hipGetLastError());
    mark(c, ST_CORNERS, 1, st);
    // volume: A * sum of segment lengths / (panels x phases traced)
    {
        const int64_t NN = (int64_t)R.N * R.N;
        int64_t traced = 0;
        const VolumeSlots vs = volume_slots(c, &traced);
        const int64_t per = sum_partials_count(NN);
        HIP_TRY(c, c->partial.ensure((size_t)per * c->nslots * 8));
        launch_sum_slots(st, R.rvol, NN, vs, c->partial.as<double>());
        launch_volume(st, c->partial.as<double>(), per, vs, (double *)(c->dres.as<unsigned long long>() + RES_VOL));
        HIP_TRY(c, hipGetLastError());
    }
    mark(c, ST_CAV, 0, st);
    if (o->fill_cavities && o->build_status) {
        rc = enqueue_cavities(c, st);
        if (rc) return rc;
    }
    mark(c, ST_CAV, 1, st);
    return enqueue_results(c, st);
}

static int enqueue_surf(nsgpu_ctx *c, hipStream_t st) {
    int rc = enqueue_trace(c, st);
    if (!rc) rc = enqueue_maps(c, st);
    return rc;
}

// host side of a completed enqueue_surf: statistics, volume, cavities
static int finish_surf(nsgpu_ctx *c) {
    take_trace_stats(c);
    c->have_trace = true;
    int64_t traced = 0;
    (void)volume_slots(c, &traced);
    c->res.total_rays = traced;
    c->res.volume = c->hres->volume();
    c->cav_size.clear();
    c->cav_filled.clear();
    c->cav_on_device = false;
    c->res.num_cavities = 0;
    if (c->opts.fill_cavities && c->opts.build_status) {
        const int *info = c->hres->cav_info();
        if (info[0] == 0) {
            const int rc = cavities_on_host(c);
            if (rc) return rc;
        } else {
            // sizes and flags stay on the device until nsgpu_read_cavities asks
            c->res.num_cavities = info[1];
            c->cav_on_device = true;
        }
    }
    return NSGPU_OK;
}

// The pinned staging buffer with room for `bytes`.  Called only when no copy
// into it is pending (every use ends with a stream synchronisation).
static int pinned(nsgpu_ctx *c, size_t bytes, void **out) {
    if (bytes > c->hpin_cap) {
        if (c->hpin) (void)hipHostFree(c->hpin);
        c->hpin = nullptr;
        c->hpin_cap = 0;
        const size_t cap = std::max<size_t>(bytes, 4096);
        HIP_TRY(c, hipHostMalloc(&c->hpin, cap, hipHostMallocDefault));
        c->hpin_cap = cap;
    }
    *out = c->hpin;
    return NSGPU_OK;
}

// count[n-1] + offs[n-1] of an exclusive scan, for each (counts, offs) pair:
// the copies go to pinned memory, so they queue behind one synchronisation
// Device scalars (up to 8 ints) and nd doubles to the host: one gather
// kernel, one copy into the pinned buffer, one synchronisation.  On return
// *ints points at the ints and *dbl at the doubles.
static int read_gathered(nsgpu_ctx *c, hipStream_t st, const GatherSrc &g, const int **ints, const double **dbl) {
    const size_t bytes = 32 + (size_t)g.nd * 8;
    void *hp = nullptr;
    int rc = pinned(c, bytes, &hp);
    if (rc) return rc;
    HIP_TRY(c, c->dgather.ensure(bytes));
    launch_gather(st, g, c->dgather.as<int>());
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, hipMemcpyAsync(hp, c->dgather.p, bytes, hipMemcpyDeviceToHost, st));
    HIP_TRY(c, hipStreamSynchronize(st));
    *ints = static_cast<const int *>(hp);
    *dbl = reinterpret_cast<const double *>(static_cast<const char *>(hp) + 32);
    return NSGPU_OK;
}

static int read_totals(nsgpu_ctx *c, int npairs, const int *const *counts, const int *const *offs, const int64_t *n,
                       int64_t *total) {
    GatherSrc gs{};
    int at[4] = {-1, -1, -1, -1};
    for (int q = 0; q < npairs && q < 4; ++q) {
        total[q] = 0;
        if (n[q] <= 0) continue;
        at[q] = gs.n;
        gs.src[gs.n++] = counts[q] + n[q] - 1;
        gs.src[gs.n++] = offs[q] + n[q] - 1;
    }
    if (gs.n == 0) return NSGPU_OK;
    const int *h = nullptr;
    const double *unused = nullptr;
    const int rc = read_gathered(c, c->stream, gs, &h, &unused);
    if (rc) return rc;
    for (int q = 0; q < npairs && q < 4; ++q)
        if (at[q] >= 0) total[q] = (int64_t)h[at[q]] + h[at[q] + 1];
    return NSGPU_OK;
}

static int read_total(nsgpu_ctx *c, const int *counts, const int *offs, int64_t n, int64_t *total) {
    return read_totals(c, 1, &counts, &offs, &n, total);
}

// the analytic vertex count once it is known on the host
static int set_nv(nsgpu_ctx *c, int64_t nv) {
    c->nv_pending = false;
    if (nv > c->nv_cap) return fail(c, NSGPU_E_STATE, "vertex count above its bound");
    c->n_analytic = nv;
    c->nverts = nv;
    return NSGPU_OK;
}

static int resolve_nv(nsgpu_ctx *c) {
    if (!c->nv_pending) return NSGPU_OK;
    int64_t nv = 0;
    const int rc = read_total(c, c->dnv.as<int>(), c->dnv.as<int>() + 1, 1, &nv);
    return rc ? rc : set_nv(c, nv);
}

// sum of per partial sums, in order, after one synchronisation
static int read_partial_sum(nsgpu_ctx *c, const double *partial, int64_t per, hipStream_t st, double *sum) {
    *sum = 0;
    if (per <= 0) return NSGPU_OK;
    void *hp = nullptr;
    int rc = pinned(c, (size_t)per * 8, &hp);
    if (rc) return rc;
    HIP_TRY(c, hipMemcpyAsync(hp, partial, (size_t)per * 8, hipMemcpyDeviceToHost, st));
    HIP_TRY(c, hipStreamSynchronize(st));
    const double *h = static_cast<const double *>(hp);
    double acc = 0;
    for (int64_t q = 0; q < per; ++q) acc += h[q];
    *sum = acc;
    return NSGPU_OK;
}

static int assemble(nsgpu_ctx *c, const int eslots[3]) {
    const DevGrid g = dev_grid(c);
    const DevRays R = dev_rays(c);
    const int64_t nvox = (int64_t)g.nx * g.ny * g.nz;
    const int64_t nrays = 3 * (int64_t)R.N * R.N;
    const int64_t nmax = nrays;
    int rc = resolve_nv(c);  // the previous pass's count, if still pending
    if (rc) return rc;
    HIP_TRY(c, c->vcount.ensure((size_t)nmax * 4));
    HIP_TRY(c, c->voffs.ensure((size_t)nmax * 4));
    HIP_TRY(c, c->scan_scratch.ensure((size_t)scan_scratch_ints(nmax) * 4));
    launch_vtx_count(c->stream, g, R, eslots, c->bits.as<uint32_t>(), c->vcount.as<int>());
    HIP_TRY(c, c->dnv.ensure(8));
    launch_scan(c->stream, c->vcount.as<int>(), nrays, c->voffs.as<int>(), c->scan_scratch.as<int>(),
                c->dnv.as<int>());
    HIP_TRY(c, hipGetLastError());
    // Every vertex comes from a depth of an edge ray (at most one per depth,
    // at most K per ray), so the edge rays' list capacity bounds the total:
    // the buffers are sized by it, the pass can be enqueued before getSurf's
    // result block is read, and the count itself is read later, with the
    // next read the host makes.
    const int64_t cap = std::max<int64_t>(3 * (int64_t)R.N * R.N * R.K, 1);
    HIP_TRY(c, c->verts.ensure((size_t)cap * 24));
    HIP_TRY(c, c->vslot.ensure((size_t)cap * 8));
    HIP_TRY(c, c->vhit.ensure((size_t)cap * 8));
    const int64_t nwords = nrays * ((R.N + 63) / 64);
    HIP_TRY(c, c->vbits.ensure((size_t)nwords * 8));
    HIP_TRY(c, c->vpre.ensure((size_t)nwords * 4));
    launch_vtx_write(c->stream, g, R, eslots, c->bits.as<uint32_t>(), c->voffs.as<int>(), c->verts.as<double>(),
                     c->vbits.as<uint64_t>(), c->vpre.as<int>(), c->vslot.as<int64_t>(), c->vhit.as<int64_t>(), cap);
    HIP_TRY(c, hipGetLastError());
    c->nv_cap = cap;
    c->nv_pending = true;
    return NSGPU_OK;
}

// Cube classification of triangulate (row-segment masks, their scan) and a
// copy of the active-cube and analytic-vertex totals into pinned memory:
// enqueued by getSurf behind the vertex pass, read after its one
// synchronisation.
static int enqueue_classify(nsgpu_ctx *c) {
    const DevGrid g = dev_grid(c);
    const int64_t nseg = mc_num_segments(g);
    HIP_TRY(c, c->segmask.ensure((size_t)std::max<int64_t>(nseg, 1) * 4));
    HIP_TRY(c, c->segcnt.ensure((size_t)std::max<int64_t>(nseg, 1) * 4));
    HIP_TRY(c, c->segoff.ensure((size_t)std::max<int64_t>(nseg, 1) * 4));
    HIP_TRY(c, c->scan_scratch.ensure((size_t)scan_scratch_ints(std::max<int64_t>(nseg, 1)) * 4));
    void *hp = nullptr;
    int rc = pinned(c, 32, &hp);
    if (rc) return rc;
    HIP_TRY(c, c->dgather.ensure(32));
    launch_mc_classify(c->stream, g, c->bits.as<uint32_t>(), c->segmask.as<uint32_t>(), c->segcnt.as<int>());
    launch_scan(c->stream, c->segcnt.as<int>(), nseg, c->segoff.as<int>(), c->scan_scratch.as<int>());
    GatherSrc gs{};
    gs.src[0] = c->dnv.as<int>();
    gs.src[1] = c->dnv.as<int>() + 1;
    gs.n = 2;
    if (nseg > 0) {
        gs.src[2] = c->segcnt.as<int>() + nseg - 1;
        gs.src[3] = c->segoff.as<int>() + nseg - 1;
        gs.n = 4;
    }
    launch_gather(c->stream, gs, c->dgather.as<int>());
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, hipMemcpyAsync(hp, c->dgather.p, 32, hipMemcpyDeviceToHost, c->stream));
    return NSGPU_OK;
}

int nsgpu_get_surf(nsgpu_ctx *c, const nsgpu_surf_opts *o, nsgpu_surf_result *res) {
    if (!c || !o) return NSGPU_E_ARG;
    int rc = begin_surf(c, o);
    if (rc) return rc;
    HIP_TRY(c, hipSetDevice(c->device));
    if (c->cast_mode == NSGPU_CAST_TILES && !c->surface_is_mesh) {
        rc = ensure_cells(c);
        if (rc) return rc;
    }
    // the edge rays' depths keep their patch ids for the vertex normals
    c->record_pid = o->accurate_triangulation != 0 && c->want_normals;
    const FillOpts fo = fill_opts(c);
    const int es[3] = {fo.eslot_x, fo.eslot_z, fo.eslot_y};
    for (int attempt = 0;; ++attempt) {
        if (attempt == 4) {
            c->record_pid = false;
            return fail(c, NSGPU_E_STATE, "ray hit capacity did not converge");
        }
        rc = enqueue_surf(c, c->stream);
        if (rc) {
            c->record_pid = false;
            return rc;
        }
        // the vertex pass goes into the same stream before the result block
        // is read: a repeated attempt clears the slots it wrote
        mark(c, ST_VERTS, 0, c->stream);
        c->pre_mc = false;
        if (o->accurate_triangulation) {
            rc = assemble(c, es);
            if (rc == NSGPU_OK) rc = enqueue_classify(c);
            if (rc) {
                c->record_pid = false;
                return rc;
            }
        }
        mark(c, ST_VERTS, 1, c->stream);
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        if (!regrow(c)) break;
    }
    c->hits_have_patches = c->record_pid;
    c->record_pid = false;
    rc = finish_surf(c);
    if (rc) return rc;
    if (o->accurate_triangulation) {
        // the totals read with the result block (enqueue_classify)
        const int *h = static_cast<const int *>(c->hpin);
        rc = set_nv(c, (int64_t)h[0] + h[1]);
        if (rc) return rc;
        c->pre_nact = mc_num_segments(dev_grid(c)) > 0 ? (int64_t)h[2] + h[3] : 0;
        c->pre_mc = true;
        // cavities numbered on the host (more components than the device
        // replay holds) changed the corner bitmap after the vertex pass ran
        if (o->fill_cavities && o->build_status && c->hres->cav_info()[0] == 0) {
            c->pre_mc = false;
            rc = assemble(c, es);
            if (rc) return rc;
        }
        c->have_vertices = true;
    }
    c->have_surf = true;
    if (res) *res = c->res;
    return NSGPU_OK;
}

int nsgpu_set_cast_mode(nsgpu_ctx *c, int mode) {
    if (!c || (mode != NSGPU_CAST_PATCHES && mode != NSGPU_CAST_TILES)) return NSGPU_E_ARG;
    if (c->cast_mode != mode) ++c->setgen;
    c->cast_mode = mode;
    return NSGPU_OK;
}

int nsgpu_cast_rays(nsgpu_ctx *c, int32_t direction, int32_t n, const double *origins, const double *coords,
                    int32_t max_hits, int32_t *counts, double *hits, double *normals) {
    if (!c || n < 0 || direction < 0 || direction > 2 || max_hits < 1) return NSGPU_E_ARG;
    if (n > 0 && (!origins || !coords || !counts || !hits)) return NSGPU_E_ARG;
    if (!c->have_grid || !c->have_patches) return fail(c, NSGPU_E_STATE, "set_grid and set_patches first");
    if (c->surface_is_mesh) return fail(c, NSGPU_E_STATE, "single rays need an SES or skin patch set");
    if (n == 0) return NSGPU_OK;
    HIP_TRY(c, hipSetDevice(c->device));
    int rc = ensure_cells(c);
    if (rc) return rc;
    // direction 0 = x, 1 = y, 2 = z; panels 0 (x), 1 (z), 2 (y) (NanoShaper::panelResolver)
    const int panel = direction == 0 ? 0 : (direction == 2 ? 1 : 2);
    const size_t slots = (size_t)n * max_hits;
    HIP_TRY(c, c->crorig.ensure((size_t)n * 24));
    HIP_TRY(c, c->crcoord.ensure((size_t)n * 8));
    HIP_TRY(c, c->crcount.ensure((size_t)n * 4));
    HIP_TRY(c, c->crhits.ensure(slots * 8));
    if (normals) HIP_TRY(c, c->crnorm.ensure(slots * 24));
    HIP_TRY(c, hipMemcpyAsync(c->crorig.p, origins, (size_t)n * 24, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->crcoord.p, coords, (size_t)n * 8, hipMemcpyHostToDevice, c->stream));
    launch_cast_rays(c->stream, dev_grid(c), dev_patches(c), cell_grid(c), c->probe, panel, n, c->crorig.as<double>(),
                     c->crcoord.as<double>(), max_hits, c->crcount.as<int>(), c->crhits.as<double>(),
                     normals ? c->crnorm.as<double>() : nullptr);
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, hipMemcpyAsync(counts, c->crcount.p, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipMemcpyAsync(hits, c->crhits.p, slots * 8, hipMemcpyDeviceToHost, c->stream));
    if (normals) HIP_TRY(c, hipMemcpyAsync(normals, c->crnorm.p, slots * 24, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return NSGPU_OK;
}

}  // extern "C"

int nsg_ctx_device(const nsgpu_ctx *c) { return c->device; }

// The input buffers (grid, patches, mesh, atoms) and the result block stay;
// every other buffer of the context is set aside for an allocation-free call.
static bool kept_buffer(nsgpu_ctx *c, const DBuf *b) {
    const DBuf *keep[] = {&c->gx, &c->gy, &c->gz, &c->pkind, &c->poff, &c->pcnt, &c->pflags, &c->pbsph,
                          &c->pdata, &c->porder, &c->patom, &c->atoms, &c->mverts, &c->mtris, &c->mrr, &c->dres};
    for (const DBuf *k : keep)
        if (k == b) return true;
    return false;
}

int nsg_get_surf_in(nsgpu_ctx *c, const nsgpu_surf_opts *o, nsgpu_surf_result *res, int32_t *epsmap,
                    uint8_t *idebmap, int32_t *status, void *scratch, size_t bytes) {
    if (!c || !o) return NSGPU_E_ARG;
    if (!scratch) return fail(c, NSGPU_E_DEVICE, "no scratch array on this device");
    if (o->accurate_triangulation)
        return fail(c, NSGPU_E_ARG, "the allocation-free getSurf has no edge-ray panels (accurate_triangulation = 0)");
    if ((o->build_epsmap && (!epsmap || !idebmap)) || (o->build_status && !status))
        return fail(c, NSGPU_E_ARG, "a map to build has no output array");
    int rc = begin_surf(c, o);
    if (rc) return rc;
    HIP_TRY(c, hipSetDevice(c->device));
    const int64_t nvox = (int64_t)c->grid.nx * c->grid.ny * c->grid.nz;
    // set the context's own buffers aside (a fixed array: no host allocation)
    constexpr int kMaxBufs = 192;
    DBuf saved[kMaxBufs];
    DBuf *which[kMaxBufs];
    int ns = 0;
    for_each_buf(c, [&](DBuf *b) {
        if (kept_buffer(c, b) || ns == kMaxBufs) return;
        which[ns] = b;
        saved[ns++] = *b;
    });
    Arena a{static_cast<char *>(scratch), bytes, 0, 0};
    c->record_pid = false;
    // the patch-centric cast: the cell grid's lists would need allocating
    const int cast_mode = c->cast_mode;
    c->cast_mode = NSGPU_CAST_PATCHES;
    for (int attempt = 0;; ++attempt) {
        if (attempt == 4) {
            rc = fail(c, NSGPU_E_STATE, "ray hit capacity did not converge");
            break;
        }
        a.used = 0;
        for (int q = 0; q < ns; ++q) {
            which[q]->p = nullptr;
            which[q]->cap = 0;
        }
        // the maps are the caller's arrays; a map not built is never read or
        // written by the kernels (the fill and the cavity labels test
        // build_eps / build_status), so it only needs a non-null address
        c->eps.p = o->build_epsmap ? (void *)epsmap : scratch;
        c->ideb.p = o->build_epsmap ? (void *)idebmap : scratch;
        c->status.p = o->build_status ? (void *)status : scratch;
        c->eps.cap = o->build_epsmap ? (size_t)nvox * 12 : SIZE_MAX;
        c->ideb.cap = o->build_epsmap ? (size_t)nvox : SIZE_MAX;
        c->status.cap = o->build_status ? (size_t)nvox * 4 : SIZE_MAX;
        g_arena = &a;
        rc = enqueue_surf(c, c->stream);
        g_arena = nullptr;
        if (rc) {
            // kernels already queued may still write the caller's arrays
            (void)hipStreamSynchronize(c->stream);
            if (a.need > a.cap) {
                char msg[160];
                std::snprintf(msg, sizeof msg, "the call needs %zu bytes of scratch; the allocation-free path holds %zu",
                              a.need, a.cap);
                rc = fail(c, NSGPU_E_NOMEM, msg);
            }
            break;
        }
        if (hipStreamSynchronize(c->stream) != hipSuccess) {
            rc = fail(c, NSGPU_E_DEVICE, "stream synchronisation failed");
            break;
        }
        if (!regrow(c)) break;  // longer ray lists: the next attempt carves them
    }
    if (!rc) {
        take_trace_stats(c);
        int64_t traced = 0;
        (void)volume_slots(c, &traced);
        c->res.total_rays = traced;
        c->res.volume = c->hres->volume();
        c->cav_size.clear();
        c->cav_filled.clear();
        c->cav_on_device = false;  // the sizes stayed in the scratch array
        c->res.num_cavities = 0;
        if (c->opts.fill_cavities && c->opts.build_status) {
            const int *info = c->hres->cav_info();
            if (info[0] == 0)
                rc = fail(c, NSGPU_E_STATE,
                          "more cavity components than the device numbering holds: use nsgpu_get_surf");
            else
                c->res.num_cavities = info[1];
        }
    }
    for (int q = 0; q < ns; ++q) *which[q] = saved[q];
    c->cast_mode = cast_mode;
    // the context's buffers do not hold this call's results
    c->have_trace = c->have_surf = c->have_vertices = c->have_mesh = false;
    if (!rc && res) *res = c->res;
    return rc;
}

extern "C" {

int nsgpu_triangulate(nsgpu_ctx *c, nsgpu_mesh_info *info) {
    if (!c) return NSGPU_E_ARG;
    if (!c->have_vertices) return fail(c, NSGPU_E_STATE, "get_surf with accurate_triangulation first");
    HIP_TRY(c, hipSetDevice(c->device));
    const DevGrid g = dev_grid(c);
    const int64_t nseg = mc_num_segments(g);
    mark(c, ST_MC, 0, c->stream);
    // 1. classify 32-cube row segments, compact the active cubes in cube order
    // (getSurf did the classification and read its count when the bitmap has
    // not changed since)
    int64_t tot[2] = {c->pre_nact, 0};
    int rc = NSGPU_OK;
    if (!c->pre_mc) {
    HIP_TRY(c, c->segmask.ensure((size_t)std::max<int64_t>(nseg, 1) * 4));
    HIP_TRY(c, c->segcnt.ensure((size_t)std::max<int64_t>(nseg, 1) * 4));
    HIP_TRY(c, c->segoff.ensure((size_t)std::max<int64_t>(nseg, 1) * 4));
    HIP_TRY(c, c->scan_scratch.ensure((size_t)scan_scratch_ints(std::max<int64_t>(nseg, 1)) * 4));
    launch_mc_classify(c->stream, g, c->bits.as<uint32_t>(), c->segmask.as<uint32_t>(), c->segcnt.as<int>());
    launch_scan(c->stream, c->segcnt.as<int>(), nseg, c->segoff.as<int>(), c->scan_scratch.as<int>());
    HIP_TRY(c, hipGetLastError());
    // the active-cube count, with the analytic vertex count if still pending
    {
        const int *cnts[2] = {c->segcnt.as<int>(), c->dnv.as<int>()};
        const int *offs[2] = {c->segoff.as<int>(), c->dnv.as<int>() + 1};
        const int64_t ns[2] = {nseg, c->nv_pending ? 1 : 0};
        rc = read_totals(c, 2, cnts, offs, ns, tot);
        if (rc) return rc;
        if (c->nv_pending) {
            rc = set_nv(c, tot[1]);
            if (rc) return rc;
        }
    }
    }
    c->pre_mc = false;
    const int64_t nact = tot[0];
    c->nverts = c->n_analytic;
    const size_t na = (size_t)std::max<int64_t>(nact, 1);
    HIP_TRY(c, c->anode.ensure(na * 4));
    HIP_TRY(c, c->acase.ensure(na));
    HIP_TRY(c, c->anv.ensure(na * 4));
    HIP_TRY(c, c->ant.ensure(na * 4));
    HIP_TRY(c, c->avoff.ensure(na * 4));
    HIP_TRY(c, c->atoff.ensure(na * 4));
    HIP_TRY(c, c->scan_scratch.ensure((size_t)scan_scratch_ints(std::max<int64_t>(std::max<int64_t>(nseg, nact), 1)) * 4));
    launch_mc_compact(c->stream, g, c->bits.as<uint32_t>(), c->segmask.as<uint32_t>(), c->segoff.as<int>(),
                      c->anode.as<int>(), c->acase.as<uint8_t>());
    // 2. per active cube: mid-point vertices added and triangles emitted
    HIP_TRY(c, c->amask.ensure(na));
    McVertexMap vm{};
    vm.offs = c->voffs.as<int>();
    vm.vbits = c->vbits.as<uint64_t>();
    vm.vpre = c->vpre.as<int>();
    vm.N = dev_rays(c).N;
    vm.segmask = c->segmask.as<uint32_t>();
    vm.segoff = c->segoff.as<int>();
    vm.amask = c->amask.as<uint8_t>();
    vm.avoff = c->avoff.as<int>();
    vm.base = (int)c->n_analytic;
    launch_mc_counts(c->stream, g, (int)nact, c->anode.as<int>(), c->acase.as<uint8_t>(), vm, c->amask.as<uint8_t>(),
                     c->anv.as<int>(), c->ant.as<int>());
    launch_scan(c->stream, c->anv.as<int>(), nact, c->avoff.as<int>(), c->scan_scratch.as<int>());
    launch_scan(c->stream, c->ant.as<int>(), nact, c->atoff.as<int>(), c->scan_scratch.as<int>());
    HIP_TRY(c, hipGetLastError());
    // A cube adds at most three mid points (edges 5, 6, 10) and emits at most
    // five triangles: the buffers are sized by those bounds, so the totals
    // are read once, at the end, with the area's partial sums.
    const int64_t nadd_cap = 3 * nact, nt_cap = 5 * nact;
    const int64_t nv_cap = c->n_analytic + nadd_cap;
    if (nv_cap >= ((int64_t)1 << 31) || nt_cap >= ((int64_t)1 << 31)) return fail(c, NSGPU_E_NOMEM, "too many vertices");
    if ((size_t)nv_cap * 24 > c->verts.cap) {
        // grow while keeping the analytic vertices
        DBuf nb;
        HIP_TRY(c, nb.ensure((size_t)nv_cap * 24));
        if (c->n_analytic)
            HIP_TRY(c, hipMemcpyAsync(nb.p, c->verts.p, (size_t)c->n_analytic * 24, hipMemcpyDeviceToDevice, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        c->verts.release();
        c->verts = nb;
    }
    if ((size_t)nv_cap * 8 > c->vslot.cap) {
        DBuf nb;
        HIP_TRY(c, nb.ensure((size_t)nv_cap * 8));
        if (c->n_analytic)
            HIP_TRY(c, hipMemcpyAsync(nb.p, c->vslot.p, (size_t)c->n_analytic * 8, hipMemcpyDeviceToDevice, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        c->vslot.release();
        c->vslot = nb;
    }
    // 3. write the mid points, then the triangles that may reference them
    HIP_TRY(c, c->vorig.ensure((size_t)std::max<int64_t>(nadd_cap, 1) * 4));
    launch_mc_vwrite(c->stream, g, (int)nact, c->anode.as<int>(), c->acase.as<uint8_t>(), vm, c->verts.as<double>(),
                     c->vslot.as<int64_t>(), c->vorig.as<int>());
    c->nact = nact;
    HIP_TRY(c, c->tris.ensure((size_t)std::max<int64_t>(nt_cap, 1) * 12));
    HIP_TRY(c, c->tarea.ensure((size_t)std::max<int64_t>(nt_cap, 1) * 8));
    launch_mc_twrite(c->stream, g, (int)nact, c->anode.as<int>(), c->acase.as<uint8_t>(), vm, c->atoff.as<int>(),
                     c->verts.as<double>(), c->tris.as<int>(), c->tarea.as<double>());
    HIP_TRY(c, hipGetLastError());
    const int64_t per = sum_partials_count(nt_cap);
    HIP_TRY(c, c->partial.ensure((size_t)std::max<int64_t>(per, 1) * 8));
    if (nact)
        launch_sum_blocks_scan(c->stream, c->tarea.as<double>(), nt_cap, c->ant.as<int>() + nact - 1,
                               c->atoff.as<int>() + nact - 1, c->partial.as<double>());
    HIP_TRY(c, hipGetLastError());
    mark(c, ST_MC, 1, c->stream);
    // one read: the mid-point and triangle totals, then the area partials
    int64_t nadd = 0, nt = 0;
    double area = 0;
    if (nact) {
        GatherSrc gs{};
        gs.src[0] = c->anv.as<int>() + nact - 1;
        gs.src[1] = c->avoff.as<int>() + nact - 1;
        gs.src[2] = c->ant.as<int>() + nact - 1;
        gs.src[3] = c->atoff.as<int>() + nact - 1;
        gs.n = 4;
        gs.dsrc = c->partial.as<double>();
        gs.nd = per;
        const int *tail = nullptr;
        const double *hpart = nullptr;
        rc = read_gathered(c, c->stream, gs, &tail, &hpart);
        if (rc) return rc;
        nadd = (int64_t)tail[0] + tail[1];
        nt = (int64_t)tail[2] + tail[3];
        for (int64_t q = 0; q < per; ++q) area += hpart[q];  // partials past the last triangle are 0
    } else {
        HIP_TRY(c, hipStreamSynchronize(c->stream));
    }
    const int64_t nv = c->n_analytic + nadd;
    c->nverts = nv;
    c->ntris = nt;
    c->area = area;
    c->have_mesh = true;
    c->normals_smoothed = false;
    if (info) {
        info->num_vertices = c->nverts;
        info->num_triangles = c->ntris;
        info->num_analytic_vertices = c->n_analytic;
        info->area = area;
    }
    return NSGPU_OK;
}

static int read_buf(nsgpu_ctx *c, const DBuf &b, void *out, size_t bytes) {
    if (!c || !out) return NSGPU_E_ARG;
    if (!c->have_surf) return fail(c, NSGPU_E_STATE, "no surface computed");
    HIP_TRY(c, hipSetDevice(c->device));
    if (bytes) HIP_TRY(c, hipMemcpyAsync(out, b.p, bytes, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return NSGPU_OK;
}

static int64_t nvox_of(const nsgpu_ctx *c) { return (int64_t)c->grid.nx * c->grid.ny * c->grid.nz; }

// ---------------------------------------------------------------------------
// Triangle-mesh surfaces: MeshSurface (a loaded mesh) and BlobbySurface (the
// blobby field triangulated, smoothed and reloaded as a mesh).

// Auxiliary 2D grid of MeshSurface::preProcessPanel (MeshSurface.cpp:338-384):
// cells of gridMul x gridMul rays, gridMul the smallest odd factor that brings
// the grid to at most aux_dim cells.  A ray tests the triangles registered in
// its cell (getRayIntersection :1779-1796), so the rays that test a triangle
// are those whose cell lies in the triangle's cell range; per axis and phase
// the rays of a cell range form one interval, tabulated here.
static int setup_aux_grid(nsgpu_ctx *c, int aux_dim, int ray_based) {
    const nsgpu_grid &G = c->grid;
    int64_t igrid = G.nx;
    int gridMul = 1;
    while (igrid > aux_dim) {
        gridMul += 2;
        int64_t digrid = G.nx;
        while (true) {
            const int64_t fixedPoint = (digrid + (gridMul - 1)) / gridMul;
            igrid = fixedPoint * gridMul;
            if (igrid % 2 == 0) {
                digrid = igrid + 1;
            } else {
                igrid = fixedPoint;
                break;
            }
        }
    }
    const double scale2d = G.scale / ((double)gridMul);
    const double side2d = ((double)gridMul) / G.scale;
    for (int d = 0; d < 3; ++d) c->m_min2d[d] = G.baricenter[d] - (igrid - 1) * 0.5 * side2d;
    c->m_scale2d = scale2d;
    c->m_n2d = (int)igrid;
    c->m_aux_dim = aux_dim;
    const int n2d = (int)igrid;
    std::vector<int> rr((size_t)3 * 2 * 2 * n2d);
    const double mins[3] = {G.xmin, G.ymin, G.zmin};
    const int ns[3] = {G.nx, G.ny, G.nz};
    for (int d = 0; d < 3; ++d)
        for (int ph = 0; ph < 2; ++ph) {
            const double delta = ph ? 1e-10 - G.hside : 0.;
            std::vector<int64_t> cell(ns[d]);
            for (int m = 0; m < ns[d]; ++m) {
                const double coord = mins[d] + m * G.side;  // the grid coordinates (DelphiShared.cpp:1135-1142)
                cell[m] = (int64_t)std::floor(((coord + delta) - c->m_min2d[d]) * scale2d + 0.5);
                // the reference indexes outside its cell table there
                if (ray_based && (cell[m] < 0 || cell[m] >= n2d))
                    return fail(c, NSGPU_E_ARG, "a grid ray falls outside the auxiliary 2D grid");
            }
            int *lo = rr.data() + (size_t)((d * 2 + ph) * 2) * n2d, *hi = lo + n2d;
            for (int q = 0; q < n2d; ++q) {
                int a = 0;
                while (a < ns[d] && cell[a] < q) ++a;
                int b = ns[d] - 1;
                while (b >= 0 && cell[b] > q) --b;
                lo[q] = a;
                hi[q] = b;
            }
        }
    HIP_TRY(c, c->mrr.ensure(rr.size() * 4));
    HIP_TRY(c, hipMemcpyAsync(c->mrr.p, rr.data(), rr.size() * 4, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return NSGPU_OK;
}

static void install_mesh(nsgpu_ctx *c, int64_t nv, int64_t nt, int ray_based) {
    c->mnv = nv;
    c->mnt = nt;
    c->mesh_ray_based = ray_based != 0;
    c->surface_is_mesh = true;
    c->have_patches = true;
    c->npatch = c->npatch_plain = c->npatch_skin = 0;
    c->patch_max_atom = -1;
    c->have_trace = c->have_surf = c->have_vertices = c->have_mesh = false;
}

int nsgpu_set_mesh(nsgpu_ctx *c, const double *verts, int32_t nv, const int32_t *tris, int32_t nt, int32_t ray_based,
                   int32_t aux_grid_dim) {
    if (!c || nv < 0 || nt < 0 || (nv && !verts) || (nt && !tris) || aux_grid_dim < 1) return NSGPU_E_ARG;
    if (!c->have_grid) return fail(c, NSGPU_E_STATE, "set_grid first");
    ++c->setgen;
    c->cg_valid = false;
    for (int64_t q = 0; q < 3 * (int64_t)nt; ++q)
        if (tris[q] < 0 || tris[q] >= nv) return fail(c, NSGPU_E_ARG, "triangle vertex index out of range");
    for (int64_t q = 0; q < 3 * (int64_t)nv; ++q)
        if (!std::isfinite(verts[q])) return fail(c, NSGPU_E_ARG, "non-finite vertex coordinate");
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, c->mverts.ensure((size_t)std::max(nv, 1) * 24));
    HIP_TRY(c, c->mtris.ensure((size_t)std::max(nt, 1) * 12));
    if (nv) HIP_TRY(c, hipMemcpyAsync(c->mverts.p, verts, (size_t)nv * 24, hipMemcpyHostToDevice, c->stream));
    if (nt) HIP_TRY(c, hipMemcpyAsync(c->mtris.p, tris, (size_t)nt * 12, hipMemcpyHostToDevice, c->stream));
    int rc = setup_aux_grid(c, aux_grid_dim, ray_based);
    if (rc) return rc;
    install_mesh(c, nv, nt, ray_based);
    c->have_raw = false;
    return NSGPU_OK;
}

int nsgpu_blobby_build(nsgpu_ctx *c, const double *xyzr, int32_t n, double B, double cutoff, int32_t aux_grid_dim,
                       nsgpu_mesh_info *info) {
    if (!c || n < 0 || (n && !xyzr) || !(cutoff > 0) || aux_grid_dim < 1) return NSGPU_E_ARG;
    if (!c->have_grid) return fail(c, NSGPU_E_STATE, "set_grid first");
    if (!(B < 0)) B = -2.5;  // BlobbySurface::setBlobbyness (BlobbySurface.cpp:46-54)
    ++c->setgen;
    c->cg_valid = false;
    HIP_TRY(c, hipSetDevice(c->device));
    c->pre_mc = false;  // the classification buffers are reused below
    const DevGrid g = dev_grid(c);
    const int64_t nvox = (int64_t)g.nx * g.ny * g.nz;
    hipStream_t st = c->stream;
    // 1. field: atom/brick pairs, brick lists, per-brick ordered sums
    BlobbyArgs a;
    a.n = n;
    a.B = B;
    a.ng = (int)(cutoff / c->grid.side + 0.5);
    blobby_bricks(g, &a.nbx, &a.nby, &a.nbz);
    const int64_t nb = (int64_t)a.nbx * a.nby * a.nbz;
    HIP_TRY(c, c->batoms.ensure((size_t)std::max(n, 1) * 32));
    if (n) HIP_TRY(c, hipMemcpyAsync(c->batoms.p, xyzr, (size_t)n * 32, hipMemcpyHostToDevice, st));
    HIP_TRY(c, c->bbox.ensure((size_t)std::max(n, 1) * 32));
    HIP_TRY(c, c->bcount.ensure((size_t)nb * 4));
    HIP_TRY(c, c->boff.ensure((size_t)nb * 4));
    HIP_TRY(c, c->bcur.ensure((size_t)nb * 4));
    HIP_TRY(c, c->scan_scratch.ensure((size_t)scan_scratch_ints(nb) * 4));
    HIP_TRY(c, c->field.ensure((size_t)nvox * 8));
    HIP_TRY(c, c->flag.ensure(8));
    a.xyzr = c->batoms.as<double>();
    a.box = c->bbox.as<int>();
    a.bcount = c->bcount.as<int>();
    a.boff = c->boff.as<int>();
    a.bcur = c->bcur.as<int>();
    a.field = c->field.as<double>();
    a.overflow = c->flag.as<int>();
    launch_fill32(st, c->bcount.p, 0u, nb);
    launch_fill32(st, c->bcur.p, 0u, nb);
    launch_fill32(st, c->flag.p, 0u, 2);
    launch_blobby_pairs(st, g, a, 0);
    launch_scan(st, a.bcount, nb, a.boff, c->scan_scratch.as<int>());
    HIP_TRY(c, hipGetLastError());
    int64_t npairs = 0;
    int rc = read_total(c, a.bcount, a.boff, nb, &npairs);
    if (rc) return rc;
    if (npairs >= INT32_MAX) return fail(c, NSGPU_E_NOMEM, "too many atom/brick pairs");
    HIP_TRY(c, c->blist.ensure((size_t)std::max<int64_t>(npairs, 1) * 4));
    a.blist = c->blist.as<int>();
    launch_blobby_pairs(st, g, a, 1);
    launch_blobby_field(st, g, a);
    HIP_TRY(c, hipGetLastError());
    int over = 0;
    HIP_TRY(c, hipMemcpyAsync(&over, c->flag.p, 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(c, hipStreamSynchronize(st));
    if (over) return fail(c, NSGPU_E_ARG, "too many atoms reach one 8^3-node brick of the blobby field");
    c->have_field = true;
    // 2. marching cubes on the field at iso 1.0 (triangulateSurface(false, false, 1.0), BlobbySurface.cpp:186)
    HIP_TRY(c, c->fbits.ensure((size_t)((nvox + 31) / 32) * 4));
    launch_field_bits(st, a.field, nvox, 1.0, c->fbits.as<uint32_t>());
    rc = resolve_nv(c);  // the previous pass's count, if still pending
    if (rc) return rc;
    c->have_trace = c->have_surf = c->have_vertices = c->have_mesh = false;
    const int64_t nseg = mc_num_segments(g);
    HIP_TRY(c, c->segmask.ensure((size_t)std::max<int64_t>(nseg, 1) * 4));
    HIP_TRY(c, c->segcnt.ensure((size_t)std::max<int64_t>(nseg, 1) * 4));
    HIP_TRY(c, c->segoff.ensure((size_t)std::max<int64_t>(nseg, 1) * 4));
    HIP_TRY(c, c->scan_scratch.ensure((size_t)scan_scratch_ints(std::max<int64_t>(nseg, 1)) * 4));
    launch_mc_classify(st, g, c->fbits.as<uint32_t>(), c->segmask.as<uint32_t>(), c->segcnt.as<int>());
    launch_scan(st, c->segcnt.as<int>(), nseg, c->segoff.as<int>(), c->scan_scratch.as<int>());
    HIP_TRY(c, hipGetLastError());
    int64_t nact = 0;
    rc = read_total(c, c->segcnt.as<int>(), c->segoff.as<int>(), nseg, &nact);
    if (rc) return rc;
    const size_t na = (size_t)std::max<int64_t>(nact, 1);
    HIP_TRY(c, c->anode.ensure(na * 4));
    HIP_TRY(c, c->acase.ensure(na));
    HIP_TRY(c, c->anv.ensure(na * 4));
    HIP_TRY(c, c->ant.ensure(na * 4));
    HIP_TRY(c, c->avoff.ensure(na * 4));
    HIP_TRY(c, c->atoff.ensure(na * 4));
    HIP_TRY(c, c->scan_scratch.ensure((size_t)scan_scratch_ints(std::max<int64_t>(std::max<int64_t>(nseg, nact), 1)) * 4));
    launch_mc_compact(st, g, c->fbits.as<uint32_t>(), c->segmask.as<uint32_t>(), c->segoff.as<int>(), c->anode.as<int>(),
                      c->acase.as<uint8_t>());
    HIP_TRY(c, c->amask.ensure(na));
    // a scalar field has no analytic vertices: every vertex is a mid point
    McVertexMap vm{};
    vm.segmask = c->segmask.as<uint32_t>();
    vm.segoff = c->segoff.as<int>();
    vm.amask = c->amask.as<uint8_t>();
    vm.avoff = c->avoff.as<int>();
    vm.base = 0;
    launch_mc_counts(st, g, (int)nact, c->anode.as<int>(), c->acase.as<uint8_t>(), vm, c->amask.as<uint8_t>(),
                     c->anv.as<int>(), c->ant.as<int>());
    launch_scan(st, c->anv.as<int>(), nact, c->avoff.as<int>(), c->scan_scratch.as<int>());
    launch_scan(st, c->ant.as<int>(), nact, c->atoff.as<int>(), c->scan_scratch.as<int>());
    HIP_TRY(c, hipGetLastError());
    int64_t tot[2] = {0, 0};  // vertices, triangles
    {
        const int *cnts[2] = {c->anv.as<int>(), c->ant.as<int>()};
        const int *offs[2] = {c->avoff.as<int>(), c->atoff.as<int>()};
        const int64_t ns[2] = {nact, nact};
        rc = read_totals(c, 2, cnts, offs, ns, tot);
        if (rc) return rc;
    }
    const int64_t nv = tot[0], nt = tot[1];
    if (nv >= INT32_MAX / 3 || nt >= INT32_MAX / 3) return fail(c, NSGPU_E_NOMEM, "blobby mesh too large");
    HIP_TRY(c, c->verts.ensure((size_t)std::max<int64_t>(nv, 1) * 24));
    HIP_TRY(c, c->vslot.ensure((size_t)std::max<int64_t>(nv, 1) * 8));
    HIP_TRY(c, c->vorig.ensure((size_t)std::max<int64_t>(nv, 1) * 4));
    launch_mc_vwrite(st, g, (int)nact, c->anode.as<int>(), c->acase.as<uint8_t>(), vm, c->verts.as<double>(),
                     c->vslot.as<int64_t>(), c->vorig.as<int>(), a.field, 1.0);
    HIP_TRY(c, c->tris.ensure((size_t)std::max<int64_t>(nt, 1) * 12));
    HIP_TRY(c, c->tarea.ensure((size_t)std::max<int64_t>(nt, 1) * 8));
    launch_mc_twrite(st, g, (int)nact, c->anode.as<int>(), c->acase.as<uint8_t>(), vm, c->atoff.as<int>(),
                     c->verts.as<double>(), c->tris.as<int>(), c->tarea.as<double>(), 0);
    const int64_t per = sum_partials_count(nt);
    HIP_TRY(c, c->partial.ensure((size_t)std::max<int64_t>(per, 1) * 8));
    launch_sum_blocks(st, c->tarea.as<double>(), nt, c->partial.as<double>());
    HIP_TRY(c, hipGetLastError());
    // 3. one smoothing pass (smoothSurface, BlobbySurface.cpp:188)
    HIP_TRY(c, c->mraw.ensure((size_t)std::max<int64_t>(nv, 1) * 24));
    HIP_TRY(c, c->mverts.ensure((size_t)std::max<int64_t>(nv, 1) * 24));
    HIP_TRY(c, c->mtris.ensure((size_t)std::max<int64_t>(nt, 1) * 12));
    SmoothArgs sm;
    sm.vin = c->verts.as<double>();
    sm.vslot = c->vslot.as<int64_t>();
    sm.nv = nv;
    sm.anode = c->anode.as<int>();
    sm.ant = c->ant.as<int>();
    sm.atoff = c->atoff.as<int>();
    sm.tris = c->tris.as<int>();
    sm.nact = (int)nact;
    sm.vout = c->mraw.as<double>();
    sm.nin = nullptr;
    sm.nout = nullptr;
    HIP_TRY(c, c->cubeidx.ensure((size_t)std::max<int64_t>(nvox, 1) * 4));
    sm.cube_index = c->cubeidx.as<int>();
    launch_smooth(st, g, sm);
    // 4. the OFF text round trip ("%.3f" written, "%f" read into floats) and the mesh
    launch_quantize(st, c->mraw.as<double>(), 3 * nv, c->mverts.as<double>());
    if (nt) HIP_TRY(c, hipMemcpyAsync(c->mtris.p, c->tris.p, (size_t)nt * 12, hipMemcpyDeviceToDevice, st));
    HIP_TRY(c, hipGetLastError());
    double area = 0;
    rc = read_partial_sum(c, c->partial.as<double>(), per, st, &area);
    if (rc) return rc;
    rc = setup_aux_grid(c, aux_grid_dim, 1);
    if (rc) return rc;
    install_mesh(c, nv, nt, 1);
    c->have_raw = true;
    c->have_field = true;
    if (info) {
        info->num_vertices = nv;
        info->num_triangles = nt;
        info->num_analytic_vertices = 0;
        info->area = area;
    }
    return NSGPU_OK;
}

int nsgpu_surface_mesh_size(nsgpu_ctx *c, int64_t *nv, int64_t *nt) {
    if (!c || !nv || !nt) return NSGPU_E_ARG;
    if (!c->surface_is_mesh) return fail(c, NSGPU_E_STATE, "no mesh surface");
    *nv = c->mnv;
    *nt = c->mnt;
    return NSGPU_OK;
}

int nsgpu_read_surface_mesh(nsgpu_ctx *c, int32_t raw, double *verts, int32_t *tris) {
    if (!c) return NSGPU_E_ARG;
    if (!c->surface_is_mesh) return fail(c, NSGPU_E_STATE, "no mesh surface");
    if (raw && !c->have_raw) return fail(c, NSGPU_E_STATE, "raw vertices exist for a blobby build only");
    HIP_TRY(c, hipSetDevice(c->device));
    if (verts && c->mnv)
        HIP_TRY(c, hipMemcpyAsync(verts, (raw ? c->mraw : c->mverts).p, (size_t)c->mnv * 24, hipMemcpyDeviceToHost,
                                  c->stream));
    if (tris && c->mnt)
        HIP_TRY(c, hipMemcpyAsync(tris, c->mtris.p, (size_t)c->mnt * 12, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return NSGPU_OK;
}

int nsgpu_read_blobby_field(nsgpu_ctx *c, double *field) {
    if (!c || !field) return NSGPU_E_ARG;
    if (!c->have_field) return fail(c, NSGPU_E_STATE, "no blobby field for this grid");
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipMemcpyAsync(field, c->field.p, (size_t)nvox_of(c) * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return NSGPU_OK;
}

int nsgpu_read_epsmap(nsgpu_ctx *c, int32_t *out) { return read_buf(c, c->eps, out, (size_t)nvox_of(c) * 12); }
int nsgpu_read_idebmap(nsgpu_ctx *c, uint8_t *out) { return read_buf(c, c->ideb, out, (size_t)nvox_of(c)); }
int nsgpu_read_status(nsgpu_ctx *c, int32_t *out) { return read_buf(c, c->status, out, (size_t)nvox_of(c) * 4); }
int nsgpu_read_corner_bits(nsgpu_ctx *c, uint32_t *out) {
    return read_buf(c, c->bits, out, (size_t)((nvox_of(c) + 31) / 32) * 4);
}

int nsgpu_read_cavities(nsgpu_ctx *c, double *sizes, uint8_t *filled) {
    if (!c) return NSGPU_E_ARG;
    if (c->cav_on_device) {
        const int ncav = c->res.num_cavities;
        c->cav_size.resize(ncav);
        c->cav_filled.resize(ncav);
        if (ncav) {
            HIP_TRY(c, hipSetDevice(c->device));
            HIP_TRY(c, hipMemcpyAsync(c->cav_size.data(), c->dcavsize.p, (size_t)ncav * 8, hipMemcpyDeviceToHost,
                                      c->stream));
            HIP_TRY(c, hipMemcpyAsync(c->cav_filled.data(), c->dcavfilled.p, (size_t)ncav, hipMemcpyDeviceToHost,
                                      c->stream));
            HIP_TRY(c, hipStreamSynchronize(c->stream));
        }
        c->cav_on_device = false;
    }
    for (size_t q = 0; q < c->cav_size.size(); ++q) {
        if (sizes) sizes[q] = c->cav_size[q];
        if (filled) filled[q] = c->cav_filled[q];
    }
    return (int)c->cav_size.size();
}

int nsgpu_read_mesh(nsgpu_ctx *c, double *verts, int32_t *tris) {
    if (!c) return NSGPU_E_ARG;
    if (!c->have_mesh) return fail(c, NSGPU_E_STATE, "no mesh computed");
    HIP_TRY(c, hipSetDevice(c->device));
    if (verts && c->nverts)
        HIP_TRY(c, hipMemcpyAsync(verts, c->verts.p, (size_t)c->nverts * 24, hipMemcpyDeviceToHost, c->stream));
    if (tris && c->ntris)
        HIP_TRY(c, hipMemcpyAsync(tris, c->tris.p, (size_t)c->ntris * 12, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return NSGPU_OK;
}

int nsgpu_read_ray_hits(nsgpu_ctx *c, int32_t slot, int32_t max_hits, int32_t *counts, double *t) {
    if (!c || !counts || !t || max_hits <= 0) return NSGPU_E_ARG;
    if (!c->have_trace) return fail(c, NSGPU_E_STATE, "no rays traced");
    if (slot < 0 || slot >= c->nslots) return fail(c, NSGPU_E_ARG, "slot out of range");
    if (max_hits < c->res.max_hits_per_ray) return fail(c, NSGPU_E_ARG, "max_hits below the longest ray");
    HIP_TRY(c, hipSetDevice(c->device));
    const int64_t N = std::max(c->grid.nx, std::max(c->grid.ny, c->grid.nz));
    const int64_t NN = N * N;
    std::vector<double> buf((size_t)NN * c->K);
    HIP_TRY(c, hipMemcpyAsync(counts, c->rcnt.as<int>() + slot * NN, NN * 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipMemcpyAsync(buf.data(), c->rt.as<double>() + slot * NN * c->K, buf.size() * 8,
                              hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    for (int64_t r = 0; r < NN; ++r)
        for (int h = 0; h < counts[r] && h < max_hits; ++h) t[r * max_hits + h] = buf[(size_t)r * c->K + h];
    return NSGPU_OK;
}

int nsgpu_set_atoms(nsgpu_ctx *c, const double *xyzr, int32_t n) {
    if (!c || n < 0 || (n > 0 && !xyzr)) return NSGPU_E_ARG;
    ++c->setgen;
    c->cg_valid = false;
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, c->atoms.ensure(std::max<size_t>((size_t)n, 1) * 32));
    if (n) HIP_TRY(c, hipMemcpyAsync(c->atoms.p, xyzr, (size_t)n * 32, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->natoms = n;
    return NSGPU_OK;
}

int nsgpu_surface_gradient(nsgpu_ctx *c, double *volume, double *area, double *dvolume, double *darea,
                           nsgpu_surf_result *res) {
    if (!c || !volume || !area) return NSGPU_E_ARG;
    if (!c->have_grid || !c->have_patches) return fail(c, NSGPU_E_STATE, "set_grid and set_patches first");
    if (c->surface_is_mesh || c->npatch_skin) return fail(c, NSGPU_E_STATE, "gradients need an SES patch set");
    if (c->npatch > 0 && c->patch_max_atom < 0) return fail(c, NSGPU_E_STATE, "the patch set carries no atom ids");
    if (c->patch_max_atom >= c->natoms) return fail(c, NSGPU_E_STATE, "set_atoms with the patch set's atoms first");
    HIP_TRY(c, hipSetDevice(c->device));
    // the three grid-ray panels only (the volume of getSurf with the epsmap)
    c->opts = nsgpu_surf_opts{1, 0, 0, 0, 0.0, 5, NSGPU_PREC_FP64};
    std::memset(&c->res, 0, sizeof c->res);
    c->have_trace = c->have_surf = c->have_vertices = c->have_mesh = false;
    c->panels[0] = 3;
    c->panels[1] = 0;
    c->nslots = 3;
    c->record_pid = true;
    const int rc = trace(c);
    c->record_pid = false;
    if (rc) return rc;
    const DevGrid g = dev_grid(c);
    DevRays R = dev_rays(c);
    R.pid = c->rpid.as<int>();
    DevPatches P;
    P.n = c->npatch;
    P.kind = c->pkind.as<int>();
    P.off = c->poff.as<int>();
    P.cnt = c->pcnt.as<int>();
    P.flags = c->pflags.as<int>();
    P.bsph = c->pbsph.as<double>();
    P.data = c->pdata.as<double>();
    P.order = c->porder.as<int>();
    P.n_plain = c->npatch_plain;
    P.n_skin = c->npatch_skin;
    P.atom = c->patom.as<int>();
    const size_t ng = std::max<size_t>((size_t)c->natoms, 1) * 4;
    const int64_t nb = adjoint_blocks(R);
    HIP_TRY(c, c->gradv.ensure(ng * 8));
    HIP_TRY(c, c->grada.ensure(ng * 8));
    HIP_TRY(c, c->asums.ensure((size_t)std::max<int64_t>(nb, 1) * 16));
    FillList fl{};
    fl.s[0] = FillSeg{static_cast<uint32_t *>(c->gradv.p), 0u, (int64_t)ng * 2};
    fl.s[1] = FillSeg{static_cast<uint32_t *>(c->grada.p), 0u, (int64_t)ng * 2};
    fl.k = 2;
    launch_fill_multi(c->stream, fl);
    AdjointArgs a;
    a.atoms = c->atoms.as<double>();
    a.probe = c->probe;
    a.area_px = c->grid.A;
    a.flag = 3.0;
    a.grad_v = c->gradv.as<double>();
    a.grad_a = c->grada.as<double>();
    a.sums = c->asums.as<double>();
    mark(c, ST_ADJOINT, 0, c->stream);
    launch_adjoint(c->stream, g, P, R, a);
    HIP_TRY(c, hipGetLastError());
    mark(c, ST_ADJOINT, 1, c->stream);
    // the block partials and both gradients go through the pinned staging
    // buffer (a pageable copy is staged and synchronised by the runtime on
    // its own): three queued copies, one synchronisation
    const size_t gbytes = (size_t)c->natoms * 32, sbytes = (size_t)nb * 16;
    void *hp = nullptr;
    const int prc = pinned(c, sbytes + 2 * gbytes, &hp);
    if (prc) return prc;
    char *hsum = static_cast<char *>(hp), *hdv = hsum + sbytes, *hda = hdv + gbytes;
    if (nb) HIP_TRY(c, hipMemcpyAsync(hsum, c->asums.p, sbytes, hipMemcpyDeviceToHost, c->stream));
    if (dvolume && gbytes) HIP_TRY(c, hipMemcpyAsync(hdv, c->gradv.p, gbytes, hipMemcpyDeviceToHost, c->stream));
    if (darea && gbytes) HIP_TRY(c, hipMemcpyAsync(hda, c->grada.p, gbytes, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (dvolume && gbytes) std::memcpy(dvolume, hdv, gbytes);
    if (darea && gbytes) std::memcpy(darea, hda, gbytes);
    const double *sums = static_cast<const double *>(hp);
    double v = 0., s = 0.;
    for (int64_t b = 0; b < nb; ++b) {
        v += sums[2 * b];
        s += sums[2 * b + 1];
    }
    *volume = c->grid.A * v / 3.0;
    *area = c->grid.A * s;
    c->res.volume = *volume;
    c->res.total_rays = 3LL * R.N * R.N;
    if (res) *res = c->res;
    return NSGPU_OK;
}

int nsgpu_set_vertex_normals(nsgpu_ctx *c, int on) {
    if (!c) return NSGPU_E_ARG;
    if (c->want_normals != (on != 0)) ++c->setgen;
    c->want_normals = on != 0;
    return NSGPU_OK;
}

}  // extern "C"

// the vertex normals of the current mesh into c->normals
static int compute_normals(nsgpu_ctx *c) {
    if (!c->hits_have_patches)
        return fail(c, NSGPU_E_STATE, "vertex normals need nsgpu_set_vertex_normals(ctx, 1) before get_surf");
    const DevGrid g = dev_grid(c);
    DevRays R = dev_rays(c);
    R.pid = c->rpid.as<int>();
    DevPatches P;
    P.n = c->npatch;
    P.kind = c->pkind.as<int>();
    P.off = c->poff.as<int>();
    P.cnt = c->pcnt.as<int>();
    P.flags = c->pflags.as<int>();
    P.bsph = c->pbsph.as<double>();
    P.data = c->pdata.as<double>();
    P.order = c->porder.as<int>();
    P.n_plain = c->npatch_plain;
    P.n_skin = c->npatch_skin;
    P.atom = c->patom.as<int>();
    HIP_TRY(c, c->normals.ensure((size_t)std::max<int64_t>(c->nverts, 1) * 24));
    NormalArgs a;
    a.verts = c->verts.as<double>();
    a.vhit = c->vhit.as<int64_t>();
    a.vorig = c->vorig.as<int>();
    a.anode = c->anode.as<int>();
    a.ant = c->ant.as<int>();
    a.atoff = c->atoff.as<int>();
    a.tris = c->tris.as<int>();
    a.n_analytic = c->n_analytic;
    a.nverts = c->nverts;
    a.nact = (int)c->nact;
    a.eslot0 = c->panels[0];
    a.probe = c->probe;
    a.normals = c->normals.as<double>();
    launch_normals(c->stream, g, P, R, a);
    HIP_TRY(c, hipGetLastError());
    return NSGPU_OK;
}

extern "C" {

int nsgpu_read_normals(nsgpu_ctx *c, double *normals) {
    if (!c || !normals) return NSGPU_E_ARG;
    if (!c->have_mesh) return fail(c, NSGPU_E_STATE, "no mesh computed");
    HIP_TRY(c, hipSetDevice(c->device));
    if (!c->normals_smoothed) {
        const int rc = compute_normals(c);
        if (rc) return rc;
    }
    if (c->nverts)
        HIP_TRY(c, hipMemcpyAsync(normals, c->normals.p, (size_t)c->nverts * 24, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return NSGPU_OK;
}

int nsgpu_smooth_surface(nsgpu_ctx *c) {
    if (!c) return NSGPU_E_ARG;
    if (!c->have_mesh) return fail(c, NSGPU_E_STATE, "no mesh computed");
    HIP_TRY(c, hipSetDevice(c->device));
    // normals are smoothed with the vertices when the context keeps them
    const bool smooth_normals = c->hits_have_patches;
    const int64_t nv = c->nverts, nvox = nvox_of(c);
    if (smooth_normals && !c->normals_smoothed) {
        const int rc = compute_normals(c);
        if (rc) return rc;
    }
    HIP_TRY(c, c->mraw.ensure((size_t)std::max<int64_t>(nv, 1) * 24));
    HIP_TRY(c, c->cubeidx.ensure((size_t)std::max<int64_t>(nvox, 1) * 4));
    if (smooth_normals) HIP_TRY(c, c->vorig.ensure((size_t)std::max<int64_t>(nv, 1) * 24));
    SmoothArgs sm;
    sm.vin = c->verts.as<double>();
    sm.vslot = c->vslot.as<int64_t>();
    sm.nv = nv;
    sm.anode = c->anode.as<int>();
    sm.ant = c->ant.as<int>();
    sm.atoff = c->atoff.as<int>();
    sm.tris = c->tris.as<int>();
    sm.nact = (int)c->nact;
    sm.vout = c->mraw.as<double>();
    sm.nin = smooth_normals ? c->normals.as<double>() : nullptr;
    sm.nout = smooth_normals ? c->vorig.as<double>() : nullptr;
    sm.cube_index = c->cubeidx.as<int>();
    launch_smooth(c->stream, dev_grid(c), sm);
    HIP_TRY(c, hipGetLastError());
    if (nv) {
        HIP_TRY(c, hipMemcpyAsync(c->verts.p, c->mraw.p, (size_t)nv * 24, hipMemcpyDeviceToDevice, c->stream));
        if (smooth_normals)
            HIP_TRY(c, hipMemcpyAsync(c->normals.p, c->vorig.p, (size_t)nv * 24, hipMemcpyDeviceToDevice, c->stream));
    }
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    // the mid-point provenance in vorig is spent: normals are final from here
    if (smooth_normals) c->normals_smoothed = true;
    return NSGPU_OK;
}

// ---------------------------------------------------------------------------
// Batches: many frames (grid + patch set each), one context per frame, all
// frames' getSurf sequences captured into one graph.  The frames run in
// `chains` concurrent chains (largest frames first, dealt round-robin), so
// small frames fill the device together; the chain count is picked by timing
// (nsgpu_batch_prepare with chains <= 0).

struct nsgpu_batch {
    int device = 0;
    std::vector<nsgpu_ctx *> frames;
    hipStream_t s0 = nullptr;
    hipStream_t s1 = nullptr;  // the torus launch of a joint cast, beside s0's plain one
    hipEvent_t fork = nullptr;
    hipEvent_t cast_fork = nullptr, cast_join = nullptr, cast_done = nullptr;
    std::vector<hipEvent_t> join;
    std::vector<hipEvent_t> head;  // each frame's clears done (joint cast)
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int chains = 0;
    nsgpu_batch_shape shape{};  // of the captured graph
    DBuf fcast;                 // FrameCast per frame, in item order (joint cast)
    DBuf queue;                 // the persistent launches' item counters
    int cus = 0;
    nsgpu_surf_opts opts{};
    bool prepared = false;
    DBuf stamps;  // frame completion clocks, then the start clock
    double clock_khz = 0;
    // each frame's ctx_generation when the graph was captured: a frame
    // changed since (new inputs, a buffer moved, other options through its
    // nsgpu_batch_frame context) makes nsgpu_batch_run capture again
    std::vector<uint64_t> gen;
    std::string err;
};

static int bfail(nsgpu_batch *b, int code, const std::string &msg) {
    b->err = msg;
    return code;
}

#define BHIP_TRY(b, expr)                                                                              \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess) return bfail((b), NSGPU_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

static void drop_graph(nsgpu_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->prepared = false;
}

int nsgpu_batch_create(int device, int32_t n_frames, nsgpu_batch **out) {
    if (!out || n_frames <= 0) return NSGPU_E_ARG;
    *out = nullptr;
    auto *b = new nsgpu_batch();
    b->device = device;
    for (int f = 0; f < n_frames; ++f) {
        nsgpu_ctx *c = nullptr;
        const int rc = nsgpu_create(device, &c);
        if (rc) {
            nsgpu_batch_destroy(b);
            return rc;
        }
        b->frames.push_back(c);
        hipEvent_t e = nullptr, h = nullptr;
        (void)hipEventCreateWithFlags(&e, hipEventDisableTiming);
        (void)hipEventCreateWithFlags(&h, hipEventDisableTiming);
        b->join.push_back(e);
        b->head.push_back(h);
    }
    int khz = 0;
    if (hipStreamCreateWithFlags(&b->s0, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&b->s1, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->cast_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->cast_join, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->cast_done, hipEventDisableTiming) != hipSuccess ||
        hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device) != hipSuccess ||
        hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
        b->stamps.ensure((size_t)(n_frames + 1) * 8) != hipSuccess || b->queue.ensure(2 * 4) != hipSuccess) {
        nsgpu_batch_destroy(b);
        return NSGPU_E_DEVICE;
    }
    b->clock_khz = khz;
    *out = b;
    return NSGPU_OK;
}

void nsgpu_batch_destroy(nsgpu_batch *b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->s0) (void)hipStreamSynchronize(b->s0);
    drop_graph(b);
    for (nsgpu_ctx *c : b->frames) nsgpu_destroy(c);
    for (hipEvent_t e : b->join)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : b->head)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : {b->fork, b->cast_fork, b->cast_join, b->cast_done})
        if (e) (void)hipEventDestroy(e);
    if (b->s0) (void)hipStreamDestroy(b->s0);
    if (b->s1) (void)hipStreamDestroy(b->s1);
    b->stamps.release();
    b->fcast.release();
    b->queue.release();
    delete b;
}

const char *nsgpu_batch_last_error(const nsgpu_batch *b) { return b ? b->err.c_str() : "null batch"; }

nsgpu_ctx *nsgpu_batch_frame(nsgpu_batch *b, int32_t f) {
    if (!b || f < 0 || f >= (int)b->frames.size()) return nullptr;
    return b->frames[f];
}

int nsgpu_batch_set_frame(nsgpu_batch *b, int32_t f, const nsgpu_grid *grid, const nsgpu_patches *p, double probe) {
    if (!b || f < 0 || f >= (int)b->frames.size() || !grid || !p) return NSGPU_E_ARG;
    nsgpu_ctx *c = b->frames[f];
    int rc = nsgpu_set_grid(c, grid);
    if (!rc) rc = nsgpu_set_patches(c, p, probe);
    if (rc) return bfail(b, rc, std::string("frame ") + std::to_string(f) + ": " + c->err);
    // the frame's generation changed: a prepared batch captures again at its next run
    return NSGPU_OK;
}

static int64_t frame_nodes(const nsgpu_ctx *c) { return (int64_t)c->grid.nx * c->grid.ny * c->grid.nz; }

// a joint cast covers SES frames traced patch by patch (no skin cells, mesh
// surface, cell-grid cast or recorded patch ids)
static bool joint_cast_ok(const nsgpu_batch *b) {
    for (const nsgpu_ctx *c : b->frames)
        if (c->surface_is_mesh || c->cast_mode != NSGPU_CAST_PATCHES || c->record_pid || dev_patches(c).n_skin > 0)
            return false;
    return true;
}

// The frames' descriptors in item order, uploaded for a joint cast.
static int upload_frames(nsgpu_batch *b, const std::vector<int> &order, BatchCast *bc) {
    std::vector<FrameCast> fc(order.size());
    int64_t items[2] = {0, 0};
    for (size_t pos = 0; pos < order.size(); ++pos) {
        nsgpu_ctx *c = b->frames[order[pos]];
        FrameCast &F = fc[pos];
        F = frame_cast(c, dev_rays(c), c->cand.as<unsigned long long>());
        F.first[0] = items[0];
        F.first[1] = items[1];
        if (F.R.nslots > 0) {
            items[0] += F.P.n_plain;
            items[1] += (int64_t)(F.P.n - F.P.n_skin - F.P.n_plain) * F.R.nslots;
        }
    }
    if (items[0] >= (int64_t)1 << 31 || items[1] >= (int64_t)1 << 31)
        return bfail(b, NSGPU_E_ARG, "a joint cast holds at most 2^31 items per class");
    BHIP_TRY(b, b->fcast.ensure(fc.size() * sizeof(FrameCast)));
    BHIP_TRY(b, hipMemcpy(b->fcast.p, fc.data(), fc.size() * sizeof(FrameCast), hipMemcpyHostToDevice));
    bc->f = b->fcast.as<FrameCast>();
    bc->nf = (int)fc.size();
    bc->items[0] = items[0];
    bc->items[1] = items[1];
    bc->queue = b->queue.as<unsigned int>();
    return NSGPU_OK;
}

// Capture every frame's getSurf into b->graph: frames dealt round-robin onto
// `chains` concurrent chains.  With a joint cast the chains first clear each
// frame's rays, then one launch per patch class intersects every frame's
// patches, then the chains sort, fill and label each frame.
static int capture(nsgpu_batch *b, const nsgpu_batch_shape &shape) {
    drop_graph(b);
    const int n = (int)b->frames.size();
    const int chains = std::max(1, std::min((int)shape.chains, n));
    const bool joint = shape.cast != NSGPU_BATCH_CAST_FRAMES;
    if (joint && !joint_cast_ok(b))
        return bfail(b, NSGPU_E_ARG, "a joint cast needs SES frames traced patch by patch");
    std::vector<int> order(n);
    for (int f = 0; f < n; ++f) order[f] = f;
    std::stable_sort(order.begin(), order.end(),
                     [&](int x, int y) { return frame_nodes(b->frames[x]) > frame_nodes(b->frames[y]); });
    BatchCast bc{};
    if (joint) {
        // begin_surf decides the frames' slots, which the descriptors carry
        for (nsgpu_ctx *c : b->frames) {
            const int rc = begin_surf(c, &b->opts);
            if (rc) return bfail(b, rc, c->err);
        }
        const int rc = upload_frames(b, order, &bc);
        if (rc) return rc;
        bc.persistent = shape.cast == NSGPU_BATCH_CAST_PERSISTENT;
        bc.chunk = shape.chunk > 0 ? shape.chunk : 16;
    }
    const int roots = torus_roots(b->frames[0]);
    int grid[2] = {0, 0};
    if (bc.persistent)
        for (int cls = 0; cls < 2; ++cls)
            grid[cls] = shape.workgroups > 0 ? shape.workgroups
                                             : std::max(1, intersect_frames_occupancy(cls, roots)) * b->cus;
    unsigned long long *stamps = b->stamps.as<unsigned long long>();
    BHIP_TRY(b, hipStreamBeginCapture(b->s0, hipStreamCaptureModeRelaxed));
    g_no_alloc = true;
    int rc = NSGPU_OK;
    auto wait = [&](hipStream_t st, hipEvent_t e) {
        if (!rc && hipStreamWaitEvent(st, e, 0) != hipSuccess) rc = bfail(b, NSGPU_E_DEVICE, "graph wait");
    };
    auto record = [&](hipEvent_t e, hipStream_t st) {
        if (!rc && hipEventRecord(e, st) != hipSuccess) rc = bfail(b, NSGPU_E_DEVICE, "graph event");
    };
    // one pass over the frames in chain order: work(c) on each frame's stream
    auto chain_pass = [&](hipEvent_t start, std::vector<hipEvent_t> &done, auto work) {
        for (int pos = 0; pos < n && !rc; ++pos) {
            const int f = order[pos];
            nsgpu_ctx *c = b->frames[f];
            wait(c->stream, pos < chains ? start : done[order[pos - chains]]);
            if (rc) break;
            c->capturing = true;
            rc = work(c);
            c->capturing = false;
            if (rc) {
                rc = bfail(b, rc, std::string("frame ") + std::to_string(f) + ": " + c->err);
                break;
            }
            if (&done == &b->join) launch_stamp(c->stream, stamps + f);
            record(done[f], c->stream);
        }
        for (int f = 0; f < n; ++f) wait(b->s0, done[f]);
    };
    launch_stamp(b->s0, stamps + n);
    record(b->fork, b->s0);
    if (!joint) {
        chain_pass(b->fork, b->join, [&](nsgpu_ctx *c) {
            int r = begin_surf(c, &b->opts);
            if (!r) r = enqueue_surf(c, c->stream);
            return r;
        });
    } else {
        chain_pass(b->fork, b->head, [&](nsgpu_ctx *c) { return trace_head(c, c->stream); });
        if (!rc) {
            // plain patches on s0, tori on s1, concurrently
            if (bc.persistent) launch_fill32(b->s0, bc.queue, 0u, 2);
            record(b->cast_fork, b->s0);
            wait(b->s1, b->cast_fork);
            if (!rc) {
                launch_intersect_frames(b->s1, bc, 1, roots, grid[1]);
                launch_intersect_frames(b->s0, bc, 0, roots, grid[0]);
                if (hipGetLastError() != hipSuccess) rc = bfail(b, NSGPU_E_DEVICE, "joint cast launch");
            }
            record(b->cast_join, b->s1);
            wait(b->s0, b->cast_join);
            record(b->cast_done, b->s0);
        }
        chain_pass(b->cast_done, b->join, [&](nsgpu_ctx *c) {
            int r = trace_tail(c, c->stream);
            if (!r) r = enqueue_maps(c, c->stream);
            return r;
        });
    }
    g_no_alloc = false;
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(b->s0, &graph);
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) return bfail(b, NSGPU_E_DEVICE, std::string("capture: ") + hipGetErrorString(e));
    b->graph = graph;
    BHIP_TRY(b, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    b->chains = chains;
    b->shape = shape;
    b->shape.chains = chains;
    b->shape.workgroups = grid[1];
    b->shape.chunk = bc.persistent ? bc.chunk : 0;
    b->prepared = true;
    b->gen.resize(n);
    for (int f = 0; f < n; ++f) b->gen[f] = ctx_generation(b->frames[f]);
    return NSGPU_OK;
}

// frames that no longer match the captured graph
static bool graph_stale(nsgpu_batch *b, std::vector<int> *stale) {
    stale->clear();
    for (size_t f = 0; f < b->frames.size(); ++f) {
        nsgpu_ctx *c = b->frames[f];
        if (f >= b->gen.size() || ctx_generation(c) != b->gen[f] ||
            std::memcmp(&c->opts, &b->opts, sizeof(nsgpu_surf_opts)) != 0)
            stale->push_back((int)f);
    }
    return !stale->empty();
}

static int launch_graph(nsgpu_batch *b, double *ms) {
    const auto t0 = std::chrono::steady_clock::now();
    BHIP_TRY(b, hipGraphLaunch(b->exec, b->s0));
    BHIP_TRY(b, hipStreamSynchronize(b->s0));
    if (ms) *ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return NSGPU_OK;
}

// Eager pass: sizes every buffer and each frame's ray capacity.
static int batch_eager(nsgpu_batch *b, const nsgpu_surf_opts *opts) {
    if (!opts) return NSGPU_E_ARG;
    if (opts->accurate_triangulation)
        return bfail(b, NSGPU_E_ARG, "batches compute maps and cavities only (accurate_triangulation = 0)");
    BHIP_TRY(b, hipSetDevice(b->device));
    b->opts = *opts;
    drop_graph(b);
    for (size_t f = 0; f < b->frames.size(); ++f) {
        const int rc = nsgpu_get_surf(b->frames[f], opts, nullptr);
        if (rc) return bfail(b, rc, std::string("frame ") + std::to_string(f) + ": " + b->frames[f]->err);
    }
    return NSGPU_OK;
}

// Autotuner: captures each candidate shape, keeps the one with the shortest
// of two launches.  Candidates: per-frame casts on 1, 2, 4, ... 64 chains;
// with at least 8 SES frames also the joint cast (a workgroup per item) and
// the persistent one (the resident limit of workgroups taking 16 items at a
// time), each on 16, 32 and 64 chains.
static int autotune(nsgpu_batch *b) {
    const int n = (int)b->frames.size();
    std::vector<nsgpu_batch_shape> cands;
    for (int ch = 1;; ch *= 2) {
        cands.push_back(nsgpu_batch_shape{NSGPU_BATCH_CAST_FRAMES, std::min(ch, n), 0, 0});
        if (ch >= n || ch >= 64) break;
    }
    if (n >= 8 && joint_cast_ok(b)) {
        const int roots = torus_roots(b->frames[0]);
        const int full = std::max(1, intersect_frames_occupancy(1, roots)) * b->cus;
        for (int ch : {16, 32, 64}) {
            if (ch > 2 * n) break;
            cands.push_back(nsgpu_batch_shape{NSGPU_BATCH_CAST_JOINT, std::min(ch, n), 0, 0});
            cands.push_back(nsgpu_batch_shape{NSGPU_BATCH_CAST_PERSISTENT, std::min(ch, n), full, 16});
        }
    }
    nsgpu_batch_shape best = cands[0];
    double best_ms = 1e30;
    for (const nsgpu_batch_shape &sh : cands) {
        int rc = capture(b, sh);
        if (rc) return rc;
        double t = 1e30;
        for (int rep = 0; rep < 2; ++rep) {
            double ms = 0;
            rc = launch_graph(b, &ms);
            if (rc) return rc;
            t = std::min(t, ms);
        }
        if (t < best_ms) {
            best_ms = t;
            best = sh;
        }
    }
    return capture(b, best);
}

int nsgpu_batch_prepare(nsgpu_batch *b, const nsgpu_surf_opts *opts, int32_t chains) {
    if (!b) return NSGPU_E_ARG;
    int rc = batch_eager(b, opts);
    if (rc) return rc;
    if (chains > 0) return capture(b, nsgpu_batch_shape{NSGPU_BATCH_CAST_FRAMES, chains, 0, 0});
    return autotune(b);
}

int nsgpu_batch_prepare_shape(nsgpu_batch *b, const nsgpu_surf_opts *opts, const nsgpu_batch_shape *shape) {
    if (!b) return NSGPU_E_ARG;
    if (shape && (shape->cast < NSGPU_BATCH_CAST_FRAMES || shape->cast > NSGPU_BATCH_CAST_PERSISTENT ||
                  shape->chains <= 0 || shape->workgroups < 0 || shape->chunk < 0))
        return bfail(b, NSGPU_E_ARG, "unknown batch shape");
    int rc = batch_eager(b, opts);
    if (rc) return rc;
    return shape ? capture(b, *shape) : autotune(b);
}

int nsgpu_batch_get_shape(const nsgpu_batch *b, nsgpu_batch_shape *shape) {
    if (!b || !shape) return NSGPU_E_ARG;
    if (!b->prepared) return NSGPU_E_STATE;
    *shape = b->shape;
    return NSGPU_OK;
}

int32_t nsgpu_batch_chains(const nsgpu_batch *b) { return b ? b->chains : 0; }

int nsgpu_batch_run(nsgpu_batch *b, double *latency_ms, int32_t *redone) {
    if (!b) return NSGPU_E_ARG;
    if (!b->prepared) return bfail(b, NSGPU_E_STATE, "nsgpu_batch_prepare first");
    BHIP_TRY(b, hipSetDevice(b->device));
    const int n = (int)b->frames.size();
    int rc;
    std::vector<int> stale;
    if (graph_stale(b, &stale)) {
        // size the changed frames' buffers eagerly, then capture again
        for (int f : stale) {
            rc = nsgpu_get_surf(b->frames[f], &b->opts, nullptr);
            if (rc) return bfail(b, rc, std::string("frame ") + std::to_string(f) + ": " + b->frames[f]->err);
        }
        rc = capture(b, b->shape);
        if (rc) return rc;
    }
    rc = launch_graph(b, nullptr);
    if (rc) return rc;
    std::vector<unsigned long long> st(n + 1);
    BHIP_TRY(b, hipMemcpy(st.data(), b->stamps.p, (size_t)(n + 1) * 8, hipMemcpyDeviceToHost));
    if (latency_ms)
        for (int f = 0; f < n; ++f) latency_ms[f] = (double)(st[f] - st[n]) / b->clock_khz;
    int again = 0;
    for (int f = 0; f < n; ++f) {
        nsgpu_ctx *c = b->frames[f];
        if (regrow(c)) {
            // the frame's rays outgrew their lists: redo it eagerly
            ++again;
            rc = nsgpu_get_surf(c, &b->opts, nullptr);
        } else {
            rc = finish_surf(c);
            c->have_surf = true;
            if (c->opts.fill_cavities && c->opts.build_status && c->hres->cav_info()[0] == 0) ++again;
        }
        if (rc) return bfail(b, rc, std::string("frame ") + std::to_string(f) + ": " + c->err);
    }
    // a frame redone with grown ray lists (or numbered on the host, which may
    // grow its key buffers) has a new generation: the next run captures again
    if (redone) *redone = again;
    return NSGPU_OK;
}

int nsgpu_batch_run_loop(nsgpu_batch *b, double *latency_ms) {
    if (!b) return NSGPU_E_ARG;
    for (size_t f = 0; f < b->frames.size(); ++f) {
        const auto t0 = std::chrono::steady_clock::now();
        const int rc = nsgpu_get_surf(b->frames[f], &b->opts, nullptr);
        if (rc) return bfail(b, rc, std::string("frame ") + std::to_string(f) + ": " + b->frames[f]->err);
        if (latency_ms)
            latency_ms[f] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    return NSGPU_OK;
}

int nsgpu_batch_result(nsgpu_batch *b, int32_t f, nsgpu_surf_result *res) {
    if (!b || !res || f < 0 || f >= (int)b->frames.size()) return NSGPU_E_ARG;
    if (!b->frames[f]->have_surf) return bfail(b, NSGPU_E_STATE, "frame not computed");
    *res = b->frames[f]->res;
    return NSGPU_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Boundary grid points (Surface.cpp:1476-1588) projected on the SES
// (Surface::projector :4657-4680, ConnollySurface::getProjection
// ConnollySurface.cpp:5822-6106), from the epsmap of the last getSurf.

// The tables of the projection that depend only on the grid, the patch set
// and the auxiliary grid size: the auxiliary grid (ConnollySurface.cpp:
// 3683-3742) with each cell's grid-node ranges and patch lists, the facets by
// atom pair and the sample angles; built once per (inputs, aux_grid_dim).
static int bgp_tables(nsgpu_ctx *c, int32_t aux_grid_dim) {
    if (c->bg_valid && c->bg_setgen == c->setgen && c->bg_aux == aux_grid_dim) return NSGPU_OK;
    c->bg_valid = false;
    const nsgpu_grid &G = c->grid;
    const DevPatches P = dev_patches(c);
    int64_t igrid = G.nx, gridMul = 1;
    while (igrid > aux_grid_dim) {
        gridMul += 2;
        int64_t digrid = G.nx;
        while (true) {
            const int64_t fixedPoint = (digrid + (gridMul - 1)) / gridMul;
            igrid = fixedPoint * gridMul;
            if (igrid % 2 == 0) {
                digrid = igrid + 1;
            } else {
                igrid = fixedPoint;
                break;
            }
        }
    }
    const int na = (int)igrid;
    if ((int64_t)na * na * na > INT32_MAX) return fail(c, NSGPU_E_ARG, "auxiliary grid too large");
    BgpArgs A{};
    A.na = na;
    A.ascale = G.scale / ((double)gridMul);
    const double aside = ((double)gridMul) / G.scale;
    for (int d = 0; d < 3; ++d) A.amin[d] = G.baricenter[d] - (igrid - 1) * 0.5 * aside;
    // grid nodes of each auxiliary cell, per axis: the node's cell is
    // rintp((x - amin) * scale) (ConnollySurface.cpp:5835-5837)
    std::vector<int> range((size_t)6 * na);
    const double mins[3] = {G.xmin, G.ymin, G.zmin};
    const int ns[3] = {G.nx, G.ny, G.nz};
    const int NN = std::max(G.nx, std::max(G.ny, G.nz));
    std::vector<int> node_aux((size_t)3 * NN, 0);
    for (int d = 0; d < 3; ++d) {
        int *rg = range.data() + (size_t)2 * d * na;
        for (int a = 0; a < na; ++a) {
            rg[2 * a] = INT32_MAX;
            rg[2 * a + 1] = -1;
        }
        for (int m = 0; m < ns[d]; ++m) {
            const double coord = mins[d] + m * G.side;
            int64_t a = (int64_t)std::floor((coord - A.amin[d]) * A.ascale + 0.5);
            a = std::min<int64_t>(std::max<int64_t>(a, 0), na - 1);
            node_aux[(size_t)d * NN + m] = (int)a;
            rg[2 * a] = std::min(rg[2 * a], m);
            rg[2 * a + 1] = std::max(rg[2 * a + 1], m);
        }
        for (int a = 0; a < na; ++a) {
            if (rg[2 * a + 1] >= 0 && rg[2 * a + 1] - rg[2 * a] + 1 > 64)
                return fail(c, NSGPU_E_ARG, "auxiliary grid too coarse: a cell spans more than 64 grid nodes");
            if (rg[2 * a + 1] < 0) rg[2 * a] = 0;
        }
    }
    // facets by atom pair, for the sampled projections
    const int n = c->npatch;
    std::vector<int32_t> kind(n), atom(3 * (size_t)n);
    if (n) {
        HIP_TRY(c, hipMemcpyAsync(kind.data(), c->pkind.p, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipMemcpyAsync(atom.data(), c->patom.p, (size_t)n * 12, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
    }
    std::vector<std::pair<long long, int>> fk;
    if (c->patch_max_atom >= 0)
        for (int i = 0; i < n; ++i) {
            if (kind[i] != NSGPU_PATCH_FACET_REG && kind[i] != NSGPU_PATCH_FACET_SING) continue;
            const int *a = &atom[3 * (size_t)i];
            const int pr[3][2] = {{a[0], a[1]}, {a[1], a[2]}, {a[0], a[2]}};
            for (auto &q : pr)
                fk.emplace_back(((long long)std::min(q[0], q[1]) << 32) | (unsigned)std::max(q[0], q[1]), i);
        }
    std::sort(fk.begin(), fk.end());
    std::vector<long long> fkey(fk.size());
    std::vector<int> fval(fk.size());
    for (size_t q = 0; q < fk.size(); ++q) {
        fkey[q] = fk[q].first;
        fval[q] = fk[q].second;
    }
    // each torus patch's facets: the run of its atom pair's key
    std::vector<int> efac(2 * (size_t)std::max(n, 1), 0);
    if (c->patch_max_atom >= 0)
        for (int i = 0; i < n; ++i) {
            if (kind[i] != NSGPU_PATCH_TORUS_REG && kind[i] != NSGPU_PATCH_TORUS_SING) continue;
            const int a = atom[3 * (size_t)i], b = atom[3 * (size_t)i + 1];
            const long long key = ((long long)std::min(a, b) << 32) | (unsigned)std::max(a, b);
            const auto r = std::equal_range(fkey.begin(), fkey.end(), key);
            efac[2 * (size_t)i] = (int)(r.first - fkey.begin());
            efac[2 * (size_t)i + 1] = (int)(r.second - r.first);
        }
    // the sample angles of getCoi (ConnollySurface.cpp:6175-6195): the angle
    // accumulated step by step, cos and sin of each
    std::vector<double> coi(2 * kCoiSamples);
    const double step = (2.0 * 3.14159265358979323846264338327950288) / kCoiSamples;
    double tcoi = 0.;
    for (int k = 0; k < kCoiSamples; ++k) {
        coi[2 * k] = std::cos(tcoi);
        coi[2 * k + 1] = std::sin(tcoi);
        tcoi += step;
    }
    const int64_t ncell = (int64_t)na * na * na;
    HIP_TRY(c, c->bgrange.ensure(range.size() * 4));
    HIP_TRY(c, c->bgccount.ensure((size_t)ncell * 4));
    HIP_TRY(c, c->bgcoff.ensure((size_t)ncell * 4));
    HIP_TRY(c, c->bgccursor.ensure((size_t)ncell * 4));
    HIP_TRY(c, c->bgfkey.ensure(efac.size() * 4));
    HIP_TRY(c, c->bgfval.ensure(std::max<size_t>(fval.size(), 1) * 4));
    HIP_TRY(c, c->bgcoi.ensure(coi.size() * 8));
    HIP_TRY(c, c->bgnaux.ensure(node_aux.size() * 4));
    HIP_TRY(c, c->bgcflag.ensure((size_t)ncell * 4));
    HIP_TRY(c, c->bgcoffa.ensure((size_t)ncell * 4));
    HIP_TRY(c, c->bgactive.ensure((size_t)ncell * 4));
    HIP_TRY(c, c->scan_scratch.ensure((size_t)scan_scratch_ints(ncell) * 4));
    HIP_TRY(c, hipMemcpyAsync(c->bgnaux.p, node_aux.data(), node_aux.size() * 4, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->bgrange.p, range.data(), range.size() * 4, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->bgcoi.p, coi.data(), coi.size() * 8, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->bgfkey.p, efac.data(), efac.size() * 4, hipMemcpyHostToDevice, c->stream));
    if (!fval.empty())
        HIP_TRY(c, hipMemcpyAsync(c->bgfval.p, fval.data(), fval.size() * 4, hipMemcpyHostToDevice, c->stream));
    // auxiliary cell lists: count, scan, fill
    A.range = c->bgrange.as<int>();
    A.ccount = c->bgccount.as<int>();
    A.coff = c->bgcoff.as<int>();
    A.ccursor = c->bgccursor.as<int>();
    launch_bgp_bin(c->stream, P, A, false);
    launch_scan(c->stream, A.ccount, ncell, A.coff, c->scan_scratch.as<int>());
    HIP_TRY(c, hipGetLastError());
    int64_t entries = 0;
    int rc = read_total(c, A.ccount, A.coff, ncell, &entries);
    if (rc) return rc;
    if (entries > INT32_MAX) return fail(c, NSGPU_E_NOMEM, "auxiliary cell lists exceed 2^31 entries");
    HIP_TRY(c, c->bgclist.ensure((size_t)std::max<int64_t>(entries, 1) * 4));
    A.clist = c->bgclist.as<int>();
    launch_bgp_bin(c->stream, P, A, true);
    HIP_TRY(c, hipGetLastError());
    A.efac = c->bgfkey.as<int>();
    A.fval = c->bgfval.as<int>();
    A.coi = c->bgcoi.as<double>();
    A.node_aux = c->bgnaux.as<int>();
    A.cflag = c->bgcflag.as<int>();
    A.coff_active = c->bgcoffa.as<int>();
    A.active = c->bgactive.as<int>();
    // the host vectors above are read by the copies before they go
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->bgA = A;
    c->bg_setgen = c->setgen;
    c->bg_aux = aux_grid_dim;
    c->bg_valid = true;
    return NSGPU_OK;
}

extern "C" {

int nsgpu_project_bgp(nsgpu_ctx *c, int32_t aux_grid_dim, int64_t *counts) {
    if (!c || aux_grid_dim < 1) return NSGPU_E_ARG;
    if (!c->have_surf || !c->opts.build_epsmap)
        return fail(c, NSGPU_E_STATE, "boundary grid points need a getSurf with the epsilon map");
    if (c->surface_is_mesh || c->npatch_skin > 0)
        return fail(c, NSGPU_E_STATE, "boundary grid point projection is built for SES patches only");
    HIP_TRY(c, hipSetDevice(c->device));
    c->have_bgp = false;
    int rc = bgp_tables(c, aux_grid_dim);
    if (rc) return rc;
    const nsgpu_grid &G = c->grid;
    const DevGrid g = dev_grid(c);
    const DevPatches P = dev_patches(c);
    const int W = (G.nx + 63) / 64;
    const int64_t nwords = (int64_t)G.ny * G.nz * W;
    HIP_TRY(c, c->bgbits.ensure((size_t)nwords * 8));
    HIP_TRY(c, c->bgwcount.ensure((size_t)nwords * 4));
    HIP_TRY(c, c->bgwoff.ensure((size_t)nwords * 4));
    HIP_TRY(c, c->bgcnt.ensure(4 * 4));
    const int64_t ncell = (int64_t)c->bgA.na * c->bgA.na * c->bgA.na;
    HIP_TRY(c, c->scan_scratch.ensure((size_t)scan_scratch_ints(std::max(nwords, ncell)) * 4));
    int *cnt = c->bgcnt.as<int>();
    BgpArgs A = c->bgA;
    A.eps = c->eps.as<int>();
    A.W = W;
    A.bits = c->bgbits.as<unsigned long long>();
    A.woff = c->bgwoff.as<int>();
    A.counters = cnt;
    launch_fill32(c->stream, cnt, 0u, 4);
    launch_fill32(c->stream, A.cflag, 0u, ncell);
    // detection (offsets in the reference's z, y, x order) and the auxiliary
    // cells holding boundary points
    launch_bgp_flags(c->stream, g, A, c->bgbits.as<unsigned long long>(), c->bgwcount.as<int>(), cnt + 2);
    launch_scan(c->stream, c->bgwcount.as<int>(), nwords, c->bgwoff.as<int>(), c->scan_scratch.as<int>());
    launch_scan(c->stream, A.cflag, ncell, A.coff_active, c->scan_scratch.as<int>());
    HIP_TRY(c, hipGetLastError());
    int64_t tot[2] = {0, 0};
    {
        const int *cs[2] = {c->bgwcount.as<int>(), A.cflag};
        const int *os[2] = {c->bgwoff.as<int>(), A.coff_active};
        const int64_t nn[2] = {nwords, ncell};
        rc = read_totals(c, 2, cs, os, nn, tot);
        if (rc) return rc;
    }
    const int64_t nb = tot[0], n_active = tot[1];
    if (nb > INT32_MAX / 3) return fail(c, NSGPU_E_NOMEM, "too many boundary grid points");
    HIP_TRY(c, c->bgib.ensure((size_t)std::max<int64_t>(nb, 1) * 12));
    HIP_TRY(c, c->bgpos.ensure((size_t)std::max<int64_t>(nb, 1) * 24));
    HIP_TRY(c, c->bgnor.ensure((size_t)std::max<int64_t>(nb, 1) * 24));
    HIP_TRY(c, c->bgfq.ensure((size_t)std::max<int64_t>(nb, 1) * 8));
    A.ibgp = c->bgib.as<int>();
    A.pos = c->bgpos.as<double>();
    A.nor = c->bgnor.as<double>();
    A.fq = c->bgfq.as<int>();
    if (nb > 0) {
        launch_bgp_cells(c->stream, A);
        launch_bgp_project(c->stream, g, P, A, c->probe, n_active);
    }
    HIP_TRY(c, hipGetLastError());
    int hc[4] = {0, 0, 0, 0};
    HIP_TRY(c, hipMemcpyAsync(hc, cnt, sizeof hc, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->nbgp = nb;
    c->have_bgp = true;
    if (counts) {
        counts[0] = nb;
        counts[1] = hc[2];
        counts[2] = hc[0];
        counts[3] = hc[1];
    }
    return NSGPU_OK;
}

int nsgpu_read_bgp(nsgpu_ctx *c, int32_t *ibgp, double *scspos, double *scsnor) {
    if (!c) return NSGPU_E_ARG;
    if (!c->have_bgp || !c->have_surf) return fail(c, NSGPU_E_STATE, "no boundary grid points computed");
    HIP_TRY(c, hipSetDevice(c->device));
    const size_t n = (size_t)c->nbgp;
    if (n && ibgp) HIP_TRY(c, hipMemcpyAsync(ibgp, c->bgib.p, n * 12, hipMemcpyDeviceToHost, c->stream));
    if (n && scspos) HIP_TRY(c, hipMemcpyAsync(scspos, c->bgpos.p, n * 24, hipMemcpyDeviceToHost, c->stream));
    if (n && scsnor) HIP_TRY(c, hipMemcpyAsync(scsnor, c->bgnor.p, n * 24, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return NSGPU_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Pocket detection (the grid part of pocketMode, main.cpp:1411-1660) on the
// status maps of a fat-probe and a small-probe surface.

extern "C" {

int nsgpu_pockets(nsgpu_ctx *fat, nsgpu_ctx *small, const nsgpu_pocket_opts *o, nsgpu_pocket_result *res) {
    if (!fat || !small || !o || fat == small) return NSGPU_E_ARG;
    nsgpu_ctx *c = fat;
    if (!fat->have_surf || !small->have_surf || !fat->opts.build_status || !small->opts.build_status)
        return fail(c, NSGPU_E_STATE, "pockets need a getSurf with the status map on both surfaces");
    if (fat->device != small->device) return fail(c, NSGPU_E_ARG, "both surfaces must be on one device");
    const nsgpu_grid &G = fat->grid;
    if (G.nx != small->grid.nx || G.ny != small->grid.ny || G.nz != small->grid.nz)
        return fail(c, NSGPU_E_ARG, "the two surfaces are on different grids");
    if (!(o->filter_probe > 0) || !(o->num_mol >= 0)) return fail(c, NSGPU_E_ARG, "bad pocket options");
    HIP_TRY(c, hipSetDevice(c->device));
    c->have_pockets = false;
    const DevGrid g = dev_grid(c);
    const int64_t nvox = (int64_t)G.nx * G.ny * G.nz;
    const int W = (G.nx + 63) / 64;
    const int64_t nwords = (int64_t)G.ny * G.nz * W;
    // the probe's digital ball (filterCavities' box and dist2 test,
    // Surface.cpp:2775-2830): per (dy, dz) the largest |dx| of a node within
    // the probe radius
    std::vector<int> ball;
    {
        const double r = o->filter_probe, side = G.side, r2 = r * r;
        const int m = (int)std::floor(r / side) + 1;
        for (int dk = -m; dk <= m; ++dk)
            for (int dj = -m; dj <= m; ++dj) {
                int hw = -1;
                for (int di = 0; di <= m; ++di) {
                    const double a = di * side, b = dj * side, cc = dk * side;
                    if (a * a + b * b + cc * cc <= r2) hw = di;
                }
                if (hw < 0) continue;
                if (hw > 63) return fail(c, NSGPU_E_ARG, "the filter probe spans more than 63 grid nodes");
                ball.insert(ball.end(), {dj, dk, hw});
            }
    }
    PocketArgs A{};
    A.W = W;
    A.nwords = nwords;
    A.cav_and_pockets = o->cav_and_pockets != 0;
    A.nball = (int)(ball.size() / 3);
    HIP_TRY(c, c->pkball.ensure(ball.size() * 4));
    HIP_TRY(c, hipMemcpyAsync(c->pkball.p, ball.data(), ball.size() * 4, hipMemcpyHostToDevice, c->stream));
    for (DBuf *b : {&c->pkD, &c->pkS1, &c->pkK, &c->pkU, &c->pkS2, &c->pkK2, &c->pkroots})
        HIP_TRY(c, b->ensure((size_t)nwords * 8));
    for (DBuf *b : {&c->pkrcount, &c->pkroff, &c->pkscount, &c->pksoff}) HIP_TRY(c, b->ensure((size_t)nwords * 4));
    HIP_TRY(c, c->pkparent.ensure((size_t)nvox * 4));
    HIP_TRY(c, c->scan_scratch.ensure((size_t)scan_scratch_ints(nwords) * 4));
    A.ball = c->pkball.as<int>();
    A.st1 = c->status.as<int>();
    A.st2 = small->status.as<int>();
    A.D = c->pkD.as<unsigned long long>();
    A.U = c->pkU.as<unsigned long long>();
    A.S2 = c->pkS2.as<unsigned long long>();
    A.K2 = c->pkK2.as<unsigned long long>();
    A.roots = c->pkroots.as<unsigned long long>();
    A.parent = c->pkparent.as<int>();
    A.rcount = c->pkrcount.as<int>();
    A.roff = c->pkroff.as<int>();
    A.scount = c->pkscount.as<int>();
    A.soff = c->pksoff.as<int>();
    unsigned long long *S1 = c->pkS1.as<unsigned long long>(), *K = c->pkK.as<unsigned long long>();
    // the small surface's map is complete: its getSurf synchronised its stream
    // difference, the shape filter on it (operator-=, Surface.cpp:15556-15616),
    // the components of what it keeps
    launch_pk_diff(c->stream, g, A);
    launch_pk_morph(c->stream, g, A, A.D, nullptr, S1, false);
    launch_pk_morph(c->stream, g, A, S1, A.D, K, true);
    launch_pk_components(c->stream, g, A, K);
    launch_scan(c->stream, A.rcount, nwords, A.roff, c->scan_scratch.as<int>());
    HIP_TRY(c, hipGetLastError());
    int64_t ncomp = 0;
    int rc = read_total(c, A.rcount, A.roff, nwords, &ncomp);
    if (rc) return rc;
    const size_t nc = (size_t)std::max<int64_t>(ncomp, 1);
    for (DBuf *b : {&c->pkccount, &c->pkcroot, &c->pkcfilled, &c->pkcid, &c->pkckept, &c->pkcpocket})
        HIP_TRY(c, b->ensure(nc * 4));
    HIP_TRY(c, c->pkcbkey.ensure(nc * 8));
    A.ccount = c->pkccount.as<int>();
    A.croot = c->pkcroot.as<int>();
    A.cfilled = c->pkcfilled.as<int>();
    A.cid = c->pkcid.as<int>();
    A.ckept = c->pkckept.as<int>();
    A.cpocket = c->pkcpocket.as<int>();
    A.cbkey = c->pkcbkey.as<long long>();
    std::vector<int> cnt(nc), root(nc);
    std::vector<long long> bkey(nc);
    if (ncomp > 0) {
        launch_fill32(c->stream, A.ccount, 0u, ncomp);
        launch_fill32(c->stream, A.cbkey, 0x7fffffffu, 2 * ncomp);
        launch_pk_stats(c->stream, g, A);
        HIP_TRY(c, hipGetLastError());
        HIP_TRY(c, hipMemcpyAsync(cnt.data(), A.ccount, (size_t)ncomp * 4, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipMemcpyAsync(root.data(), A.croot, (size_t)ncomp * 4, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipMemcpyAsync(bkey.data(), A.cbkey, (size_t)ncomp * 8, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
    }
    // ids in the alternating scan order of getCavities(STATUS_POINT_OUT)
    // (Surface.cpp:1920-1981): even passes take the component met first
    // scanning k upwards (smallest root), odd passes scanning k downwards
    std::vector<int> by_f(ncomp), by_b(ncomp), id((size_t)ncomp), comp_of_id((size_t)ncomp);
    for (int q = 0; q < ncomp; ++q) by_f[q] = by_b[q] = q;
    std::sort(by_f.begin(), by_f.end(), [&](int a, int b) { return root[a] < root[b]; });
    std::sort(by_b.begin(), by_b.end(), [&](int a, int b) { return bkey[a] < bkey[b]; });
    {
        std::vector<char> done((size_t)ncomp, 0);
        size_t pf = 0, pb = 0;
        for (int t = 0; t < ncomp; ++t) {
            int q;
            if (t % 2 == 0) {
                while (done[by_f[pf]]) ++pf;
                q = by_f[pf];
            } else {
                while (done[by_b[pb]]) ++pb;
                q = by_b[pb];
            }
            done[q] = 1;
            id[q] = NSGPU_STATUS_FIRST_CAV + t;
            comp_of_id[t] = q;
        }
    }
    // volume threshold (fillCavities(11.4 * Num_Wat_Pocket), :2355-2487)
    const double cubeVol = G.side * G.side * G.side;
    const double vmax = 11.4 * o->num_mol;
    std::vector<int> filled((size_t)ncomp);
    c->pk_vol.assign((size_t)ncomp, 0.0);
    for (int t = 0; t < ncomp; ++t) {
        const int q = comp_of_id[t];
        c->pk_vol[t] = cnt[q] * cubeVol;
        filled[q] = c->pk_vol[t] <= vmax;
    }
    if (ncomp > 0) {
        HIP_TRY(c, hipMemcpyAsync(A.cid, id.data(), (size_t)ncomp * 4, hipMemcpyHostToDevice, c->stream));
        HIP_TRY(c, hipMemcpyAsync(A.cfilled, filled.data(), (size_t)ncomp * 4, hipMemcpyHostToDevice, c->stream));
        launch_fill32(c->stream, A.ckept, 0u, ncomp);
        launch_fill32(c->stream, A.cpocket, 0u, ncomp);
    }
    // the filter on the kept components (main.cpp:1616), the final labels
    launch_pk_union_kept(c->stream, g, A);
    launch_pk_morph(c->stream, g, A, A.U, nullptr, A.S2, false);
    launch_pk_morph(c->stream, g, A, A.S2, A.U, A.K2, true);
    launch_pk_label(c->stream, g, A);
    launch_scan(c->stream, A.scount, nwords, A.soff, c->scan_scratch.as<int>());
    HIP_TRY(c, hipGetLastError());
    int64_t nsup = 0;
    rc = read_total(c, A.scount, A.soff, nwords, &nsup);
    if (rc) return rc;
    HIP_TRY(c, c->pksupn.ensure((size_t)std::max<int64_t>(nsup, 1) * 8));
    HIP_TRY(c, c->pksupc.ensure((size_t)std::max<int64_t>(nsup, 1) * 4));
    A.sup_node = c->pksupn.as<long long>();
    A.sup_comp = c->pksupc.as<int>();
    launch_pk_support_list(c->stream, g, A);
    HIP_TRY(c, hipGetLastError());
    std::vector<int> kept((size_t)ncomp), pocket((size_t)ncomp), supc((size_t)nsup);
    std::vector<long long> supn((size_t)nsup);
    if (ncomp > 0) {
        HIP_TRY(c, hipMemcpyAsync(kept.data(), A.ckept, (size_t)ncomp * 4, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipMemcpyAsync(pocket.data(), A.cpocket, (size_t)ncomp * 4, hipMemcpyDeviceToHost, c->stream));
    }
    if (nsup > 0) {
        HIP_TRY(c, hipMemcpyAsync(supn.data(), A.sup_node, (size_t)nsup * 8, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipMemcpyAsync(supc.data(), A.sup_comp, (size_t)nsup * 4, hipMemcpyDeviceToHost, c->stream));
    }
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    // a component the filter removed entirely counts as filled
    // (filterCavities, Surface.cpp:2996-3006); results in id order, support
    // points of each kept component in flat order
    c->pk_filled.assign((size_t)ncomp, 0);
    c->pk_pocket.assign((size_t)ncomp, 0);
    std::vector<int> kept_rank((size_t)ncomp, -1);
    int nk = 0;
    for (int t = 0; t < ncomp; ++t) {
        const int q = comp_of_id[t];
        c->pk_filled[t] = filled[q] || kept[q] == 0;
        c->pk_pocket[t] = pocket[q] != 0;
        if (!c->pk_filled[t]) kept_rank[q] = nk++;
    }
    c->pk_off.assign((size_t)nk + 1, 0);
    for (int64_t q = 0; q < nsup; ++q)
        if (kept_rank[supc[q]] >= 0) c->pk_off[kept_rank[supc[q]] + 1]++;
    for (int k = 0; k < nk; ++k) c->pk_off[k + 1] += c->pk_off[k];
    c->pk_sup.assign((size_t)c->pk_off[nk], 0);
    {
        std::vector<int64_t> cur(c->pk_off.begin(), c->pk_off.end() - 1);
        for (int64_t q = 0; q < nsup; ++q)
            if (kept_rank[supc[q]] >= 0) c->pk_sup[cur[kept_rank[supc[q]]]++] = supn[q];
    }
    c->pkA = A;
    c->have_pockets = true;
    if (res) {
        res->num_components = (int32_t)ncomp;
        res->num_kept = nk;
        res->support_points = c->pk_off[nk];
    }
    return NSGPU_OK;
}

int nsgpu_read_pocket_labels(nsgpu_ctx *c, int32_t *labels) {
    if (!c || !labels) return NSGPU_E_ARG;
    if (!c->have_pockets) return fail(c, NSGPU_E_STATE, "no pockets computed");
    HIP_TRY(c, hipSetDevice(c->device));
    const int64_t nvox = (int64_t)c->grid.nx * c->grid.ny * c->grid.nz;
    HIP_TRY(c, c->pklab.ensure((size_t)nvox * 4));
    launch_pk_comp_labels(c->stream, dev_grid(c), c->pkA, c->pklab.as<int>());
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, hipMemcpyAsync(labels, c->pklab.p, (size_t)nvox * 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return NSGPU_OK;
}

int nsgpu_read_pockets(nsgpu_ctx *c, double *volumes, uint8_t *filled, uint8_t *is_pocket, int64_t *offsets,
                       int64_t *support) {
    if (!c) return NSGPU_E_ARG;
    if (!c->have_pockets) return fail(c, NSGPU_E_STATE, "no pockets computed");
    const size_t n = c->pk_vol.size();
    if (volumes && n) std::memcpy(volumes, c->pk_vol.data(), n * sizeof(double));
    if (filled && n) std::memcpy(filled, c->pk_filled.data(), n);
    if (is_pocket && n) std::memcpy(is_pocket, c->pk_pocket.data(), n);
    if (offsets) std::memcpy(offsets, c->pk_off.data(), c->pk_off.size() * sizeof(int64_t));
    if (support && !c->pk_sup.empty()) std::memcpy(support, c->pk_sup.data(), c->pk_sup.size() * sizeof(int64_t));
    return NSGPU_OK;
}

}  // extern "C"
