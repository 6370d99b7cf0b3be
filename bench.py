"""Benchmark of the grid ray-casting SES builder on one MI355X.

Workload: 5,000-atom seeded chain, probe 1.4 A, grid 4.0 points/A (perfil 80),
epsmap + idebmap + status, 3 grid-ray + 3 edge-ray panels, cavity detection
and filling, analytic marching cubes.  A step is one Surface::getSurf plus
triangulateSurface over the device-resident patch set; host stages (grid,
jitter, patch construction) run once before timing.

Prints one JSON line: rays/s (all six panels) as the value, with voxels/s,
stage times, the roofline of the dominant kernel and the CPU baseline (the
oracle on the same workload, on all available host cores with the
reference's thread decomposition and on one core) beside it.  Three extensions
of the plain path are measured alongside:
* "mixed_precision": the same step with FP32-seeded, FP64-refined torus roots;
* "batched": 256 frames of 500-8000 atoms (log-uniform, sigma 0.3 A
  perturbations), 2 points/A, maps + cavities, one captured graph launch vs
  the per-frame launch loop;
* "differentiable": 2,000 atoms at 3 points/A, three-panel volume and area
  estimate with their gradients for all atoms (trace + adjoint kernel);
* "blobby": the blobby surface of 20,000 atoms at 2 points/A, built and ray
  cast as a mesh on the device, with the oracle's time for the same work;
* "skin": the skin surface (s = 0.45) of the same atoms, its mixed cells
  built on the host and cast on the device, with the oracle's time;
* "bgp": the boundary grid points of the benchmark surface projected on the
  SES (Project_boundary_grid_points), with the oracle's time;
* "pockets": pocket detection (the grid part of pocket mode) on the same
  atoms at 2 points/A, with the oracle's time.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

import nanoshaper_amd as ns  # noqa: E402

# MI355X published peaks
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0

# FP64 operations per path of the ray/patch candidate set, counted statically
# in the intersection kernel (add/sub/mul/div/sqrt/cbrt/acos/cos = 1 op):
# bounding-box candidate: ray origin (3) + disc cull (6); disc pass: ray/sphere
# up to the discriminant (23); sphere hit: roots (5); torus: frame change and
# quartic coefficients (55) + resolvent cubic and Ferrari roots (80) + root
# checks (25); tested root: hit point (2); feasibility item: one plane or
# clipping-sphere test (6).
OPS = {"candidates": 9, "disc_pass": 23, "sphere_hits": 5, "torus_quartics": 160, "roots_tested": 2,
       "feas_items": 6}

# the two launches of the intersection stage (spheres and facets; tori)
INTERSECT_KERNELS = ("nsg::k_intersect_plain<false>", "nsg::k_intersect_torus<false, 0>")

CONFIGS = {
    # name: atoms, confinement radius, grid scale, perfil
    "c2": (5000, 28.0, 4.0, 80.0),
    "c1": (330, 12.0, 2.0, 90.0),
    "c5": (2000, 21.0, 3.0, 80.0),
}


def build_inputs(cfg):
    n, radius, scale, perfil = CONFIGS[cfg]
    t0 = time.perf_counter()
    xyzr, r_used = ns.generate_chain(n, radius, seed=0)
    grid = ns.build_grid(xyzr, scale, perfil)
    jit, _ = ns.jitter(xyzr)
    ps = ns.ses_patches(jit, 1.4)
    return xyzr, r_used, grid, ps, time.perf_counter() - t0


def host_info():
    """The host cores the CPU legs may use: the process's CPU set, capped at
    16 (one GPU's share of the box), and the CPU model."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpus_available": avail, "threads": max(1, min(16, avail)), "cpu_model": model}


def _time_oracle(grid, ps, threads, budget_s, min_runs=1):
    """Median seconds of getSurf + triangulate of the oracle with `threads`
    threads (the reference's decomposition), repeated while the budget lasts."""
    from oracle import binding
    binding.set_threads(threads)
    try:
        orc = binding.OracleSurface(grid, ps)
        times, rays = [], 0
        t_all = time.perf_counter()
        while True:
            t = time.perf_counter()
            r = orc.get_surf()
            orc.triangulate()
            times.append(time.perf_counter() - t)
            rays = r["total_rays"]
            if len(times) >= min_runs and time.perf_counter() - t_all + times[-1] > budget_s:
                break
        orc.close()
    finally:
        binding.set_threads(1)
    return float(np.median(times)), len(times), rays


def cpu_baseline(grid, ps, budget_s):
    """The oracle (C restatement of the reference's path) on the full
    workload: on all available host cores with the reference's thread
    decomposition (patches strided over threads, 4-row ray blocks, z-strided
    marching cubes), and on one core."""
    info = host_info()
    T = info["threads"]
    sec_t, runs_t, rays = _time_oracle(grid, ps, T, 0.5 * budget_s)
    sec_1, runs_1, _ = _time_oracle(grid, ps, 1, 0.5 * budget_s)
    return {"value": rays / sec_t, "unit": "rays/s", "cores": T, "kind": "port",
            "sample": "full workload: %d thread(s) median %.3f s per step over %d run(s); 1 thread median %.3f s "
                      "over %d run(s)" % (T, sec_t, runs_t, sec_1, runs_1),
            "single_core": {"value": rays / sec_1, "unit": "rays/s", "cores": 1, "seconds_per_step": sec_1},
            "seconds_per_step": sec_t, **info}


def bench_mixed(dev, steps, warmup):
    """C2 step with the mixed-precision torus roots."""
    for _ in range(warmup):
        dev.get_surf(precision=1)
        dev.triangulate()
    dev.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        dev.get_surf(precision=1)
        dev.triangulate()
    dev.synchronize()
    return (time.perf_counter() - t0) * 1e3 / steps


def mixed_deviation(dev, grid):
    """Largest depth difference (A) between the mixed-precision and the FP64
    torus roots over every ray of every slot, and the rays whose hit counts
    differ."""
    L = (grid.nx - 1) * grid.side
    dev.get_surf(precision=0)
    ref = [dev.ray_hits(s) for s in range(6)]
    dev.get_surf(precision=1)
    worst, ndiff = 0.0, 0
    for s in range(6):
        ca, ta = dev.ray_hits(s)
        cb, tb = ref[s]
        ndiff += int((ca != cb).sum())
        k = min(ta.shape[1], tb.shape[1])
        mask = (ca == cb)[:, None] & (np.arange(k)[None, :] < ca[:, None])
        d = np.abs(ta[:, :k] - tb[:, :k])[mask]
        if d.size:
            worst = max(worst, float(d.max()) * L)
    dev.get_surf(precision=0)
    return worst, ndiff


def bench_cast_tiles(dev, steps, warmup):
    """The C2 step with the ray-tile cast over the cell grid instead of the
    patch-centric cast: step time, and the intersection stage time per mode."""
    out = {}
    for name, mode in (("patches", dev.CAST_PATCHES), ("tiles", dev.CAST_TILES)):
        dev.set_cast_mode(mode)
        dev.get_surf()  # builds the cell grid (tiles) outside the timed region
        dev.triangulate()
        for _ in range(warmup):
            dev.get_surf()
            dev.triangulate()
        dev.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            dev.get_surf()
            dev.triangulate()
        dev.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        dev.set_profiling(True)
        inter = 0.0
        for _ in range(steps):
            dev.get_surf()
            dev.triangulate()
            inter += dev.stage_times()["intersect"] / steps
        dev.set_profiling(False)
        out[name] = {"ms_per_step": ms, "intersect_ms": inter}
    dev.set_cast_mode(dev.CAST_PATCHES)
    return out


def bench_batched(n_frames=256, repeats=3):
    """Uneven frames, one captured graph launch vs the per-frame loop."""
    t0 = time.perf_counter()
    frames = ns.perturbed_frames(n_frames, seed=0)
    inputs = []
    for x in frames:
        g = ns.build_grid(x, 2.0, 80.0)
        j, _ = ns.jitter(x)
        inputs.append((g, ns.ses_patches(j, 1.4)))
    host_s = time.perf_counter() - t0
    b = ns.Batch(n_frames)
    for f, (g, ps) in enumerate(inputs):
        b.set_frame(f, g, ps)
    chains = b.prepare(chains=0)
    best, lat_best, redone = 1e30, None, 0
    for _ in range(repeats):
        t = time.perf_counter()
        lat, redone = b.run()
        t = time.perf_counter() - t
        if t < best:
            best, lat_best = t, lat
    shape = b.shape()
    loop_best, loop_lat = 1e30, None
    for _ in range(repeats):
        t = time.perf_counter()
        lat = b.run_loop()
        t = time.perf_counter() - t
        if t < loop_best:
            loop_best, loop_lat = t, lat
    # each cast form at 64 chains, for comparison with the autotuned shape
    by_cast = {}
    for cast in ("frames", "joint", "persistent"):
        b.prepare(chains=64, cast=cast)
        t_best = 1e30
        for _ in range(repeats):
            t = time.perf_counter()
            b.run()
            t_best = min(t_best, time.perf_counter() - t)
        by_cast[cast] = {"ms_per_batch": t_best * 1e3, "workgroups": b.shape()["workgroups"]}
    nvox = sum(g.nx * g.ny * g.nz for g, _ in inputs)
    b.close()
    return {"workload": "%d frames, %d-%d atoms (log-uniform), sigma 0.3 A, probe 1.4 A, 2.0 points/A, "
                        "epsmap+idebmap+status+cavities, 3 panels" % (n_frames, min(map(len, frames)),
                                                                       max(map(len, frames))),
            "frames_per_s": n_frames / best, "ms_per_batch": best * 1e3,
            "p50_ms": float(np.percentile(lat_best, 50)), "p99_ms": float(np.percentile(lat_best, 99)),
            "latency": "completion of each frame after the batch starts: the graph's from the device clock, "
                       "the loop's as the running sum of its synchronous per-frame calls",
            "loop_frames_per_s": n_frames / loop_best,
            "loop_p50_ms": float(np.percentile(np.cumsum(loop_lat), 50)),
            "loop_p99_ms": float(np.percentile(np.cumsum(loop_lat), 99)),
            "loop_call_p50_ms": float(np.percentile(loop_lat, 50)),
            "loop_call_p99_ms": float(np.percentile(loop_lat, 99)), "speedup_vs_loop": loop_best / best,
            "chains": chains, "shape": shape, "ms_by_cast_at_64_chains": by_cast, "frames_redone": redone,
            "voxels": nvox, "host_build_s": host_s}


def bench_differentiable(steps=10):
    """2,000 atoms at 3 points/A: forward + adjoint, device part and through autograd."""
    xyzr, _ = ns.generate_chain(2000, 21.0, seed=0)
    grid = ns.build_grid(xyzr, 3.0, 80.0)
    jit, _ = ns.jitter(xyzr)
    dev = ns.DeviceSurface(0)
    dev.set_grid(grid)
    dev.set_patches(ns.ses_patches(jit, 1.4))
    dev.set_atoms(jit)
    out = dev.surface_gradient()
    t0 = time.perf_counter()
    for _ in range(steps):
        out = dev.surface_gradient()
    dev_ms = (time.perf_counter() - t0) * 1e3 / steps
    # stage breakdown from a second, profiled pass (its events and reads are
    # kept out of the timed loop above)
    dev.set_profiling(True)
    st = {"intersect": 0.0, "sort": 0.0, "adjoint": 0.0}
    for _ in range(steps):
        dev.surface_gradient()
        for k, v in dev.stage_times().items():
            if k in st:
                st[k] += v / steps
    dev.set_profiling(False)
    res = {"workload": "2000-atom chain, probe 1.4 A, 3.0 points/A (%d^3), FP64; volume and area estimate, "
                       "gradients for %d coordinates + %d radii" % (grid.nx, 3 * len(jit), len(jit)),
           "fwd_bwd_ms": dev_ms, "stage_ms": st, "volume": out["volume"], "area": out["area"]}
    try:
        import torch
        from nanoshaper_amd.autograd import surface_volume_area
        x = torch.tensor(jit, dtype=torch.float64, requires_grad=True)
        t0 = time.perf_counter()
        for _ in range(3):
            x.grad = None
            v, a = surface_volume_area(x, grid, 1.4, dev)
            (v + a).backward()
        res["autograd_fwd_bwd_ms_incl_host_patches"] = (time.perf_counter() - t0) * 1e3 / 3
    except ImportError:
        pass
    dev.close()
    return res


def counter_roofline(profile, kernel, seconds=None):
    """Roofline entry of a config-3 kernel from its committed counter profile
    (profiles/*_traffic.json): FP64 lane operations issued (wave instructions
    x 64, an FMA counting 2) over the kernel's time (the live stage time when
    given, else the profile's rocprof average) against the FP64 peak, and its
    calibrated HBM bytes against the HBM peak.  Issued, not algorithmic:
    inactive lanes of an issued instruction are counted too."""
    path = os.path.join(ROOT, "profiles", profile)
    if not os.path.exists(path):
        return None
    k = json.load(open(path))["kernels"].get(kernel)
    if not k or "fp64_lane_ops" not in k:
        return None
    t = seconds if seconds else k.get("avg_us", 0) * 1e-6
    if not t:
        return None
    traffic = k["fetch_bytes"] + k["write_bytes"]
    fl = k["fp64_lane_ops"] / t / 1e12
    return {"kernel": kernel, "bound": "mfma", "compute": "FP64 VALU (no matrix form)",
            "achieved": fl, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s", "frac": fl / PEAK_FP64_TFLOPS,
            "ops": "FP64 lane operations issued, from counters", "traffic": traffic,
            "hbm_GBs": traffic / t / 1e9, "hbm_frac": traffic / t / 1e9 / PEAK_HBM_GBS,
            "time_us": t * 1e6, "time": "live stage time" if seconds else "rocprof average (profile)",
            "profile": "profiles/" + profile}


def bench_blobby(steps=5, cpu=True):
    """Blobby surface (B = -2.5, cutoff 6 A): 20,000 atoms (compacted radius
    48 A), 2.0 points/A, perfil 80.  A step is BlobbySurface::build (field,
    marching cubes, smoothing, OFF round trip) plus getSurf on the mesh
    (ray-based, maps + cavities + 3 edge panels) and triangulateSurface."""
    xyzr, r_used = ns.generate_chain(20000, 48.0, seed=0)
    grid = ns.build_grid(xyzr, 2.0, 80.0)
    dev = ns.DeviceSurface(0)
    dev.set_grid(grid)
    opts = dict(build_epsmap=True, build_status=True, accurate_triangulation=True, fill_cavities=True, inside=5)
    info = dev.blobby_build(xyzr)
    r = dev.get_surf(**opts)
    m = dev.triangulate()
    dev.synchronize()
    tb = ts = tt = 0.0
    for _ in range(steps):
        t0 = time.perf_counter()
        info = dev.blobby_build(xyzr)
        dev.synchronize()
        t1 = time.perf_counter()
        r = dev.get_surf(**opts)
        t2 = time.perf_counter()
        m = dev.triangulate()
        dev.synchronize()
        t3 = time.perf_counter()
        tb += (t1 - t0) / steps
        ts += (t2 - t1) / steps
        tt += (t3 - t2) / steps
    dev.close()
    nvox = grid.nx * grid.ny * grid.nz
    step = tb + ts + tt
    res = {"workload": "blobby B -2.5, cutoff 6 A: %d-atom chain (R %.1f A), 2.0 points/A perfil 80 -> %d^3; "
                       "build + getSurf (ray-based, maps + cavities + 6 panels) + triangulate"
                       % (len(xyzr), r_used, grid.nx),
           "ms_per_step": step * 1e3, "build_ms": tb * 1e3, "get_surf_ms": ts * 1e3, "triangulate_ms": tt * 1e3,
           "rays_per_s": r["total_rays"] / step, "voxels_per_s": nvox / step,
           "blobby_mesh": {"vertices": info["num_vertices"], "triangles": info["num_triangles"]},
           "result": {"volume": r["volume"], "hits": r["total_hits"], "odd_rays": r["odd_rays"],
                      "cavities": r["num_cavities"], "triangles": m["num_triangles"], "area": m["area"]},
           "roofline": {"field": counter_roofline("blobby_c3_20000atoms_traffic.json", "nsg::k_blobby_field"),
                        "mesh_cast": counter_roofline("blobby_c3_20000atoms_traffic.json",
                                                      "nsg::k_intersect_mesh<true>")}}
    if cpu:
        # the oracle (single thread) on the same workload, once
        from oracle import binding
        t0 = time.perf_counter()
        b = binding.blobby_build(grid, xyzr)
        o = binding.OracleSurface.from_mesh(grid, b["verts"], b["tris"], ray_based=True)
        o.get_surf(**opts)
        o.triangulate()
        cpu_s = time.perf_counter() - t0
        o.close()
        res["cpu_baseline"] = {"seconds_per_step": cpu_s, "cores": 1, "kind": "port",
                               "sample": "full workload, 1 run (the blobby build and the mesh cast of the oracle "
                                         "are single threaded)"}
        res["speedup_vs_cpu"] = cpu_s / step
    return res


def bench_skin(steps=10, cpu=True):
    """Skin surface (s = 0.45): 20,000 atoms (compacted radius 48 A), 2.0
    points/A, perfil 80.  Host: jitter, regular triangulation and mixed
    cells, once.  A step is getSurf on the cells (maps + cavities + 6 panels)
    plus triangulateSurface."""
    xyzr, r_used = ns.generate_chain(20000, 48.0, seed=0)
    grid = ns.build_grid(xyzr, 2.0, 80.0)
    t0 = time.perf_counter()
    jit, _ = ns.jitter(xyzr)
    ps = ns.skin_patches(jit, 0.45)
    host_s = time.perf_counter() - t0
    dev = ns.DeviceSurface(0)
    dev.set_grid(grid)
    dev.set_patches(ps)
    opts = dict(build_epsmap=True, build_status=True, accurate_triangulation=True, fill_cavities=True, inside=5)
    for _ in range(2):
        r = dev.get_surf(**opts)
        m = dev.triangulate()
    dev.set_profiling(True)
    st = {k: 0.0 for k in ns.STAGES}
    dev.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        r = dev.get_surf(**opts)
        m = dev.triangulate()
        for k, v in dev.stage_times().items():
            st[k] += v / steps
    dev.synchronize()
    step = (time.perf_counter() - t0) / steps
    dev.close()
    kinds = np.bincount(ps.kind, minlength=9)[5:].tolist()
    res = {"workload": "skin s 0.45: %d-atom chain (R %.1f A), 2.0 points/A perfil 80 -> %d^3; getSurf (maps + "
                       "cavities + 6 panels) + triangulate" % (len(xyzr), r_used, grid.nx),
           "ms_per_step": step * 1e3, "rays_per_s": r["total_rays"] / step,
           "voxels_per_s": grid.nx * grid.ny * grid.nz / step, "stage_ms": st,
           "cells": dict(zip(ns.surface.SKIN_KINDS, kinds)), "host_build_s": host_s,
           "roofline": counter_roofline("skin_c3_20000atoms_traffic.json", "nsg::k_intersect_skin",
                                        st["intersect"] * 1e-3),
           "result": {"volume": r["volume"], "hits": r["total_hits"], "odd_rays": r["odd_rays"],
                      "cavities": r["num_cavities"], "triangles": m["num_triangles"], "area": m["area"]}}
    if cpu:
        # the oracle on all available cores (skin cells strided over threads,
        # ray rows in 4-row blocks, z-strided marching cubes) and on one core
        from oracle import binding
        info = host_info()
        secs = {}
        for threads in (info["threads"], 1):
            binding.set_threads(threads)
            try:
                o = binding.OracleSurface(grid, ps)
                t0 = time.perf_counter()
                o.get_surf(**opts)
                o.triangulate()
                secs[threads] = time.perf_counter() - t0
                o.close()
            finally:
                binding.set_threads(1)
        T = info["threads"]
        res["cpu_baseline"] = {"seconds_per_step": secs[T], "cores": T, "kind": "port",
                               "sample": "full workload, 1 run per thread count",
                               "single_core_seconds_per_step": secs[1], **info}
        res["speedup_vs_cpu"] = secs[T] / step
        res["speedup_vs_cpu_1core"] = secs[1] / step
    return res


def bench_bgp(dev, grid, ps, steps=10, cpu=True):
    """Project_boundary_grid_points on the benchmark surface: detection from
    the epsmap, the auxiliary cell lists and the projections, timed per call
    of the device stage (results left on the device), beside the oracle's
    single-threaded restatement on the same maps."""
    r = dev.get_surf(build_epsmap=True, build_status=True, accurate_triangulation=False, fill_cavities=True)
    counts = dev.project_bgp_device()
    dev.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        counts = dev.project_bgp_device()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    n = int(counts[0])
    out = {"workload": "the benchmark SES (%d^3), maps + cavities, auxiliary grid <= 50 cells per side" % grid.nx,
           "ms_per_call": ms, "points": n, "points_per_s": n / (ms * 1e-3),
           "counts": {"points": n, "no_face_outside": int(counts[1]), "sampled": int(counts[2]),
                      "left_at_node": int(counts[3])}}
    if cpu:
        from oracle.binding import OracleSurface
        orc = OracleSurface(grid, ps)
        orc.get_surf(build_epsmap=True, build_status=True, accurate_triangulation=False, fill_cavities=True)
        t0 = time.perf_counter()
        bo = orc.project_bgp()
        cs = time.perf_counter() - t0
        orc.close()
        out["cpu_baseline"] = {"seconds_per_call": cs, "cores": 1, "kind": "port",
                               "sample": "full workload, 1 run (the oracle's projection is single threaded)",
                               "points_equal": bool(int(bo["counts"][0]) == n)}
        out["speedup_vs_cpu"] = cs / (ms * 1e-3)
    return out


def bench_pockets(steps=10, cpu=True, n_atoms=5000, scale=2.0):
    """Pocket detection (the grid part of pocket mode) on the 5,000-atom
    chain at 2 points/A: the fat-probe (3.0 A) and small-probe (1.4 A)
    status maps computed once, then nsgpu_pockets timed per call (difference,
    two probe-ball filters, components, labels, support lists), beside the
    oracle's restatement of the reference's loops on the same maps."""
    import math
    xyzr, _ = ns.generate_chain(n_atoms, CONFIGS["c2"][1], seed=0)
    grid = ns.build_grid(xyzr, scale, 80.0)
    j1, st = ns.jitter(xyzr)
    j2, _ = ns.jitter(xyzr, seed=st)
    p1, p2 = ns.ses_patches(j1, 3.0), ns.ses_patches(j2, 1.4)
    fat_o = dict(build_epsmap=False, build_status=True, accurate_triangulation=False, fill_cavities=True,
                 cav_vol=math.inf)
    small_o = dict(build_epsmap=False, build_status=True, accurate_triangulation=False, fill_cavities=True,
                   cav_vol=0.0)
    fat, small = ns.DeviceSurface(0), ns.DeviceSurface(0)
    for d, ps, o in ((fat, p1, fat_o), (small, p2, small_o)):
        d.set_grid(grid)
        d.set_patches(ps)
    small.get_surf(**small_o)
    ms = []
    for it in range(steps + 1):
        fat.get_surf(**fat_o)  # the pocket labels replace its status map
        fat.synchronize()
        t0 = time.perf_counter()
        pk = fat.pockets(small)
        ms.append((time.perf_counter() - t0) * 1e3)
    ms = float(np.median(ms[1:]))
    nvox = grid.nx * grid.ny * grid.nz
    out = {"workload": "pocket detection: %d-atom chain, probes 3.0 / 1.4 A, 2.0 points/A perfil 80 -> %d^3"
                       % (n_atoms, grid.nx),
           "ms_per_call": ms, "voxels_per_s": nvox / (ms * 1e-3), "components": len(pk["volumes"]),
           "kept": len(pk["support"]), "pockets": int(pk["is_pocket"][~pk["filled"]].sum()),
           "support_points": int(sum(len(x) for x in pk["support"]))}
    if cpu:
        from oracle.binding import OracleSurface
        of, os_ = OracleSurface(grid, p1), OracleSurface(grid, p2)
        of.get_surf(**fat_o)
        os_.get_surf(**small_o)
        t0 = time.perf_counter()
        po = of.pockets(os_)
        cs = time.perf_counter() - t0
        out["cpu_baseline"] = {"seconds_per_call": cs, "cores": 1, "kind": "port",
                               "sample": "full workload, 1 run (the oracle's filter loops are single threaded)",
                               "components_equal": bool(len(po["volumes"]) == len(pk["volumes"]))}
        out["speedup_vs_cpu"] = cs / (ms * 1e-3)
        of.close()
        os_.close()
    fat.close()
    small.close()
    return out


def copy_reference(nbytes, repeats=5):
    """Plain device-to-device copy of known size: the calibration kernel for
    the FETCH_SIZE / WRITE_SIZE counters (a copy reads and writes nbytes)."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    for _ in range(repeats):
        assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_reference_bytes": nbytes, "repeats": repeats}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--copy-ref", type=int, default=0,
                    help="only run a device-to-device copy of this many bytes (counter calibration)")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="c2", choices=sorted(CONFIGS))
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--precision", type=int, default=0)
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the mixed, batched, differentiable and blobby lines")
    ap.add_argument("--blobby-only", action="store_true", help="only the blobby line (profiling)")
    ap.add_argument("--skin-only", action="store_true", help="only the skin line (profiling)")
    ap.add_argument("--bgp-only", action="store_true", help="only the boundary-grid-point line (profiling)")
    ap.add_argument("--pockets-only", action="store_true", help="only the pocket-detection line (profiling)")
    ap.add_argument("--batched-only", action="store_true", help="only the batched-frames line (profiling)")
    args = ap.parse_args()
    if args.copy_ref:
        copy_reference(args.copy_ref)
        return
    if args.blobby_only:
        print(json.dumps({"blobby": bench_blobby(args.steps, cpu=not args.no_cpu)}))
        return
    if args.skin_only:
        print(json.dumps({"skin": bench_skin(args.steps, cpu=not args.no_cpu)}))
        return
    if args.pockets_only:
        print(json.dumps({"pockets": bench_pockets(args.steps, cpu=not args.no_cpu)}))
        return
    if args.batched_only:
        print(json.dumps({"batched": bench_batched()}))
        return
    if args.bgp_only:
        xyzr, r_used, grid, ps, host_s = build_inputs(args.config)
        dev = ns.DeviceSurface(0)
        dev.set_grid(grid)
        dev.set_patches(ps)
        print(json.dumps({"bgp": bench_bgp(dev, grid, ps, args.steps, cpu=not args.no_cpu)}))
        dev.close()
        return
    if args.gpus != 1:
        print("bench.py runs on exactly one GPU: this build has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    xyzr, r_used, grid, ps, host_s = build_inputs(args.config)
    print("inputs: %d atoms (R %.2f A), grid %d^3, %d patches %s, host build %.3f s" %
          (len(xyzr), r_used, grid.nx, len(ps), ps.counts(), host_s), file=sys.stderr, flush=True)
    dev = ns.DeviceSurface(0)
    dev.set_grid(grid)
    dev.set_patches(ps)
    opts = dict(precision=args.precision)
    for _ in range(args.warmup):
        dev.get_surf(**opts)
        dev.triangulate()
    dev.synchronize()
    print("warmup done", file=sys.stderr, flush=True)

    dev.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r = dev.get_surf(**opts)
        m = dev.triangulate()
    dev.synchronize()
    elapsed = time.perf_counter() - t0
    # stage breakdown from a second pass of the same steps with stage events
    # (kept out of the timed loop: the events and their reads cost ~3 %)
    dev.set_profiling(True)
    stage_sum = {s: 0.0 for s in ns.STAGES}
    for _ in range(args.steps):
        dev.get_surf(**opts)
        dev.triangulate()
        st = dev.stage_times()
        for k, v in st.items():
            stage_sum[k] += v
    dev.set_profiling(False)
    stage_avg = {k: v / args.steps for k, v in stage_sum.items()}
    ms_per_step = elapsed * 1e3 / args.steps
    nvox = grid.nx * grid.ny * grid.nz
    rays = r["total_rays"]

    # roofline of the dominant kernel (the ray/patch intersection)
    pc = dev.intersect_path_counts(**opts)
    flops = sum(OPS[k] * pc[k] for k in OPS)
    t_int = stage_avg["intersect"] * 1e-3
    achieved = flops / t_int / 1e12 if t_int > 0 else 0.0
    # HBM bytes per launch from the FETCH_SIZE / WRITE_SIZE counter runs
    # (calibrated against a copy kernel) committed under profiles/
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "ses_c2_5000atoms_traffic.json")
    if args.config == "c2" and args.precision == 0 and os.path.exists(tfile):
        kern = json.load(open(tfile)).get("kernels", {})
        parts = [kern.get(k) for k in INTERSECT_KERNELS]
        if all(parts):
            traffic = sum(t["fetch_bytes"] + t["write_bytes"] for t in parts)
    # compute-bound ("mfma" is the contract's name for the compute roof): FP64
    # vector arithmetic, priced against MI355X's FP64 peak, the same 78.6
    # TFLOP/s for vector and matrix FP64; the kernel has no matrix form
    roofline = {"kernel": "k_intersect_plain + k_intersect_torus (one stage)", "bound": "mfma",
                "compute": "FP64 VALU: the roof is the compute roof at MI355X's FP64 rate (78.6 TFLOP/s, "
                           "vector and matrix alike); the kernels have no matrix form and use no MFMA",
                "achieved": achieved, "peak": PEAK_FP64_TFLOPS,
                "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS, "traffic": traffic,
                "flops_per_launch": flops, "path_counts": pc}
    fill_bytes = 17.0 * nvox
    t_fill = stage_avg["fill"] * 1e-3
    fill_roof = {"kernel": "k_fill_bits", "bound": "hbm", "bytes_per_launch": fill_bytes,
                 "achieved_GBs": fill_bytes / t_fill / 1e9 if t_fill > 0 else 0.0, "peak_GBs": PEAK_HBM_GBS}
    fill_roof["frac"] = fill_roof["achieved_GBs"] / PEAK_HBM_GBS

    extras = {}
    if not args.no_extras and args.config == "c2":
        print("mixed precision ...", file=sys.stderr, flush=True)
        mixed_ms = bench_mixed(dev, args.steps, args.warmup)
        dev_a, nd = mixed_deviation(dev, grid)
        extras["mixed_precision"] = {"ms_per_step": mixed_ms, "value": rays / (mixed_ms * 1e-3),
                                     "max_depth_deviation_vs_fp64_A": dev_a, "rays_with_other_hit_counts": nd,
                                     "speedup_vs_fp64": ms_per_step / mixed_ms,
                                     "hit_error": "<= 1e-6 A against the true roots of the same quartics "
                                                  "(tests/test_gpu_parity.py::test_mixed_precision_matches_oracle)"}
        print("cell grid ray tiles ...", file=sys.stderr, flush=True)
        extras["cast_modes"] = bench_cast_tiles(dev, args.steps, args.warmup)
        print("boundary grid points ...", file=sys.stderr, flush=True)
        extras["bgp"] = bench_bgp(dev, grid, ps, cpu=not args.no_cpu)
    dev.close()
    if not args.no_extras:
        print("batched frames ...", file=sys.stderr, flush=True)
        extras["batched"] = bench_batched()
        print("differentiable ...", file=sys.stderr, flush=True)
        extras["differentiable"] = bench_differentiable()
        print("blobby ...", file=sys.stderr, flush=True)
        extras["blobby"] = bench_blobby(cpu=not args.no_cpu)
        print("skin ...", file=sys.stderr, flush=True)
        extras["skin"] = bench_skin(cpu=not args.no_cpu)
        print("pockets ...", file=sys.stderr, flush=True)
        extras["pockets"] = bench_pockets(cpu=not args.no_cpu)

    cpu = None
    if not args.no_cpu:
        print("cpu baseline ...", file=sys.stderr, flush=True)
        cpu = cpu_baseline(grid, ps, args.cpu_budget)
    value = rays * args.steps / elapsed
    out = {
        "metric": "SES rays/s & voxels/s at 1e-6 A hit err; batched frames/s; fwd+bwd ms",
        "value": value,
        "unit": "rays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == 0 else "f32-cull/f64",
        "data": "synthetic",
        "config": {"workload": "plain SES %s: %d-atom seeded chain (R %.2f A), probe 1.4 A, grid %.1f/A "
                               "perfil %.0f -> %d^3, maps+cavities+MC, 6 panels" %
                               (args.config, len(xyzr), r_used, CONFIGS[args.config][2], CONFIGS[args.config][3],
                                grid.nx),
                   "atoms": len(xyzr), "grid": grid.nx, "patches": len(ps), "rays_per_step": rays,
                   "voxels": nvox},
        "voxels_per_s": nvox * args.steps / elapsed,
        "stage_ms": stage_avg,
        "roofline": roofline,
        "roofline_fill": fill_roof,
        "cpu_baseline": cpu,
        "speedup_vs_cpu": (value / cpu["value"]) if cpu else None,
        "speedup_vs_cpu_1core": (value / cpu["single_core"]["value"]) if cpu else None,
        "result": {"volume": r["volume"], "area": m["area"], "cavities": r["num_cavities"],
                   "triangles": m["num_triangles"], "hits": r["total_hits"], "odd_rays": r["odd_rays"]},
    }
    out.update(extras)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
