"""GPU: batches of uneven frames in one captured graph launch give exactly
the per-frame get_surf results, on every replay."""
import numpy as np
import pytest

import nanoshaper_amd as ns

pytestmark = pytest.mark.gpu

OPTS = dict(build_epsmap=True, build_status=True, accurate_triangulation=False, fill_cavities=True)


def _inputs(frames, scale=2.0):
    out = []
    for xyzr in frames:
        grid = ns.build_grid(xyzr, scale, 80.0)
        jit, _ = ns.jitter(xyzr)
        out.append((grid, ns.ses_patches(jit, 1.4)))
    return out


def _reference(inputs, precision=0):
    ref = []
    for grid, ps in inputs:
        dev = ns.DeviceSurface(0)
        dev.set_grid(grid)
        dev.set_patches(ps)
        r = dev.get_surf(precision=precision, **OPTS)
        ref.append((r, dev.epsmap(), dev.idebmap(), dev.status(), dev.cavities()))
        dev.close()
    return ref


def _check(batch, ref):
    for f, (r, eps, ideb, st, cav) in enumerate(ref):
        print("frame %d" % f, flush=True)
        rb = batch.result(f)
        for k in ("total_rays", "odd_rays", "total_hits", "candidates", "num_cavities", "max_hits_per_ray", "volume"):
            assert rb[k] == r[k], (f, k, rb[k], r[k])
        v = batch.frame(f)
        assert np.array_equal(v.epsmap(), eps)
        assert np.array_equal(v.idebmap(), ideb)
        assert np.array_equal(v.status(), st)
        sa, fa = v.cavities()
        assert np.array_equal(sa, cav[0]) and np.array_equal(fa, cav[1])


@pytest.mark.parametrize("chains", [0, 2])
def test_batch_graph_matches_per_frame_get_surf(device_ok, chains):
    frames = ns.perturbed_frames(6, n_min=150, n_max=1500, seed=3)
    inputs = _inputs(frames)
    ref = _reference(inputs)
    b = ns.Batch(len(inputs))
    for f, (grid, ps) in enumerate(inputs):
        b.set_frame(f, grid, ps)
    used = b.prepare(chains=chains, **OPTS)
    assert used >= 1
    # several replays with eager work in between: every replay is checked
    for rep in range(3):
        lat, redone = b.run()
        assert redone == 0
        assert (lat > 0).all()
        _check(b, ref)
        b.run_loop()
        _check(b, ref)
    b.close()


def test_batch_frame_with_cavities(device_ok):
    # a frame with buried cavities keeps the reference's labels in the graph
    xyzr, _ = ns.generate_chain(2000, 21.0, seed=0)
    inputs = _inputs([xyzr], scale=3.0)
    ref = _reference(inputs)
    assert ref[0][0]["num_cavities"] > 0
    b = ns.Batch(1)
    b.set_frame(0, *inputs[0])
    b.prepare(chains=1, **OPTS)
    b.run()
    _check(b, ref)
    b.close()


def test_batch_captures_again_after_a_frame_changes(device_ok):
    """Calls through a frame's own context that move its buffers or change
    its options (a larger grid, get_surf with edge rays) make the next run
    capture the graph again instead of launching it onto freed memory."""
    frames = ns.perturbed_frames(3, n_min=150, n_max=600, seed=5)
    inputs = _inputs(frames)
    b = ns.Batch(len(inputs))
    for f, (grid, ps) in enumerate(inputs):
        b.set_frame(f, grid, ps)
    b.prepare(chains=2, **OPTS)
    b.run()
    _check(b, _reference(inputs))
    # frame 0 on a finer grid: every map buffer of that context grows
    big = ns.build_grid(frames[0], 3.0, 80.0)
    v0 = b.frame(0)
    v0.set_grid(big)
    v0.set_patches(inputs[0][1])
    inputs[0] = (big, inputs[0][1])
    b.grids[0] = big
    # frame 1 with other options (edge rays, vertex pass): its host state changes
    b.frame(1).get_surf(build_epsmap=True, build_status=True, accurate_triangulation=True, fill_cavities=True)
    print("run after the frames changed", flush=True)
    lat, redone = b.run()
    assert redone == 0
    _check(b, _reference(inputs))
    # and again with nothing changed: the recaptured graph is reused
    b.run()
    _check(b, _reference(inputs))
    b.close()


@pytest.mark.parametrize("cast,chains,workgroups,chunk,precision", [("joint", 2, 0, 0, 0),
                                                                    ("persistent", 3, 0, 0, 0),
                                                                    ("persistent", 2, 7, 1, 0),
                                                                    ("persistent", 2, 5, 13, 0),
                                                                    ("persistent", 4, 0, 4, 1),
                                                                    ("joint", 1, 0, 0, 2)])
def test_joint_cast_matches_per_frame_get_surf(device_ok, cast, chains, workgroups, chunk, precision):
    """One intersection launch per patch class over every frame (a workgroup
    per item, or persistent workgroups taking chunks of items from a queue;
    5 or 7 of them take hundreds of chunks each, 13-item chunks straddle the
    frames' item ranges) gives the per-frame results on every replay, with
    each torus solver."""
    frames = ns.perturbed_frames(6, n_min=150, n_max=1500, seed=4)
    inputs = _inputs(frames)
    ref = _reference(inputs, precision)
    b = ns.Batch(len(inputs))
    for f, (grid, ps) in enumerate(inputs):
        b.set_frame(f, grid, ps)
    print("shape", cast, chains, workgroups, chunk, "precision", precision, flush=True)
    b.prepare(chains=chains, cast=cast, workgroups=workgroups, chunk=chunk, precision=precision, **OPTS)
    sh = b.shape()
    assert sh["cast"] == cast and sh["chains"] == chains
    if cast == "persistent":
        assert sh["workgroups"] == (workgroups or sh["workgroups"]) > 0
        assert sh["chunk"] == (chunk or 16)
    for rep in range(3):
        lat, redone = b.run()
        assert redone == 0 and (lat > 0).all()
        _check(b, ref)
        b.run_loop()
        _check(b, ref)
    b.close()


def test_joint_cast_after_a_frame_changes(device_ok):
    # new patches for one frame: the frame descriptors are uploaded again
    frames = ns.perturbed_frames(4, n_min=150, n_max=600, seed=6)
    inputs = _inputs(frames)
    b = ns.Batch(len(inputs))
    for f, (grid, ps) in enumerate(inputs):
        b.set_frame(f, grid, ps)
    b.prepare(chains=2, cast="persistent", **OPTS)
    b.run()
    _check(b, _reference(inputs))
    moved = ns.perturbed_frames(1, n_min=700, n_max=700, seed=9)[0]
    inputs[2] = _inputs([moved])[0]
    b.set_frame(2, *inputs[2])
    print("run after frame 2 changed", flush=True)
    lat, redone = b.run()
    assert redone == 0
    _check(b, _reference(inputs))
    b.close()


def test_autotuned_shape_and_joint_cast_refusal(device_ok):
    frames = ns.perturbed_frames(4, n_min=150, n_max=600, seed=8)
    inputs = _inputs(frames)
    b = ns.Batch(len(inputs))
    for f, (grid, ps) in enumerate(inputs):
        b.set_frame(f, grid, ps)
    b.prepare(chains=0, **OPTS)
    assert b.shape()["cast"] in ("frames", "joint", "persistent")
    b.run()
    _check(b, _reference(inputs))
    # a skin frame: only per-frame casts apply
    jit, _ = ns.jitter(frames[1])
    b.set_frame(1, inputs[1][0], ns.skin_patches(jit, 0.45))
    with pytest.raises(ns.NsgpuError):
        b.prepare(chains=2, cast="joint", **OPTS)
    b.prepare(chains=0, **OPTS)
    assert b.shape()["cast"] == "frames"
    b.close()
