"""GPU parity: libmrfmap_hip (through its C ABI) against the CPU oracle.

Traversal lists must match bit for bit; marginals must agree with the FP64
BP within 1e-3 max-abs (the tolerance of the mixed-precision target; FP32
messages land far inside it).
"""
import ctypes

import numpy as np
import pytest

from mrfmap_amd import GVDBInference, _lib, synth

import oracle
from helpers import decimate, gpu_for, oracle_for

pytestmark = pytest.mark.gpu

MARG_TOL = 1e-3


def _compare_traversal(g, o, ncam):
    tot = 0
    for c in range(ncam):
        gr, gi = g.get_traversal(c)
        orr, oi = o.traversal(c)
        assert gr.shape == orr.shape, f"cam {c}: ray count {gr.shape} vs {orr.shape}"
        assert gi.shape == oi.shape, f"cam {c}: record count {gi.shape} vs {oi.shape}"
        bad = np.nonzero((gr != orr).any(axis=1))[0]
        assert len(bad) == 0, f"cam {c}: {len(bad)} ray records differ, first {bad[:3]}: {gr[bad[:1]]} vs {orr[bad[:1]]}"
        badi = np.nonzero((gi != oi).any(axis=1))[0]
        assert len(badi) == 0, f"cam {c}: {len(badi)} incidence records differ, first {badi[:3]}"
        tot += len(gi)
    return tot


@pytest.mark.parametrize("activation", [_lib.ACT_FULL_RAY, _lib.ACT_ENDPOINT])
def test_traversal_bit_exact_small(activation):
    p, frames, _ = synth.make_workload("small", activation=activation)
    g = gpu_for(p, frames)
    o = oracle_for(p, frames)
    g.prepare()
    o.prepare()
    assert g.num_bricks() == o.num_bricks()
    np.testing.assert_array_equal(g.get_occupied_bricks(), (o.brick_coords() * 8 + 4).astype(np.float32))
    n = _compare_traversal(g, o, len(frames))
    assert n > 1_000_000


def test_uneven_keyframes_match_oracle():
    """The batched generator at reduced size: 6 keyframes drawn from
    320x240, 640x480 and 848x480 (each with its own intrinsics; 848x480 with
    the 24/20-pixel border mask, GVDBImage.cpp:35-40) in the 10x10x3 m scene,
    8 cm voxels, every 4th pixel row and column: traversal lists bit-exact
    per camera and marginals after 3 iterations."""
    p, frames, _ = synth.make_workload("batched", frames=6, res=0.08)
    shapes = {f[2].shape for f in frames}
    assert len(shapes) >= 2, shapes
    frames = [(pose, K, decimate(depth, 4)) for pose, K, depth in frames]
    g = gpu_for(p, frames)
    o = oracle_for(p, frames)
    g.prepare()
    o.prepare()
    assert g.num_bricks() == o.num_bricks()
    n = _compare_traversal(g, o, len(frames))
    assert n > 100_000
    st = g.perform_inference(max_iters=3)
    so = o.perform_inference(3)
    assert st["updates_per_iter"] == so["updates_per_iter"]
    err = np.abs(g.get_alphas().astype(np.float64) - o.alphas())
    print(f"uneven keyframes {sorted(shapes)}: {n} records, max-abs marginal error {err.max():.3e}")
    assert err.max() <= MARG_TOL


def test_uneven_keyframes_at_1cm_match_oracle():
    """The batched workload at its own resolution (1 cm, 10x10x3 m, depth up
    to 10 m, so rays of more than a thousand voxels): the first keyframe of
    each image size among its 64 (848x480 with the border mask, 320x240,
    640x480, each with its own intrinsics), every 8th pixel row and column.
    Traversal bit-exact per camera, marginals after 3 iterations within
    1e-3 of the FP64 oracle."""
    p, frames, _ = synth.make_workload("batched")
    first = {}
    for i, (_, _, d) in enumerate(frames):
        first.setdefault(d.shape, i)
    assert len(first) == 3, first
    frames = [(frames[i][0], frames[i][1], decimate(frames[i][2], 8)) for i in sorted(first.values())]
    g = gpu_for(p, frames)
    o = oracle_for(p, frames)
    g.prepare()
    o.prepare()
    assert g.num_bricks() == o.num_bricks()
    n = _compare_traversal(g, o, len(frames))
    longest = max(int(o.traversal(c)[0][:, 1].max()) for c in range(len(frames)))
    assert n > 5_000_000 and longest > 1000, (n, longest)
    st = g.perform_inference(max_iters=3)
    so = o.perform_inference(3)
    assert st["updates_per_iter"] == so["updates_per_iter"]
    err = np.abs(g.get_alphas().astype(np.float64) - o.alphas())
    print(f"uneven keyframes at 1 cm: {n} records, longest ray {longest} records, max-abs marginal error "
          f"{err.max():.3e}")
    assert err.max() <= MARG_TOL


def test_traversal_bit_exact_real_shape():
    """640x480, f=525, 1 cm voxels: two keyframes, every 3rd pixel."""
    p, frames, _ = synth.make_workload("real", frames=13)
    frames = [(pose, K, decimate(d, 3)) for (pose, K, d) in frames[:13:6]]
    g = gpu_for(p, frames)
    o = oracle_for(p, frames)
    g.prepare()
    o.prepare()
    assert g.num_bricks() == o.num_bricks()
    n = _compare_traversal(g, o, len(frames))
    assert n > 5_000_000


def test_bp_marginals_real_shape():
    """640x480, f=525, 1 cm: two keyframes, every 3rd pixel.  Every ray of a
    camera crosses the voxels at its centre, so those segments hold tens of
    thousands of incidences and span many tiles of the segment kernels."""
    p, frames, _ = synth.make_workload("real", frames=13)
    frames = [(pose, K, decimate(d, 3)) for (pose, K, d) in frames[:13:6]]
    g = gpu_for(p, frames)
    o = oracle_for(p, frames)
    st = g.perform_inference(max_iters=3)
    so = o.perform_inference(3)
    assert st["updates_per_iter"] == so["updates_per_iter"]
    # the third iteration uses the kept sums of the static front chunks
    assert 0 < st["static_front_updates"] <= st["updates_per_iter"] - st["window_updates_per_iter"]
    # padded slice regions hold every record
    assert st["window_slots"] >= st["window_updates_per_iter"]
    assert st["front_slots"] + st["window_slots"] >= st["list_records"]
    err = np.abs(g.get_alphas().astype(np.float64) - o.alphas())
    print(f"real-shape max-abs marginal error {err.max():.3e} rms {np.sqrt((err ** 2).mean()):.3e}")
    assert err.max() <= MARG_TOL


@pytest.mark.timeout(600)
def test_full_resolution_keyframe_matches_oracle():
    """One keyframe of the 13-keyframe workload at full size (640x480,
    f=525, 1 cm; about 150 M incidences): the per-camera unit of the
    benchmark at its real shape.  Traversal bit-exact; marginals after three
    iterations (the third reads the kept sums of the static front chunks)
    within 1e-4 of the FP64 oracle.  The oracle takes about 45 s."""
    p, frames, _ = synth.make_workload("real", frames=13)
    frames = frames[:1]
    g = gpu_for(p, frames)
    o = oracle_for(p, frames)
    g.prepare()
    o.prepare()
    assert g.num_bricks() == o.num_bricks()
    n = _compare_traversal(g, o, 1)
    assert n > 100_000_000
    st = g.perform_inference(max_iters=3)
    so = o.perform_inference(3)
    assert st["updates_per_iter"] == so["updates_per_iter"]
    assert st["static_front_updates"] > 0
    err = np.abs(g.get_alphas().astype(np.float64) - o.alphas())
    print(f"full-resolution keyframe: {n} records, max-abs marginal error {err.max():.3e}")
    assert err.max() <= 1e-4


def _check_cum_lengths(g, o, ncam):
    near = 0
    for c in range(ncam):
        gc = g.get_cum_lengths(c).astype(np.float64)
        oc = o.cum_lengths(c)
        # the device sums exactly and rounds once to float (one ulp), moving
        # by one more ulp only where rounding would land on the 0.1f gate
        # (gvdbkernels.cu:408): on the same side of the gate as the FP64 sum
        ulp = np.spacing(np.abs(oc).astype(np.float32)).astype(np.float64)
        bad = np.abs(gc - oc) > 2 * ulp
        assert not bad.any(), f"cam {c}: {bad.sum()} voxels, first {np.argwhere(bad)[:3]}"
        np.testing.assert_array_equal(gc > np.float32(0.1), oc > float(np.float32(0.1)))
        near += int((np.abs(oc - float(np.float32(0.1))) < 1e-6).sum())
    return near


def test_cum_lengths_match_oracle():
    """Per (camera, voxel) summed chord length C, the reference's CUM_LENS
    channel (gvdbkernels.cu:519-571, 1276-1302): small workload, and two
    640x480 / 1 cm keyframes (every 3rd pixel) whose C include values within
    1e-6 of the 0.1f gate (gvdbkernels.cu:408)."""
    p, frames, _ = synth.make_workload("small")
    g, o = gpu_for(p, frames), oracle_for(p, frames)
    _check_cum_lengths(g, o, len(frames))
    p, frames, _ = synth.make_workload("real", frames=13)
    frames = [(pose, K, decimate(d, 3)) for (pose, K, d) in frames[:13:6]]
    g, o = gpu_for(p, frames), oracle_for(p, frames)
    near = _check_cum_lengths(g, o, len(frames))
    print(f"pairs within 1e-6 of the gate: {near}")
    assert near > 0
    with pytest.raises(_lib.MrfError) as e:
        g.get_cum_lengths(len(frames))
    assert e.value.code == -5
    # through the C ABI with a wrong shape: MRF_E_SHAPE, nothing written
    n = g.num_bricks()
    bad = np.full((n + 1, 512), 7.0, dtype=np.float32)
    rc = g._L.mrf_get_cum_lengths(g._h, 0, bad.ctypes.data_as(ctypes.c_void_p), n + 1, 512)
    assert rc == -4
    assert (bad == 7.0).all()


@pytest.mark.parametrize("activation", [_lib.ACT_FULL_RAY, _lib.ACT_ENDPOINT])
def test_bp_marginals_small(activation):
    p, frames, _ = synth.make_workload("small", activation=activation)
    g = gpu_for(p, frames)
    o = oracle_for(p, frames)
    st = g.perform_inference(max_iters=5)
    so = o.perform_inference(5)
    assert st["iters"] == 5 and so["iters"] == 5
    assert st["updates_per_iter"] == so["updates_per_iter"]
    ga, oa = g.get_alphas(), o.alphas()
    assert ga.shape == oa.shape
    err = np.abs(ga.astype(np.float64) - oa)
    print(f"max-abs marginal error {err.max():.3e} rms {np.sqrt((err ** 2).mean()):.3e}")
    assert err.max() <= MARG_TOL
    # voxels never updated read 0 on both sides
    np.testing.assert_array_equal(ga == 0, oa == 0)
    assert abs(st["residual"] - so["residual"]) < 1e-3


@pytest.mark.parametrize("seed", [3, 10, 17])
def test_infinity_term_decided_as_oracle(seed):
    """Rays that ran out take the infinity term only when d > s_n + bias(s_n)
    (gvdbkernels.cu:1196-1208). Seed 10 holds a ray whose depth lies 1.3e-7
    beyond s_n: an FP32 s_n puts it on the other side and its voxels' first
    marginals move by ~5e-2. After one iteration every marginal must agree
    with the FP64 oracle to 1e-4 (FP32 messages land near 1e-6)."""
    p, frames, _ = synth.make_workload("small", seed=seed)
    g = gpu_for(p, frames)
    o = oracle_for(p, frames)
    g.perform_inference(max_iters=1)
    o.perform_inference(1)
    err = np.abs(g.get_alphas().astype(np.float64) - o.alphas())
    print(f"seed {seed}: first-iteration max-abs {err.max():.3e}")
    assert err.max() <= 1e-4


def test_bp_repeated_calls_keep_beliefs():
    """Beliefs persist across perform_inference calls, messages are reset
    (GVDBInference.cpp:128-134)."""
    p, frames, _ = synth.make_workload("small")
    g = gpu_for(p, frames)
    o = oracle_for(p, frames)
    for _ in range(2):
        g.perform_inference(max_iters=2)
        o.perform_inference(2)
    err = np.abs(g.get_alphas().astype(np.float64) - o.alphas()).max()
    assert err <= MARG_TOL


def test_get_alphas_shape_error():
    p, frames, _ = synth.make_workload("small", frames=1)
    g = gpu_for(p, frames)
    g.perform_inference(max_iters=1)
    n = g.num_bricks()
    with pytest.raises(_lib.MrfError) as e:
        g.get_indices_and_alphas(np.zeros((n + 1, 512), dtype=np.float32))
    assert e.value.code == -4


def test_bad_camera_id():
    p, frames, _ = synth.make_workload("small", frames=1)
    g = gpu_for(p, frames)
    with pytest.raises(_lib.MrfError) as e:
        g.set_pose(5, np.eye(4))
    assert e.value.code == -5


def test_bp_bitwise_reproducible():
    """No atomics in the BP iteration: two runs give identical bits."""
    p, frames, _ = synth.make_workload("small")
    outs = []
    for _ in range(2):
        g = gpu_for(p, frames)
        g.perform_inference(max_iters=3)
        outs.append(g.get_beliefs())
    np.testing.assert_array_equal(outs[0].view(np.uint32), outs[1].view(np.uint32))


# 16-bit storage of the ray->voxel messages on the small workload: bounds a
# margin above the errors measured against the FP64 oracle (max-abs after 1
# iteration; max-abs and RMS after 5).  FP16 measured 1.5e-3 / 3.4e-3 /
# 3.9e-5, BF16 1.1e-2 / 8.3e-3 / 1.5e-4, the same with the per-pair values
# kept in FP32 (mixed); FP32 storage stays within 5e-5.  The error comes from
# the stored p * l: the marginal depends on the camera sum S / C through
# logit, which multiplies a relative error of S by 1 / (1 - S / C), so 16-bit
# storage does not meet 1e-3 where a voxel's messages are near 1.
MIXED = {
    _lib.MSG_F16: dict(it1=2.5e-3, max5=6e-3, rms5=8e-5),
    _lib.MSG_BF16: dict(it1=2e-2, max5=1.6e-2, rms5=3e-4),
    _lib.MSG_MIXED_F16: dict(it1=2.5e-3, max5=6e-3, rms5=8e-5),
    _lib.MSG_MIXED_BF16: dict(it1=2e-2, max5=1.6e-2, rms5=3e-4),
}


@pytest.mark.parametrize("msg", sorted(MIXED))
def test_bp_mixed_precision_small(msg):
    bound = MIXED[msg]
    p, frames, _ = synth.make_workload("small", msg_dtype=msg)
    errs = {}
    for iters in (1, 5):
        g = gpu_for(p, frames)
        o = oracle_for(p, frames)
        g.perform_inference(max_iters=iters)
        o.perform_inference(iters)
        err = np.abs(g.get_alphas().astype(np.float64) - o.alphas())
        rms = float(np.sqrt((err ** 2).mean()))
        errs[iters] = (float(err.max()), rms)
        print(f"msg {msg} iters {iters}: max-abs {err.max():.3e} rms {rms:.3e}")
    assert errs[1][0] <= bound["it1"]
    assert errs[5][0] <= bound["max5"] and errs[5][1] <= bound["rms5"]


@pytest.mark.parametrize("sched", sorted(_lib.RAY_SCHEDULES))
def test_every_ray_schedule_matches_oracle(sched, monkeypatch):
    """Each ray-kernel schedule (lanes per ray x loads in flight) the
    autotuner may pick gives marginals within 1e-3 of the FP64 oracle."""
    monkeypatch.setenv("MRF_RAY_SCHED", str(sched))
    p, frames, _ = synth.make_workload("small")
    g = gpu_for(p, frames)
    o = oracle_for(p, frames)
    st = g.perform_inference(max_iters=5)
    o.perform_inference(5)
    assert st["ray_sched"] == sched
    err = np.abs(g.get_alphas().astype(np.float64) - o.alphas())
    print(f"schedule {sched} {_lib.RAY_SCHEDULES[sched]}: max-abs {err.max():.3e}")
    assert err.max() <= MARG_TOL


def test_cameras_added_between_calls():
    """Keyframes added after an inference (the viewer's worker loop,
    Viewer.cpp:56-57): buffers grow with the larger camera set and the
    result matches an oracle fed the same call sequence."""
    p, frames, _ = synth.make_workload("small")
    g = gpu_for(p, frames[:1])
    o = oracle_for(p, frames[:1])
    g.perform_inference(max_iters=2)
    o.perform_inference(2)
    for pose, K, depth in frames[1:]:
        g.add_camera_with_depth(pose, depth, K)
        o.add_camera(pose, depth, K)
    g.perform_inference(max_iters=2)
    o.perform_inference(2)
    assert g.num_bricks() == o.num_bricks()
    err = np.abs(g.get_alphas().astype(np.float64) - o.alphas()).max()
    print(f"cameras added between calls: max-abs {err:.3e}")
    assert err <= MARG_TOL


def test_set_pose_then_inference_matches_oracle():
    """set_pose (GVDBInference.cpp:66-82) then perform_inference, the viewer's
    keyframe-pose update (Viewer.cpp:961-966): voxels no other camera reaches
    return to the prior at once, the moved camera is traversed again."""
    p, frames, _ = synth.make_workload("small")
    g = gpu_for(p, frames)
    o = oracle_for(p, frames)
    g.perform_inference(max_iters=2)
    o.perform_inference(2)
    moved = frames[1][0].copy()
    moved[0, 3] += 0.05
    moved[1, 3] -= 0.03
    g.set_pose(1, moved)
    o.set_pose(1, moved)
    assert g.num_bricks() == o.num_bricks()
    err0 = np.abs(g.get_alphas().astype(np.float64) - o.alphas())
    np.testing.assert_array_equal(g.get_alphas() == 0, o.alphas() == 0)
    g.perform_inference(max_iters=2)
    o.perform_inference(2)
    _compare_traversal(g, o, len(frames))
    err = np.abs(g.get_alphas().astype(np.float64) - o.alphas())
    print(f"set_pose: after reset max-abs {err0.max():.3e}, after inference {err.max():.3e}")
    assert err0.max() <= MARG_TOL and err.max() <= MARG_TOL


@pytest.mark.parametrize("shift", [0.0, 0.08])
def test_map_evaluation_matches_oracle(shift):
    """Log-likelihood, expected depth, accuracy and simple accuracy images of
    the map after 3 BP iterations, seen from keyframe 0 (shift 0: the
    camera itself) or 8 cm beside it, against the FP64 oracle.  The map
    itself differs by <= 1e-4 (FP32 BP), so the continuous images agree to
    a few 1e-3 and the threshold classes flip for at most a few pixels."""
    p, frames, _ = synth.make_workload("small")
    g = gpu_for(p, frames)
    o = oracle_for(p, frames)
    g.perform_inference(max_iters=3)
    o.perform_inference(3)
    pose, K, depth = frames[0]
    if shift:
        pose = pose.copy()
        pose[0, 3] += shift
        rg = g.evaluate(pose=pose, depth=depth, K=K)
    else:
        rg = g.evaluate(0)
    ro = o.evaluate(pose, depth, K, acc_thresh=p.acc_thresh)
    dll = np.abs(rg["loglik"].astype(np.float64) - ro["loglik"])
    ded = np.abs(rg["expected_depth"].astype(np.float64) - ro["expected_depth"])
    flips_acc = float(np.mean(rg["accuracy_image"] != ro["accuracy_image"]))
    flips_s = float(np.mean(rg["simple_accuracy_image"] != ro["simple_accuracy_image"]))
    print(f"shift {shift}: loglik max {dll.max():.2e} p99.9 {np.percentile(dll, 99.9):.2e}; expected depth max "
          f"{ded.max():.2e}; class flips {flips_acc:.2e} / {flips_s:.2e}; sums {rg['log_likelihood']:.6g} vs "
          f"{ro['log_likelihood']:.6g}")
    assert np.percentile(dll, 99.9) <= 1e-2 and np.median(dll) <= 1e-4
    assert np.percentile(ded, 99.9) <= 2e-3
    assert flips_acc <= 2e-3 and flips_s <= 2e-3
    assert abs(rg["log_likelihood"] - ro["log_likelihood"]) <= 1e-4 * abs(ro["log_likelihood"]) + 1.0
    assert abs(rg["accuracy"] - ro["accuracy"]) <= 2e-3


def test_loaded_map_evaluation_matches_oracle():
    """An external map (the synthetic room's ground-truth occupancy at the
    voxel centres, 0.95 inside objects and walls) loaded with load_map
    (GVDBMapLikelihoodEstimator.cpp:168-240) and scored from keyframe 1."""
    from mrfmap_amd import GVDBMapLikelihoodEstimator

    p, frames, scene = synth.make_workload("small")
    n = [int(round(d / p.res)) for d in p.dims]
    ii, jj, kk = np.meshgrid(*[np.arange(m) for m in n], indexing="ij")
    vox = np.stack([ii.ravel(), jj.ravel(), kk.ravel()], axis=1).astype(np.float32) + 0.5
    world = (vox - np.array([n[0] / 2, n[1] / 2, 0], np.float32)) * np.float32(p.res)
    occ = scene.occupancy(world)
    pts = vox[occ > 0.5]
    vals = np.full(len(pts), 0.95, np.float32)
    est = GVDBMapLikelihoodEstimator(params=p)
    o = oracle.Oracle(p.K, p.rows, p.cols, p.res, p.prior, p.dims, p.const_sigmasq_poly, p.const_bias_poly,
                      lookup_n=p.lookup_n, activation=p.activation, brick_local_vis=p.brick_local_vis)
    est.load_map(pts, vals)
    o.load_map(pts, vals)
    assert est.num_bricks() == o.num_bricks()
    pose, K, depth = frames[1]
    rg = est.evaluate(pose=pose, depth=depth, K=K)
    ro = o.evaluate(pose, depth, K, acc_thresh=p.acc_thresh)
    dll = np.abs(rg["loglik"].astype(np.float64) - ro["loglik"])
    ded = np.abs(rg["expected_depth"].astype(np.float64) - ro["expected_depth"])
    flips = float(np.mean(rg["simple_accuracy_image"] != ro["simple_accuracy_image"]))
    print(f"loaded map: {len(pts)} points, loglik max {dll.max():.2e}, expected depth max {ded.max():.2e}, "
          f"simple flips {flips:.2e}, accuracy {rg['accuracy']:.4f} vs {ro['accuracy']:.4f}")
    assert np.percentile(dll, 99.9) <= 1e-2 and ded.max() <= 2e-3 and flips <= 2e-3
    assert abs(rg["log_likelihood"] - ro["log_likelihood"]) <= 1e-4 * abs(ro["log_likelihood"]) + 1.0


def test_graph_replays_match_eager_launches(monkeypatch):
    """The BP iteration replayed from a captured HIP graph gives the same bits
    as the same kernels launched one by one, call after call (each call's
    message reset, marginal refresh and residual readback run between the
    replays), also after a keyframe is added (a new capture).  Eager launches
    are selected both through the API and through MRF_NO_GRAPH=1."""
    p, frames, _ = synth.make_workload("small")
    runs = {}
    for mode in ("graph", "timed", "env"):
        if mode == "env":
            monkeypatch.setenv("MRF_NO_GRAPH", "1")
        g = gpu_for(p, frames[:3])
        if mode == "timed":
            g.set_launch_mode(_lib.LAUNCH_TIMED)
        want_graph = 1 if mode == "graph" else 0
        outs = []
        for k in range(3):
            st = g.perform_inference(max_iters=2 + k)
            assert st["graph"] == want_graph
            assert (st["t_ray_kernel_ms"] > 0) == (not want_graph)
            assert st["t_iters_ms"] > 0
            outs.append(g.get_beliefs().copy())
        g.add_camera_with_depth(frames[3][0], frames[3][2], frames[3][1])
        st = g.perform_inference(max_iters=3)
        assert st["graph"] == want_graph
        outs.append(g.get_beliefs().copy())
        runs[mode] = outs
        monkeypatch.delenv("MRF_NO_GRAPH", raising=False)
    for mode in ("timed", "env"):
        for a, b in zip(runs["graph"], runs[mode]):
            np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))


def _edge_frames():
    """A normal keyframe, two without valid depth (NaN, 0), a 1x1 image and a
    75x101 image whose 8x8 tiles overhang the border, each with its own
    intrinsics."""
    p, frames, scene = synth.make_workload("small")
    pose, K, depth = frames[0]
    K2 = synth.intrinsics(120, 160).copy()
    K2[0, 2], K2[1, 2] = 50.0, 37.0
    K1 = np.array([[60, 0, 0.0], [0, 60, 0.0], [0, 0, 1]], np.float32)
    fr = [frames[0], (frames[1][0], K, np.full_like(depth, np.nan)), (frames[2][0], K, np.zeros_like(depth)),
          (frames[3][0], K1, np.full((1, 1), 2.0, np.float32)),
          (frames[1][0], K2, scene.render(frames[1][0], K2, 75, 101))]
    return p, fr


def test_edge_case_keyframes_match_oracle():
    """Empty keyframes, a 1x1 image and ragged tiles: traversal bit-exact per
    camera (the empty ones have no rays), marginals after 3 iterations."""
    p, fr = _edge_frames()
    g = gpu_for(p, fr)
    o = oracle_for(p, fr)
    g.prepare()
    o.prepare()
    assert g.num_bricks() == o.num_bricks()
    for cam in (1, 2):
        rays, recs = g.get_traversal(cam)
        assert len(rays) == 0 and len(recs) == 0
    n = _compare_traversal(g, o, len(fr))
    assert n > 2_000_000
    st = g.perform_inference(max_iters=3)
    so = o.perform_inference(3)
    assert st["updates_per_iter"] == so["updates_per_iter"]
    err = np.abs(g.get_alphas().astype(np.float64) - o.alphas())
    print(f"edge-case keyframes: max-abs marginal error {err.max():.3e}")
    assert err.max() <= MARG_TOL


def test_full_ray_activation_on_tied_rays():
    """Full-ray activation walks the brick grid and sizes each ray's list by a
    bound of its voxel count.  Cameras centred on brick corners with
    axis-aligned rotations: rays along an axis (two axes never move), through
    voxel edges and brick corners at every step (two or three axes tied), at
    half slopes (ties every other step), truncated inside the grid or leaving
    it through a face, and rays that enter the grid from a camera outside it:
    bricks, lists and marginals as the oracle's."""
    p, _, _ = synth.make_workload("small")
    n = 129
    K = np.array([[64, 0, 64.0], [0, 64, 64.0], [0, 0, 1]], np.float32)
    depth = np.full((n, n), 2.0, np.float32)
    depth[:, 96:] = 9.0  # these rays leave the grid
    off = np.array([p.dims[0] / p.res / 2, p.dims[1] / p.res / 2, 0.0])  # world -> voxel offset

    def pose(R, corner):
        T = np.eye(4, dtype=np.float32)
        T[:3, :3] = R
        T[:3, 3] = (np.asarray(corner, np.float64) - off) * p.res
        return T

    Rx = np.array([[1, 0, 0], [0, 0, -1], [0, 1, 0]], np.float32)  # optical axis along -y
    # the third camera sits below the grid: its rays enter through the z = 0
    # face; its depth puts its surface on the first camera's (z = 66 voxels),
    # since contradicting surfaces make the FP32 and FP64 maps drift apart
    # over the iterations (9.7e-4 after three with the first camera's depth)
    depth3 = np.full((n, n), 3.28, np.float32)
    depth3[:, 96:] = 9.0
    frames = [(pose(np.eye(3), (48, 48, 16)), K, depth), (pose(Rx, (40, 64, 24)), K, depth),
              (pose(np.eye(3), (56, 40, -16)), K, depth3)]
    g = gpu_for(p, frames)
    o = oracle_for(p, frames)
    g.prepare()
    o.prepare()
    assert g.num_bricks() == o.num_bricks()
    np.testing.assert_array_equal(g.get_occupied_bricks(), (o.brick_coords() * 8 + 4).astype(np.float32))
    assert _compare_traversal(g, o, len(frames)) > 100_000
    g.perform_inference(max_iters=3)
    o.perform_inference(3)
    err = np.abs(g.get_alphas().astype(np.float64) - o.alphas())
    print(f"tied rays: {g.num_bricks()} bricks, max-abs marginal error {err.max():.3e}")
    assert err.max() <= MARG_TOL


def test_reused_device_blocks_give_the_same_map():
    """Contexts built one after another reuse the large device blocks (lists of
    1.2 GB per 640x480 keyframe at 1 cm) that the previous one freed, stale
    contents included: the maps are bitwise equal, and again after the kept
    blocks are given back to the runtime."""
    p, frames, _ = synth.make_workload("real", frames=2)
    maps = []
    for k in range(3):
        if k == 2:
            _lib.release_device_cache()
        g = gpu_for(p, frames)
        g.perform_inference(max_iters=3)
        maps.append(g.get_alphas().copy())
        g.close()
    assert maps[0].shape[0] > 0
    np.testing.assert_array_equal(maps[1], maps[0])
    np.testing.assert_array_equal(maps[2], maps[0])


def test_empty_map_and_no_cameras():
    """Only keyframes without valid depth: no bricks, empty readback, no
    features, a warm start over no voxels; a context without keyframes runs
    its iterations on an empty map."""
    p, fr = _edge_frames()
    g = gpu_for(p, fr[1:3])
    st = g.perform_inference(max_iters=2)
    assert st["iters"] == 2 and st["updates_per_iter"] == 0 and st["residual"] == 0.0
    assert g.num_bricks() == 0
    assert g.get_alphas().shape == (0, 512)
    assert g.get_occupied_bricks().shape == (0, 3)
    X, vox = g.features()
    assert X.shape == (0, 32) and vox.shape == (0,)
    st = g.perform_inference_warm(np.zeros(_lib.MLP_PARAMS, np.float32), max_iters=3)
    assert st["iters"] == 3 and st["warm_voxels"] == 0
    e = GVDBInference(params=p)
    st = e.perform_inference(max_iters=2)
    assert st["iters"] == 2 and st["updates_per_iter"] == 0


@pytest.mark.parametrize("vis", [_lib.VIS_CARRY, _lib.VIS_BRICK_LOCAL])
def test_large_scene_limit_cycle_matches_oracle(vis):
    """On the 10x10x3 m scene (16 keyframes of the batched generator, 4 cm
    voxels, every 4th pixel row and column) Gauss-Seidel BP converges
    slowly or not at all.  With the carried pre-pass visibility some voxels
    alternate from the 8th iteration on and the residual stays at 0.199 in
    FP64 as in FP32 (a limit cycle); with the reference's brick-local
    visibility the residual falls slowly (0.30 after 12 iterations).  Either
    way the device follows the oracle's residual iteration by iteration and
    ends on the same map."""
    p, frames, _ = synth.make_workload("batched", frames=16, res=0.04)
    p.brick_local_vis = vis
    frames = [(pose, K, decimate(depth, 4)) for pose, K, depth in frames]
    g = gpu_for(p, frames)
    g.perform_inference(max_iters=12)
    rg, _ = g.get_trace()
    o = oracle_for(p, frames)
    ro = [o.perform_inference(1)["residual"]]
    for _ in range(11):
        ro.append(o.continue_inference(1)["residual"])
    print("residual gpu   ", " ".join(f"{x:.4f}" for x in rg))
    print("residual oracle", " ".join(f"{x:.4f}" for x in ro))
    np.testing.assert_allclose(rg, ro, rtol=0, atol=1e-3)
    if vis == _lib.VIS_CARRY:
        assert min(ro[-4:]) > 0.15 and max(ro[-4:]) - min(ro[-4:]) < 1e-3  # the cycle
    err = np.abs(g.get_alphas().astype(np.float64) - o.alphas())
    print(f"limit cycle: max-abs marginal error {err.max():.3e}")
    assert err.max() <= MARG_TOL


class _GpuScene:
    """A scene of bp_loop_cases on the device, through the C ABI."""

    def __init__(self, cams):
        from mrfmap_amd import gvdb_params
        from bp_loop_cases import DIMS, K_UNIT, PRIOR, RES, SIG, camera

        p = gvdb_params()
        p.res, p.prior, p.dims = RES, PRIOR, np.asarray(DIMS, np.float32)
        p.rows = p.cols = 1
        p.K = K_UNIT.reshape(3, 3).copy()
        p.lookup_n, p.const_sigmasq_poly, p.const_bias_poly = 3, [SIG, 0.0, 0.0], [0.0, 0.0, 0.0]
        p.activation = _lib.ACT_FULL_RAY
        self.g = GVDBInference(params=p)
        for x0, d in cams:
            self.g.add_camera_with_depth(*camera(x0, d))

    def perform_inference(self, iters):
        return self.g.perform_inference(max_iters=iters)

    def beliefs(self):
        return self.g.get_beliefs()

    def alphas(self):
        return self.g.get_alphas()

    def num_bricks(self):
        return self.g.num_bricks()


@pytest.mark.parametrize("name", ["one_camera_one_iteration", "cum_length_gate", "gauss_seidel",
                                  "fixed_point_of_one_camera", "second_call_resets_messages"])
def test_bp_loop_by_hand_on_gpu(name):
    """The hand-computed BP-loop cases (camera message logit(clamp(S/C)),
    B += m - M_c, the C > 0.1 gate, Gauss-Seidel order, the per-call reset)
    on the device: beliefs within FP32 rounding of the hand values."""
    from bp_loop_cases import CHECKS

    err = CHECKS[name](_GpuScene)
    print(f"{name}: max-abs error {err:.3e}")
    assert err < 2e-5


@pytest.mark.parametrize("ln", [2, 3])
def test_per_cell_sensor_model_matches_oracle(ln):
    """use_polys tables (GVDBParams.h:134-162) on the small workload: every
    20x20-pixel cell has its own sigma and bias polynomial, looked up in the
    vertically flipped image (gvdbkernels.cu:46-55).  lookup_n = 3 scales the
    quadratic sigma and bias polynomials of coffin_world_640.yaml:33-35 per
    cell; lookup_n = 2 a linear pair.  Traversal (truncation depends on the
    cell's polynomials) bit-exact, marginals after 5 iterations within 1e-3."""
    p, frames, _ = synth.make_workload("small")
    rng = np.random.default_rng(40 + ln)
    cells = p.lookup_cells()
    base_sig = [0.01, -0.00095238, 0.00857143] if ln == 3 else [0.02, 0.012]
    base_bia = [0.02375, -0.0425, 0.04833333] if ln == 3 else [0.005, -0.002]
    p.use_polys, p.lookup_n = True, ln
    p.depth_sigma_lookup = np.outer(rng.uniform(0.7, 1.6, cells), base_sig).astype(np.float32).ravel()
    p.depth_bias_lookup = np.outer(rng.uniform(0.0, 1.5, cells), base_bia).astype(np.float32).ravel()
    g = gpu_for(p, frames)
    o = oracle_for(p, frames)
    g.prepare()
    o.prepare()
    assert g.num_bricks() == o.num_bricks()
    n = _compare_traversal(g, o, len(frames))
    assert n > 500_000
    st = g.perform_inference(max_iters=5)
    so = o.perform_inference(5)
    assert st["updates_per_iter"] == so["updates_per_iter"]
    err = np.abs(g.get_alphas().astype(np.float64) - o.alphas())
    print(f"per-cell sensor model, lookup_n={ln}: {n} records, max-abs marginal error {err.max():.3e}")
    assert err.max() <= MARG_TOL


@pytest.mark.parametrize("vis", [_lib.VIS_CARRY, _lib.VIS_BRICK_LOCAL])
@pytest.mark.parametrize("activation", [_lib.ACT_ENDPOINT, _lib.ACT_FULL_RAY])
def test_reference_frames_match_oracle(activation, vis):
    """The reference's recorded keyframes 0 and 6 (data/depth_k.bin,
    pose_k.bin composed with cam_in_body) under coffin_world_640.yaml's
    parameters (res 0.02, K 381.36, sigma 0.05, prior 0.1, 3 iterations),
    with endpoint activation (the reference's) and full-ray activation, and
    with both pre-pass visibilities: traversal bit-exact, marginals after the
    config's 3 iterations within 1e-3 of the FP64 oracle."""
    from helpers import coffin_frames

    p, frames = coffin_frames(activation, vis)
    g = gpu_for(p, frames)
    o = oracle_for(p, frames)
    g.prepare()
    o.prepare()
    assert g.num_bricks() == o.num_bricks()
    n = _compare_traversal(g, o, len(frames))
    assert n > (5_000_000 if activation == _lib.ACT_ENDPOINT else 50_000_000)
    st = g.perform_inference(max_iters=p.max_iters)
    so = o.perform_inference(p.max_iters)
    assert st["updates_per_iter"] == so["updates_per_iter"]
    a = o.alphas()
    err = np.abs(g.get_alphas().astype(np.float64) - a)
    print(f"reference frames, activation {activation}, vis {vis}: {n} records, {g.num_bricks()} bricks, "
          f"max marginal {a.max():.3f}, max-abs marginal error {err.max():.3e}")
    assert err.max() <= MARG_TOL


@pytest.mark.parametrize("vis", [_lib.VIS_BRICK_LOCAL, _lib.VIS_CARRY])
@pytest.mark.parametrize("activation", [_lib.ACT_FULL_RAY, _lib.ACT_ENDPOINT])
def test_pre_pass_visibility_small_matches_oracle(activation, vis):
    """Both pre-pass visibilities on the small workload -- the reference's
    (restarted at every brick entry until a tau survives FP32 flush-to-zero,
    gvdbkernels.cu:708-715; the default) and the carried one: after one
    iteration within 1e-5 of the oracle, after five more within 1e-3; the
    two maps differ."""
    p, frames, _ = synth.make_workload("small", activation=activation)
    p.brick_local_vis = vis
    g = gpu_for(p, frames)
    o = oracle_for(p, frames)
    g.perform_inference(max_iters=1)
    o.perform_inference(1)
    e1 = np.abs(g.get_alphas().astype(np.float64) - o.alphas()).max()
    g.perform_inference(max_iters=5)
    o.perform_inference(5)
    a = o.alphas()
    e5 = np.abs(g.get_alphas().astype(np.float64) - a).max()
    pc = p.copy()
    pc.brick_local_vis = _lib.VIS_CARRY if vis == _lib.VIS_BRICK_LOCAL else _lib.VIS_BRICK_LOCAL
    oc = oracle_for(pc, frames)
    oc.perform_inference(1)
    oc.perform_inference(5)
    d = np.abs(oc.alphas() - a).max()
    print(f"pre-pass visibility {vis}, activation {activation}: 1 iteration {e1:.3e}, 1+5 iterations {e5:.3e}; "
          f"the two visibilities' maps differ by {d:.3f}")
    assert e1 <= 1e-5
    assert e5 <= MARG_TOL
    assert d > 1e-2


@pytest.mark.timeout(600)
def test_converged_map_matches_fp64_anchor():
    """The FP64 anchor (tests/golden/c2_anchor.npz, written by
    tests/golden/make_c2_anchor.py): the 640x480 / 1 cm / 13-keyframe
    workload with every 3rd pixel row and column, iterated by the FP64
    oracle until the max marginal change of an iteration fell below 1e-6.
    The GPU's FP32 BP, iterated from the prior until its own residual is
    below 1e-5 (or 400 iterations), must land within 1e-3 of that map at
    every voxel the anchor stores (marginal above its threshold, rounded to
    1/65535) and below the anchor's bound plus 1e-3 everywhere else."""
    import json
    import os

    from helpers import c2_anchor_workload

    z = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "c2_anchor.npz"))
    meta = json.loads(str(z["meta"]))
    p, frames = c2_anchor_workload(meta["step"])
    assert int(p.brick_local_vis) == meta["brick_local_vis"]
    g = gpu_for(p, frames)
    g.prepare()
    assert g.num_bricks() == meta["num_bricks"]
    np.testing.assert_array_equal(g.get_occupied_bricks(), (z["bricks"].astype(np.float32) * 8 + 4))
    st = g.perform_inference(max_iters=400, tol=1e-5)
    a = g.get_alphas().astype(np.float64).ravel()
    idx = np.cumsum(z["deltas"].astype(np.int64))
    ref = z["vals"].astype(np.float64) / 65535.0
    err = np.abs(a[idx] - ref)
    rest = np.ones(a.size, bool)
    rest[idx] = False
    print(f"FP64 anchor (step {meta['step']}, {meta['iters']} FP64 iterations, residual {meta['residual']:.1e}): "
          f"GPU {st['iters']} iterations, residual {st['residual']:.1e}; {len(idx)} stored voxels max-abs "
          f"{err.max():.3e} rms {np.sqrt((err ** 2).mean()):.3e}; others max {a[rest].max():.4f} "
          f"(anchor bound {meta['rest_max']:.4f})")
    assert err.max() <= MARG_TOL
    assert a[rest].max() <= meta["rest_max"] + MARG_TOL
