"""Pins the oracle's plane-sweep cost (motion_stereo/calc_cost.cu:11-210) on the CPU."""
import json
import os

import numpy as np
import pytest

import oracle
from oracle.sequence import homography

HERE = os.path.dirname(os.path.abspath(__file__))


def _load(name):
    with open(os.path.join(HERE, "golden", name)) as f:
        return json.load(f)


@pytest.mark.parametrize("case", range(2))
@pytest.mark.parametrize("sources", [1, 2])
def test_identity_pose_closed_form(case, sources):
    """H = I, t = 0, measurement == reference: the hand-derived closed form, including the
    garbled right / up-right warps and the swapped up/down reference rows."""
    c = _load("cost_identity.json")[case]
    img = np.array(c["image"], np.uint8)
    hom = (np.eye(3, dtype=np.float32), np.zeros(3, np.float32))
    got = oracle.calc_cost(img, [img] * sources, [hom] * sources, c["idep_min"],
                           c["dep_sample"], c["planes"], "f64")
    want = np.array(c["cost"])
    for d in range(c["planes"]):
        np.testing.assert_allclose(got[:, :, d], want, rtol=1e-13, atol=1e-13)


def test_float_instantiation_agrees(small_scene):
    s = small_scene
    homs = [homography(s["scene"].K, *s["ref_pose"], R, T) for R, T in s["src_poses"]]
    ds = (s["idep_max"] - s["idep_min"]) / (s["planes"] - 1)
    c64 = oracle.calc_cost(s["ref"], s["srcs"], homs, s["idep_min"], ds, s["planes"], "f64")
    c32 = oracle.calc_cost(s["ref"], s["srcs"], homs, s["idep_min"], ds, s["planes"], "f32")
    both = (c64 >= 0) & (c32 >= 0)
    assert ((c64 >= 0) != (c32 >= 0)).mean() < 1e-4
    assert np.quantile(np.abs(c64[both] - c32[both]) / np.maximum(c64[both], 1), 0.999) < 1e-4


def test_source_mean_is_linear(small_scene):
    """The cost over {a, b} is the mean of the single-source costs where both are valid."""
    s = small_scene
    homs = [homography(s["scene"].K, *s["ref_pose"], R, T) for R, T in s["src_poses"]]
    ds = (s["idep_max"] - s["idep_min"]) / (s["planes"] - 1)
    args = (s["idep_min"], ds, s["planes"], "f64")
    ab = oracle.calc_cost(s["ref"], s["srcs"], homs, *args)
    a = oracle.calc_cost(s["ref"], s["srcs"][:1], homs[:1], *args)
    b = oracle.calc_cost(s["ref"], s["srcs"][1:], homs[1:], *args)
    both = (a >= 0) & (b >= 0)
    np.testing.assert_allclose(ab[both], 0.5 * (a[both] + b[both]), rtol=1e-12)
    only_a = (a >= 0) & (b < 0)
    np.testing.assert_allclose(ab[only_a], a[only_a], rtol=1e-12)
    assert np.all(ab[(a < 0) & (b < 0)] == -1)


def test_behind_camera_and_no_sources(small_scene):
    s = small_scene
    flip = (np.diag([1.0, 1.0, -1.0]).astype(np.float32), np.zeros(3, np.float32))
    c = oracle.calc_cost(s["ref"], s["srcs"][:1], [flip], 0.2, 0.025, 6, "f64")
    assert np.all(c == -1)
    c = oracle.calc_cost(s["ref"], [], [], 0.2, 0.025, 6, "f64")
    assert np.all(c == -1)


def test_true_depth_plane_has_lowest_regular_terms(small_scene):
    """Warping the reference pixel at its true inverse depth lands on the analytic projection
    of the 3-D point, so a fronto-parallel plane rendered at depth 1/idep_k is matched best at
    plane k (pins the homography sign conventions of stereo_mapper.cpp:79-80)."""
    from probabilistic_mapping_amd import synth
    K = np.array([[200.0, 0, 48.0], [0, 200.0, 32.0], [0, 0, 1.0]])
    D, idep_min, ds = 9, 0.3, 0.05
    k_true = 4
    z = 1.0 / (idep_min + k_true * ds)
    patch = synth._patch([0, 0, z], [1, 0, 0], [0, 1, 0], 0, 0, 7, freq=12.0, amp=170)
    sc = synth.Scene(96, 64, K, np.array([patch]))
    ref, _ = sc.render(np.eye(3), np.zeros(3))
    R = synth.axis_angle([0.2, 1.0, 0.1], np.radians(1.0))
    T = np.array([0.06, 0.02, 0.0])
    src, _ = sc.render(R, T)
    hom = homography(K, np.eye(3), np.zeros(3), R, T)
    c = oracle.calc_cost(ref, [src], [hom], idep_min, ds, D, "f64")
    inner = c[8:-8, 12:-12]
    valid = inner.min(axis=2) >= 0
    best = np.argmin(inner, axis=2)[valid]
    assert valid.mean() > 0.5
    # the sampled texture is itself bilinear-resampled u8, so neighbouring planes can tie on
    # smooth texture; a sign or convention error would put the minimum far away instead
    assert (best == k_true).mean() > 0.75
    assert (np.abs(best - k_true) <= 1).mean() > 0.9


def test_validity_margin_identity_closed_form():
    """H = I, t = 0 puts the nine warps of pixel (u, v) at fixed points for every plane:
    (u, v), (u, v -+ 1), (u -+ 1, v), (u -+ 1, v -+ 1), (u +- 1, v -+ 1) and the two garbled
    triples (calc_cost.cu:50-51, 70-71) at (1 + 1/u, 1) and (1, u / (u + 1)).  The margin is
    the smallest distance of any of them to the bounds [0, W-1] x [0, H-1], in absolute value;
    column 0 has w = 0 in the garbled warps and no margin."""
    W, H, D = 9, 7, 3
    img = np.random.default_rng(2).integers(0, 256, (H, W), dtype=np.uint8)
    hom = (np.eye(3, dtype=np.float32), np.zeros(3, np.float32))
    cost, mg = oracle.calc_cost(img, [img], [hom], 0.2, 0.1, D, "f64", margin=True)
    for v in range(H):
        for u in range(W):
            if u == 0:
                assert np.all(cost[v, u] == -1)
                continue
            pts = [(u, v), (u, v - 1), (u, v + 1), (u - 1, v), (1 + 1 / u, 1.0),
                   (u - 1, v - 1), (u + 1, v + 1), (u - 1, v + 1), (1.0, u / (u + 1))]
            m = min(min(x, W - 1 - x, y, H - 1 - y) for x, y in pts)
            np.testing.assert_allclose(mg[v, u], abs(m), rtol=1e-6, atol=1e-7)
            assert np.all((cost[v, u] >= 0) == (m >= 0))


def test_validity_margin_bounds_source_set_changes():
    """Moving every warp by a tiny amount (the translation scaled by 1 + 1e-5, i.e. well under
    2e-3 px here) can change an entry's validity, or its set of contributing sources (a jump of
    the mean far beyond the smooth change, which is below 0.6 cost units for such a move),
    only where the margin is below that move."""
    from probabilistic_mapping_amd import synth
    s = synth.trajectory_scene(2, width=320, height=200, n_frames=7, n_sources=6, planes=32,
                               fx=200.0)
    K = s["scene"].K
    i = 3
    srcs = s["sources"][i]
    homs = [homography(K, *s["poses"][i], *s["poses"][j]) for j in srcs]
    moved = [(h, (np.asarray(t, np.float64) * (1 + 1e-5))) for h, t in homs]
    ds = (s["idep_max"] - s["idep_min"]) / 31
    imgs = [s["images"][j] for j in srcs]
    args = (s["idep_min"], ds, 32, "f64")
    c0, mg = oracle.calc_cost(s["images"][i], imgs, homs, *args, margin=True)
    c1 = oracle.calc_cost(s["images"][i], imgs, moved, *args)
    # the largest warp move: |dU| = |dt1 idep| / w over every source, plane and pixel
    flip = (c0 >= 0) != (c1 >= 0)
    both = (c0 >= 0) & (c1 >= 0)
    jump = np.zeros_like(flip)
    jump[both] = np.abs(c0[both] - c1[both]) > 0.6
    assert flip.sum() + jump.sum() >= 5          # the sample has changes to explain
    assert np.all(mg[flip | jump] < 2e-3)
    assert (mg < 2e-3).mean() < 1e-3
