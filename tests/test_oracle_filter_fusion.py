"""Pins the oracle's filter (motion_stereo/filter_cost.cu:3-55), median, and Gaussian x Beta
fusion (motion_stereo/depth_fusion.cu:8-225) on the CPU."""
import json
import os

import numpy as np
import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))


def _load(name):
    with open(os.path.join(HERE, "golden", name)) as f:
        return json.load(f)


def test_filter_golden():
    g = _load("filter_parabola.json")
    agg = np.array([c["agg"] for c in g["cases"]], np.float64)[None]   # [1][n][D]
    got = oracle.filter_cost(agg, g["var_scale"], g["idep_min"], g["dep_sample"], "f64")[0]
    for c, v in zip(g["cases"], got):
        assert v == pytest.approx(c["expect"], rel=1e-12, abs=0), c["rule"]


def test_filter_subpixel_offset_bounded():
    """Rule 3 keeps the curvature positive, so the fitted offset stays within half a plane."""
    rng = np.random.default_rng(1)
    agg = rng.uniform(1, 100, size=(40, 50, 24))
    idep_min, ds = 0.1, 0.05
    d = oracle.filter_cost(agg, 1.02, idep_min, ds, "f64")
    m = np.argmin(agg, axis=2)
    ok = (d > 0) & (d < 1000)
    s = (1.0 / d[ok] - idep_min) / ds
    assert np.all(np.abs(s - m[ok]) <= 0.5 + 1e-9)


def test_median_brute_force():
    rng = np.random.default_rng(7)
    d = rng.uniform(0.5, 5.0, size=(13, 17))
    d[rng.random(d.shape) < 0.25] = 0.0
    d[rng.random(d.shape) < 0.1] = 1000.0
    got = oracle.median3x3(d, "f64")
    H, W = d.shape
    for y in range(H):
        for x in range(W):
            if not (0 < d[y, x] < 1000):
                assert got[y, x] == d[y, x]
                continue
            w = d[max(y - 1, 0):y + 2, max(x - 1, 0):x + 2].ravel()
            w = np.sort(w[(w > 0) & (w < 1000)])
            assert got[y, x] == w[(len(w) - 1) // 2]


def _state(H, W, a=0.0, b=0.0, mu=0.0, sig=0.0):
    return {"alpha": np.full((H, W), a), "beta": np.full((H, W), b),
            "mu": np.full((H, W), mu), "sig": np.full((H, W), sig)}


CAM = (100.0, 100.0, 3.0, 2.0)
EYE = (np.eye(3, dtype=np.float32), np.zeros(3, np.float32))


@pytest.mark.parametrize("case", range(5))
def test_update_matches_posterior_moments(case):
    """depthUpdate's moment matching against brute-force quadrature of the posterior."""
    c = _load("fusion_posterior.json")[case]
    a, b, mu, s2 = c["prior"]
    H, W = 5, 6
    # identity pose: the transform keeps the depth and adds 0.05 mu to the variance
    pre = _state(H, W, a, b, mu, s2 - 0.05 * mu)
    cur = _state(H, W)
    depth = np.full((H, W), c["measurement"])
    oracle.depth_fuse(depth, pre, cur, *EYE, CAM, "f64")
    p = c["posterior"]
    y, x = 2, 3
    assert cur["mu"][y, x] == pytest.approx(p["mu"], rel=1e-7)
    assert cur["sig"][y, x] == pytest.approx(p["sig"], rel=1e-6)
    assert cur["alpha"][y, x] == pytest.approx(p["alpha"], rel=1e-6)
    assert cur["beta"][y, x] == pytest.approx(p["beta"], rel=1e-6)
    # output gate on the PRIOR inlier ratio (depth_fusion.cu:219)
    expect = p["mu"] if a / (a + b) > 0.6 else 0.0
    assert depth[y, x] == pytest.approx(expect, rel=1e-7)


def test_identity_pose_table_and_variance_growth():
    H, W = 6, 7
    rng = np.random.default_rng(0)
    mu = rng.uniform(1, 4, size=(H, W))
    pre = {"alpha": np.full((H, W), 10.0), "beta": np.full((H, W), 5.0), "mu": mu.copy(),
           "sig": np.full((H, W), 0.2)}
    cur = _state(H, W)
    depth = np.zeros((H, W))          # no measurement: only prediction and the gated output
    table = oracle.depth_fuse(depth, pre, cur, *EYE, CAM, "f64")
    idx = np.arange(H * W).reshape(H, W)
    # every pixel maps to itself, except pixel 0 (the empty sentinel) which is filled from
    # its only interior neighbour (1, 1)
    assert table[0, 0] == W + 1
    assert np.array_equal(table.ravel()[1:], idx.ravel()[1:])
    np.testing.assert_allclose(pre["sig"], 0.2 + 0.05 * mu, rtol=1e-15)
    np.testing.assert_allclose(cur["mu"].ravel()[1:], mu.ravel()[1:], rtol=0)
    # no measurement and prior ratio 10/15 > 0.6: the output is the predicted mean
    np.testing.assert_allclose(depth.ravel()[1:], mu.ravel()[1:], rtol=0)


def test_hole_filling_golden():
    g = _load("hole_fill.json")
    W, H = g["width"], g["height"]
    t = np.array(g["table"])
    pre = _state(H, W)
    src = t != 0
    pre["alpha"][src], pre["beta"][src], pre["mu"][src], pre["sig"][src] = 10, 10, 2.0, 1.0
    cur = _state(H, W)
    depth = np.zeros((H, W))
    table = oracle.depth_fuse(depth, pre, cur, *EYE, CAM, "f64")
    assert np.array_equal(table, np.array(g["filled"]))


def test_zbuffer_keeps_nearest_and_lowest_index():
    """Two sources landing on one target: the nearer transformed depth wins; on a tie the
    lower source index wins."""
    H, W = 4, 8
    cam = (1.0, 1.0, 0.0, 0.0)    # x2 = trunc(px2/pz2 + 0.5) with px = x * depth
    pre = _state(H, W)
    # with a pure x translation t and depth z, source x lands on x + t/z
    # sources (1,1) at z=1 and (2,1) at z=2: t = 2 moves them to 3 and 3
    for x, z in [(1, 1.0), (2, 2.0)]:
        pre["alpha"][1, x], pre["beta"][1, x], pre["mu"][1, x] = 10, 10, z
    r = np.eye(3, dtype=np.float32)
    t = np.array([2.0, 0.0, 0.0], np.float32)
    table = oracle.depth_fuse(np.zeros((H, W)), pre, _state(H, W), r, t, cam, "f64")
    assert table[1, 3] == 1 * W + 1
    # equal depths: both land on x = 4 with t = 2 from x = 2 at z = 1 ... use two rows
    pre = _state(H, W)
    for y in (1, 2):
        pre["alpha"][y, 2], pre["beta"][y, 2], pre["mu"][y, 2] = 10, 10, 1.0
    r = np.array([[1, 0, 0], [0, 0, 0], [0, 0, 1]], np.float32)   # flattens y to 0
    table = oracle.depth_fuse(np.zeros((H, W)), pre, _state(H, W), r,
                              np.zeros(3, np.float32), cam, "f64")
    assert table[0, 2] == 1 * W + 2


def test_gates():
    """alpha < FLOAT_EPS or inlier ratio < 0.4 never propagates (depth_fusion.cu:26-32)."""
    H, W = 5, 5
    pre = _state(H, W, 10.0, 20.0, 2.0, 1.0)      # ratio 1/3 < 0.4
    table = oracle.depth_fuse(np.zeros((H, W)), pre, _state(H, W), *EYE, CAM, "f64")
    assert np.all(table == 0)
    cur = _state(H, W)
    depth = np.full((H, W), 1.5)
    oracle.depth_fuse(depth, _state(H, W), cur, *EYE, CAM, "f64")
    # nothing predicted: every pixel initialises to (10, 10, d, 2500) and outputs d
    assert np.all(cur["alpha"] == 10) and np.all(cur["beta"] == 10)
    assert np.all(cur["mu"] == 1.5) and np.all(cur["sig"] == 2500) and np.all(depth == 1.5)


# ---- the CPU rules of outputFusionCPU (CASE >= 3), stereo_mapper.cpp:246-587, 719-800 ------

def _enc(y, x):
    return (y << 16) | x


@pytest.mark.parametrize("case", range(5))
def test_cpu_rules_update_matches_posterior_moments(case):
    """depthUpdateCPU's moment matching against brute-force quadrature with the reference's own
    weights (tau^2 = 64, outlier density 1/16, normpdfCPU's sqrt factor); output gated on the
    POSTERIOR inlier ratio > 0.51."""
    c = _load("fusion_posterior_cpu.json")[case]
    a, b, mu, s2 = c["prior"]
    H, W = 9, 9
    pre = _state(H, W, a, b, mu, s2 - 0.05 * mu)    # identity pose adds 0.05 mu
    cur = _state(H, W)
    depth = np.full((H, W), c["measurement"])
    oracle.depth_fuse_cpu_rules(depth, pre, cur, *EYE, CAM, "f64")
    p = c["posterior"]
    y, x = 4, 4
    assert cur["mu"][y, x] == pytest.approx(p["mu"], rel=1e-7)
    assert cur["sig"][y, x] == pytest.approx(p["sig"], rel=1e-6)
    assert cur["alpha"][y, x] == pytest.approx(p["alpha"], rel=1e-6)
    assert cur["beta"][y, x] == pytest.approx(p["beta"], rel=1e-6)
    expect = p["mu"] if p["inlier"] > 0.51 else 0.0
    assert depth[y, x] == pytest.approx(expect, rel=1e-7)


def test_cpu_rules_identity_pose_table():
    """Every eligible pixel maps to itself, encoded (y << 16) | x; pixel (0, 0) encodes to the
    empty sentinel and is filled from the first interior source of its 7x7 window, (3, 3)."""
    H, W = 9, 11
    rng = np.random.default_rng(1)
    mu = rng.uniform(1, 4, size=(H, W))
    pre = {"alpha": np.full((H, W), 10.0), "beta": np.full((H, W), 5.0), "mu": mu.copy(),
           "sig": np.full((H, W), 0.2)}
    cur = _state(H, W)
    depth = np.zeros((H, W))
    table = oracle.depth_fuse_cpu_rules(depth, pre, cur, *EYE, CAM, "f64")
    ys, xs = np.mgrid[0:H, 0:W]
    want = (ys << 16) | xs
    want[0, 0] = _enc(3, 3)
    assert np.array_equal(table, want)
    np.testing.assert_allclose(pre["sig"], 0.2 + 0.05 * mu, rtol=1e-15)
    # no measurement: beta + 1, output the predicted mean if 10 / 16 > 0.51, then >= 0.1 kept
    assert np.all(cur["beta"].ravel()[1:] == 6.0)
    np.testing.assert_allclose(depth.ravel()[1:], mu.ravel()[1:], rtol=0)


def test_cpu_rules_zbuffer_compares_transformed_with_original_depth():
    """transformMapCPU keeps the first arrival unless its TRANSFORMED depth exceeds the
    newcomer's ORIGINAL depth (stereo_mapper.cpp:315-325).  A forward translation t_z = 1 at
    depth 1 halves offsets from the principal point: x = 4 and x = 5 (cx = 3) both land on
    x = 4; the first (x = 4) has transformed depth 2 > 1, so x = 5 takes the cell, where the
    GPU rules (nearest transformed depth, then lowest index) keep x = 4."""
    H, W = 9, 9
    cam = (1.0, 1.0, 3.0, 4.0)
    pre = _state(H, W)
    for x in (4, 5):
        pre["alpha"][4, x], pre["beta"][4, x], pre["mu"][4, x] = 10, 1, 1.0
    t = np.array([0.0, 0.0, 1.0], np.float32)
    table = oracle.depth_fuse_cpu_rules(np.zeros((H, W)), pre, _state(H, W),
                                        np.eye(3, dtype=np.float32), t, cam, "f64")
    assert table[4, 4] == _enc(4, 5)
    assert pre["mu"][4, 4] == 1.0      # the displaced source's mean is overwritten (:322)
    pre2 = _state(H, W)
    for x in (4, 5):
        pre2["alpha"][4, x], pre2["beta"][4, x], pre2["mu"][4, x] = 10, 1, 1.0
    gpu = oracle.depth_fuse(np.zeros((H, W)), pre2, _state(H, W),
                            np.eye(3, dtype=np.float32), t, cam, "f64")
    assert gpu[4, 4] == 4 * W + 4


def test_cpu_rules_hole_fill_first_source_in_raster_order():
    """holeFillingCPU (7x7, in place with visited marks): an empty cell takes the entry of the
    first non-empty cell, in raster order, among the interior sources (3 <= y < H-3,
    3 <= x < W-3) of its 7x7 window; the rule is checked here by direct construction."""
    H, W = 12, 14
    src = {(3, 4): 2.0, (5, 8): 3.0, (8, 3): 4.0, (9, 10): 5.0, (1, 1): 6.0}   # (1,1): not interior
    pre = _state(H, W)
    for (y, x), z in src.items():
        pre["alpha"][y, x], pre["beta"][y, x], pre["mu"][y, x], pre["sig"][y, x] = 10, 1, z, 1
    table = oracle.depth_fuse_cpu_rules(np.zeros((H, W)), pre, _state(H, W), *EYE,
                                        (100.0, 100.0, 6.0, 5.0), "f64")
    orig = {(y, x) for (y, x) in src}
    want = np.zeros((H, W), np.int64)
    for (y, x) in orig:
        want[y, x] = _enc(y, x)
    for y in range(H):
        for x in range(W):
            if (y, x) in orig:
                continue
            for sy in range(y - 3, y + 4):
                hit = [(sy, sx) for sx in range(x - 3, x + 4)
                       if (sy, sx) in orig and 3 <= sy < H - 3 and 3 <= sx < W - 3]
                if hit:
                    want[y, x] = _enc(*hit[0])
                    break
    assert np.array_equal(table, want)


def test_cpu_rules_clamps_gates_and_init():
    """Measurements above 100 are dropped and their output masked, above 10 dropped; pixels
    with an inlier ratio below 0.5 never propagate; a new pixel starts at (10, 10, d, 256)
    with output 0; outputs below 0.1 become 0."""
    H, W = 9, 9
    pre = _state(H, W, 10.0, 10.5, 2.0, 1.0)      # ratio < 0.5: nothing propagates
    cur = _state(H, W)
    depth = np.full((H, W), 1.5)
    depth[2, 2], depth[3, 3], depth[4, 4] = 150.0, 20.0, 0.05
    table = oracle.depth_fuse_cpu_rules(depth, pre, cur, *EYE, CAM, "f64")
    assert np.all(table == 0)
    assert cur["alpha"][0, 0] == 10 and cur["beta"][0, 0] == 10
    assert cur["mu"][0, 0] == 1.5 and cur["sig"][0, 0] == 256
    assert np.all(depth == 0)                      # a new pixel outputs 0
    assert cur["alpha"][2, 2] == 0 and cur["mu"][2, 2] == 0   # dropped measurement: mu = 0 only
    assert cur["alpha"][4, 4] == 10 and cur["mu"][4, 4] == 0.05
    # second step: identity pose, every pixel now propagates (ratio 0.5)
    depth2 = np.full((H, W), 1.5)
    depth2[2, 2] = 150.0
    pre2, cur2 = cur, _state(H, W)
    oracle.depth_fuse_cpu_rules(depth2, pre2, cur2, *EYE, CAM, "f64")
    assert depth2[2, 2] == 0.0                     # masked
    assert depth2[5, 5] > 0.1
