"""Pins the oracle's semi-global aggregation (motion_stereo/sgm.cu:5-135) on the CPU.

The brute force below is the textbook SGM formulation -- L(p,d) = C(p,d) +
min_d' (L(q,d') + pen(|d-d'|)) - min_d' L(q,d') with pen(0)=0, pen(1)=P1, pen(>=2)=P2, walked
pixel by pixel from every entry pixel of each direction -- not the reference's
min(L, L+-1 + P1, m + P2) expression; the two agree whenever P1 <= P2.
"""
import itertools

import numpy as np
import pytest

import oracle

DIRS = [(1, 0), (-1, 0), (0, 1), (0, -1), (1, 1), (-1, -1), (1, -1), (-1, 1)]


def brute_sgm(ref, cost, kn, paths):
    H, W, D = cost.shape
    invalid = cost.min(axis=2) < 0
    cp = np.where(invalid[..., None], 0.0, cost)
    agg = np.zeros_like(cost)
    for dx, dy in DIRS[:paths]:
        starts = [(x, y) for y in range(H) for x in range(W)
                  if not (0 <= x - dx < W and 0 <= y - dy < H)]
        for x0, y0 in starts:
            prev = None
            x, y = x0, y0
            while 0 <= x < W and 0 <= y < H:
                if prev is None:
                    L = cp[y, x].copy()
                else:
                    qx, qy = x - dx, y - dy
                    g = abs(float(ref[y, x]) - float(ref[qy, qx]))
                    P1, P2 = kn["pi1"], kn["pi2"]
                    if g < kn["tau_so"]:
                        P1, P2 = P1 * kn["sgm_q1"], P2 * kn["sgm_q2"]
                    m = prev.min()
                    L = np.empty(D)
                    for d in range(D):
                        best = np.inf
                        for dd in range(D):
                            pen = 0.0 if dd == d else (P1 if abs(dd - d) == 1 else P2)
                            best = min(best, prev[dd] + pen)
                        L[d] = cp[y, x, d] + best - m
                agg[y, x] = 0.0 if invalid[y, x] else agg[y, x] + L
                prev = L
                x, y = x + dx, y + dy
    return agg


KN = dict(pi1=1.5, pi2=4.0, tau_so=30.0, sgm_q1=2.0, sgm_q2=3.0, var_scale=1.02)


@pytest.mark.parametrize("paths", [4, 8])
@pytest.mark.parametrize("shape", [(5, 6, 5), (4, 3, 7), (1, 5, 4)])
def test_brute_force(paths, shape):
    rng = np.random.default_rng(sum(shape) * paths)
    H, W, D = shape
    ref = rng.integers(0, 256, size=(H, W), dtype=np.uint8)
    cost = rng.uniform(0, 20, size=shape)
    cost[rng.random((H, W)) < 0.15, rng.integers(0, D)] = -1.0   # some invalid pixels
    got = oracle.sgm(ref, cost, KN, paths, "f64")
    want = brute_sgm(ref, cost, KN, paths)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("paths", [4, 8])
def test_zero_penalties_give_path_count_times_cost(paths):
    rng = np.random.default_rng(5)
    cost = rng.uniform(0, 50, size=(9, 11, 6))
    ref = rng.integers(0, 256, size=(9, 11), dtype=np.uint8)
    kn = dict(KN, pi1=0.0, pi2=0.0)
    # (c + m) - m rounds at 1 ulp of the running minimum m
    np.testing.assert_allclose(oracle.sgm(ref, cost, kn, paths, "f64"), paths * cost,
                               rtol=1e-13, atol=1e-12)


def test_constant_cost_is_a_fixed_point():
    """With C = c everywhere each path's L stays c (m = c, val = c), so agg = P c."""
    ref = np.random.default_rng(2).integers(0, 256, size=(7, 8), dtype=np.uint8)
    cost = np.full((7, 8, 5), 3.25)
    for paths in (4, 8):
        np.testing.assert_allclose(oracle.sgm(ref, cost, KN, paths, "f64"), paths * cost,
                                   rtol=1e-14)


def test_invalid_pixels_aggregate_to_zero():
    rng = np.random.default_rng(9)
    cost = rng.uniform(0, 10, size=(6, 6, 4))
    cost[2, 3, 1] = -1
    cost[0, 0, 3] = -1
    agg = oracle.sgm(np.zeros((6, 6), np.uint8), cost, KN, 8, "f64")
    assert np.all(agg[2, 3] == 0) and np.all(agg[0, 0] == 0)
    assert np.all(agg[1, 1] > 0)


def test_float_instantiation_agrees():
    rng = np.random.default_rng(4)
    cost = rng.uniform(0, 40, size=(24, 31, 16))
    ref = rng.integers(0, 256, size=(24, 31), dtype=np.uint8)
    a = oracle.sgm(ref, cost, KN, 8, "f64")
    b = oracle.sgm(ref, cost.astype(np.float32), KN, 8, "f32")
    assert np.max(np.abs(a - b) / np.maximum(a, 1)) < 1e-5


def test_scanline_enumeration_covers_every_pixel_once():
    """Each direction's scanlines partition the image (the coverage the GPU launch relies on)."""
    for (W, H) in [(7, 5), (1, 4), (6, 1), (3, 3)]:
        for dx, dy in DIRS:
            seen = np.zeros((H, W), int)
            for x0, y0 in itertools.product(range(W), range(H)):
                if 0 <= x0 - dx < W and 0 <= y0 - dy < H:
                    continue
                x, y = x0, y0
                while 0 <= x < W and 0 <= y < H:
                    seen[y, x] += 1
                    x, y = x + dx, y + dy
            assert np.all(seen == 1)
            n_starts = sum(1 for x0, y0 in itertools.product(range(W), range(H))
                           if not (0 <= x0 - dx < W and 0 <= y0 - dy < H))
            expect = W + H - 1 if (dx and dy) else (H if dx else W)
            assert n_starts == expect


@pytest.mark.parametrize("shape,paths", [((13, 17, 9), 8), ((7, 11, 64), 8), ((9, 5, 5), 4)])
def test_float32_scanline_replay_sums_to_oracle(shape, paths):
    """The float32 scanline replay that the GPU tests compare stored path volumes with
    (``parity.replay_scanline``), run over every scanline of every direction and summed per
    pixel in path order, equals the float32 oracle's aggregate bit for bit and the FP64
    oracle's to float rounding; invalid pixels are 0.  Each scanline of a direction is visited
    once (``parity.scanline`` covers every pixel exactly once), and the per-pixel edge switch of
    P1/P2 is exercised by a reference image with both small and large steps."""
    from tests import parity
    rng = np.random.default_rng(shape[0] * 100 + shape[2] + paths)
    H, W, D = shape
    ref = rng.integers(0, 12, size=(H, W), dtype=np.uint8) * np.uint8(rng.integers(1, 4))
    ref[rng.random((H, W)) < 0.3] = 200
    cost = rng.uniform(0.0, 40.0, size=(H, W, D)).astype(np.float32)
    cost[rng.random((H, W)) < 0.1, 0] = -1.0
    kn = dict(pi1=3.0, pi2=11.0, tau_so=4.0, sgm_q1=2.0, sgm_q2=3.0, var_scale=1.02)
    invalid = cost.min(axis=2) < 0
    total = np.zeros((H, W, D), np.float32)
    step = parity.excess_step(kn)
    for p in range(paths):
        dx, dy = parity.PATH_DX[p], parity.PATH_DY[p]
        nlines = W + H - 1 if dx and dy else (H if dx else W)
        seen = np.zeros((H, W), np.int32)
        for i in range(nlines):
            pts = parity.scanline(W, H, dx, dy, i)
            _, lv = parity.replay_scanline(cost, ref, pts, dx, dy, kn, step, True)
            for k, (x, y) in enumerate(pts):
                total[y, x] = (total[y, x] + lv[k]).astype(np.float32)
                seen[y, x] += 1
        assert np.all(seen == 1)
    want32 = oracle.sgm(ref, cost, kn, paths, precision="f32")
    assert np.array_equal(total.view(np.uint32), want32.view(np.uint32))
    want64 = oracle.sgm(ref, cost.astype(np.float64), kn, paths, precision="f64")
    assert np.all(total[invalid] == 0) and np.all(want64[invalid] == 0)
    rel = np.abs(total - want64) / np.maximum(np.abs(want64), 1.0)
    assert rel.max() <= 1e-5
