"""Keyframe / source selection driver (host logic, no GPU).

The library's ``pm::FrameDriver`` plan (which image is the reference, which are its sources)
is checked against ``oracle/driver.py``, a plain restatement of evaluate.cc's message lists
and processImg; the oracle itself is pinned by the closed form of an aligned stream.  The
quaternion conversion is pinned by scipy and two closed forms.  The ``pm_evaluate``
executable's file parsing is checked through its plan-only mode.
"""
import os

import numpy as np
import pytest

from oracle.driver import DriverOracle

C = pytest.importorskip("probabilistic_mapping_amd._capi")


@pytest.fixture(scope="module")
def drv():
    from probabilistic_mapping_amd import driver
    C.load()
    return driver


def test_quaternion_to_rotation(drv):
    from scipy.spatial.transform import Rotation
    rng = np.random.default_rng(3)
    for _ in range(50):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        R = drv.quaternion_to_rotation(q)
        assert np.allclose(R, Rotation.from_quat(q).as_matrix(), atol=1e-15, rtol=0)
    assert np.array_equal(drv.quaternion_to_rotation([0, 0, 0, 1.0]), np.eye(3))
    s = np.sqrt(0.5)   # 90 degrees about z
    assert np.allclose(drv.quaternion_to_rotation([0, 0, s, s]),
                       [[0, -1, 0], [1, 0, 0], [0, 0, 1]], atol=1e-15)


def _run(plan, stream):
    """Feeds ('p'|'i', stamp) messages to a planner; returns the steps it runs."""
    steps = []
    for kind, t in stream:
        if kind == "p":
            plan.grab_pose(t)
        else:
            plan.grab_image(t)
        r = plan.process()
        if r is not None:
            steps.append(r)
    return steps


class _Lib:
    """The library's planning driver behind the oracle's interface."""

    def __init__(self, drv, agg, interval):
        self.d = drv.FrameDriver(agg_num=agg, interval=interval, with_mapper=False)
        self.img = np.zeros((4, 4), np.uint8)

    def grab_pose(self, t):
        self.d.grab_pose(t, (0, 0, 0), (0, 0, 0, 1))

    def grab_image(self, t):
        self.d.grab_image(t, self.img)

    def process(self):
        r = self.d.process()
        return None if r is None else (r[0], r[1], r[2])


def test_oracle_aligned_stream_closed_form():
    """Pose then image per frame: the first step runs once 19 frames are buffered, then every
    second frame (each step drops one and raises the bar by one); the reference is the newest
    frame and the sources are 6, 12 and 18 frames older."""
    stream = []
    for f in range(60):
        stream += [("p", 1000 + f), ("i", 1000 + f)]
    steps = _run(DriverOracle(3, 6), stream)
    refs = [s[1] for s in steps]
    assert refs == list(range(18, 60, 2))
    for t, ref, ref_pose, srcs, src_poses in steps:
        assert t == 1000 + ref and ref_pose == ref
        assert srcs == [ref - 6, ref - 12, ref - 18] and src_poses == srcs


def _scenarios():
    aligned = []
    for f in range(50):
        aligned += [("p", f * 10), ("i", f * 10)]
    missing = [m for m in aligned if m != ("p", 200)]                 # frame 20 has no pose
    late = []
    for f in range(50):                                                # each pose after its image
        late += [("i", f * 10), ("p", f * 10)]
    odd = []
    for f in range(50):                                                # poses start late
        if f >= 5:
            odd.append(("p", f * 10))
        odd.append(("i", f * 10))
    dup = aligned[:30] + [("p", 145), ("i", 150), ("p", 150)] + aligned[30:]
    return {"aligned": aligned, "missing_pose": missing, "late_poses": late,
            "poses_start_late": odd, "extra_pose": dup}


@pytest.mark.parametrize("name", sorted(_scenarios()))
@pytest.mark.parametrize("agg,interval", [(3, 6), (2, 3), (1, 1)])
def test_plan_matches_oracle(drv, name, agg, interval):
    stream = _scenarios()[name]
    want = [(s[0], s[1], s[3]) for s in _run(DriverOracle(agg, interval), stream)]
    lib = _Lib(drv, agg, interval)
    got = _run(lib, stream)
    assert got == want
    assert len(got) > 0


@pytest.mark.parametrize("agg,interval", [(3, 6), (2, 3), (1, 1), (0, 4)])
def test_driver_frees_pixels_outside_the_window(drv, agg, interval):
    """The message lists grow as the reference's do (one more than they shrink per step), but
    only the newest agg_num * interval + 1 images keep their pixels; the plan is unchanged."""
    lib = _Lib(drv, agg, interval)
    o = DriverOracle(agg, interval)
    stream = _scenarios()["aligned"] * 1
    got, want = [], []
    for kind, t in stream:
        for plan, out in ((lib, got), (o, want)):
            (plan.grab_pose if kind == "p" else plan.grab_image)(t)
            r = plan.process()
            if r is not None:
                out.append(r if plan is lib else (r[0], r[1], r[3]))
        n_img, n_pose, with_px = lib.d.buffered()
        assert n_img == len(o.images) and n_pose == len(o.poses)
        assert with_px == min(n_img, agg * interval + 1)
    assert got == want and len(got) > 0


def test_plan_skips_without_source_poses(drv):
    """Images buffered, but the sources' poses not yet: no step (the reference would read past
    its pose list), counted as skipped, and the oracle agrees."""
    stream = [("i", f) for f in range(25)] + [("p", 24)]
    o = DriverOracle(3, 6)
    assert _run(o, stream) == [] and o.skipped == 1
    lib = _Lib(drv, 3, 6)
    assert _run(lib, stream) == [] and lib.d.skipped == 1


def test_evaluate_plan_only(drv, tmp_path):
    """pm_evaluate reads the frame and pose lists (TUM layouts, PGM frames), replays them in
    stamp order and plans the same steps as the oracle."""
    if not os.path.exists(drv.EVALUATE):
        pytest.skip("pm_evaluate is not built")
    rng = np.random.default_rng(0)
    n = 34
    stamps = [1305031102_175304000 + 33_333_333 * f for f in range(n)]
    images = [rng.integers(0, 256, (12, 16), dtype=np.uint8) for _ in range(n)]
    poses = [((0.01 * f, 0.0, 0.0), (0.0, 0.0, 0.0, 1.0)) for f in range(n)]
    img_list, pose_list = drv.write_dataset(str(tmp_path), stamps, images, poses)
    steps = drv.run_evaluate(img_list, pose_list, str(tmp_path), plan_only=True)
    o = DriverOracle(3, 6)
    stream = []
    for t in stamps:
        stream += [("p", t), ("i", t)]
    want = _run(o, stream)
    assert [(s[1], s[2]) for s in steps] == [(w[1], w[3]) for w in want]
    assert [s[0] for s in steps] == [drv.stamp_text(w[0]) for w in want]
