"""Keyframe / source selection driver end to end (needs a GPU).

The driver's depth maps must be the StereoMapper's for the same calls: the test replays the
driver's plan through a second, directly driven StereoMapper (initReference, update per
source, outputFusionCPU, with the poses the driver decodes from the quaternions) and compares
bit for bit, in CASE 0 (the reference node's default) and CASE 3 (fusion).  The
``pm_evaluate`` executable then gives the same depth maps from the files.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

PLANES = 32


@pytest.fixture(scope="module")
def seq():
    from scipy.spatial.transform import Rotation
    from probabilistic_mapping_amd import synth
    s = synth.trajectory_scene(5, width=160, height=120, n_frames=30, n_sources=3,
                               planes=PLANES, fx=110.0)
    stamps = [1_700_000_000_000_000_000 + 50_000_000 * f for f in range(len(s["images"]))]
    poses = []
    for R, T in s["poses"]:
        q = Rotation.from_matrix(np.asarray(R, np.float64)).as_quat()   # (x, y, z, w)
        poses.append((np.asarray(T, np.float64).reshape(3), q))
    K = s["scene"].K
    cfg = dict(agg_num=3, interval=4, planes=PLANES, fx=float(K[0, 0]), fy=float(K[1, 1]),
               cx=float(K[0, 2]), cy=float(K[1, 2]), dep_sample=2.0 / (PLANES - 1))
    return {"stamps": stamps, "images": s["images"], "poses": poses, "cfg": cfg}


def _drive(seq):
    from probabilistic_mapping_amd import driver
    d = driver.FrameDriver(**seq["cfg"])
    out = []
    for t, img, (tv, q) in zip(seq["stamps"], seq["images"], seq["poses"]):
        d.grab_pose(t, tv, q)
        d.grab_image(t, img)
        r = d.process()
        if r is not None:
            out.append(r)
    d.close()
    return out


@pytest.mark.parametrize("case", [0, 3])
def test_driver_matches_direct_mapper(seq, case):
    from probabilistic_mapping_amd import driver, mapper
    old = mapper.get_knobs()
    mapper.set_knobs(case_mode=case)
    try:
        steps = _drive(seq)
        assert len(steps) >= 3
        cfg = seq["cfg"]
        H, W = seq["images"][0].shape
        m = mapper.StereoMapper(W, H, PLANES)
        K = np.array([[cfg["fx"], 0, cfg["cx"]], [0, cfg["fy"], cfg["cy"]], [0, 0, 1.0]])
        m.initIntrinsic(K, 0.02 * cfg["fx"], cfg["dep_sample"])
        for t, ref, srcs, depth in steps:
            tv, q = seq["poses"][ref]
            R_l = driver.quaternion_to_rotation(q)
            m.initReference(seq["images"][ref])
            for j in srcs:
                tj, qj = seq["poses"][j]
                m.update(seq["images"][j], R_l, tv, driver.quaternion_to_rotation(qj), tj)
            want = m.outputFusionCPU(R_l, tv, cfg["fx"], cfg["fy"], cfg["cx"], cfg["cy"])
            print(f"[driver] CASE {case} step ref {ref} srcs {srcs}: valid "
                  f"{float(np.mean(want > 0)):.3f}", flush=True)
            assert np.array_equal(depth, want), (case, ref)
        m.close()
    finally:
        mapper.set_knobs(case_mode=old["case_mode"])


def test_evaluate_executable_matches_driver(seq, tmp_path):
    from probabilistic_mapping_amd import driver, mapper
    if not os.path.exists(driver.EVALUATE):
        pytest.skip("pm_evaluate is not built")
    old = mapper.get_knobs()
    mapper.set_knobs(case_mode=3)
    try:
        steps = _drive(seq)
    finally:
        mapper.set_knobs(case_mode=old["case_mode"])
    img_list, pose_list = driver.write_dataset(str(tmp_path), seq["stamps"], seq["images"],
                                               seq["poses"])
    out = tmp_path / "out"
    out.mkdir()
    got = driver.run_evaluate(img_list, pose_list, str(out), case=3, **seq["cfg"])
    assert len(got) == len(steps)
    for (st, ref, srcs, depth), (t, ref2, srcs2, depth2) in zip(got, steps):
        assert st == driver.stamp_text(t) and ref == ref2 and srcs == srcs2
        assert np.array_equal(depth, depth2), ref
    # uploading every frame at every use (--no-frame-ring) gives the same files
    out2 = tmp_path / "out_per_use"
    out2.mkdir()
    got2 = driver.run_evaluate(img_list, pose_list, str(out2), case=3, frame_ring=False,
                               **seq["cfg"])
    assert len(got2) == len(got)
    for a, b in zip(got, got2):
        assert a[:3] == b[:3] and np.array_equal(a[3], b[3])
