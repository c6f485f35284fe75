"""Upload-once frame ring (needs a GPU).

A mapper fed through the ring must give bit-identical cost, aggregate and depth to one fed
host images (StereoMapper::update's upload path, stereo_mapper.cpp:73-77), for single and
mixed staging; evicted frames are refused on the host before anything is launched; and the
driver uploads each used frame once instead of at every use.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

D = 32


@pytest.fixture(scope="module")
def scene():
    from probabilistic_mapping_amd import synth
    s = synth.trajectory_scene(3, width=200, height=136, n_frames=8, n_sources=3, planes=D,
                               fx=130.0)
    return s


def _mapper(s):
    from probabilistic_mapping_amd.mapper import StereoMapper
    H, W = s["images"][0].shape
    m = StereoMapper(W, H, D)
    m.initIntrinsic(s["scene"].K, 0.0, (s["idep_max"] - s["idep_min"]) / (D - 1))
    return m


def _outputs(m):
    dep = m.output()
    return dep, m.download("raw_cost"), m.download("sgm_cost")


@pytest.mark.parametrize("mixed", [False, True])
def test_ring_matches_host_images(scene, mixed):
    from probabilistic_mapping_amd.mapper import FrameRing, set_knobs
    s = scene
    H, W = s["images"][0].shape
    old = set_knobs(idep_min=s["idep_min"], sgm_paths=8, **s["knobs"])
    try:
        ring = FrameRing(W, H, len(s["images"]))
        for f, img in enumerate(s["images"]):
            ring.put(f, img)
        assert ring.puts == len(s["images"])
        a, b = _mapper(s), _mapper(s)
        for i in (0, 4, 7):
            R, T = s["poses"][i]
            a.initReference(s["images"][i])
            b.initReferenceFrame(ring, i)
            for k, j in enumerate(s["sources"][i]):
                a.update(s["images"][j], R, T, *s["poses"][j])
                if mixed and k == 1:
                    b.update(s["images"][j], R, T, *s["poses"][j])
                else:
                    b.updateFrame(ring, j, R, T, *s["poses"][j])
            da, ca, ga = _outputs(a)
            db, cb, gb = _outputs(b)
            print(f"[ring] keyframe {i} sources {s['sources'][i]} mixed={mixed}: valid "
                  f"{float(np.mean(da > 0)):.3f}", flush=True)
            assert np.array_equal(ca, cb)
            assert np.array_equal(ga, gb)
            assert np.array_equal(da, db)
            assert np.array_equal(a.download("img_l"), b.download("img_l"))
        a.close()
        b.close()
        ring.close()
    finally:
        set_knobs(**old)


def test_ring_eviction_is_refused(scene):
    from probabilistic_mapping_amd import _capi as C
    from probabilistic_mapping_amd.mapper import FrameRing
    s = scene
    H, W = s["images"][0].shape
    ring = FrameRing(W, H, 2)
    ring.put(0, s["images"][0])
    ring.put(1, s["images"][1])
    assert ring.holds(0) and ring.holds(1) and not ring.holds(2)
    ring.put(2, s["images"][2])               # slot 0: frame 0 is gone
    assert not ring.holds(0) and ring.holds(2)
    m = _mapper(s)
    with pytest.raises(C.PmError) as e:
        m.initReferenceFrame(ring, 0)
    assert e.value.status == C.PM_ERR_INVALID_STATE
    R, T = s["poses"][2]
    m.initReferenceFrame(ring, 2)
    with pytest.raises(C.PmError):
        m.updateFrame(ring, 0, R, T, *s["poses"][0])
    m.updateFrame(ring, 1, R, T, *s["poses"][1])
    ring.put(3, s["images"][3])               # evicts frame 1 before the output call
    with pytest.raises(C.PmError) as e:
        m.output()
    assert e.value.status == C.PM_ERR_INVALID_STATE
    # the mapper keeps working after the refusal
    m.initReferenceFrame(ring, 2)
    m.updateFrame(ring, 3, R, T, *s["poses"][3])
    assert np.isfinite(m.output()).all()
    with pytest.raises(C.PmError):
        ring.put(4, np.zeros((H, W + 1), np.uint8))
    m.close()
    ring.close()


def test_driver_ring_uploads_each_frame_once():
    """The driver with and without the ring gives the same depths; with it, each used frame is
    uploaded once (the reference uploads aggNum + 1 frames per step)."""
    from scipy.spatial.transform import Rotation
    from probabilistic_mapping_amd import driver, mapper, synth
    s = synth.trajectory_scene(5, width=160, height=120, n_frames=30, n_sources=3, planes=D,
                               fx=110.0)
    K = s["scene"].K
    cfg = dict(agg_num=3, interval=4, planes=D, fx=float(K[0, 0]), fy=float(K[1, 1]),
               cx=float(K[0, 2]), cy=float(K[1, 2]), dep_sample=2.0 / (D - 1))
    old = mapper.get_knobs()
    mapper.set_knobs(case_mode=3)
    try:
        out = {}
        for ring in (True, False):
            d = driver.FrameDriver(frame_ring=ring, **cfg)
            steps = []
            for f, (img, (R, T)) in enumerate(zip(s["images"], s["poses"])):
                q = Rotation.from_matrix(np.asarray(R, np.float64)).as_quat()
                d.grab_pose(1000 + f, np.asarray(T, np.float64).reshape(3), q)
                d.grab_image(1000 + f, img)
                r = d.process()
                if r is not None:
                    steps.append(r)
            out[ring] = (steps, d.ring_puts)
            d.close()
        (s1, puts1), (s0, puts0) = out[True], out[False]
        assert len(s1) == len(s0) >= 3
        for a, b in zip(s1, s0):
            assert a[:3] == b[:3]
            assert np.array_equal(a[3], b[3])
        used = {r for st in s1 for r in [st[1], *st[2]]}
        print(f"[ring] driver: {len(s1)} steps, {len(used)} frames used, {puts1} uploads "
              f"(per-use uploads: {len(s1) * (cfg['agg_num'] + 1)})", flush=True)
        assert puts1 == len(used)
        assert puts0 == 0
    finally:
        mapper.set_knobs(case_mode=old["case_mode"])
