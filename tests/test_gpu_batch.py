"""Batched form: uneven keyframes per launch (needs a GPU).

The batch runs the same stage bodies as the single-keyframe entry points, so its outputs must
be bit-identical to them for every launch shape the autotuner can pick and through hipGraph
replay.  The FP16 / BF16 storage forms are then measured against the FP64 oracle.
"""
import json

import numpy as np
import pytest

from oracle.sequence import homography as np_homography
from tests import parity

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

KNOBS = dict(pi1=8.0, pi2=24.0, tau_so=4.0, sgm_q1=2.0, sgm_q2=3.0, var_scale=1.02)


def _report(name, stats):
    print(f"[parity] {name}: {json.dumps(stats)}", flush=True)


@pytest.fixture(scope="module")
def uneven():
    """Four keyframes of different size, plane count and source count."""
    from probabilistic_mapping_amd import pipeline, synth
    specs = []
    for seed, (W, H, D, S, fx) in enumerate([(320, 240, 32, 2, 250.0), (160, 120, 96, 3, 110.0),
                                              (200, 100, 128, 4, 130.0), (256, 128, 64, 5, 170.0)]):
        sc = synth.trajectory_scene(seed + 7, width=W, height=H, n_frames=S + 1, n_sources=S,
                                    planes=D, fx=fx)
        K = sc["scene"].K
        imgs = [torch.from_numpy(im).cuda() for im in sc["images"]]
        quads = [pipeline.make_quad(im) for im in imgs]
        homs = [np_homography(K, *sc["poses"][0], *sc["poses"][j]) for j in sc["sources"][0]]
        specs.append({"ref": imgs[0], "quads": [quads[j] for j in sc["sources"][0]],
                      "homs": homs, "planes": D, "idep_min": sc["idep_min"],
                      "dep_sample": (sc["idep_max"] - sc["idep_min"]) / (D - 1)})
    return specs


def _single(pipeline, kf, cost_type, paths, storage="exact"):
    cost = pipeline.calc_cost(kf["ref"], kf["quads"], kf["homs"], kf["idep_min"],
                              kf["dep_sample"], kf["planes"], cost_type)
    if storage == "compact":
        agg, dep = pipeline.sgm_filter_excess(kf["ref"], cost, KNOBS, paths, kf["idep_min"],
                                              kf["dep_sample"])
    else:
        agg = pipeline.sgm(kf["ref"], cost, KNOBS, paths)
        dep = pipeline.filter_cost(agg, KNOBS["var_scale"], kf["idep_min"], kf["dep_sample"])
    return {"cost": cost, "agg": agg, "depth": dep}


def _assert_equal(batch, expect):
    for i, (got, want) in enumerate(zip(batch.outputs, expect)):
        for k in ("cost", "agg", "depth"):
            assert torch.equal(got[k], want[k]), (i, k)


@pytest.mark.parametrize("cost_type,paths,storage", [("f32", 8, "exact"), ("f16", 8, "exact"),
                                                      ("bf16", 4, "exact"),
                                                      ("f16", 8, "compact"),
                                                      ("bf16", 4, "compact"),
                                                      ("f32", 8, "compact")])
def test_batch_equals_single_keyframe(uneven, cost_type, paths, storage):
    from probabilistic_mapping_amd import pipeline
    expect = [_single(pipeline, kf, cost_type, paths, storage) for kf in uneven]
    batch = pipeline.Batch(uneven, cost_type, KNOBS, paths, storage)
    batch.run()
    torch.cuda.synchronize()
    _assert_equal(batch, expect)
    # every launch shape gives the same bits (sgm_rows: the compact paths four scanlines per
    # wave for the 16-bit costs of 32, 64, 96 and 128 planes, or one scanline per wave)
    for rows in (2, 4, 8):
        for planes in ((4, 8) if rows == 8 else (8, 16)):
            for pref in (8, 16):
                for bpc, split, sgm_rows in ((0, 0, 1), (4, 1, 0), (0, 1, 0), (4, 0, 1)):
                    for o in batch.outputs:
                        for v in o.values():
                            v.fill_(7)
                    batch.tuning = dict(cost_tile_rows=rows, cost_stage_planes=planes,
                                        sgm_prefetch=pref, cost_blocks_per_cu=bpc,
                                        sgm_blocks_per_cu=bpc, cost_plane_split=split,
                                        sgm_rows=sgm_rows)
                    batch.run()
                    torch.cuda.synchronize()
                    _assert_equal(batch, expect)


@pytest.mark.parametrize("storage", ["exact", "compact"])
def test_batch_autotune_and_graph_replay(uneven, storage):
    from probabilistic_mapping_amd import pipeline
    expect = [_single(pipeline, kf, "f16", 8, storage) for kf in uneven]
    batch = pipeline.Batch(uneven, "f16", KNOBS, 8, storage)
    best = batch.autotune(reps=2)
    print(f"[autotune] {json.dumps(best)}", flush=True)
    assert best == batch.tuning
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        batch.capture()
    torch.cuda.synchronize()
    for rep in range(3):
        for o in batch.outputs:          # eager work between replays
            for v in o.values():
                v.fill_(rep + 1)
        torch.cuda.synchronize()
        with torch.cuda.stream(side):
            batch.launch()
        side.synchronize()
        _assert_equal(batch, expect)


@pytest.mark.parametrize("cost_type,paths", [("f16", 8), ("f32", 8), ("bf16", 4), ("f16", 4)])
def test_batch_sweep_against_exact(uneven, cost_type, paths):
    """Sweep SGM storage against the exact batch on the same keyframes (D = 32, 96, 128, 64:
    both lane classes, full and partial).  Every path runs the same FP32 recurrence; the sweep
    aggregate P C + step (q_LR + q_RL) + np step q_down + (upward excess) differs from the FP32
    sum of the paths' L by the level rounding (step / 2 for each horizontal path, np step / 2
    for the downward sum) and FP32 rounding of the sums.  Invalid pixels are 0 in both, and the
    filter decisions agree."""
    from probabilistic_mapping_amd import pipeline
    ex = pipeline.Batch(uneven, cost_type, KNOBS, paths, "exact")
    ex.run()
    sw = pipeline.Batch(uneven, cost_type, KNOBS, paths, "sweep")
    sw.run()
    dn = pipeline.Batch(uneven, cost_type, KNOBS, paths, "sweep", write_agg=False)
    dn.run()
    torch.cuda.synchronize()
    step = parity.excess_step(KNOBS)
    n_down = 3 if paths == 8 else 1
    for i, (a, b, c) in enumerate(zip(ex.outputs, sw.outputs, dn.outputs)):
        assert torch.equal(a["cost"], b["cost"])
        ge = a["agg"].cpu().numpy().astype(np.float64)
        gs = b["agg"].cpu().numpy().astype(np.float64)
        ulp = np.spacing(np.abs(ge).astype(np.float32)).astype(np.float64)
        bound = (2 + n_down) * step / 2 + 4 * paths * ulp
        err = np.abs(gs - ge)
        _report(f"sweep vs exact kf{i} {cost_type} P={paths}",
                {"max_abs": float(err.max()), "max_over_bound": float((err / bound).max())})
        assert np.all(err <= bound), i
        inval = a["cost"].float().cpu().numpy().min(axis=2) < 0
        assert np.all(gs[inval] == 0.0)
        st = parity.depth_parity(b["depth"].cpu().numpy(), a["depth"].cpu().numpy())
        assert st["category_mismatch"] <= 1e-3, (i, st)
        assert st["rel"]["p999"] <= 1e-4, (i, st)
        assert torch.equal(b["depth"], c["depth"])   # depth-only run: same depth


def test_batch_sweep_small_keyframes():
    """Sweep storage on keyframes narrower than one band and shorter than one row block (and
    one just past each), where the halo runs off both image edges."""
    from probabilistic_mapping_amd import pipeline, synth
    specs = []
    for seed, (W, H, D) in enumerate([(40, 12, 32), (70, 17, 64), (65, 33, 40), (130, 16, 64)]):
        sc = synth.trajectory_scene(seed + 31, width=W, height=H, n_frames=3, n_sources=2,
                                    planes=D, fx=0.6 * W)
        K = sc["scene"].K
        imgs = [torch.from_numpy(im).cuda() for im in sc["images"]]
        quads = [pipeline.make_quad(im) for im in imgs]
        homs = [np_homography(K, *sc["poses"][0], *sc["poses"][j]) for j in sc["sources"][0]]
        specs.append({"ref": imgs[0], "quads": [quads[j] for j in sc["sources"][0]],
                      "homs": homs, "planes": D, "idep_min": sc["idep_min"],
                      "dep_sample": (sc["idep_max"] - sc["idep_min"]) / (D - 1)})
    ex = pipeline.Batch(specs, "f16", KNOBS, 8, "exact")
    ex.run()
    sw = pipeline.Batch(specs, "f16", KNOBS, 8, "sweep")
    sw.run()
    torch.cuda.synchronize()
    bound_step = (2 + 3) * parity.excess_step(KNOBS) / 2
    for i, (a, b) in enumerate(zip(ex.outputs, sw.outputs)):
        ge = a["agg"].cpu().numpy().astype(np.float64)
        gs = b["agg"].cpu().numpy().astype(np.float64)
        ulp = np.spacing(np.abs(ge).astype(np.float32)).astype(np.float64)
        assert np.all(np.abs(gs - ge) <= bound_step + 32 * ulp), i
        st = parity.depth_parity(b["depth"].cpu().numpy(), a["depth"].cpu().numpy())
        n = ge.shape[0] * ge.shape[1]
        assert st["category_mismatch"] * n <= max(3, 1e-3 * n), (i, st)   # a few ties


@pytest.mark.parametrize("cost_type", ["f16", "bf16"])
def test_batch_compact_ragged_keyframes(cost_type):
    """Compact storage with 16-bit costs on keyframes of 1 to 9 rows or 1 to 5 columns (paths
    with fewer scanlines than one 16-scanline item, diagonals from length 1) beside two of
    ordinary size: the batch, with the paths four scanlines per wave and one, equals the
    single-keyframe calls bit for bit.  Keyframes under 4 rows or columns have no valid cost
    (the 9-term pattern needs the neighbouring rows and columns), so they run the invalid-pixel
    route; the others mix valid and invalid pixels."""
    from probabilistic_mapping_amd import pipeline, synth
    specs = []
    for seed, (W, H, D) in enumerate([(40, 3, 32), (1, 24, 32), (33, 1, 64), (130, 7, 96),
                                      (5, 9, 128), (2, 2, 32), (19, 6, 128), (90, 41, 64),
                                      (47, 70, 96)]):
        sc = synth.trajectory_scene(seed + 61, width=W, height=H, n_frames=4, n_sources=3,
                                    planes=D, fx=max(0.612 * W, 4.0))
        K = sc["scene"].K
        imgs = [torch.from_numpy(im).cuda() for im in sc["images"]]
        quads = [pipeline.make_quad(im) for im in imgs]
        homs = [np_homography(K, *sc["poses"][1], *sc["poses"][j]) for j in (0, 2, 3)]
        specs.append({"ref": imgs[1], "quads": [quads[j] for j in (0, 2, 3)],
                      "homs": homs, "planes": D, "idep_min": sc["idep_min"],
                      "dep_sample": (sc["idep_max"] - sc["idep_min"]) / (D - 1)})
    expect = [_single(pipeline, kf, cost_type, 8, "compact") for kf in specs]
    valid = [float((e["cost"].float().amin(dim=2) >= 0).float().mean()) for e in expect]
    _report(f"ragged batch {cost_type} valid pixel fractions", {"valid": valid})
    assert max(valid) > 0.3
    batch = pipeline.Batch(specs, cost_type, KNOBS, 8, "compact")
    try:
        for sgm_rows in (1, 0):
            for pref in (8, 16):
                for o in batch.outputs:
                    for v in o.values():
                        v.fill_(5)
                batch.tuning = dict(sgm_rows=sgm_rows, sgm_prefetch=pref)
                batch.run()
                torch.cuda.synchronize()
                _assert_equal(batch, expect)
    finally:
        batch.close()


def test_batch_sweep_graph_replay(uneven):
    """The sweep form's ~2 x (row blocks) launches captured in one hipGraph replay to the
    eager run's bits."""
    from probabilistic_mapping_amd import pipeline
    batch = pipeline.Batch(uneven, "f16", KNOBS, 8, "sweep")
    batch.run()
    torch.cuda.synchronize()
    expect = [{k: v.clone() for k, v in o.items()} for o in batch.outputs]
    best = batch.autotune(reps=1)
    assert best == batch.tuning
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        batch.capture()
    torch.cuda.synchronize()
    for rep in range(3):
        for o in batch.outputs:
            for v in o.values():
                v.fill_(rep + 1)
        torch.cuda.synchronize()
        with torch.cuda.stream(side):
            batch.launch()
        side.synchronize()
        _assert_equal(batch, expect)


def test_batch_rejects_bad_input(uneven):
    from probabilistic_mapping_amd import _capi as C, pipeline
    bad = dict(uneven[0], planes=0)
    with pytest.raises(C.PmError):
        pipeline.Batch([bad], "f32", KNOBS, 8)
    with pytest.raises(C.PmError):
        pipeline.Batch(uneven, "f32", KNOBS, 6)
    b = pipeline.Batch(uneven[:1], "f32", KNOBS, 4)
    with pytest.raises(C.PmError):
        b.tuning = dict(cost_tile_rows=3)
    with pytest.raises(C.PmError):
        b.tuning = dict(cost_tile_rows=8, cost_stage_planes=16)
    with pytest.raises(C.PmError):
        b.tuning = dict(cost_plane_split=2)
    with pytest.raises(C.PmError):
        b.tuning = dict(sgm_rows=2)
    with pytest.raises(C.PmError):
        b.launch()   # nothing captured
    with pytest.raises(C.PmError):   # the sweep form takes at most 128 planes
        pipeline.Batch([dict(uneven[0], planes=129)], "f16", KNOBS, 8, "sweep")


def _with_sources(kf, count):
    """The keyframe with its sources repeated up to ``count``, each repeat with its own
    translation scale."""
    kf = dict(kf)
    n = len(kf["quads"])
    kf["quads"] = [kf["quads"][i % n] for i in range(count)]
    kf["homs"] = [(h, (t * (1.0 + 0.03 * i)).astype(np.float32))
                  for i, (h, t) in enumerate(kf["homs"][i % n] for i in range(count))]
    return kf


def test_batch_max_sources_tile_rows(uneven):
    """A keyframe with 16 sources: 2-row cost tiles hold the warp records of at most 8
    sources, so that shape is refused (and skipped by the autotuner); 4 and 8 rows run it.
    With 9 sources 2 rows are refused too; with 8 they run and give the same bits."""
    from probabilistic_mapping_amd import _capi as C, pipeline
    kf = _with_sources(uneven[3], 16)
    expect = [_single(pipeline, kf, "f16", 8)]
    b = pipeline.Batch([kf], "f16", KNOBS, 8)
    with pytest.raises(C.PmError):
        b.tuning = dict(cost_tile_rows=2, cost_stage_planes=8)
    for rows, planes in ((4, 8), (8, 4)):
        b.tuning = dict(cost_tile_rows=rows, cost_stage_planes=planes)
        b.run()
        torch.cuda.synchronize()
        _assert_equal(b, expect)
    best = b.autotune(reps=1)
    assert best["cost_tile_rows"] != 2
    with pytest.raises(C.PmError):
        pipeline.Batch([_with_sources(uneven[3], 9)], "f16", KNOBS, 8).tuning = dict(
            cost_tile_rows=2, cost_stage_planes=8)
    kf8 = _with_sources(uneven[3], 8)
    b8 = pipeline.Batch([kf8], "f16", KNOBS, 8)
    b8.tuning = dict(cost_tile_rows=2, cost_stage_planes=8)
    b8.run()
    torch.cuda.synchronize()
    _assert_equal(b8, [_single(pipeline, kf8, "f16", 8)])


def test_batched_sequence_groups_match_one_batch(trajectory_scene):
    """Keyframe groups (several batches, each captured and replayed) give the one-batch
    depths and fused state bit for bit, through graph replay; the cache-sized default groups
    the 752x480x64 FP16 sequence by five."""
    from probabilistic_mapping_amd.pipeline import BatchedSequence, BatchGroups
    s = trajectory_scene
    n = 10
    args = (s["images"][:n], s["poses"][:n], [[j for j in src if j < n] or [1]
                                               for src in s["sources"][:n]],
            s["scene"].K, s["planes"], s["idep_min"], s["idep_max"], s["knobs"], s["paths"])
    one = BatchedSequence(*args, cost_type="f16", sgm_storage="compact", group=0)
    grp = BatchedSequence(*args, cost_type="f16", sgm_storage="compact", group=3)
    auto = BatchedSequence(*args, cost_type="f16", sgm_storage="compact")
    assert isinstance(grp.batch, BatchGroups) and len(grp.batch.batches) == 4
    assert [len(b.outputs) for b in auto.batch.batches] == [5, 5]
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        for q in (one, grp, auto):
            q.batch.capture()
        for rep in range(2):
            for q in (one, grp, auto):
                q.reset()
                q.batch.launch()
                for i in range(n):
                    q.fusion_step(i)
            torch.cuda.synchronize()
            for q in (grp, auto):
                for i in range(n):
                    assert torch.equal(q.depth_of(i), one.depth_of(i)), (rep, i)
                assert torch.equal(q.fused, one.fused)
                for k in one.state:
                    assert torch.equal(q.state[k], one.state[k]), (rep, k)


def test_overlapped_sequence_equals_serial(trajectory_scene):
    """The sequence with each group's cost launch beside the previous group's SGM pass and
    fusion (two streams), run eagerly and as one captured hipGraph replayed three times with
    eager work in between, gives the serial run's depths, fused depth and state bit for bit
    after every run."""
    from probabilistic_mapping_amd.pipeline import BatchedSequence
    s = trajectory_scene
    n = 12
    args = (s["images"][:n], s["poses"][:n], [[j for j in src if j < n] or [1]
                                               for src in s["sources"][:n]],
            s["scene"].K, s["planes"], s["idep_min"], s["idep_max"], s["knobs"], s["paths"])
    ref = BatchedSequence(*args, cost_type="f16", sgm_storage="compact", median=True,
                          lookahead=True, group=4)
    ovl = BatchedSequence(*args, cost_type="f16", sgm_storage="compact", median=True,
                          lookahead=True, group=4)
    assert len(ovl.batch.batches) == 3
    ref.run()
    torch.cuda.synchronize()
    want_dep = [ref.depth_of(i).clone() for i in range(n)]
    want_fused = ref.fused.clone()
    want_state = {k: v.clone() for k, v in ref.state.items()}

    def check(tag):
        torch.cuda.synchronize()
        for i in range(n):
            assert torch.equal(ovl.depth_of(i), want_dep[i]), (tag, i)
        assert torch.equal(ovl.fused, want_fused), tag
        for k, v in want_state.items():
            assert torch.equal(ovl.state[k], v), (tag, k)

    ovl.run_overlapped()
    check("eager")
    ovl.capture_overlapped()
    for rep in range(3):
        for i in range(n):
            ovl.depth_of(i).fill_(rep + 5)
        ovl.fused.fill_(-1.0)
        ovl.launch_overlapped()
        check(f"replay {rep}")
        ref.run()          # eager work between the replays
        torch.cuda.synchronize()


@pytest.mark.parametrize("storage", ["exact", "compact"])
def test_batched_sequence_with_median_matches_plain(trajectory_scene, storage):
    """The batched sequence's median -> fusion steps give the plain sequence's fused depth and
    state bit for bit (FP16 cost, 4 keyframes of the 752x480x64 workload)."""
    from probabilistic_mapping_amd.pipeline import BatchedSequence, KeyframeSequence
    s = trajectory_scene
    n = 4
    args = (s["images"][:n], s["poses"][:n], [[j for j in src if j < n] or [1]
                                               for src in s["sources"][:n]],
            s["scene"].K, s["planes"], s["idep_min"], s["idep_max"], s["knobs"], s["paths"])
    g = BatchedSequence(*args, cost_type="f16", sgm_storage=storage, median=True)
    p = KeyframeSequence(*args, cost_type="f16", sgm_storage=storage, median=True)
    g.reset()
    g.batch.run()
    p.reset()
    for i in range(n):
        g.fusion_step(i)
        p.run_keyframe(i)
        torch.cuda.synchronize()
        assert torch.equal(g.dep, p.dep), i
        assert torch.equal(g.fused, p.fused), i
        for k in g.state:
            assert torch.equal(g.state[k], p.state[k]), (i, k)
    assert not torch.equal(g.dep, g.depth_of(n - 1))   # the median changed the depth
