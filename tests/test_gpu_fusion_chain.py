"""The fusion chains of the sequence forms against pm_depth_fuse, bit for bit (needs a GPU).

pm_depth_fuse runs one outputFusion step as the reference's depth_fuse does
(motion_stereo/depth_fusion.cu:8-280): forward warp, then hole fill, predict and update; the
3x3 median runs before it as its own launch.  The sequences instead call
pm_depth_fuse_warp + pm_depth_fuse_step, which computes the median inside the update's
launch, and with ``lookahead`` the step of keyframe i also runs keyframe i+1's warp, so that
every later keyframe is one launch.  These tests run the 752x480 trajectory through all three
and require, at every keyframe, the same bits in the median depth, the fused depth, the
propagation table and the state (with lookahead, alpha and beta at every keyframe and the
whole state after the last one: in between, mu and sig already hold the next keyframe's warped
values, which the two-launch chain writes at the start of the next step).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _args(s):
    return (s["images"], s["poses"], s["sources"], s["scene"].K, s["planes"], s["idep_min"],
            s["idep_max"], s["knobs"], s["paths"])


class _TwoKernelChain:
    """The reference-shaped chain: median3x3 launch, then pm_depth_fuse on alternating
    z-buffers, fed with another sequence's filter output."""

    def __init__(self, s, H, W):
        from probabilistic_mapping_amd import pipeline as pl
        self.pl = pl
        self.poses = s["poses"]
        K = np.asarray(s["scene"].K, np.float64)
        self.cam = (K[0, 0], K[1, 1], K[0, 2], K[1, 2])
        dev = torch.device("cuda")
        self.states = [pl.new_state(H, W, dev), pl.new_state(H, W, dev)]
        self.zbufs = [torch.empty((H, W), dtype=torch.int64, device=dev) for _ in range(2)]
        for z in self.zbufs:
            pl.zbuf_reset(z)
        self.dep = torch.empty((H, W), dtype=torch.float32, device=dev)
        self.fused = torch.zeros((H, W), dtype=torch.float32, device=dev)
        self.table = torch.zeros((H, W), dtype=torch.int32, device=dev)
        self.pre, self.zpar, self.prev = 0, 0, None

    def step(self, i, raw):
        pl = self.pl
        pl.median3x3(raw, out=self.dep)
        R, t = self.poses[i]
        if self.prev is None:
            for st in self.states:
                st["alpha"].zero_()
            self.fused.zero_()
        else:
            r, tt = pl.pose_p2c(R, t, *self.prev)
            pl.depth_fuse(self.dep, self.zbufs[self.zpar], self.states[self.pre],
                          self.states[self.pre ^ 1], r, tt, self.cam, table=self.table,
                          fused=self.fused, zbuf_next=self.zbufs[self.zpar ^ 1])
            self.pre ^= 1
            self.zpar ^= 1
        self.prev = (R, t)

    @property
    def state(self):
        return self.states[self.pre]


def _same(a, b):
    return torch.equal(a, b) or bool(torch.equal(torch.isnan(a), torch.isnan(b)) and
                                     torch.equal(a.nan_to_num(), b.nan_to_num()))


def test_plain_chain_and_lookahead_equal_depth_fuse(trajectory_scene):
    """The plain FP32 sequence (warp + step with the median folded in, as the bench times
    it) and the same sequence with lookahead, against the two-launch pm_depth_fuse chain."""
    from probabilistic_mapping_amd.pipeline import KeyframeSequence
    s = trajectory_scene
    g = KeyframeSequence(*_args(s), cost_type="f32", median=True)
    g2 = KeyframeSequence(*_args(s), cost_type="f32", median=True, lookahead=True)
    ref = _TwoKernelChain(s, g.H, g.W)
    n = len(s["images"])
    for i in range(n):
        g.run_keyframe(i)
        g2.run_keyframe(i)
        ref.step(i, g.dep_raw)
        torch.cuda.synchronize()
        assert torch.equal(g.dep_raw, g2.dep_raw), i
        assert torch.equal(g.dep, ref.dep) and torch.equal(g2.dep, ref.dep), i
        assert torch.equal(g.fused, ref.fused) and torch.equal(g2.fused, ref.fused), i
        if i == 0:
            continue
        assert torch.equal(g.table, ref.table) and torch.equal(g2.table, ref.table), i
        for k in ("alpha", "beta", "mu", "sig"):
            assert _same(g.state[k], ref.state[k]), (i, k)
        for k in ("alpha", "beta") if i + 1 < n else ("alpha", "beta", "mu", "sig"):
            assert _same(g2.state[k], ref.state[k]), (i, k)
    print(f"[chain] plain: {n} keyframes bit-identical", flush=True)


def test_batched_lookahead_equals_plain_chain(trajectory_scene):
    """The FP16 batched sequence with lookahead (the bench's timed form: one fusion launch
    per keyframe after the first two) against the same batch without it."""
    from probabilistic_mapping_amd.pipeline import BatchedSequence
    s = trajectory_scene
    a = BatchedSequence(*_args(s), cost_type="f16", median=True, sgm_storage="compact",
                        write_agg=False)
    b = BatchedSequence(*_args(s), cost_type="f16", median=True, sgm_storage="compact",
                        write_agg=False, lookahead=True)
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        for q in (a, b):
            q.reset()
            q.batch.run()
    torch.cuda.synchronize()
    n = len(s["images"])
    for i in range(n):
        with torch.cuda.stream(side):
            a.fusion_step(i)
            b.fusion_step(i)
        torch.cuda.synchronize()
        assert torch.equal(a.depth_of(i), b.depth_of(i)), i
        assert torch.equal(a.dep, b.dep) and torch.equal(a.fused, b.fused), i
        if i == 0:
            continue
        assert torch.equal(a.table, b.table), i
        for k in ("alpha", "beta") if i + 1 < n else ("alpha", "beta", "mu", "sig"):
            assert _same(a.state[k], b.state[k]), (i, k)
    a.batch.close()
    b.batch.close()
    print(f"[chain] batched f16 lookahead: {n} keyframes bit-identical", flush=True)


def test_lookahead_refuses_out_of_order_steps(trajectory_scene):
    """With lookahead the step of keyframe i has already warped keyframe i+1; running another
    keyframe next is refused before anything is launched."""
    from probabilistic_mapping_amd.pipeline import KeyframeSequence
    s = trajectory_scene
    g = KeyframeSequence(*_args(s), cost_type="f32", median=True, lookahead=True)
    for i in range(2):
        g.run_keyframe(i)
    g.depth_stage(3, median=False)
    with pytest.raises(ValueError):
        g.fusion_stage(3)
