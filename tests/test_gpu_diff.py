"""Differentiable soft-argmin readout on the GPU (needs a GPU).

The GPU forward and hand-written adjoint (through the C ABI) are compared with the torch FP64
restatement (oracle/diff_ref.py), whose gradient is itself pinned against FP64 central
differences on the CPU.  The kernel adjoint is compared on identical inputs: the restatement
takes the same float-rounded homographies as leaves, so dL/d(H, t) and dL/dI_ref are checked
directly; the se3 chain is then checked through the torch.autograd.Function.
"""
import json

import numpy as np
import pytest

import oracle
from oracle import diff_ref

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
F64 = torch.float64


def _report(name, stats):
    print(f"[diff] {name}: {json.dumps(stats)}", flush=True)


def _setup(s):
    from probabilistic_mapping_amd import diff, pipeline
    D = s["planes"]
    ds = (s["idep_max"] - s["idep_min"]) / (D - 1)
    ref = torch.from_numpy(s["ref"]).cuda()
    quads = [pipeline.make_quad(torch.from_numpy(im).cuda()) for im in s["srcs"]]
    homs = []
    for R, t in s["rel_poses"]:
        H, th = diff.se3_homography(s["K"], R, t, np.zeros(6))
        homs.append((H.astype(np.float32), th.astype(np.float32)))
    gt = torch.from_numpy(s["gt_depth"].astype(np.float32)).cuda()
    return dict(ref=ref.float(), quads=quads, homs=homs, gt=gt, ds=ds)


def _make_op(s, recompute=False):
    from probabilistic_mapping_amd import diff
    H, W = s["ref"].shape
    return diff.SoftArgminDepth(W, H, s["planes"], len(s["srcs"]), s["knobs"], paths=8,
                                temperature=s["temperature"], cost_scale=s["cost_scale"],
                                recompute=recompute)


def _signs(depth_gpu, s):
    """sign(depth_hat - gt) of the GPU forward: the L1 branch its adjoint differentiates."""
    return np.sign(depth_gpu.astype(np.float64) - s["gt_depth"].astype(np.float64))


def _reference(s, homs, signs=None):
    """Restatement forward + autograd with the float-rounded homographies as leaves, on the L1
    branch given by ``signs``."""
    D = s["planes"]
    idep = torch.tensor(np.float32(s["idep_min"]), dtype=F64) + \
        torch.arange(D, dtype=F64) * float(np.float32((s["idep_max"] - s["idep_min"]) / (D - 1)))
    ref = torch.tensor(s["ref"], dtype=F64, requires_grad=True)
    hl = [(torch.tensor(h.astype(np.float64).reshape(3, 3), requires_grad=True),
           torch.tensor(t.astype(np.float64), requires_grad=True)) for h, t in homs]
    cost = diff_ref.calc_cost(ref, s["srcs"], hl, idep)
    cost.retain_grad()
    agg, inv = diff_ref.sgm(ref, cost, s["knobs"], 8)
    agg.retain_grad()
    dh, loss, M = diff_ref.readout(agg, inv, idep, s["gt_depth"].astype(np.float64),
                                   s["temperature"], s["cost_scale"], signs)
    gt = s["gt_depth"].astype(np.float64)
    l1 = float(np.abs(dh.detach().numpy() - gt)[M.numpy()].mean())
    loss.backward()
    ghom = np.stack([np.concatenate([h.grad.numpy().reshape(9), t.grad.numpy()]) for h, t in hl])
    return dict(cost=cost.detach().numpy(), agg=agg.detach().numpy(), depth=dh.detach().numpy(),
                loss=l1, mask=M.numpy(), gref=ref.grad.numpy(), ghom=ghom,
                gcost=cost.grad.numpy(), gagg=agg.grad.numpy())


@pytest.fixture(scope="module")
def medium():
    from probabilistic_mapping_amd import synth
    torch.set_num_threads(16)
    s = synth.diff_scene(1, width=128, height=96, n_sources=2, planes=32)
    g = _setup(s)
    ref = _reference(s, g["homs"], _signs(_run(s, g)["depth"], s))
    return s, g, ref


def _run(s, g, recompute=False):
    op = _make_op(s, recompute)
    depth, loss = op.forward(g["ref"], g["quads"], g["homs"], s["idep_min"], g["ds"], g["gt"])
    gref, ghom = op.backward(1.0)
    torch.cuda.synchronize()
    out = dict(depth=depth.cpu().numpy(), loss=float(loss.cpu()[0]), gref=gref.cpu().numpy(),
               ghom=ghom.cpu().numpy(), memory=op.memory_bytes)
    op.close()
    return out


def test_forward_matches_restatement(medium):
    s, g, ref = medium
    o = _run(s, g)
    M = ref["mask"]
    rel = np.abs(o["depth"][M] - ref["depth"][M]) / ref["depth"][M]
    st = {"loss_gpu": o["loss"], "loss_ref": ref["loss"],
          "loss_rel": abs(o["loss"] - ref["loss"]) / ref["loss"],
          "depth_rel_max": float(rel.max()), "depth_rel_p999": float(np.quantile(rel, 0.999)),
          "masked_frac": float(M.mean())}
    _report("forward 128x96x32 S=2", st)
    assert M.mean() > 0.8
    # FP32 aggregate (hundreds of cost units) feeding logits of agg / 12.75: 1e-4 relative
    assert st["loss_rel"] <= 1e-4
    assert st["depth_rel_p999"] <= 1e-4


def test_adjoint_matches_reference_gradient(medium):
    s, g, ref = medium
    o = _run(s, g)
    gr, rr = o["gref"].astype(np.float64), ref["gref"]
    ref_l2 = float(np.linalg.norm(gr - rr) / np.linalg.norm(rr))
    hom_err = np.abs(o["ghom"] - ref["ghom"])
    # per source: H entries and t entries have different scales (pixel vs unit coordinates)
    h_rel = float(max((hom_err[i, :9] / np.abs(ref["ghom"][i, :9]).max()).max()
                      for i in range(len(s["srcs"]))))
    t_rel = float(max((hom_err[i, 9:] / np.abs(ref["ghom"][i, 9:]).max()).max()
                      for i in range(len(s["srcs"]))))
    st = {"gref_rel_l2": ref_l2, "ghom_H_rel": h_rel, "ghom_t_rel": t_rel,
          "gref_norm": float(np.linalg.norm(rr)),
          "ghom_gpu": o["ghom"].round(6).tolist(), "ghom_ref": ref["ghom"].round(6).tolist()}
    _report("adjoint 128x96x32 S=2", st)
    # end to end against FP64: bounded by FP32-vs-FP64 SGM branch divergence (test_adjoint_stages)
    assert ref_l2 <= 2e-2
    assert h_rel <= 1e-2 and t_rel <= 1e-2


def _l2(a, b):
    return float(np.linalg.norm(a.astype(np.float64) - b) / max(np.linalg.norm(b), 1e-300))


def _cost_backward(s, homs, upstream):
    """FP64 restatement of the cost adjoint: dL/dI_ref and dL/d(H, t) for dL/dC = upstream."""
    D = s["planes"]
    idep = torch.tensor(np.float32(s["idep_min"]), dtype=F64) + \
        torch.arange(D, dtype=F64) * float(np.float32((s["idep_max"] - s["idep_min"]) / (D - 1)))
    ref = torch.tensor(s["ref"], dtype=F64, requires_grad=True)
    hl = [(torch.tensor(h.astype(np.float64).reshape(3, 3), requires_grad=True),
           torch.tensor(t.astype(np.float64), requires_grad=True)) for h, t in homs]
    cost = diff_ref.calc_cost(ref, s["srcs"], hl, idep)
    (cost * torch.from_numpy(upstream.astype(np.float64))).sum().backward()
    ghom = np.stack([np.concatenate([h.grad.numpy().reshape(9), t.grad.numpy()]) for h, t in hl])
    return ref.grad.numpy(), ghom


@pytest.mark.parametrize("width,height,planes", [(96, 64, 32), (96, 64, 64), (128, 96, 32)])
def test_adjoint_stages(width, height, planes):
    """Stage by stage: dL/dagg against the FP64 restatement; dL/dC against the C++ oracle's SGM
    adjoint replayed on the GPU's own FP32 cost (same routing decisions); dL/dI_ref and
    dL/d(H, t) against the FP64 cost adjoint fed with the GPU's dL/dC.  End to end against the
    FP64 gradient, which differs where FP32 and FP64 forwards take different SGM branches
    (near-ties), measured here on the CPU alone as fp32_vs_fp64_branches."""
    from probabilistic_mapping_amd import synth
    torch.set_num_threads(16)
    s = synth.diff_scene(2, width=width, height=height, n_sources=2, planes=planes)
    g = _setup(s)
    ref = _reference(s, g["homs"], _signs(_run(s, g)["depth"], s))
    op = _make_op(s)
    op.forward(g["ref"], g["quads"], g["homs"], s["idep_min"], g["ds"], g["gt"])
    cost = op.download("cost")
    gref, ghom = op.backward(1.0)
    gagg = op.download("agg")
    gC = op.download("grad_cost")
    gref = gref.cpu().numpy()
    ghom = ghom.cpu().numpy()
    op.close()
    replay = oracle.sgm_adjoint(s["ref"], cost, s["knobs"], 8, gagg, precision="f32")
    replay64 = oracle.sgm_adjoint(s["ref"], ref["cost"], s["knobs"], 8, gagg, precision="f64")
    cref, chom = _cost_backward(s, g["homs"], gC)
    st = {"gagg_vs_fp64": _l2(gagg, ref["gagg"]),
          "gcost_vs_replay": _l2(gC, replay),
          "gref_vs_fp64_cost_adjoint": _l2(gref, cref),
          "ghom_vs_fp64_cost_adjoint": float(np.abs(ghom - chom).max() / np.abs(chom).max()),
          "fp32_vs_fp64_branches": _l2(replay, replay64),
          "end_to_end_gref": _l2(gref, ref["gref"]),
          "end_to_end_ghom": float(np.abs(ghom - ref["ghom"]).max() / np.abs(ref["ghom"]).max())}
    _report(f"adjoint stages {width}x{height}x{planes}", st)
    assert st["gagg_vs_fp64"] <= 1e-3
    assert st["gcost_vs_replay"] <= 1e-5
    assert st["gref_vs_fp64_cost_adjoint"] <= 1e-3
    assert st["ghom_vs_fp64_cost_adjoint"] <= 1e-3
    assert st["end_to_end_gref"] <= 2e-2 and st["end_to_end_ghom"] <= 1e-2


def test_store_and_recompute_agree(medium):
    s, g, _ = medium
    a = _run(s, g, recompute=False)
    b = _run(s, g, recompute=True)
    _report("memory 128x96x32", {"store": a["memory"], "recompute": b["memory"]})
    assert b["memory"] < a["memory"]
    for k in ("depth", "gref", "ghom"):
        assert np.array_equal(a[k], b[k]), k
    assert a["loss"] == b["loss"]


def test_backward_after_backward(medium):
    """Store mode: the backward consumes the forward's volumes (the aggregate becomes dL/dagg,
    the path volumes the path adjoints), so a second backward is refused until the next
    forward, which then gives the same gradients again.  Recompute mode rebuilds them: a second
    backward runs and repeats the first bit for bit."""
    from probabilistic_mapping_amd import _capi as C
    s, g, _ = medium
    args = (g["ref"], g["quads"], g["homs"], s["idep_min"], g["ds"], g["gt"])
    op = _make_op(s, recompute=False)
    op.forward(*args)
    first = [t.cpu().numpy() for t in op.backward(1.0)]
    with pytest.raises(C.PmError):
        op.backward(1.0)
    op.forward(*args)
    again = [t.cpu().numpy() for t in op.backward(1.0)]
    op.close()
    for a, b in zip(first, again):
        assert np.array_equal(a, b)
    op = _make_op(s, recompute=True)
    op.forward(*args)
    r1 = [t.cpu().numpy() for t in op.backward(1.0)]
    r2 = [t.cpu().numpy() for t in op.backward(1.0)]
    op.close()
    for a, b, c in zip(r1, r2, first):
        assert np.array_equal(a, b) and np.array_equal(a, c)


def test_autograd_function(medium):
    """DepthLoss through torch.autograd: the pose gradient matches the restatement's autograd
    through the FP64 matrix exponential; the intensity gradient is the adjoint's."""
    from probabilistic_mapping_amd import diff
    s, g, ref = medium
    op = _make_op(s)
    xi = torch.zeros((len(s["srcs"]), 6), dtype=F64, requires_grad=True)
    refp = g["ref"].clone().requires_grad_(True)
    loss = diff.DepthLoss.apply(xi, refp, op, g["quads"], s["K"], s["rel_poses"], s["idep_min"],
                                g["ds"], g["gt"])
    loss.backward()
    # restatement, differentiated with respect to xi through exact FP64 homographies
    D = s["planes"]
    idep = torch.tensor(np.float32(s["idep_min"]), dtype=F64) + \
        torch.arange(D, dtype=F64) * float(np.float32(g["ds"]))
    x = torch.zeros((len(s["srcs"]), 6), dtype=F64, requires_grad=True)
    out = diff_ref.forward(x, torch.tensor(s["ref"], dtype=F64), s["srcs"], s["K"],
                           s["rel_poses"], idep, s["gt_depth"].astype(np.float64), s["knobs"], 8,
                           s["temperature"], s["cost_scale"],
                           _signs(op._keep[3].cpu().numpy(), s))
    out["loss"].backward()
    gx, rx = xi.grad.numpy(), x.grad.numpy()
    rel = float((np.abs(gx - rx) / np.abs(rx).max(axis=1, keepdims=True)).max())
    _report("autograd.Function pose gradient", {"gpu": gx.round(5).tolist(),
                                                "ref": rx.round(5).tolist(), "rel": rel})
    assert rel <= 1e-3
    op.close()
    # the intensity gradient through autograd is the adjoint's, bit for bit
    assert np.array_equal(refp.grad.cpu().numpy(), _run(s, g)["gref"])


def test_full_size_differentiation_workload():
    """640x480, 4 sources, 64 planes, 8 paths, T = 0.05: store and recompute modes agree bit for
    bit; the loss matches the FP64 oracle pipeline; the pose gradient's directional
    derivative matches FP64 central differences of the oracle forward, at the configured step
    1e-4 and at 1e-8 (the 1e-4 step moves warps by ~1e-2 px, so some bilinear cells and
    min-routing choices change inside it: a looser bar, reported beside the converged one);
    and dL/dI_ref matches FP64 central differences (step 1e-4) of the oracle forward at the
    pixels of largest gradient away from the |dI| = tau_so switch."""
    from probabilistic_mapping_amd import diff, synth
    s = synth.diff_scene(0)
    g = _setup(s)
    a = _run(s, g, recompute=False)
    b = _run(s, g, recompute=True)
    for k in ("depth", "gref", "ghom"):
        assert np.array_equal(a[k], b[k]), k
    D = s["planes"]
    idep = torch.tensor(np.float32(s["idep_min"]), dtype=F64) + \
        torch.arange(D, dtype=F64) * float(np.float32(g["ds"]))
    oracle.set_threads(min(16, oracle.max_threads()))

    signs = _signs(a["depth"], s)

    def oracle_loss(homs, sg=None, ref=None):
        ref = s["ref"] if ref is None else ref
        cost = oracle.calc_cost(ref, s["srcs"], homs, float(np.float32(s["idep_min"])),
                                float(np.float32(g["ds"])), D)
        agg = oracle.sgm(ref, cost, dict(s["knobs"], var_scale=1.0), 8)
        inv = torch.from_numpy((cost < 0).any(axis=2))
        _, loss, _ = diff_ref.readout(torch.from_numpy(agg), inv, idep,
                                      s["gt_depth"].astype(np.float64), s["temperature"],
                                      s["cost_scale"], sg)
        return float(loss)

    l0 = oracle_loss(g["homs"])
    # directional derivative along a random se3 direction of every source
    rng = np.random.default_rng(5)
    e = rng.normal(size=(len(s["srcs"]), 6))
    e /= np.linalg.norm(e)
    gx = np.stack([diff.se3_pose_grad(s["K"], R, t, np.zeros(6), a["ghom"][i])
                   for i, (R, t) in enumerate(s["rel_poses"])])
    h = 1e-8   # FP64 FD converges to the branch derivative well below kink spacing

    def homs_at(scale):
        return [diff.se3_homography(s["K"], R, t, scale * e[i]) for i, (R, t) in
                enumerate(s["rel_poses"])]

    # on the GPU forward's L1 branch (see _signs)
    fd = (oracle_loss(homs_at(h), signs) - oracle_loss(homs_at(-h), signs)) / (2 * h)
    h4 = 1e-4
    fd4 = (oracle_loss(homs_at(h4), signs) - oracle_loss(homs_at(-h4), signs)) / (2 * h4)
    dd = float((gx * e).sum())
    # dL/dI_ref at the pixels of largest |gradient| whose scanline neighbours are not at the
    # P1/P2 switch |I(p) - I(q)| = tau_so (a step there changes the penalties)
    gref = a["gref"]
    img = s["ref"].astype(np.int64)
    tau = s["knobs"]["tau_so"]
    picked = []
    for i in np.argsort(-np.abs(gref).ravel()):
        y, x = divmod(int(i), img.shape[1])
        nb = img[max(y - 1, 0):y + 2, max(x - 1, 0):x + 2]
        if not (np.abs(nb - img[y, x]) == tau).any():
            picked.append((y, x))
        if len(picked) == 4:
            break
    ref_fd = []
    for y, x in picked:
        rp, rm = s["ref"].astype(np.float64), s["ref"].astype(np.float64)
        rp[y, x] += h4
        rm[y, x] -= h4
        f = (oracle_loss(g["homs"], signs, rp) - oracle_loss(g["homs"], signs, rm)) / (2 * h4)
        ref_fd.append({"y": y, "x": x, "gpu": float(gref[y, x]), "fd": f,
                       "rel": abs(float(gref[y, x]) - f) / abs(f)})
    # the SGM adjoint at full size: GPU against the oracle replay on the GPU's own cost
    op = _make_op(s)
    op.forward(g["ref"], g["quads"], g["homs"], s["idep_min"], g["ds"], g["gt"])
    cost = op.download("cost")
    op.backward(1.0)
    gagg, gC = op.download("agg"), op.download("grad_cost")
    op.close()
    replay = oracle.sgm_adjoint(s["ref"], cost, s["knobs"], 8, gagg, precision="f32")
    st = {"loss_gpu": a["loss"], "loss_oracle": l0, "loss_rel": abs(a["loss"] - l0) / l0,
          "dir_deriv_gpu": dd, "dir_deriv_fd": fd, "dir_rel": abs(dd - fd) / abs(fd),
          "dir_deriv_fd_h1e-4": fd4, "dir_rel_h1e-4": abs(dd - fd4) / abs(fd4),
          "ref_fd_h1e-4": ref_fd, "ref_rel_max": max(r["rel"] for r in ref_fd),
          "grad_xi_norm": float(np.linalg.norm(gx)), "gcost_vs_replay": _l2(gC, replay),
          "memory_store": a["memory"], "memory_recompute": b["memory"]}
    _report("640x480x64 S=4", st)
    assert st["loss_rel"] <= 1e-4
    assert st["gcost_vs_replay"] <= 1e-5
    assert st["dir_rel"] <= 1e-2
    assert st["dir_rel_h1e-4"] <= 5e-2
    assert st["ref_rel_max"] <= 2e-2
    assert b["memory"] < 0.5 * a["memory"]
