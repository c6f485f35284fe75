"""The FP64 reference gradient of the differentiable readout, and the host se3 chain rule.

* The torch restatement (oracle/diff_ref.py) reproduces the C++ oracle's cost and aggregate.
* Its autograd gradient equals FP64 central finite differences: with respect to the se3 poses
  FD converges to it as the step shrinks (at the 1e-4 step the error is dominated by samples
  that cross bilinear / |.| / min kinks inside the step; reported, loosely bounded), and with
  respect to reference intensities at pixels away from the |dI| = tau_so penalty switch.
* pm_se3_homography / pm_se3_pose_grad (host code of libpm_hip.so; no GPU needed) agree with
  torch's matrix exponential and autograd.
"""
import json

import numpy as np
import pytest
import torch

import oracle
from oracle import diff_ref

F64 = torch.float64


@pytest.fixture(scope="module")
def tiny():
    from probabilistic_mapping_amd import synth
    s = synth.diff_scene(0, width=64, height=48, n_sources=2, planes=16)
    D = s["planes"]
    s["idep"] = torch.tensor(s["idep_min"] + np.arange(D) * (s["idep_max"] - s["idep_min"])
                             / (D - 1), dtype=F64)
    return s


def _loss(s, xi, ref):
    return diff_ref.forward(xi, ref, s["srcs"], s["K"], s["rel_poses"], s["idep"], s["gt_depth"],
                            s["knobs"], 8, s["temperature"], s["cost_scale"])["loss"]


def test_argmin_takes_first_minimum():
    x = torch.tensor([[3.0, 1.0, 1.0, 2.0], [0.0, 0.0, 0.0, 0.0]], dtype=F64)
    assert torch.argmin(x, dim=1).tolist() == [1, 0]


def test_restatement_matches_cpp_oracle(tiny):
    s = tiny
    D = s["planes"]
    ds = (s["idep_max"] - s["idep_min"]) / (D - 1)
    homs = []
    for R, t in s["rel_poses"]:
        H, th = diff_ref.relative_homography(s["K"], R, t, torch.zeros(6, dtype=F64))
        homs.append((H.numpy().astype(np.float32), th.numpy().astype(np.float32)))
    c_or = oracle.calc_cost(s["ref"], s["srcs"], homs, s["idep_min"], ds, D)
    homs64 = [(torch.tensor(h, dtype=F64).reshape(3, 3), torch.tensor(t, dtype=F64))
              for h, t in homs]
    c_t = diff_ref.calc_cost(torch.tensor(s["ref"], dtype=F64), s["srcs"], homs64, s["idep"])
    assert np.array_equal(c_t.numpy() >= 0, c_or >= 0)
    assert np.abs(c_t.numpy() - c_or).max() <= 1e-9
    assert (c_or >= 0).mean() > 0.9
    a_or = oracle.sgm(s["ref"], c_or, dict(s["knobs"], var_scale=1.0), 8)
    a_t, _ = diff_ref.sgm(torch.tensor(s["ref"], dtype=F64), torch.tensor(c_or), s["knobs"], 8)
    assert np.abs(a_t.numpy() - a_or).max() <= 1e-9 * np.abs(a_or).max()


def test_pose_gradient_equals_central_differences(tiny):
    s = tiny
    ref = torch.tensor(s["ref"], dtype=F64)
    xi = torch.zeros((2, 6), dtype=F64, requires_grad=True)
    _loss(s, xi, ref).backward()
    g = xi.grad.numpy()
    assert np.all(np.abs(g).max(axis=1) > 1.0)
    errs = {}
    with torch.no_grad():
        for h in (1e-4, 1e-6, 1e-8):
            fd = np.zeros_like(g)
            for si in range(2):
                for k in range(6):
                    e = torch.zeros((2, 6), dtype=F64)
                    e[si, k] = h
                    fd[si, k] = (float(_loss(s, e, ref)) - float(_loss(s, -e, ref))) / (2 * h)
            errs[h] = float((np.abs(fd - g) / np.abs(g).max(axis=1, keepdims=True)).max())
    print("[fd] pose:", json.dumps({str(k): v for k, v in errs.items()}), flush=True)
    assert errs[1e-8] <= 1e-6     # converges to the gradient of the piece holding xi = 0
    assert errs[1e-6] <= 1e-2
    assert errs[1e-4] <= 0.1      # kink crossings inside the step


def test_reference_gradient_equals_central_differences(tiny):
    s = tiny
    ref = torch.tensor(s["ref"], dtype=F64, requires_grad=True)
    xi = torch.zeros((2, 6), dtype=F64)
    _loss(s, xi, ref).backward()
    g = ref.grad.numpy()
    img = s["ref"].astype(np.int64)
    H, W = img.shape
    tau = s["knobs"]["tau_so"]

    def at_switch(y, x):   # |I(p) - I(q)| == tau_so for a scanline neighbour q
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                yy, xx = y + dy, x + dx
                if (dy or dx) and 0 <= yy < H and 0 <= xx < W and abs(img[yy, xx] - img[y, x]) == tau:
                    return True
        return False

    order = [int(i) for i in np.argsort(-np.abs(g).ravel())]
    picked = [divmod(i, W) for i in order if not at_switch(*divmod(i, W))][:6]
    h = 1e-4
    worst = 0.0
    with torch.no_grad():
        for y, x in picked:
            rp, rm = ref.detach().clone(), ref.detach().clone()
            rp[y, x] += h
            rm[y, x] -= h
            fd = (float(_loss(s, xi, rp)) - float(_loss(s, xi, rm))) / (2 * h)
            worst = max(worst, abs(fd - g[y, x]) / abs(g[y, x]))
    print(f"[fd] reference intensities: max rel err {worst:.3e}", flush=True)
    assert worst <= 1e-5


@pytest.mark.parametrize("scale", [0.0, 1e-3, 0.05, 0.7])
def test_se3_host_chain(scale):
    from probabilistic_mapping_amd import diff, synth
    rng = np.random.default_rng(int(scale * 1000) + 3)
    K = np.array([[500.0, 0, 320.0], [0, 480.0, 240.0], [0, 0, 1.0]])
    R0 = synth.axis_angle(rng.normal(size=3), 0.1)
    t0 = rng.normal(size=3) * 0.1
    xi = rng.normal(size=6) * scale
    H, t = diff.se3_homography(K, R0, t0, xi)
    Ht, tt = diff_ref.relative_homography(K, R0, t0, torch.tensor(xi, dtype=F64))
    assert np.allclose(H, Ht.numpy(), rtol=1e-12, atol=1e-9)
    assert np.allclose(t, tt.numpy(), rtol=1e-12, atol=1e-12)
    G = rng.normal(size=12)
    x = torch.tensor(xi, dtype=F64, requires_grad=True)
    Hx, tx = diff_ref.relative_homography(K, R0, t0, x)
    f = (Hx.reshape(9) * torch.tensor(G[:9])).sum() + (tx * torch.tensor(G[9:])).sum()
    f.backward()
    got = diff.se3_pose_grad(K, R0, t0, xi, G)
    assert np.allclose(got, x.grad.numpy(), rtol=1e-9, atol=1e-9 * np.abs(x.grad.numpy()).max())


@pytest.mark.parametrize("paths", [4, 8])
def test_oracle_sgm_adjoint_equals_autograd(tiny, paths):
    """The C++ oracle's SGM adjoint (the checker of the GPU's) against torch autograd of the
    restatement, for a random upstream gradient."""
    s = tiny
    D = s["planes"]
    ds = (s["idep_max"] - s["idep_min"]) / (D - 1)
    homs = []
    for R, t in s["rel_poses"]:
        H, th = diff_ref.relative_homography(s["K"], R, t, torch.zeros(6, dtype=F64))
        homs.append((H.numpy().astype(np.float32), th.numpy().astype(np.float32)))
    c = oracle.calc_cost(s["ref"], s["srcs"], homs, s["idep_min"], ds, D)
    cost = torch.tensor(c, requires_grad=True)
    agg, inv = diff_ref.sgm(torch.tensor(s["ref"], dtype=F64), cost, s["knobs"], paths)
    G = np.random.default_rng(3).normal(size=agg.shape)
    (agg * torch.from_numpy(G)).sum().backward()
    got = oracle.sgm_adjoint(s["ref"], c, s["knobs"], paths, G)
    want = cost.grad.numpy()
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
    assert np.abs(want).max() > 0


def test_real_valued_reference_in_the_cpp_oracle(tiny):
    """The C++ oracle with a float64 reference image (used for intensity finite differences
    at full size) equals its uint8 path on integer intensities, and its FP64 central
    differences in one intensity equal the torch restatement's autograd gradient."""
    s = tiny
    K = s["K"]
    homs = []
    for R, t in s["rel_poses"]:
        H, th = diff_ref.relative_homography(K, R, t, torch.zeros(6, dtype=F64))
        homs.append((H.detach().numpy().astype(np.float32), th.detach().numpy().astype(np.float32)))
    D = s["planes"]
    ds = (s["idep_max"] - s["idep_min"]) / (D - 1)
    c8 = oracle.calc_cost(s["ref"], s["srcs"], homs, s["idep_min"], ds, D)
    cf = oracle.calc_cost(s["ref"].astype(np.float64), s["srcs"], homs, s["idep_min"], ds, D)
    assert np.array_equal(c8, cf)
    kn = dict(s["knobs"], var_scale=1.0)   # the readout needs no filter knob
    a8 = oracle.sgm(s["ref"], c8, kn, 8)
    af = oracle.sgm(s["ref"].astype(np.float64), cf, kn, 8)
    assert np.array_equal(a8, af)
    # FD through the C++ oracle against the restatement's gradient at the same float-rounded
    # homographies, on the pixel of largest gradient away from the tau_so switch
    ref = torch.tensor(s["ref"], dtype=F64, requires_grad=True)
    _loss(s, torch.zeros((2, 6), dtype=F64), ref).backward()
    g = ref.grad.numpy()
    idep = s["idep"]

    def cpp_loss(refd):
        cost = oracle.calc_cost(refd, s["srcs"], homs, s["idep_min"], ds, D)
        agg = oracle.sgm(refd, cost, kn, 8)
        inv = torch.from_numpy((cost < 0).any(axis=2))
        _, loss, _ = diff_ref.readout(torch.from_numpy(agg), inv, idep,
                                      s["gt_depth"].astype(np.float64), s["temperature"],
                                      s["cost_scale"])
        return float(loss)

    img = s["ref"].astype(np.int64)
    tau = s["knobs"]["tau_so"]
    Hh, Ww = img.shape
    for i in np.argsort(-np.abs(g).ravel()):
        y, x = divmod(int(i), Ww)
        nb = img[max(y - 1, 0):y + 2, max(x - 1, 0):x + 2]
        if not (np.abs(nb - img[y, x]) == tau).any():
            break
    h = 1e-4
    rp, rm = s["ref"].astype(np.float64), s["ref"].astype(np.float64)
    rp[y, x] += h
    rm[y, x] -= h
    fd = (cpp_loss(rp) - cpp_loss(rm)) / (2 * h)
    assert abs(fd - g[y, x]) <= 1e-5 * abs(g[y, x])
