"""FP64 restatement of the differentiable depth pipeline in torch (CPU) -- TEST INFRASTRUCTURE ONLY.

Used by the tests as the reference gradient for the GPU adjoint (torch float64 autograd through
this forward) and for FP64 central finite differences.  It is never called by the product.

Forward, term for term as the C++ oracle (pm_oracle.hpp) restates the reference:
  * cost: ADCalcCostKernel, motion_stereo/calc_cost.cu:11-210 -- nine warp triples per pixel
    with the garbled right / up-right warps (calc_cost.cu:49-51, 69-71) and swapped up/down
    reference samples (calc_cost.cu:106, 119); a source counts only if all nine warps are valid
    (calc_cost.cu:89-91); exact bilinear sampling with border 0; mean over counting sources,
    -1 where none counts (stereo_mapper.cpp:69, calc_cost.cu:200-207);
  * SGM: sgm2<>, motion_stereo/sgm.cu:5-113, on 4 or 8 paths; invalid pixels (any plane < 0)
    use cost 0 and get aggregate 0 (sgm.cu:33-38, 101-104); penalties scaled by q1, q2 where
    |I(p) - I(q)| < tau_so (sgm.cu:75-82);
  * readout (no reference counterpart): depth_hat = sum_d softmax(-agg / (cost_scale T)) / idep_d,
    loss = mean over M of |depth_hat - gt|, M = valid pixels with gt > 0.
Discrete choices are made explicit so autograd routes gradients the way the GPU adjoint does:
every min is an argmin (first minimum among [L(p,d), L(p,d-1)+P1, L(p,d+1)+P1, m_p+P2]) plus a
gather, m_p = L(p, argmin_d) (lowest index), |x|' = sign(x) with sign(0) = 0, and the bilinear
cell is that of floor(U), floor(V).

Pose convention: xi = (rho, phi), (R, t) = Exp(xi) (R0, t0) with Exp the matrix exponential of
the twist [[phi^, rho], [0, 0]]; H = K R K^-1, t_h = K t.
"""
from __future__ import annotations

import torch

PATH_DX = (1, -1, 0, 0, 1, -1, 1, -1)
PATH_DY = (0, 0, 1, -1, 1, -1, -1, 1)
F64 = torch.float64


def hat(v):
    z = torch.zeros((), dtype=v.dtype)
    return torch.stack([torch.stack([z, -v[2], v[1]]), torch.stack([v[2], z, -v[0]]),
                        torch.stack([-v[1], v[0], z])])


def se3_exp(xi):
    """4x4 Exp of xi = (rho, phi) through the matrix exponential of the twist."""
    top = torch.cat([hat(xi[3:]), xi[:3, None]], 1)
    T = torch.cat([top, torch.zeros((1, 4), dtype=xi.dtype)], 0)
    return torch.linalg.matrix_exp(T)


def relative_homography(K, R0, t0, xi):
    """(H, t_h) of the pose Exp(xi) (R0, t0), in FP64."""
    K = torch.as_tensor(K, dtype=F64)
    E = se3_exp(xi)
    R = E[:3, :3] @ torch.as_tensor(R0, dtype=F64)
    t = E[:3, :3] @ torch.as_tensor(t0, dtype=F64).reshape(3) + E[:3, 3]
    return K @ R @ torch.linalg.inv(K), K @ t


def _bilinear(img_pad, W, H, X, Y, w):
    """Sample of a border-0 image at (X/w, Y/w) and the warp's validity; img_pad is the image
    with one zero row and column appended."""
    with torch.no_grad():
        U0, V0 = X / w, Y / w
        valid = (w >= 0) & (U0 >= 0) & (U0 <= W - 1) & (V0 >= 0) & (V0 <= H - 1)
    one = torch.ones((), dtype=F64)
    ws = torch.where(valid, w, one)
    U = torch.where(valid, X, 0 * one) / ws
    V = torch.where(valid, Y, 0 * one) / ws
    x0 = torch.floor(U).detach()
    y0 = torch.floor(V).detach()
    ax, ay = U - x0, V - y0
    xi, yi = x0.long(), y0.long()
    i00 = img_pad[yi, xi]
    i10 = img_pad[yi, xi + 1]
    i01 = img_pad[yi + 1, xi]
    i11 = img_pad[yi + 1, xi + 1]
    s = (i00 + ax * (i10 - i00)) + ay * ((i01 - i00) + ax * (i11 - i10 - i01 + i00))
    return s, valid


def calc_cost(ref, srcs, homs, idep):
    """ref: [H, W] float64 (may require grad); srcs: list of [H, W] images; homs: list of
    (H 3x3, t 3) float64 tensors; idep: [D].  Returns cost [H, W, D] (-1 = no source)."""
    H, W = ref.shape
    D = idep.shape[0]
    rp = torch.nn.functional.pad(ref, (1, 1, 1, 1))

    def R(dx, dy):   # reference intensity at (u+dx, v+dy), border 0
        return rp[1 + dy:1 + dy + H, 1 + dx:1 + dx + W]

    r = [R(0, 0), R(0, 1), R(0, -1), R(-1, 0), R(1, 0), R(-1, -1), R(1, 1), R(-1, 1), R(1, -1)]
    u = torch.arange(W, dtype=F64)[None, :]
    v = torch.arange(H, dtype=F64)[:, None]
    total = torch.zeros((D, H, W), dtype=F64)
    count = torch.zeros((D, H, W), dtype=F64)
    for img, (Hm, tv) in zip(srcs, homs):
        ip = torch.zeros((H + 1, W + 1), dtype=F64)
        ip[:H, :W] = torch.as_tensor(img, dtype=F64)
        h11, h12, h13, h21, h22, h23, h31, h32, h33 = Hm.reshape(9)
        x = h11 * u + h12 * v + h13
        y = h21 * u + h22 * v + h23
        z = h31 * u + h32 * v + h33
        xr = x + h11
        X = [x, x - h12, x + h12, x - h11, xr, x - h12 - h11, x + h12 + h11, x - h11 + h12,
             xr - h12]
        Y = [y, y - h22, y + h22, y - h21, x + h21, y - h22 - h21, y + h22 + h21, y - h21 + h22,
             xr - h22]
        Z = [z, z - h32, z + h32, z - h31, x + h31, z - h32 - h31, z + h32 + h31, z - h31 + h32,
             xr - h32]
        idd = idep[:, None, None]
        tmp = torch.zeros((D, H, W), dtype=F64)
        ok = torch.ones((D, H, W), dtype=torch.bool)
        for k in range(9):
            w = Z[k][None] + tv[2] * idd
            s, valid = _bilinear(ip, W, H, X[k][None] + tv[0] * idd, Y[k][None] + tv[1] * idd, w)
            tmp = tmp + torch.abs(r[k][None] - s)
            ok = ok & valid
        total = total + torch.where(ok, tmp / 9.0, torch.zeros((), dtype=F64))
        count = count + ok.to(F64)
    n = torch.where(count > 0, count, torch.ones((), dtype=F64))
    cost = torch.where(count > 0, total / n, -torch.ones((), dtype=F64))
    return cost.permute(1, 2, 0)


def _step(Cp, Lp, P1, P2):
    """One SGM step for a batch of pixels: Cp [n, D] costs, Lp [n, D] predecessor L,
    P1, P2 [n] penalties.  Explicit first-minimum routing."""
    inf = torch.full((Lp.shape[0], 1), float("inf"), dtype=F64)
    im = torch.argmin(Lp, dim=1, keepdim=True)
    m = torch.gather(Lp, 1, im)
    c0 = Lp
    c1 = torch.cat([inf, Lp[:, :-1]], 1) + P1[:, None]
    c2 = torch.cat([Lp[:, 1:], inf], 1) + P1[:, None]
    c3 = (m + P2[:, None]).expand_as(Lp)
    cand = torch.stack([c0, c1, c2, c3], 2)
    idx = torch.argmin(cand, dim=2, keepdim=True)
    val = torch.gather(cand, 2, idx)[:, :, 0]
    return Cp + val - m


def sgm(ref, cost, knobs, paths):
    """Aggregate [H, W, D] of `paths` scanline directions (first-minimum routing)."""
    H, W, D = cost.shape
    inv = (cost < 0).any(dim=2)
    Cp = torch.where(inv[:, :, None], torch.zeros((), dtype=F64), cost)
    refd = ref.detach().to(F64)
    pi1, pi2, tau = float(knobs["pi1"]), float(knobs["pi2"]), float(knobs["tau_so"])
    q1, q2 = float(knobs["sgm_q1"]), float(knobs["sgm_q2"])

    def pen(Ia, Ib):
        sm = torch.abs(Ia - Ib) < tau
        P1 = torch.where(sm, torch.tensor(pi1 * q1, dtype=F64), torch.tensor(pi1, dtype=F64))
        P2 = torch.where(sm, torch.tensor(pi2 * q2, dtype=F64), torch.tensor(pi2, dtype=F64))
        return P1, P2

    agg = None
    for path in range(paths):
        dx, dy = PATH_DX[path], PATH_DY[path]
        if dy == 0:   # march over columns
            order = range(W) if dx > 0 else range(W - 1, -1, -1)
            cols, prev = {}, None
            for x in order:
                if prev is None:
                    L = Cp[:, x, :]
                else:
                    P1, P2 = pen(refd[:, x], refd[:, x - dx])
                    L = _step(Cp[:, x, :], prev, P1, P2)
                cols[x] = L
                prev = L
            Lr = torch.stack([cols[x] for x in range(W)], 1)
        else:          # march over rows; the predecessor of (y, x) is (y - dy, x - dx)
            order = range(H) if dy > 0 else range(H - 1, -1, -1)
            rows, prev = {}, None
            xs = torch.arange(W)
            has = ((xs - dx) >= 0) & ((xs - dx) < W)
            src = (xs - dx).clamp(0, W - 1)
            for y in order:
                if prev is None:
                    L = Cp[y]
                else:
                    Lp = prev[src]
                    P1, P2 = pen(refd[y], refd[y - dy][src])
                    Ls = _step(Cp[y], Lp, P1, P2)
                    L = torch.where(has[:, None], Ls, Cp[y])
                rows[y] = L
                prev = L
            Lr = torch.stack([rows[y] for y in range(H)], 0)
        agg = Lr if agg is None else agg + Lr
    return torch.where(inv[:, :, None], torch.zeros((), dtype=F64), agg), inv


def readout(agg, inv, idep, gt, temperature, cost_scale, signs=None):
    """Soft-argmin depth [H, W] and the masked mean L1 loss.

    ``signs`` (optional, [H, W] of +-1/0) fixes the branch of |depth_hat - gt| per pixel:
    loss = mean over M of signs (depth_hat - gt), the L1 loss on the piece where
    sign(depth_hat - gt) = signs.  With the signs of another forward pass (the FP32 GPU one)
    this differentiates the same branch that pass differentiates; pixels whose depth lies
    within rounding of gt otherwise flip the whole pixel's gradient."""
    z = -agg / (cost_scale * temperature)
    sig = torch.softmax(z, dim=2)
    dh = (sig * (1.0 / idep)[None, None, :]).sum(2)
    gt = torch.as_tensor(gt, dtype=F64)
    M = (~inv) & (gt > 0) & torch.isfinite(gt)
    n = int(M.sum())
    if signs is None:
        loss = torch.abs(dh - gt)[M].sum() / max(n, 1)
    else:
        sg = torch.as_tensor(signs, dtype=F64)
        loss = (sg * (dh - gt))[M].sum() / max(n, 1)
    return dh, loss, M


def forward(xi, ref, srcs, K, poses, idep, gt, knobs, paths, temperature, cost_scale,
            signs=None):
    """The whole chain: xi [S, 6], ref [H, W] (float64 tensors; either may require grad)."""
    homs = [relative_homography(K, R0, t0, xi[s]) for s, (R0, t0) in enumerate(poses)]
    cost = calc_cost(ref, srcs, homs, idep)
    agg, inv = sgm(ref, cost, knobs, paths)
    dh, loss, M = readout(agg, inv, idep, gt, temperature, cost_scale, signs)
    return {"cost": cost, "agg": agg, "depth": dh, "loss": loss, "mask": M}
