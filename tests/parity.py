"""Parity metrics between the GPU path and the FP64 oracle.

Bars (stated where they are asserted):
  * discrete decisions (validity of a cost entry, WTA plane index, rejection category of a
    depth pixel) must agree on >= 99.9 % of elements;
  * continuous values are compared as relative error on elements where both sides made the
    same discrete decision.

It also holds a float32 replay of one SGM scanline (``replay_scanline``) that the GPU tests
compare the stored per-path volumes with bit for bit; tests/test_oracle_sgm.py pins the replay
against the oracle's aggregation.
"""
from __future__ import annotations

import numpy as np


def rel_err(a, b, floor=0.0):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.abs(a - b) / np.maximum(np.abs(b), floor)


def summary(x) -> dict:
    x = np.asarray(x, np.float64).ravel()
    if x.size == 0:
        return {"n": 0, "max": 0.0, "p999": 0.0, "p9999": 0.0, "rms": 0.0}
    return {"n": int(x.size), "max": float(x.max()), "p999": float(np.quantile(x, 0.999)),
            "p9999": float(np.quantile(x, 0.9999)), "rms": float(np.sqrt(np.mean(x * x)))}


# Entries whose oracle validity margin (oracle.calc_cost(margin=True)) is below this many
# pixels may change their set of contributing sources under FP32 warp rounding: a source
# counted on one side only moves the mean by a whole term, and an entry can flip between -1
# and a value.  The FP32 warp positions are good to ~1e-4 px at |U| <= 1280.
MARGIN_PX = 1e-2


def cost_parity(gpu, orc, floor=1.0, margin=None) -> dict:
    """Cost volumes: validity (-1) agreement and relative error (floor in cost units).  With
    the oracle's validity margin, entries within ``MARGIN_PX`` of a change of their source
    set are counted (``near_boundary``) and kept out of ``rel``; ``validity_mismatch_far``
    counts validity disagreements away from that boundary (expected: none)."""
    gv, ov = gpu >= 0, orc >= 0
    both = gv & ov
    out = {"validity_mismatch": float((gv != ov).mean())}
    if margin is not None:
        near = margin < MARGIN_PX
        out["near_boundary"] = float(near.mean())
        out["validity_mismatch_far"] = int(((gv != ov) & ~near).sum())
        out["rel_near"] = summary(rel_err(gpu[both & near], orc[both & near], floor))
        both = both & ~near
    out["rel"] = summary(rel_err(gpu[both], orc[both], floor))
    return out


def wta_index(agg):
    return np.argmin(agg, axis=2)


def depth_category(d):
    """0 = rejected (0), 1 = non-overlap (DEP_INF), 2 = a depth value."""
    d = np.asarray(d)
    cat = np.full(d.shape, 2, np.int8)
    cat[d == 0] = 0
    cat[d >= 1000.0] = 1
    return cat


def depth_parity(gpu, orc) -> dict:
    cg, co = depth_category(gpu), depth_category(orc)
    same = (cg == co) & (co == 2)
    return {"category_mismatch": float((cg != co).mean()),
            "valid_frac": float((co == 2).mean()),
            "rel": summary(rel_err(gpu[same], orc[same]))}


def state_parity(gs: dict, os_: dict, mask=None) -> dict:
    """Filter state: pixels initialised on both sides (alpha >= 1e-5), and inside ``mask`` if
    given, must match; compares mean, variance and the inlier ratio alpha / (alpha + beta).
    ``init_mismatch`` is the fraction of the pixels inside ``mask`` initialised on one side
    only."""
    ga, oa = gs["alpha"], os_["alpha"]
    gi, oi = ga >= 1e-5, oa >= 1e-5
    both = gi & oi
    sel = np.ones_like(gi) if mask is None else np.asarray(mask, bool)
    both = both & sel
    out = {"init_mismatch": float((gi != oi)[sel].mean()) if sel.any() else 0.0,
           "initialised": float(oi.mean()), "compared": float(both.mean())}
    for k in ("mu", "sig"):
        out[k] = summary(rel_err(gs[k][both], os_[k][both]))
    with np.errstate(invalid="ignore", divide="ignore"):
        gr = ga / (ga + gs["beta"])
        orr = oa / (oa + os_["beta"])
    out["inlier"] = summary(rel_err(gr[both], orr[both]))
    return out


class History:
    """Which pixels' fusion history matches the oracle's, step by step.

    The fused state of a pixel depends on discrete choices made at earlier steps: the
    measurement's category (rejected / non-overlap / value) and WTA plane, those of the
    pixels in its median window, and which previous pixel the forward warp and hole fill hand
    it (the propagation table).  A pixel *diverges* at a step when its table entry differs
    from the oracle's, when its measurement was made from different discrete decisions, or
    when the source pixel it gathers had already diverged; it stays diverged while later steps
    gather from diverged sources.  Continuous parity bars apply to the pixels that have not
    diverged, and the diverged fraction is a discrete-decision rate reported beside them.
    """

    def __init__(self, shape):
        self.div = np.zeros(shape, bool)

    def step(self, g_table, o_table, undecided):
        """Advance one fusion step.  Tables: int32 linear source index, 0 = none.
        ``undecided``: pixels whose measurement came from different discrete decisions."""
        ot = np.asarray(o_table).ravel()
        prev = self.div.ravel()
        inherited = np.where(ot != 0, prev[ot], False)
        div = (np.asarray(g_table).ravel() != ot) | inherited
        div |= np.asarray(undecided, bool).ravel()
        self.div = div.reshape(self.div.shape)
        return self.div

    def reset(self, undecided):
        """The first update only initialises: every pixel starts a new history."""
        self.div = np.asarray(undecided, bool).copy()


def dilate3(mask):
    """3x3 dilation of a boolean image (a pixel is set if any pixel of its window is)."""
    m = np.asarray(mask, bool)
    p = np.pad(m, 1)
    out = np.zeros_like(m)
    H, W = m.shape
    for dy in range(3):
        for dx in range(3):
            out |= p[dy:dy + H, dx:dx + W]
    return out


def subpixel_conditioning(g_agg, o_agg, g_dep, o_dep, idep_min, dep_sample, tol=2e-6):
    """Checks every pixel where both sides kept a depth value with the same WTA index against
    the parabola's conditioning.  The offset is s = m - (c+ - c-) / (2 a), a = c- - 2 c0 + c+
    (filter_cost.cu:45-51).  If the three aggregates carry errors of at most delta, then to
    first order |ds| <= (delta / a) (1 + 2 |c+ - c-| / a) <= 3 delta / a, because an accepted
    pixel has |s - m| <= 1/2, i.e. |c+ - c-| <= a.  With depth = 1 / (idep_min + s dep_sample)
    the relative depth error is then at most 3 (delta / a) dep_sample depth.  delta is measured
    per pixel (GPU minus FP64 aggregate on the three planes); pixels with a <= 4 delta are
    outside the first-order regime (the aggregate error is a quarter of the curvature: the
    rule-3 decision itself is in play) and are only counted."""
    ga = np.asarray(g_agg, np.float64)
    oa = np.asarray(o_agg, np.float64)
    D = oa.shape[2]
    m = np.argmin(oa, axis=2)
    same = (np.argmin(ga, axis=2) == m) & (depth_category(g_dep) == 2) & \
        (depth_category(o_dep) == 2) & (m > 0) & (m < D - 1)
    yy, xx = np.nonzero(same)
    mm = m[yy, xx]
    oc = [oa[yy, xx, mm + k] for k in (-1, 0, 1)]
    gc = [ga[yy, xx, mm + k] for k in (-1, 0, 1)]
    delta = np.max([np.abs(g - o) for g, o in zip(gc, oc)], axis=0)
    a = oc[0] - 2 * oc[1] + oc[2]
    od = np.asarray(o_dep, np.float64)[yy, xx]
    gd = np.asarray(g_dep, np.float64)[yy, xx]
    rel = np.abs(gd - od) / od
    ok = a > 4 * delta
    bound = 3 * delta[ok] / a[ok] * dep_sample * od[ok] + tol
    ratio = rel[ok] / bound
    return {"checked": int(ok.sum()), "ill_conditioned": int((~ok).sum()),
            "violations": int((ratio > 1).sum()),
            "max_ratio": float(ratio.max()) if ratio.size else 0.0,
            "rel_max_checked": float(rel[ok].max()) if ok.any() else 0.0,
            "rel_max_ill": float(rel[~ok].max()) if (~ok).any() else 0.0}


def median_lipschitz(g_raw, o_raw, g_med, o_med, decided):
    """The 3x3 median picks an order statistic of the valid values in its window, a
    1-Lipschitz function of them in the max norm.  So on pixels whose window made the same
    discrete decisions on both sides (``decided``: no category or WTA difference in the 3x3
    window), |median_gpu - median_oracle| <= max over the window of |raw_gpu - raw_oracle|.
    Returns the number of pixels checked and of violations (beyond a 1e-6 relative slack)."""
    gr, orr = np.asarray(g_raw, np.float64), np.asarray(o_raw, np.float64)
    err = np.where((depth_category(gr) == 2) & (depth_category(orr) == 2), np.abs(gr - orr), 0.0)
    p = np.pad(err, 1)
    H, W = err.shape
    wmax = np.max([p[dy:dy + H, dx:dx + W] for dy in range(3) for dx in range(3)], axis=0)
    gm, om = np.asarray(g_med, np.float64), np.asarray(o_med, np.float64)
    chk = decided & (depth_category(om) == 2)
    viol = np.abs(gm[chk] - om[chk]) > wmax[chk] + 1e-6 * np.abs(om[chk])
    return {"checked": int(chk.sum()), "violations": int(viol.sum())}


def excess_step(knobs):
    """Level step of the compact SGM storage: max(P2, P2 q2) / 65534 in float32, so that an
    excess up to the largest P2 of the knobs fits the 16-bit levels 0..65534."""
    p2 = max(np.float32(knobs["pi2"]), np.float32(knobs["pi2"]) * np.float32(knobs["sgm_q2"]))
    return float(np.float32(p2) / np.float32(65534.0))


# The 8 path directions in aggregation order: L->R, R->L, T->B, B->T (sgm.cu:119-131), then
# the diagonals.
PATH_DX = (1, -1, 0, 0, 1, -1, 1, -1)
PATH_DY = (0, 0, 1, -1, 1, -1, -1, 1)


def scanline(W, H, dx, dy, i):
    """Pixels of scanline i of direction (dx, dy), in step order (the lines of a direction:
    rows or columns, or for the diagonals the lines from the left or right edge, then from the
    top or bottom edge)."""
    if dx != 0:
        if i < H:
            x, y = (0 if dx > 0 else W - 1), i
        else:
            j = i - H
            x, y = (1 + j if dx > 0 else j), (0 if dy > 0 else H - 1)
    else:
        x, y = i, (0 if dy > 0 else H - 1)
    pts = []
    while 0 <= x < W and 0 <= y < H:
        pts.append((x, y))
        x, y = x + dx, y + dy
    return pts


def replay_scanline(c, ref, pts, dx, dy, knobs, step, want_l=False):
    """The excess levels of one scanline, replayed in float32 with the operations of the GPU
    recurrence in the same order (sgm.cu:75-98): m = min_d L; val = min(L(d), m + P2,
    min(L(d-1), L(d+1)) + P1); L' = (C' + val) - m with C' = 0 at a pixel with a negative cost;
    level = E / step rounded to the nearest integer, ties to even (one rounding of the exact
    quotient), with E = val - m, at most 65534; 0 at the first step."""
    f32 = np.float32
    P1n, P2n = f32(knobs["pi1"]), f32(knobs["pi2"])
    P1s, P2s = f32(P1n * f32(knobs["sgm_q1"])), f32(P2n * f32(knobs["sgm_q2"]))
    tau = int(np.ceil(knobs["tau_so"]))
    inv = f32(1.0) / f32(step) if step > 0 else f32(0.0)
    D = c.shape[2]
    out = np.zeros((len(pts), D), np.uint16)
    lout = np.zeros((len(pts), D), f32)   # L, 0 at invalid pixels (the exact form's volume)
    L = None
    for k, (x, y) in enumerate(pts):
        cc = c[y, x]
        invalid = bool((cc < 0).any())
        cw = np.zeros(D, f32) if invalid else cc
        if k == 0:
            L = cw.copy()
            lout[k] = 0.0 if invalid else L
            continue
        m = L.min()
        smooth = abs(int(ref[y, x]) - int(ref[y - dy, x - dx])) < tau
        P1, P2 = (P1s, P2s) if smooth else (P1n, P2n)
        nb = np.full(D, np.inf, f32)
        nb[1:] = L[:-1]
        nb[:-1] = np.minimum(nb[:-1], L[1:])
        val = np.minimum(np.minimum(L, f32(m + P2)), (nb + P1).astype(f32))
        e = (val - m).astype(f32)
        q = np.rint(e.astype(np.float64) * np.float64(inv))   # exact product, one rounding
        out[k] = np.minimum(65534, q).astype(np.uint16)
        L = ((cw + val).astype(f32) - m).astype(f32)
        lout[k] = 0.0 if invalid else L
    return (out, lout) if want_l else out
