"""Parity metrics between the GPU path and the FP64 oracle.

Bars (stated where they are asserted):
  * discrete decisions (validity of a cost entry, WTA plane index, rejection category of a
    depth pixel) must agree on >= 99.9 % of elements;
  * continuous values are compared as relative error on elements where both sides made the
    same discrete decision.
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
        return {"n": 0, "max": 0.0, "p999": 0.0, "rms": 0.0}
    return {"n": int(x.size), "max": float(x.max()), "p999": float(np.quantile(x, 0.999)),
            "rms": float(np.sqrt(np.mean(x * x)))}


def cost_parity(gpu, orc, floor=1.0) -> dict:
    """Cost volumes: validity (-1) agreement and relative error (floor in cost units)."""
    gv, ov = gpu >= 0, orc >= 0
    both = gv & ov
    return {"validity_mismatch": float((gv != ov).mean()),
            "rel": summary(rel_err(gpu[both], orc[both], floor))}


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


def state_parity(gs: dict, os_: dict) -> dict:
    """Filter state: pixels initialised on both sides (alpha >= 1e-5) must match; compares
    mean, variance and the inlier ratio alpha / (alpha + beta)."""
    ga, oa = gs["alpha"], os_["alpha"]
    gi, oi = ga >= 1e-5, oa >= 1e-5
    both = gi & oi
    out = {"init_mismatch": float((gi != oi).mean()), "initialised": float(oi.mean())}
    for k in ("mu", "sig"):
        out[k] = summary(rel_err(gs[k][both], os_[k][both]))
    gr = ga / (ga + gs["beta"])
    orr = oa / (oa + os_["beta"])
    out["inlier"] = summary(rel_err(gr[both], orr[both]))
    return out
