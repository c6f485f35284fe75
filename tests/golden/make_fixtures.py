"""Writes the golden fixtures under tests/golden/ from closed forms derived by hand.

Nothing here calls the oracle or the library: each fixture is an input plus the output that
the mathematics fixes, so the oracle is checked against it rather than against itself.

* cost_identity.json  -- plane-sweep cost with H = I, t = 0 and the measurement frame equal to
  the reference.  Working the reference's 9 warps through by hand (calc_cost.cu:28-198):
  terms 0, 3, 5, 6, 7 sample their own reference texel (difference 0); terms 1 and 2 warp
  (u, v-1) / (u, v+1) but read the reference at (u, v+1) / (u, v-1), each giving
  |I(u,v+1) - I(u,v-1)|; the garbled right warp (x+h11, x+h21, x+h31) = (u+1, u, u) lands at
  (1 + 1/u, 1) and the garbled up-right warp (xr-h12, xr-h22, xr-h32) = (u+1, u, u+1) at
  (1, u/(u+1)).  All 9 are valid only for 1 <= u <= W-2, 1 <= v <= H-2.  So
      cost = (2|I(u,v+1) - I(u,v-1)| + |I(u+1,v) - B(1+1/u, 1)| + |I(u+1,v-1) - B(1, u/(u+1))|)/9
  for every plane inside, and -1 elsewhere; B is exact bilinear, evaluated with fractions.
* filter_parabola.json -- aggregate vectors built as exact parabolas a (d - s)^2 + b, plus
  the three rejection rules, with the depth the fit must return.
* fusion_posterior.json -- the Gaussian x Beta update as a posterior-moment problem: the
  moments E[z], Var[z], E[pi], E[pi^2] of
      p(z, pi | d) ~ [pi N(d; z, tau^2) + (1 - pi) / 50] N(z; mu, s2) Beta(pi; a, b)
  computed by brute-force 2-D quadrature.  The Beta(a', b') matching E[pi], E[pi^2] and
  (mu', s2') = (E[z], Var[z]) are what depthUpdate must produce (depth_fusion.cu:195-216).
* fusion_posterior_cpu.json -- the same quadrature for the CPU rules of outputFusionCPU
  (stereo_mapper.cpp:439-545): tau^2 = 64, outlier density 1/16, and an inlier weight that
  normpdfCPU MULTIPLIES by sqrt(2 pi (tau^2 + s2)) instead of dividing, i.e. the inlier
  integral N(d; mu, tau^2 + s2) scaled by 2 pi (tau^2 + s2).  The moments are matched the same
  way; the output gate is on the posterior inlier ratio (> 0.51).
* hole_fill.json -- a hand-made 6x5 propagation table and its filled form under the
  deterministic rule (empty pixels take the entry of their lowest-linear-index non-empty
  interior 3x3 neighbour, read from the pre-fill table), applied by hand here.
"""
from __future__ import annotations

import json
import os
from fractions import Fraction

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def _bilinear_exact(img, U: Fraction, V: Fraction) -> Fraction:
    H, W = img.shape

    def tx(x, y):
        return Fraction(int(img[y, x])) if 0 <= x < W and 0 <= y < H else Fraction(0)
    x0, y0 = U.numerator // U.denominator, V.numerator // V.denominator
    ax, ay = U - x0, V - y0
    return ((1 - ax) * (1 - ay) * tx(x0, y0) + ax * (1 - ay) * tx(x0 + 1, y0) +
            (1 - ax) * ay * tx(x0, y0 + 1) + ax * ay * tx(x0 + 1, y0 + 1))


def cost_identity():
    rng = np.random.default_rng(20240611)
    cases = []
    for (H, W) in [(6, 7), (5, 9)]:
        img = rng.integers(0, 256, size=(H, W), dtype=np.uint8)
        cost = np.full((H, W), -1.0)
        for v in range(H):
            for u in range(W):
                if not (1 <= u <= W - 2 and 1 <= v <= H - 2):
                    continue
                I = lambda x, y: Fraction(int(img[y, x]))  # noqa: E731
                c = 2 * abs(I(u, v + 1) - I(u, v - 1))
                c += abs(I(u + 1, v) - _bilinear_exact(img, 1 + Fraction(1, u), Fraction(1)))
                c += abs(I(u + 1, v - 1) - _bilinear_exact(img, Fraction(1), Fraction(u, u + 1)))
                cost[v, u] = float(c / 9)
        cases.append({"height": H, "width": W, "image": img.tolist(), "planes": 5,
                      "idep_min": 0.1, "dep_sample": 0.05, "cost": cost.tolist()})
    return cases


def filter_parabola():
    D = 16
    idep_min, ds, vs = 0.125, 0.03, 1.02
    cases = []
    for s, a, b in [(5.3, 2.0, 40.0), (7.75, 0.5, 12.0), (10.2, 9.0, 300.0), (3.5, 1.0, 1.0)]:
        c = [a * (d - s) ** 2 + b for d in range(D)]
        cases.append({"agg": c, "expect": 1.0 / (idep_min + s * ds), "rule": "parabola"})
    # rule 1: a minimum of 0 (non-overlap)
    cases.append({"agg": [5.0] * 7 + [0.0] + [5.0] * 8, "expect": 1000.0, "rule": "min~0"})
    # rule 2: argmin on the first or last plane
    cases.append({"agg": [1.0 + d for d in range(D)], "expect": 0.0, "rule": "first"})
    cases.append({"agg": [float(D - d) for d in range(D)], "expect": 0.0, "rule": "last"})
    # rule 3: curvature too flat: neighbours sum to less than 2 c0 var_scale
    flat = [9.0] * D
    flat[6], flat[7], flat[8] = 10.0, 10.0 - 0.01, 10.0
    cases.append({"agg": flat, "expect": 0.0, "rule": "flat"})
    # tie: lowest index wins (then the fit runs around it)
    tie = [float((d - 4) ** 2 + 3) for d in range(D)]
    tie[9] = 3.0
    cases.append({"agg": tie, "expect": 1.0 / (idep_min + 4 * ds), "rule": "tie"})
    return {"planes": D, "idep_min": idep_min, "dep_sample": ds, "var_scale": vs,
            "cases": cases}


def fusion_posterior():
    out = []
    cases = [(10.0, 10.0, 2.0, 0.09, 2.05), (10.0, 10.0, 2.0, 0.09, 2.6), (25.0, 4.0, 3.1, 0.02, 3.0),
             (6.0, 9.0, 1.2, 0.5, 1.5), (40.0, 2.0, 5.0, 0.3, 5.2)]
    z = np.linspace(-30, 60, 180001)
    dz = z[1] - z[0]
    p = np.linspace(0, 1, 200001)[1:-1]
    dp = p[1] - p[0]
    for a, b, mu, s2, d in cases:
        tau2 = 0.01 * d * d
        gz = np.exp(-(z - mu) ** 2 / (2 * s2)) / np.sqrt(2 * np.pi * s2)
        like = np.exp(-(d - z) ** 2 / (2 * tau2)) / np.sqrt(2 * np.pi * tau2)
        # integrate z numerically: A = int N(d;z,tau2) N(z;mu,s2) dz, moments of z under it
        A0 = np.sum(like * gz) * dz
        A1 = np.sum(z * like * gz) * dz
        A2 = np.sum(z * z * like * gz) * dz
        B0 = np.sum(gz) * dz / 50.0
        B1 = np.sum(z * gz) * dz / 50.0
        B2 = np.sum(z * z * gz) * dz / 50.0
        logbeta = (a - 1) * np.log(p) + (b - 1) * np.log(1 - p)
        w = np.exp(logbeta - logbeta.max())
        w /= np.sum(w) * dp
        E = lambda f: np.sum(f * w) * dp  # noqa: E731  (expectation under the Beta prior)
        Z = E(p) * A0 + E(1 - p) * B0
        Ez = (E(p) * A1 + E(1 - p) * B1) / Z
        Ez2 = (E(p) * A2 + E(1 - p) * B2) / Z
        Epi = (E(p * p) * A0 + E(p * (1 - p)) * B0) / Z
        Epi2 = (E(p ** 3) * A0 + E(p * p * (1 - p)) * B0) / Z
        # Beta(a', b') with mean Epi and second moment Epi2
        m, m2 = Epi, Epi2
        ab = (m - m2) / (m2 - m * m)
        out.append({"prior": [a, b, mu, s2], "measurement": d,
                    "posterior": {"mu": Ez, "sig": Ez2 - Ez * Ez, "alpha": m * ab,
                                  "beta": (1 - m) * ab, "inlier": m}})
    return out


def fusion_posterior_cpu():
    out = []
    cases = [(10.0, 10.0, 2.0, 256.0, 2.5), (10.0, 10.0, 2.0, 4.0, 2.05), (25.0, 4.0, 3.1, 0.5, 3.0),
             (9.0, 6.0, 1.2, 30.0, 7.5), (40.0, 2.0, 5.0, 1.0, 9.9)]
    p = np.linspace(0, 1, 200001)[1:-1]
    dp = p[1] - p[0]
    for a, b, mu, s2, d in cases:
        tau2 = 64.0
        span = 12 * np.sqrt(max(s2, tau2))
        z = np.linspace(mu - span, mu + span, 400001)
        dz = z[1] - z[0]
        gz = np.exp(-(z - mu) ** 2 / (2 * s2)) / np.sqrt(2 * np.pi * s2)
        like = np.exp(-(d - z) ** 2 / (2 * tau2)) / np.sqrt(2 * np.pi * tau2)
        kappa = 2 * np.pi * (tau2 + s2)           # normpdfCPU's sqrt factor, twice
        A0 = kappa * np.sum(like * gz) * dz
        A1 = kappa * np.sum(z * like * gz) * dz
        A2 = kappa * np.sum(z * z * like * gz) * dz
        B0 = np.sum(gz) * dz / 16.0
        B1 = np.sum(z * gz) * dz / 16.0
        B2 = np.sum(z * z * gz) * dz / 16.0
        logbeta = (a - 1) * np.log(p) + (b - 1) * np.log(1 - p)
        w = np.exp(logbeta - logbeta.max())
        w /= np.sum(w) * dp
        E = lambda f: np.sum(f * w) * dp  # noqa: E731
        Z = E(p) * A0 + E(1 - p) * B0
        Ez = (E(p) * A1 + E(1 - p) * B1) / Z
        Ez2 = (E(p) * A2 + E(1 - p) * B2) / Z
        m = (E(p * p) * A0 + E(p * (1 - p)) * B0) / Z
        m2 = (E(p ** 3) * A0 + E(p * p * (1 - p)) * B0) / Z
        ab = (m - m2) / (m2 - m * m)
        out.append({"prior": [a, b, mu, s2], "measurement": d,
                    "posterior": {"mu": Ez, "sig": Ez2 - Ez * Ez, "alpha": m * ab,
                                  "beta": (1 - m) * ab, "inlier": m}})
    return out


def hole_fill():
    """Sources at their own pixel (what an identity pose produces), so the pre-fill table holds
    t[y, x] = y*W + x at the chosen pixels; (5, 0) is on the border and cannot scatter, and
    pixel (0, 0) would be index 0, the empty sentinel."""
    W, H = 6, 5
    t = np.zeros((H, W), np.int64)
    for x, y in [(1, 1), (3, 2), (4, 3), (5, 0), (2, 3)]:
        t[y, x] = y * W + x
    filled = t.copy()
    for y in range(H):
        for x in range(W):
            if t[y, x] != 0:
                continue
            for dy in (-1, 0, 1):
                done = False
                for dx in (-1, 0, 1):
                    nx, ny = x + dx, y + dy
                    if 1 <= nx <= W - 2 and 1 <= ny <= H - 2 and t[ny, nx] != 0:
                        filled[y, x] = t[ny, nx]
                        done = True
                        break
                if done:
                    break
    return {"width": W, "height": H, "table": t.tolist(), "filled": filled.tolist()}


def main():
    fixtures = {"cost_identity.json": cost_identity(), "filter_parabola.json": filter_parabola(),
                "fusion_posterior.json": fusion_posterior(),
                "fusion_posterior_cpu.json": fusion_posterior_cpu(), "hole_fill.json": hole_fill()}
    for name, obj in fixtures.items():
        with open(os.path.join(HERE, name), "w") as f:
            json.dump(obj, f, indent=1)
        print("wrote", name)


if __name__ == "__main__":
    main()
