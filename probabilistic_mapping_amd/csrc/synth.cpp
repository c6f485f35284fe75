// Seeded synthetic scene renderer: textured rectangles (planes and box faces) seen by a
// pinhole camera.  Produces the uint8 grayscale frames and ground-truth depth the depth
// pipeline is measured on.  Pure host code; deterministic (integer hashing, no libm
// randomness), so identical inputs come out on every x86-64 machine.
//
// Patch record (16 doubles):
//   [0..2]  centre c (world)      [3..5] unit axis e1       [6..8] unit axis e2
//   [9]     half extent along e1  [10]   half extent along e2  (<= 0 means unbounded)
//   [11]    texture seed          [12]   base frequency (cycles / m)
//   [13]    octaves               [14]   mean intensity     [15] intensity amplitude
#include <cmath>
#include <cstdint>

namespace {

inline uint64_t mix64(uint64_t z) {  // splitmix64 finaliser
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

inline double lattice(uint64_t seed, int octave, int64_t ix, int64_t iy) {
  uint64_t h = mix64(seed ^ mix64((uint64_t)octave * 0x632BE59BD9B4E019ull));
  h = mix64(h ^ (uint64_t)ix);
  h = mix64(h ^ ((uint64_t)iy * 0x8CB92BA72F3D8DD7ull));
  return (double)(h >> 11) * (1.0 / 9007199254740992.0);  // [0, 1)
}

inline double smooth(double a) { return a * a * (3.0 - 2.0 * a); }

// Multi-octave value noise in [0, 1].
double value_noise(uint64_t seed, double s, double t, double freq, int octaves) {
  double acc = 0.0, norm = 0.0, amp = 1.0, f = freq;
  for (int o = 0; o < octaves; ++o) {
    const double x = s * f, y = t * f;
    const double fx = std::floor(x), fy = std::floor(y);
    const int64_t ix = (int64_t)fx, iy = (int64_t)fy;
    const double ax = smooth(x - fx), ay = smooth(y - fy);
    const double v00 = lattice(seed, o, ix, iy), v10 = lattice(seed, o, ix + 1, iy);
    const double v01 = lattice(seed, o, ix, iy + 1), v11 = lattice(seed, o, ix + 1, iy + 1);
    const double top = v00 + ax * (v10 - v00), bot = v01 + ax * (v11 - v01);
    acc += amp * (top + ay * (bot - top));
    norm += amp;
    amp *= 0.5;
    f *= 2.0;
  }
  return acc / norm;
}

}  // namespace

extern "C" {

// Renders one frame.  K = {fx, fy, cx, cy}; R = camera->world rotation (row-major 3x3);
// T = camera centre in world.  img: [H][W] uint8; depth (optional): [H][W] camera-z of the
// visible surface, 0 where nothing is hit.  Returns the number of pixels with no hit.
long pm_synth_render(const double* patches, int n, const double* K, const double* R,
                     const double* T, int W, int H, uint8_t* img, float* depth) {
  long misses = 0;
#pragma omp parallel for reduction(+ : misses) schedule(static)
  for (int v = 0; v < H; ++v) {
    for (int u = 0; u < W; ++u) {
      // ray direction in camera (z = 1) and in world
      const double dc[3] = {(u - K[2]) / K[0], (v - K[3]) / K[1], 1.0};
      double dw[3];
      for (int i = 0; i < 3; ++i)
        dw[i] = R[i * 3 + 0] * dc[0] + R[i * 3 + 1] * dc[1] + R[i * 3 + 2] * dc[2];
      double best = INFINITY;
      int hit = -1;
      double hs = 0, ht = 0;
      for (int k = 0; k < n; ++k) {
        const double* P = patches + 16 * k;
        const double nx = P[4] * P[8] - P[5] * P[7];
        const double ny = P[5] * P[6] - P[3] * P[8];
        const double nz = P[3] * P[7] - P[4] * P[6];
        const double den = nx * dw[0] + ny * dw[1] + nz * dw[2];
        if (std::fabs(den) < 1e-12) continue;
        const double lam =
            (nx * (P[0] - T[0]) + ny * (P[1] - T[1]) + nz * (P[2] - T[2])) / den;
        if (!(lam > 1e-6) || lam >= best) continue;
        const double q[3] = {T[0] + lam * dw[0] - P[0], T[1] + lam * dw[1] - P[1],
                             T[2] + lam * dw[2] - P[2]};
        const double s = q[0] * P[3] + q[1] * P[4] + q[2] * P[5];
        const double t = q[0] * P[6] + q[1] * P[7] + q[2] * P[8];
        if (P[9] > 0 && std::fabs(s) > P[9]) continue;
        if (P[10] > 0 && std::fabs(t) > P[10]) continue;
        best = lam;
        hit = k;
        hs = s;
        ht = t;
      }
      const long idx = (long)v * W + u;
      if (hit < 0) {
        ++misses;
        img[idx] = 0;
        if (depth) depth[idx] = 0.0f;
        continue;
      }
      const double* P = patches + 16 * hit;
      const double nval = value_noise((uint64_t)P[11], hs, ht, P[12], (int)P[13]);
      double val = P[14] + P[15] * (2.0 * nval - 1.0);
      val = val < 0 ? 0 : (val > 255 ? 255 : val);
      img[idx] = (uint8_t)std::lround(val);
      if (depth) depth[idx] = (float)best;  // ray has camera z = 1, so lam is the depth
    }
  }
  return misses;
}

}  // extern "C"
