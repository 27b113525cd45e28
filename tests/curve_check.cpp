// Checks the spatial ordering's Hilbert index (fpsx_curve.hpp) on the host.
// A Hilbert curve visits every aligned 128^3 sub-cube of the 1024^3 grid as
// one contiguous, aligned run of 2^21 indices with unit steps; this is checked
// on 12 sub-cubes (corners, edges, interior), and injectivity on 4 M random
// grid points.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "fpsx_curve.hpp"

static int check_cube(uint32_t cx, uint32_t cy, uint32_t cz) {
  const uint32_t S = 128, N = S * S * S;
  std::vector<uint32_t> inv(N, 0xffffffffu);
  uint32_t lo = 0xffffffffu;
  for (uint32_t x = 0; x < S; ++x)
    for (uint32_t y = 0; y < S; ++y)
      for (uint32_t z = 0; z < S; ++z) lo = std::min(lo, fpsx::hilbert10(cx + x, cy + y, cz + z));
  if (lo % N != 0) {
    std::printf("FAIL cube (%u,%u,%u): run starts at %u, not aligned\n", cx, cy, cz, lo);
    return 1;
  }
  for (uint32_t x = 0; x < S; ++x)
    for (uint32_t y = 0; y < S; ++y)
      for (uint32_t z = 0; z < S; ++z) {
        const uint32_t h = fpsx::hilbert10(cx + x, cy + y, cz + z) - lo;
        if (h >= N || inv[h] != 0xffffffffu) {
          std::printf("FAIL cube (%u,%u,%u): not a contiguous bijection\n", cx, cy, cz);
          return 1;
        }
        inv[h] = (x << 14) | (y << 7) | z;
      }
  for (uint32_t h = 1; h < N; ++h) {
    int d = 0;
    for (int s = 0; s < 21; s += 7) {
      const int u = (int)((inv[h - 1] >> s) & 127u), v = (int)((inv[h] >> s) & 127u);
      d += u > v ? u - v : v - u;
    }
    if (d != 1) {
      std::printf("FAIL cube (%u,%u,%u): step %u is not a unit step\n", cx, cy, cz, h);
      return 1;
    }
  }
  return 0;
}

int main() {
  const uint32_t cubes[][3] = {{0, 0, 0},       {896, 896, 896}, {0, 0, 896},     {896, 0, 0},
                               {0, 896, 0},     {384, 512, 640}, {512, 512, 512}, {384, 384, 384},
                               {128, 768, 256}, {768, 128, 640}, {256, 0, 768},   {640, 896, 128}};
  for (const auto &c : cubes)
    if (check_cube(c[0], c[1], c[2])) return 1;
  std::mt19937 rng(7);
  std::vector<uint32_t> pts(1u << 22), hs(1u << 22);
  for (auto &p : pts) p = rng() & ((1u << 30) - 1u);
  std::sort(pts.begin(), pts.end());
  pts.erase(std::unique(pts.begin(), pts.end()), pts.end());
  hs.resize(pts.size());
  for (size_t i = 0; i < pts.size(); ++i) {
    hs[i] = fpsx::hilbert10(pts[i] >> 20, (pts[i] >> 10) & 1023u, pts[i] & 1023u);
    if (hs[i] >= (1u << 30)) {
      std::printf("FAIL index out of range\n");
      return 1;
    }
  }
  std::sort(hs.begin(), hs.end());
  if (std::adjacent_find(hs.begin(), hs.end()) != hs.end()) {
    std::printf("FAIL two grid points share an index\n");
    return 1;
  }
  std::printf("OK\n");
  return 0;
}
