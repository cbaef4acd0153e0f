// Synthetic mesh generators for the benchmark workloads (ours; the reference has no
// icosphere generator): subdivided icosphere, geodesic icosahedron of any frequency, and
// seeded 3-D value noise for radial displacement.
#include "gen.h"

#include <cmath>
#include <cstring>
#include <map>
#include <stdexcept>
#include <tuple>

namespace lrt {
namespace {

// splitmix64 finaliser
inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// lattice value in [-1, 1)
inline double lattice(int64_t x, int64_t y, int64_t z, uint64_t seed) {
  uint64_t h = mix64(seed);
  h = mix64(h ^ uint64_t(x));
  h = mix64(h ^ uint64_t(y));
  h = mix64(h ^ uint64_t(z));
  return double(h >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

inline double fade(double t) { return t * t * (3.0 - 2.0 * t); }

void base_icosahedron(std::vector<double>& v, std::vector<uint32_t>& f) {
  const double t = (1.0 + std::sqrt(5.0)) / 2.0;
  const double p[12][3] = {{-1, t, 0}, {1, t, 0},  {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                           {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1},   {-t, 0, -1}, {-t, 0, 1}};
  v.clear();
  for (auto& q : p) {
    const double l = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
    v.push_back(q[0] / l); v.push_back(q[1] / l); v.push_back(q[2] / l);
  }
  const uint32_t faces[20][3] = {{0, 11, 5}, {0, 5, 1},  {0, 1, 7},   {0, 7, 10}, {0, 10, 11},
                                 {1, 5, 9},  {5, 11, 4}, {11, 10, 2}, {10, 7, 6}, {7, 1, 8},
                                 {3, 9, 4},  {3, 4, 2},  {3, 2, 6},   {3, 6, 8},  {3, 8, 9},
                                 {4, 9, 5},  {2, 4, 11}, {6, 2, 10},  {8, 6, 7},  {9, 8, 1}};
  f.clear();
  for (auto& q : faces) { f.push_back(q[0]); f.push_back(q[1]); f.push_back(q[2]); }
}

}  // namespace

double value_noise3(double x, double y, double z, uint64_t seed) {
  const double fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
  const int64_t ix = int64_t(fx), iy = int64_t(fy), iz = int64_t(fz);
  const double tx = fade(x - fx), ty = fade(y - fy), tz = fade(z - fz);
  double c[2][2][2];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int k = 0; k < 2; ++k) c[a][b][k] = lattice(ix + a, iy + b, iz + k, seed);
  const double x00 = c[0][0][0] + (c[1][0][0] - c[0][0][0]) * tx;
  const double x10 = c[0][1][0] + (c[1][1][0] - c[0][1][0]) * tx;
  const double x01 = c[0][0][1] + (c[1][0][1] - c[0][0][1]) * tx;
  const double x11 = c[0][1][1] + (c[1][1][1] - c[0][1][1]) * tx;
  const double y0 = x00 + (x10 - x00) * ty;
  const double y1 = x01 + (x11 - x01) * ty;
  return y0 + (y1 - y0) * tz;
}

// Unit icosphere by recursive 4-way subdivision: 20*4^k triangles, 10*4^k+2 vertices.
void make_icosphere(uint32_t subdiv, std::vector<double>& v, std::vector<uint32_t>& f) {
  base_icosahedron(v, f);
  for (uint32_t s = 0; s < subdiv; ++s) {
    std::map<uint64_t, uint32_t> mid;
    std::vector<uint32_t> nf;
    nf.reserve(f.size() * 4);
    auto midpoint = [&](uint32_t a, uint32_t b) -> uint32_t {
      const uint64_t key = a < b ? (uint64_t(a) << 32 | b) : (uint64_t(b) << 32 | a);
      auto it = mid.find(key);
      if (it != mid.end()) return it->second;
      double m[3];
      for (int i = 0; i < 3; ++i) m[i] = 0.5 * (v[3 * a + i] + v[3 * b + i]);
      const double l = std::sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
      const uint32_t id = uint32_t(v.size() / 3);
      v.push_back(m[0] / l); v.push_back(m[1] / l); v.push_back(m[2] / l);
      mid.emplace(key, id);
      return id;
    };
    for (size_t t = 0; t < f.size(); t += 3) {
      const uint32_t a = f[t], b = f[t + 1], c = f[t + 2];
      const uint32_t ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
      const uint32_t tri[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
      nf.insert(nf.end(), tri, tri + 12);
    }
    f.swap(nf);
  }
}

// Geodesic icosahedron of frequency n (class I): every base face split into n^2 triangles,
// vertices projected to the unit sphere; 20*n^2 triangles, 10*n^2+2 vertices.
void make_geodesic(uint32_t n, std::vector<double>& v, std::vector<uint32_t>& f) {
  if (n == 0) throw std::invalid_argument("geodesic frequency must be >= 1");
  std::vector<double> bv;
  std::vector<uint32_t> bf;
  base_icosahedron(bv, bf);
  v.clear();
  f.clear();
  std::map<std::tuple<int64_t, int64_t, int64_t>, uint32_t> ids;  // dedup by quantised position
  auto vid = [&](const double p[3]) -> uint32_t {
    const double l = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
    const double q[3] = {p[0] / l, p[1] / l, p[2] / l};
    const auto key = std::make_tuple(int64_t(std::llround(q[0] * 1e9)), int64_t(std::llround(q[1] * 1e9)),
                                     int64_t(std::llround(q[2] * 1e9)));
    auto it = ids.find(key);
    if (it != ids.end()) return it->second;
    const uint32_t id = uint32_t(v.size() / 3);
    v.push_back(q[0]); v.push_back(q[1]); v.push_back(q[2]);
    ids.emplace(key, id);
    return id;
  };
  for (size_t t = 0; t < bf.size(); t += 3) {
    const double* A = &bv[3 * bf[t]];
    const double* B = &bv[3 * bf[t + 1]];
    const double* C = &bv[3 * bf[t + 2]];
    std::vector<uint32_t> grid((n + 1) * (n + 2) / 2);
    auto gi = [&](uint32_t i, uint32_t j) { return i * (i + 1) / 2 + j; };  // row i has i+1 points
    for (uint32_t i = 0; i <= n; ++i)
      for (uint32_t j = 0; j <= i; ++j) {
        const double a = double(n - i) / n, b = double(i - j) / n, c = double(j) / n;
        const double p[3] = {a * A[0] + b * B[0] + c * C[0], a * A[1] + b * B[1] + c * C[1],
                             a * A[2] + b * B[2] + c * C[2]};
        grid[gi(i, j)] = vid(p);
      }
    for (uint32_t i = 0; i < n; ++i)
      for (uint32_t j = 0; j <= i; ++j) {
        f.push_back(grid[gi(i, j)]); f.push_back(grid[gi(i + 1, j)]); f.push_back(grid[gi(i + 1, j + 1)]);
        if (j < i) { f.push_back(grid[gi(i, j)]); f.push_back(grid[gi(i + 1, j + 1)]); f.push_back(grid[gi(i, j + 1)]); }
      }
  }
}

void displace_radial(std::vector<double>& v, double amplitude, double freq, uint64_t seed) {
  for (size_t i = 0; i < v.size(); i += 3) {
    const double n = value_noise3(v[i] * freq, v[i + 1] * freq, v[i + 2] * freq, seed);
    const double s = 1.0 + amplitude * n;
    v[i] *= s; v[i + 1] *= s; v[i + 2] *= s;
  }
}

}  // namespace lrt
