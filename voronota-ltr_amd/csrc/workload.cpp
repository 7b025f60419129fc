// Seeded synthetic workloads for tests and benchmarks (plain C++, no HIP).
//
// Balls are placed by random sequential addition (RSA): a candidate centre is
// drawn uniformly in the bounding cube, kept if it lies inside the region and
// at least `min_dist` from every accepted centre. Radii are drawn uniformly
// from {1.52, 1.55, 1.70, 1.80} A and the radius index is the ball's type.
// The random stream is SplitMix64, so every machine produces identical inputs.
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <unordered_map>
#include <vector>

namespace {

struct SplitMix64 {
  uint64_t s;
  explicit SplitMix64(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  // uniform in [0, 1) with 53 random bits
  double uniform() { return static_cast<double>(next() >> 11) * 0x1.0p-53; }
};

constexpr double kRadii[4] = {1.52, 1.55, 1.70, 1.80};

// Hash grid over accepted centres with cell edge = min_dist.
struct AcceptGrid {
  double cell;
  std::unordered_map<uint64_t, std::vector<uint32_t>> cells;
  explicit AcceptGrid(double c) : cell(c) {}
  static uint64_t key(int64_t x, int64_t y, int64_t z) {
    return (static_cast<uint64_t>(x + (1 << 20)) << 42) | (static_cast<uint64_t>(y + (1 << 20)) << 21) |
           static_cast<uint64_t>(z + (1 << 20));
  }
};

}  // namespace

extern "C" {

// Fills balls[4*n] (x,y,z,r) and types[n]; returns the number placed (== n
// unless max_attempts ran out). Region: the ball of radius
// (3n / (4 pi density))^(1/3) centred at the origin.
size_t vltr_workload_rsa_sphere(uint64_t seed, size_t n, double density, double min_dist, double* balls,
                                int32_t* types, uint64_t max_attempts) {
  if (n == 0) return 0;
  const double R = std::cbrt(3.0 * static_cast<double>(n) / (4.0 * M_PI * density));
  SplitMix64 rng(seed);
  AcceptGrid grid(min_dist);
  std::vector<double> acc;
  acc.reserve(3 * n);
  const double md2 = min_dist * min_dist;
  size_t placed = 0;
  for (uint64_t att = 0; att < max_attempts && placed < n; ++att) {
    const double x = (2.0 * rng.uniform() - 1.0) * R;
    const double y = (2.0 * rng.uniform() - 1.0) * R;
    const double z = (2.0 * rng.uniform() - 1.0) * R;
    if (x * x + y * y + z * z > R * R) continue;
    const int64_t cx = static_cast<int64_t>(std::floor(x / grid.cell));
    const int64_t cy = static_cast<int64_t>(std::floor(y / grid.cell));
    const int64_t cz = static_cast<int64_t>(std::floor(z / grid.cell));
    bool ok = true;
    for (int dx = -1; dx <= 1 && ok; ++dx)
      for (int dy = -1; dy <= 1 && ok; ++dy)
        for (int dz = -1; dz <= 1 && ok; ++dz) {
          auto it = grid.cells.find(AcceptGrid::key(cx + dx, cy + dy, cz + dz));
          if (it == grid.cells.end()) continue;
          for (uint32_t j : it->second) {
            const double ex = acc[3 * j] - x, ey = acc[3 * j + 1] - y, ez = acc[3 * j + 2] - z;
            if (ex * ex + ey * ey + ez * ez < md2) {
              ok = false;
              break;
            }
          }
        }
    if (!ok) continue;
    grid.cells[AcceptGrid::key(cx, cy, cz)].push_back(static_cast<uint32_t>(placed));
    acc.push_back(x);
    acc.push_back(y);
    acc.push_back(z);
    const int t = static_cast<int>(rng.next() >> 62);
    balls[4 * placed] = x;
    balls[4 * placed + 1] = y;
    balls[4 * placed + 2] = z;
    balls[4 * placed + 3] = kRadii[t];
    if (types) types[placed] = t;
    ++placed;
  }
  return placed;
}

// Periodic variant: centres in the cube [0, L)^3, minimum-image distances.
size_t vltr_workload_rsa_box(uint64_t seed, size_t n, double L, double min_dist, double* balls, int32_t* types,
                             uint64_t max_attempts) {
  if (n == 0) return 0;
  SplitMix64 rng(seed);
  const int64_t nc = std::max<int64_t>(1, static_cast<int64_t>(std::floor(L / min_dist)));
  const double cell = L / static_cast<double>(nc);
  std::vector<std::vector<uint32_t>> cells(static_cast<size_t>(nc * nc * nc));
  std::vector<double> acc;
  acc.reserve(3 * n);
  const double md2 = min_dist * min_dist;
  auto wrap = [nc](int64_t c) { return ((c % nc) + nc) % nc; };
  auto mi = [L](double d) { return d - L * std::round(d / L); };
  size_t placed = 0;
  for (uint64_t att = 0; att < max_attempts && placed < n; ++att) {
    const double x = rng.uniform() * L, y = rng.uniform() * L, z = rng.uniform() * L;
    const int64_t cx = std::min<int64_t>(nc - 1, static_cast<int64_t>(x / cell));
    const int64_t cy = std::min<int64_t>(nc - 1, static_cast<int64_t>(y / cell));
    const int64_t cz = std::min<int64_t>(nc - 1, static_cast<int64_t>(z / cell));
    bool ok = true;
    for (int dx = -1; dx <= 1 && ok; ++dx)
      for (int dy = -1; dy <= 1 && ok; ++dy)
        for (int dz = -1; dz <= 1 && ok; ++dz) {
          const auto& v = cells[static_cast<size_t>((wrap(cx + dx) * nc + wrap(cy + dy)) * nc + wrap(cz + dz))];
          for (uint32_t j : v) {
            const double ex = mi(acc[3 * j] - x), ey = mi(acc[3 * j + 1] - y), ez = mi(acc[3 * j + 2] - z);
            if (ex * ex + ey * ey + ez * ez < md2) {
              ok = false;
              break;
            }
          }
        }
    if (!ok) continue;
    cells[static_cast<size_t>((cx * nc + cy) * nc + cz)].push_back(static_cast<uint32_t>(placed));
    acc.push_back(x);
    acc.push_back(y);
    acc.push_back(z);
    const int t = static_cast<int>(rng.next() >> 62);
    balls[4 * placed] = x;
    balls[4 * placed + 1] = y;
    balls[4 * placed + 2] = z;
    balls[4 * placed + 3] = kRadii[t];
    if (types) types[placed] = t;
    ++placed;
  }
  return placed;
}

// Structure sizes of the mixed-size batch: round(exp(U[ln lo, ln hi])).
void vltr_workload_batch_sizes(uint64_t seed, size_t n_structs, double lo, double hi, uint64_t* sizes) {
  SplitMix64 rng(seed);
  const double a = std::log(lo), b = std::log(hi);
  for (size_t s = 0; s < n_structs; ++s) sizes[s] = static_cast<uint64_t>(std::llround(std::exp(a + (b - a) * rng.uniform())));
}

// Seeded symmetric 4x4 matrix with entries uniform in [-1, 1].
void vltr_workload_type_matrix(uint64_t seed, double* W16) {
  SplitMix64 rng(seed);
  for (int i = 0; i < 4; ++i)
    for (int j = i; j < 4; ++j) {
      const double w = 2.0 * rng.uniform() - 1.0;
      W16[4 * i + j] = w;
      W16[4 * j + i] = w;
    }
}

// SplitMix64 stream, exposed so callers can draw seeded perturbations.
void vltr_workload_uniform(uint64_t seed, size_t n, double* out) {
  SplitMix64 rng(seed);
  for (size_t i = 0; i < n; ++i) out[i] = rng.uniform();
}

}  // extern "C"
