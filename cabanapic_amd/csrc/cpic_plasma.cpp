// cpic_plasma.cpp -- host-side synthetic plasma loader (deck support).
//
// Counter-based, so any particle's initial state depends only on (seed, p):
// the result is independent of the thread count and of the order in which
// particles are generated.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/cabanapic_amd.h"

namespace {

inline uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

inline double unit_open_right(uint64_t r) { return (double)(r >> 11) * 0x1.0p-53; }  // [0,1)
inline double unit_open_left(uint64_t r) { return (double)((r >> 11) + 1) * 0x1.0p-53; }  // (0,1]

template <class R>
void fill(const cpic_plasma_spec& s, int64_t lo, int64_t hi, R* x, R* y, R* z, R* ux, R* uy,
          R* uz, R* w, int32_t* cell) {
  const uint64_t key = splitmix64(s.seed);
  const double two_pi = 6.283185307179586476925286766559;
  const int64_t X = s.nx + 2, Y = s.ny + 2;
  const double dxc = s.len_x / (double)s.nx;
  for (int64_t p = lo; p < hi; ++p) {
    const uint64_t base = key + 8ull * (uint64_t)p;
    double r[7];
    uint64_t bits[7];
    for (int k = 0; k < 7; ++k) bits[k] = splitmix64(base + (uint64_t)k);
    for (int k = 0; k < 3; ++k) r[k] = 2.0 * unit_open_right(bits[k]) - 1.0;
    const double rad0 = std::sqrt(-2.0 * std::log(unit_open_left(bits[3])));
    const double th0 = two_pi * unit_open_right(bits[4]);
    const double rad1 = std::sqrt(-2.0 * std::log(unit_open_left(bits[5])));
    const double th1 = two_pi * unit_open_right(bits[6]);
    const double n[3] = {rad0 * std::cos(th0), rad0 * std::sin(th0), rad1 * std::cos(th1)};
    const int64_t c = p / s.nppc;
    const int64_t ix = 1 + c % s.nx, iy = 1 + (c / s.nx) % s.ny, iz = 1 + c / (s.nx * s.ny);
    const double sign = (s.counter_streaming && (p & 1)) ? -1.0 : 1.0;
    double mod = 1.0;
    if (s.pert_amp != 0.0 && s.len_x > 0.0) {
      const double xphys = ((double)(ix - 1) + 0.5 * (r[0] + 1.0)) * dxc;
      mod = 1.0 + s.pert_amp * std::sin(two_pi * (double)s.pert_mode * xphys / s.len_x);
    }
    x[p] = (R)r[0];
    y[p] = (R)r[1];
    z[p] = (R)r[2];
    ux[p] = (R)(sign * s.drift[0] * mod + s.uth[0] * n[0]);
    uy[p] = (R)(sign * s.drift[1] * mod + s.uth[1] * n[1]);
    uz[p] = (R)(sign * s.drift[2] * mod + s.uth[2] * n[2]);
    w[p] = (R)s.weight;
    cell[p] = (int32_t)(ix + X * (iy + Y * iz));
  }
}

template <class R>
void fill_parallel(const cpic_plasma_spec& s, int64_t n, void* x, void* y, void* z, void* ux,
                   void* uy, void* uz, void* w, int32_t* cell) {
  unsigned nt = std::thread::hardware_concurrency();
  if (nt == 0) nt = 1;
  if (nt > 16) nt = 16;
  if (n < 100000) nt = 1;
  std::vector<std::thread> pool;
  for (unsigned t = 0; t < nt; ++t) {
    const int64_t lo = n * t / nt, hi = n * (t + 1) / nt;
    pool.emplace_back([&, lo, hi] {
      fill<R>(s, lo, hi, (R*)x, (R*)y, (R*)z, (R*)ux, (R*)uy, (R*)uz, (R*)w, cell);
    });
  }
  for (auto& th : pool) th.join();
}

}  // namespace

extern "C" int cpic_init_plasma(const cpic_plasma_spec* s, int32_t real_size, int64_t n, void* x,
                                void* y, void* z, void* ux, void* uy, void* uz, void* w,
                                int32_t* cell) {
  if (!s || !x || !y || !z || !ux || !uy || !uz || !w || !cell) return CPIC_ERR_INVALID;
  if (real_size != 4 && real_size != 8) return CPIC_ERR_INVALID;
  if (s->nx < 1 || s->ny < 1 || s->nz < 1 || s->nppc < 1) return CPIC_ERR_INVALID;
  if (n != s->nx * s->ny * s->nz * s->nppc) return CPIC_ERR_INVALID;
  if (real_size == 4)
    fill_parallel<float>(*s, n, x, y, z, ux, uy, uz, w, cell);
  else
    fill_parallel<double>(*s, n, x, y, z, ux, uy, uz, w, cell);
  return CPIC_OK;
}
