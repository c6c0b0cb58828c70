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

// Radial field of an infinite column of charge density
// rho(r) = charge exp(-((r - r0) / sigma)^2) (Gauss's law with eps0 = 1):
// E_r(r) = charge / r * int_0^r exp(-((s - r0) / sigma)^2) s ds, in closed
// form with s = r0 + sigma t: sigma r0 sqrt(pi)/2 [erf t] - sigma^2/2 [e^{-t^2}].
double column_radial_field(double r, double r0, double sigma, double charge) {
  if (!(r > 0.0)) return 0.0;
  const double t0 = -r0 / sigma, t1 = (r - r0) / sigma;
  const double half_sqrt_pi = 0.88622692545275801364908374167057;
  const double integral = sigma * r0 * half_sqrt_pi * (std::erf(t1) - std::erf(t0)) -
                          0.5 * sigma * sigma * (std::exp(-t1 * t1) - std::exp(-t0 * t0));
  return charge * integral / r;
}

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
    double mod = 1.0, weight = s.weight, drift[3] = {s.drift[0], s.drift[1], s.drift[2]};
    if (s.profile == 1) {
      // electron column: weight from the ring profile, E x B drift
      const double dyc = s.len_y / (double)s.ny;
      const double X = ((double)(ix - 1) + 0.5 * (r[0] + 1.0)) * dxc - 0.5 * s.len_x;
      const double Y = ((double)(iy - 1) + 0.5 * (r[1] + 1.0)) * dyc - 0.5 * s.len_y;
      const double rad = std::sqrt(X * X + Y * Y);
      const double t = (rad - s.ring_radius) / s.ring_width;
      weight *= std::exp(-t * t) *
                (1.0 + s.pert_amp * std::cos((double)s.pert_mode * std::atan2(Y, X)));
      const double er = column_radial_field(rad, s.ring_radius, s.ring_width, s.column_charge);
      const double ex = rad > 0.0 ? er * X / rad : 0.0, ey = rad > 0.0 ? er * Y / rad : 0.0;
      // v = E x B / B^2 with B = (0, 0, bz)
      const double vx = ey / s.guide_bz, vy = -ex / s.guide_bz;
      const double gam = 1.0 / std::sqrt(1.0 - (vx * vx + vy * vy));
      drift[0] += gam * vx;
      drift[1] += gam * vy;
    } else if (s.pert_amp != 0.0 && s.len_x > 0.0) {
      const double xphys = ((double)(ix - 1) + 0.5 * (r[0] + 1.0)) * dxc;
      mod = 1.0 + s.pert_amp * std::sin(two_pi * (double)s.pert_mode * xphys / s.len_x);
    }
    x[p] = (R)r[0];
    y[p] = (R)r[1];
    z[p] = (R)r[2];
    ux[p] = (R)(sign * drift[0] * mod + s.uth[0] * n[0]);
    uy[p] = (R)(sign * drift[1] * mod + s.uth[1] * n[1]);
    uz[p] = (R)(sign * drift[2] * mod + s.uth[2] * n[2]);
    w[p] = (R)weight;
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
  if (s->profile != 0 && s->profile != 1) return CPIC_ERR_INVALID;
  if (s->profile == 1 && !(s->len_x > 0.0 && s->len_y > 0.0 && s->ring_width > 0.0 &&
                           s->guide_bz != 0.0))
    return CPIC_ERR_INVALID;
  if (real_size == 4)
    fill_parallel<float>(*s, n, x, y, z, ux, uy, uz, w, cell);
  else
    fill_parallel<double>(*s, n, x, y, z, ux, uy, uz, w, cell);
  return CPIC_OK;
}

namespace {

template <class R>
void column_fields(const cpic_plasma_spec& s, R* f) {
  const int64_t X = s.nx + 2, Y = s.ny + 2, Z = s.nz + 2, N = X * Y * Z;
  const double dx = s.len_x / (double)s.nx, dy = s.len_y / (double)s.ny;
  for (int64_t iz = 0; iz < Z; ++iz)
    for (int64_t iy = 0; iy < Y; ++iy)
      for (int64_t ix = 0; ix < X; ++ix) {
        const int64_t i = ix + X * (iy + Y * iz);
        // Yee positions relative to the box centre: ex on the x edge at the
        // voxel's (y-, z-) corner, ey on the y edge at its (x-, z-) corner
        const double xc = ((double)ix - 0.5) * dx - 0.5 * s.len_x;
        const double xl = ((double)ix - 1.0) * dx - 0.5 * s.len_x;
        const double yc = ((double)iy - 0.5) * dy - 0.5 * s.len_y;
        const double yl = ((double)iy - 1.0) * dy - 0.5 * s.len_y;
        double r = std::sqrt(xc * xc + yl * yl);
        double er = column_radial_field(r, s.ring_radius, s.ring_width, s.column_charge);
        const double ex = r > 0.0 ? er * xc / r : 0.0;
        r = std::sqrt(xl * xl + yc * yc);
        er = column_radial_field(r, s.ring_radius, s.ring_width, s.column_charge);
        const double ey = r > 0.0 ? er * yc / r : 0.0;
        for (int k = 0; k < 9; ++k) f[k * N + i] = (R)0;
        f[0 * N + i] = (R)ex;
        f[1 * N + i] = (R)ey;
        f[5 * N + i] = (R)s.guide_bz;
      }
}

}  // namespace

extern "C" int cpic_init_column_fields(const cpic_plasma_spec* s, int32_t real_size,
                                       void* fields) {
  if (!s || !fields) return CPIC_ERR_INVALID;
  if (real_size != 4 && real_size != 8) return CPIC_ERR_INVALID;
  if (s->nx < 1 || s->ny < 1 || s->nz < 1 || s->profile != 1) return CPIC_ERR_INVALID;
  if (!(s->len_x > 0.0 && s->len_y > 0.0 && s->ring_width > 0.0)) return CPIC_ERR_INVALID;
  if (real_size == 4)
    column_fields<float>(*s, (float*)fields);
  else
    column_fields<double>(*s, (double*)fields);
  return CPIC_OK;
}
