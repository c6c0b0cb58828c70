// cpic_plasma.cpp -- host-side synthetic plasma loader (deck support).
//
// Counter-based, so any particle's initial state depends only on (seed, p):
// the result is independent of the thread count and of the order in which
// particles are generated.
//
// The electron column (profile 1) starts from the electrostatic field of its
// own charge in the periodic box: node charge densities of the ring profile
// with the mean removed (the immobile ions neutralize the box), the periodic
// discrete Poisson equation on the Yee grid solved by a 2D DFT, and E as
// minus the difference of the node potentials along each edge.  The field
// is the same in every z layer.  Each electron then drifts at E x B / B^2 of
// that field, interpolated at its position as the push interpolates it.
#include <cmath>
#include <complex>
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

const double kTwoPi = 6.283185307179586476925286766559;

// Ring density of the column relative to its peak at box-centred (X, Y):
// exp(-((r - r0) / sigma)^2) (1 + amp cos(m theta)).
double ring_profile(const cpic_plasma_spec& s, double X, double Y) {
  const double t = (std::sqrt(X * X + Y * Y) - s.ring_radius) / s.ring_width;
  return std::exp(-t * t) * (1.0 + s.pert_amp * std::cos((double)s.pert_mode * std::atan2(Y, X)));
}

// In-place DFT of n complex values spaced `stride` apart (sign -1 forward,
// +1 inverse, unnormalized), by the direct sum over a twiddle table.
void dft(std::complex<double>* a, int n, long stride, int sign,
         const std::vector<std::complex<double>>& tw, std::vector<std::complex<double>>& tmp) {
  for (int k = 0; k < n; ++k) {
    std::complex<double> acc = 0;
    for (int j = 0; j < n; ++j) {
      const std::complex<double> w = tw[(long)k * j % n];
      acc += a[j * stride] * (sign < 0 ? std::conj(w) : w);
    }
    tmp[k] = acc;
  }
  for (int k = 0; k < n; ++k) a[k * stride] = tmp[k];
}

// The column's field on the interior edges: ex[(i) + nx j] on the x edge
// from node (i, j) to node (i+1, j), ey on the y edge from node (i, j) to
// (i, j+1), for i < nx, j < ny.  Node (i, j) is the low corner of voxel
// (i+1, j+1), at box-centred (i dx - Lx/2, j dy - Ly/2).
struct ColumnField {
  int nx = 0, ny = 0;
  std::vector<double> ex, ey;
  double at_ex(long ix, long iy) const {  // voxel indices, ghosts wrap
    return ex[(size_t)((ix - 1 + nx) % nx) + (size_t)nx * ((iy - 1 + ny) % ny)];
  }
  double at_ey(long ix, long iy) const {
    return ey[(size_t)((ix - 1 + nx) % nx) + (size_t)nx * ((iy - 1 + ny) % ny)];
  }
};

ColumnField column_field(const cpic_plasma_spec& s) {
  const int nx = (int)s.nx, ny = (int)s.ny;
  const double dx = s.len_x / nx, dy = s.len_y / ny;
  std::vector<std::complex<double>> a((size_t)nx * ny);
  double mean = 0;
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      const double rho = s.column_charge * ring_profile(s, i * dx - 0.5 * s.len_x,
                                                        j * dy - 0.5 * s.len_y);
      a[(size_t)i + (size_t)nx * j] = rho;
      mean += rho;
    }
  mean /= (double)nx * ny;
  for (auto& v : a) v -= mean;
  // -lap phi = rho: forward DFT, divide by the eigenvalues of the 5-point
  // Laplacian, inverse DFT
  auto twiddles = [](int n) {
    std::vector<std::complex<double>> t(n);
    for (int k = 0; k < n; ++k) t[k] = std::polar(1.0, kTwoPi * k / n);
    return t;
  };
  const auto twx = twiddles(nx), twy = twiddles(ny);
  std::vector<std::complex<double>> tmp(std::max(nx, ny));
  for (int sign : {-1, 1}) {
    for (int j = 0; j < ny; ++j) dft(a.data() + (size_t)nx * j, nx, 1, sign, twx, tmp);
    for (int i = 0; i < nx; ++i) dft(a.data() + i, ny, nx, sign, twy, tmp);
    if (sign < 0)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
          const double sx = 2.0 * std::sin(0.5 * kTwoPi * i / nx) / dx;
          const double sy = 2.0 * std::sin(0.5 * kTwoPi * j / ny) / dy;
          const double lam = sx * sx + sy * sy;
          a[(size_t)i + (size_t)nx * j] = lam > 0 ? a[(size_t)i + (size_t)nx * j] / lam : 0.0;
        }
  }
  const double norm = 1.0 / ((double)nx * ny);
  ColumnField f;
  f.nx = nx;
  f.ny = ny;
  f.ex.resize((size_t)nx * ny);
  f.ey.resize((size_t)nx * ny);
  auto phi = [&](int i, int j) { return a[(size_t)(i % nx) + (size_t)nx * (j % ny)].real() * norm; };
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      f.ex[(size_t)i + (size_t)nx * j] = -(phi(i + 1, j) - phi(i, j)) / dx;
      f.ey[(size_t)i + (size_t)nx * j] = -(phi(i, j + 1) - phi(i, j)) / dy;
    }
  return f;
}

template <class R>
void fill(const cpic_plasma_spec& s, const ColumnField* col, int64_t lo, int64_t hi, R* x, R* y,
          R* z, R* ux, R* uy, R* uz, R* w, int32_t* cell) {
  const uint64_t key = splitmix64(s.seed);
  const int64_t X = s.nx + 2, Y = s.ny + 2;
  const double dxc = s.len_x / (double)s.nx;
  for (int64_t p = lo; p < hi; ++p) {
    const uint64_t base = key + 8ull * (uint64_t)p;
    double r[7];
    uint64_t bits[7];
    for (int k = 0; k < 7; ++k) bits[k] = splitmix64(base + (uint64_t)k);
    for (int k = 0; k < 3; ++k) r[k] = 2.0 * unit_open_right(bits[k]) - 1.0;
    const double rad0 = std::sqrt(-2.0 * std::log(unit_open_left(bits[3])));
    const double th0 = kTwoPi * unit_open_right(bits[4]);
    const double rad1 = std::sqrt(-2.0 * std::log(unit_open_left(bits[5])));
    const double th1 = kTwoPi * unit_open_right(bits[6]);
    const double n[3] = {rad0 * std::cos(th0), rad0 * std::sin(th0), rad1 * std::cos(th1)};
    const int64_t c = p / s.nppc;
    const int64_t ix = 1 + c % s.nx, iy = 1 + (c / s.nx) % s.ny, iz = 1 + c / (s.nx * s.ny);
    const double sign = (s.counter_streaming && (p & 1)) ? -1.0 : 1.0;
    double mod = 1.0, weight = s.weight, drift[3] = {s.drift[0], s.drift[1], s.drift[2]};
    if (col) {
      // electron column: weight from the ring profile, E x B drift in the
      // grid field interpolated as the push does (ex linear in y between the
      // x edges at iy and iy+1, ey linear in x; uniform along z)
      const double dyc = s.len_y / (double)s.ny;
      const double Xc = ((double)(ix - 1) + 0.5 * (r[0] + 1.0)) * dxc - 0.5 * s.len_x;
      const double Yc = ((double)(iy - 1) + 0.5 * (r[1] + 1.0)) * dyc - 0.5 * s.len_y;
      weight *= ring_profile(s, Xc, Yc);
      const double ex = 0.5 * ((1.0 - r[1]) * col->at_ex(ix, iy) + (1.0 + r[1]) * col->at_ex(ix, iy + 1));
      const double ey = 0.5 * ((1.0 - r[0]) * col->at_ey(ix, iy) + (1.0 + r[0]) * col->at_ey(ix + 1, iy));
      // v = E x B / B^2 with B = (0, 0, bz)
      const double vx = ey / s.guide_bz, vy = -ex / s.guide_bz;
      const double gam = 1.0 / std::sqrt(1.0 - (vx * vx + vy * vy));
      drift[0] += gam * vx;
      drift[1] += gam * vy;
    } else if (s.pert_amp != 0.0 && s.len_x > 0.0) {
      const double xphys = ((double)(ix - 1) + 0.5 * (r[0] + 1.0)) * dxc;
      mod = 1.0 + s.pert_amp * std::sin(kTwoPi * (double)s.pert_mode * xphys / s.len_x);
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
  ColumnField col;
  if (s.profile == 1) col = column_field(s);
  const ColumnField* cp = s.profile == 1 ? &col : nullptr;
  unsigned nt = std::thread::hardware_concurrency();
  if (nt == 0) nt = 1;
  if (nt > 16) nt = 16;
  if (n < 100000) nt = 1;
  std::vector<std::thread> pool;
  pool.reserve(nt);
  try {
    for (unsigned t = 0; t < nt; ++t) {
      const int64_t lo = n * t / nt, hi = n * (t + 1) / nt;
      pool.emplace_back([&, lo, hi] {
        fill<R>(s, cp, lo, hi, (R*)x, (R*)y, (R*)z, (R*)ux, (R*)uy, (R*)uz, (R*)w, cell);
      });
    }
  } catch (...) {
    // a thread that could not start: finish the started ones, then report
    for (auto& th : pool) th.join();
    throw;
  }
  for (auto& th : pool) th.join();
}

bool column_spec_ok(const cpic_plasma_spec& s) {
  return s.len_x > 0.0 && s.len_y > 0.0 && s.ring_width > 0.0 && s.guide_bz != 0.0 &&
         s.nx <= 4096 && s.ny <= 4096;
}

template <class R>
void column_fields(const cpic_plasma_spec& s, R* f) {
  const ColumnField col = column_field(s);
  const int64_t X = s.nx + 2, Y = s.ny + 2, Z = s.nz + 2, N = X * Y * Z;
  for (int64_t iz = 0; iz < Z; ++iz)
    for (int64_t iy = 0; iy < Y; ++iy)
      for (int64_t ix = 0; ix < X; ++ix) {
        const int64_t i = ix + X * (iy + Y * iz);
        for (int k = 0; k < 9; ++k) f[k * N + i] = (R)0;
        f[0 * N + i] = (R)col.at_ex(ix, iy);
        f[1 * N + i] = (R)col.at_ey(ix, iy);
        f[5 * N + i] = (R)s.guide_bz;
      }
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
  if (s->profile == 1 && !column_spec_ok(*s)) return CPIC_ERR_INVALID;
  // host allocation and thread start-up failures are returned, not thrown
  try {
    if (real_size == 4)
      fill_parallel<float>(*s, n, x, y, z, ux, uy, uz, w, cell);
    else
      fill_parallel<double>(*s, n, x, y, z, ux, uy, uz, w, cell);
  } catch (...) {
    return CPIC_ERR_INVALID;
  }
  return CPIC_OK;
}

extern "C" int cpic_init_column_fields(const cpic_plasma_spec* s, int32_t real_size,
                                       void* fields) {
  if (!s || !fields) return CPIC_ERR_INVALID;
  if (real_size != 4 && real_size != 8) return CPIC_ERR_INVALID;
  if (s->nx < 1 || s->ny < 1 || s->nz < 1 || s->profile != 1) return CPIC_ERR_INVALID;
  if (!column_spec_ok(*s)) return CPIC_ERR_INVALID;
  try {
    if (real_size == 4)
      column_fields<float>(*s, (float*)fields);
    else
      column_fields<double>(*s, (double*)fields);
  } catch (...) {
    return CPIC_ERR_INVALID;
  }
  return CPIC_OK;
}
