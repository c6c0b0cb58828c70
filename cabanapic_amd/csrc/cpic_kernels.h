// cpic_kernels.h -- CDNA4 (gfx950) kernels of the PIC hot path.
//
// Every kernel is a template on the real type R (float or double).  The
// arithmetic of each stage follows the reference operation for operation
// (operand order, literal types, sqrtf narrowing in the double build) so that
// the per-particle and per-cell results are bit-identical to the reference's
// serial evaluation; the library is compiled without floating-point
// contraction for that reason.  Only the order in which the scatter-add into
// the accumulators is summed differs (the reference's device ScatterView is
// atomic as well).
//
// Device layout in HBM:
//   particles      SoA: x,y,z,ux,uy,uz,w (R) + cell (int32)
//   fields         9 x Ng (R), component-major (as the host layout)
//   interpolators  Ng x 20 (R), cell-major: the 18 coefficients of
//                  src/types.h:64-84 padded to 20 so a cell is five 4-wide
//                  vector loads
//   accumulators   Ng x 12 (R), cell-major jx[4] jy[4] jz[4]
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cpic {

constexpr int kInterpStride = 20;
constexpr int kBlock = 256;
// Upper bound on move_p streaks per particle per step.  With a Courant-
// limited dt a particle crosses at most one face per axis (4 streaks); the
// bound only guarantees that every wave terminates on corrupt input.
constexpr int kMaxStreaks = 16;

enum DeviceError : int {
  kErrBadCell = 1,      // particle cell outside the voxel array
  kErrStreakLimit = 2,  // move_p did not finish within kMaxStreaks
};

template <class R> struct Vec4T;
template <> struct Vec4T<float> { typedef float type __attribute__((ext_vector_type(4))); };
template <> struct Vec4T<double> { typedef double type __attribute__((ext_vector_type(4))); };

struct Grid {
  int nx, ny, nz;
  int sy, sz;  // voxel strides along y and z (x stride is 1)
  int ncells;  // Ng
};

__device__ __forceinline__ int voxel(const Grid& g, int x, int y, int z) {
  return x + g.sy * y + g.sz * z;
}

// ---------------------------------------------------------------------------
// load_interpolator_array (src/interpolator.cpp:4-124): one thread per
// interior cell, x fastest so neighbouring lanes read neighbouring voxels.
template <class R>
__global__ void __launch_bounds__(kBlock)
k_load_interpolator(const R* __restrict__ fields, R* __restrict__ interp, Grid g) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int nint = g.nx * g.ny * g.nz;
  if (t >= nint) return;
  const int x = 1 + t % g.nx;
  const int y = 1 + (t / g.nx) % g.ny;
  const int z = 1 + t / (g.nx * g.ny);
  const int i = voxel(g, x, y, z);
  const R* ex = fields;
  const R* ey = fields + (size_t)g.ncells;
  const R* ez = fields + 2 * (size_t)g.ncells;
  const R* bx = fields + 3 * (size_t)g.ncells;
  const R* by = fields + 4 * (size_t)g.ncells;
  const R* bz = fields + 5 * (size_t)g.ncells;
  const R fourth = 1.0 / 4.0;
  const R half = 1.0 / 2.0;
  typedef typename Vec4T<R>::type V4;
  V4 o0, o1, o2, o3, o4;
  {
    const R a = ex[i], b = ex[i + g.sy], c = ex[i + g.sz], d = ex[i + g.sy + g.sz];
    o0.x = fourth * ((d + a) + (b + c));
    o0.y = fourth * ((d - a) + (b - c));
    o0.z = fourth * ((d - a) - (b - c));
    o0.w = fourth * ((d + a) - (b + c));
  }
  {
    const R a = ey[i], b = ey[i + g.sz], c = ey[i + 1], d = ey[i + 1 + g.sz];
    o1.x = fourth * ((d + a) + (b + c));
    o1.y = fourth * ((d - a) + (b - c));
    o1.z = fourth * ((d - a) - (b - c));
    o1.w = fourth * ((d + a) - (b + c));
  }
  {
    const R a = ez[i], b = ez[i + 1], c = ez[i + g.sy], d = ez[i + 1 + g.sy];
    o2.x = fourth * ((d + a) + (b + c));
    o2.y = fourth * ((d - a) + (b - c));
    o2.z = fourth * ((d - a) - (b - c));
    o2.w = fourth * ((d + a) - (b + c));
  }
  {
    const R a = bx[i], b = bx[i + 1];
    o3.x = half * (b + a);
    o3.y = half * (b - a);
    const R c = by[i], d = by[i + g.sy];
    o3.z = half * (d + c);
    o3.w = half * (d - c);
    const R e = bz[i], f = bz[i + g.sz];
    o4.x = half * (f + e);
    o4.y = half * (f - e);
    o4.z = 0;
    o4.w = 0;
  }
  V4* dst = reinterpret_cast<V4*>(interp + (size_t)i * kInterpStride);
  dst[0] = o0;
  dst[1] = o1;
  dst[2] = o2;
  dst[3] = o3;
  dst[4] = o4;
}

// ---------------------------------------------------------------------------
// Zero fill (clear_accumulator_array, src/accumulator.cpp:5-43, and context
// initialisation).  A kernel rather than a memset so the step stays
// capturable in a HIP graph.
template <class R>
__global__ void __launch_bounds__(kBlock) k_fill_zero(R* __restrict__ p, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0;
}

// ---------------------------------------------------------------------------
// The push (src/push.h:65-295) with move_p (src/move_p.h:59-374).

template <class R> struct PushArgs {
  R* __restrict__ x;
  R* __restrict__ y;
  R* __restrict__ z;
  R* __restrict__ ux;
  R* __restrict__ uy;
  R* __restrict__ uz;
  const R* __restrict__ w;
  int* __restrict__ cell;
  const R* __restrict__ interp;
  R* __restrict__ acc;
  int np;
  Grid g;
  R qdt_2mc, cdt_dx, cdt_dy, cdt_dz, qsp;
  int* err;
  unsigned int* movers;
};

// The correction term of a streak in move_p: a real_t product scaled by the
// double constant 1/3 and rounded back (src/move_p.h:154).
__device__ __forceinline__ float streak_v5(float q, float a, float b, float c) {
  return (float)((double)(q * a * b * c) * (1. / 3.));
}
__device__ __forceinline__ double streak_v5(double q, double a, double b, double c) {
  return q * a * b * c * (1. / 3.);
}

// 1/sqrt argument handling of push.h:148,169: sqrtf narrows a double
// argument to float in the double build.
__device__ __forceinline__ float sqrtf_narrow(float v) { return sqrtf(v); }
__device__ __forceinline__ double sqrtf_narrow(double v) { return (double)sqrtf((float)v); }

// One current component's four quadrant contributions for a streak with
// displacement dA along the component axis and midpoint (mB, mC) on the two
// transverse axes (push.h:218-232, move_p.h:156-170), added atomically.
template <class R>
__device__ __forceinline__ void deposit_component(R* a, R q, R dA, R mB, R mC, R v5) {
  R t = q * dA;
  R s1 = t * mB;
  R s0 = t - s1;
  s1 += t;
  R c = R(1) + mC;
  const R s2 = s0 * c;
  const R s3 = s1 * c;
  c = R(1) - mC;
  s0 *= c;
  s1 *= c;
  s0 += v5;
  s1 -= v5;
  unsafeAtomicAdd(a + 0, s0);
  unsafeAtomicAdd(a + 1, s1);
  unsafeAtomicAdd(a + 2, s2 - v5);
  unsafeAtomicAdd(a + 3, s3 + v5);
}

template <class R>
__device__ __forceinline__ void deposit_streak(R* acc, int cell, R q, R dx, R dy, R dz, R mx,
                                               R my, R mz, R v5) {
  R* a = acc + (size_t)cell * 12;
  deposit_component(a + 0, q, dx, my, mz, v5);
  deposit_component(a + 4, q, dy, mz, mx, v5);
  deposit_component(a + 8, q, dz, mx, my, v5);
}

// move_p: split the remaining track at cell faces, deposit each streak,
// periodic wrap at the domain faces.  pos/rem are in/out.
template <class R>
__device__ void move_p(const PushArgs<R>& a, int p, R pos[3], int& cell, R q, R rem[3]) {
  const Grid& g = a.g;
  for (int it = 0;; ++it) {
    if (it == kMaxStreaks) {
      atomicOr(a.err, kErrStreakLimit);
      break;
    }
    R dir[3], len3[3], dsp[3], mid[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      dsp[k] = rem[k];
      mid[k] = pos[k];
      dir[k] = (dsp[k] > 0) ? R(1) : R(-1);
      len3[k] = (dsp[k] == 0) ? R(3.4e38) : (dir[k] - mid[k]) / dsp[k];
    }
    R len = 2;
    int axis = 3;
    if (len3[0] < len) len = len3[0], axis = 0;
    if (len3[1] < len) len = len3[1], axis = 1;
    if (len3[2] < len) len = len3[2], axis = 2;
    len *= R(0.5);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      dsp[k] *= len;
      mid[k] += dsp[k];
    }
    const R v5 = streak_v5(q, dsp[0], dsp[1], dsp[2]);
    deposit_streak(a.acc, cell, q, dsp[0], dsp[1], dsp[2], mid[0], mid[1], mid[2], v5);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      rem[k] -= dsp[k];
      pos[k] += dsp[k] + dsp[k];
    }
    if (axis == 3) break;

    const R side = dir[axis];
    // decode the voxel, step across the face, wrap periodically
    int ix = cell % g.sy;
    const int rest = cell / g.sy;
    int iy = rest % (g.ny + 2);
    int iz = rest / (g.ny + 2);
    const int step = (side > 0) ? 1 : -1;
    if (axis == 0) {
      ix += step;
      if (ix == 0) ix = g.nx;
      else if (ix == g.nx + 1) ix = 1;
    } else if (axis == 1) {
      iy += step;
      if (iy == 0) iy = g.ny;
      else if (iy == g.ny + 1) iy = 1;
    } else {
      iz += step;
      if (iz == 0) iz = g.nz;
      else if (iz == g.nz + 1) iz = 1;
    }
    cell = voxel(g, ix, iy, iz);
    if ((unsigned)cell >= (unsigned)g.ncells) {
      atomicOr(a.err, kErrBadCell);
      cell = 0;
      break;
    }
    pos[axis] = -side;
  }
}

template <class R>
__global__ void __launch_bounds__(kBlock) k_push(PushArgs<R> a) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.np) return;
  int ii = a.cell[p];
  if ((unsigned)ii >= (unsigned)a.g.ncells) {
    atomicOr(a.err, kErrBadCell);
    return;
  }
  typedef typename Vec4T<R>::type V4;
  const V4* f = reinterpret_cast<const V4*>(a.interp + (size_t)ii * kInterpStride);
  const V4 e0 = f[0], e1 = f[1], e2 = f[2], b0 = f[3], b1 = f[4];
  const R one = 1.;
  const R one_third = 1. / 3.;
  const R two_fifteenths = 2. / 15.;
  const R qdt_2mc = a.qdt_2mc;

  const R dx = a.x[p], dy = a.y[p], dz = a.z[p];
  // field gather, push.h:124-138
  const R hax = qdt_2mc * ((e0.x + dy * e0.y) + dz * (e0.z + dy * e0.w));
  const R hay = qdt_2mc * ((e1.x + dz * e1.y) + dx * (e1.z + dz * e1.w));
  const R haz = qdt_2mc * ((e2.x + dx * e2.y) + dy * (e2.z + dx * e2.w));
  const R cbx = b0.x + dx * b0.y;
  const R cby = b0.z + dy * b0.w;
  const R cbz = b1.x + dz * b1.y;
  R ux = a.ux[p], uy = a.uy[p], uz = a.uz[p];
  ux += hax;
  uy += hay;
  uz += haz;
  // Boris rotation with the tan series, push.h:148-163
  R v0 = qdt_2mc / sqrtf_narrow(one + (ux * ux + (uy * uy + uz * uz)));
  R v1 = cbx * cbx + (cby * cby + cbz * cbz);
  R v2 = (v0 * v0) * v1;
  R v3 = v0 * (one + v2 * (one_third + v2 * two_fifteenths));
  R v4 = v3 / (one + v1 * (v3 * v3));
  v4 += v4;
  v0 = ux + v3 * (uy * cbz - uz * cby);
  v1 = uy + v3 * (uz * cbx - ux * cbz);
  v2 = uz + v3 * (ux * cby - uy * cbx);
  ux += v4 * (v1 * cbz - v2 * cby);
  uy += v4 * (v2 * cbx - v0 * cbz);
  uz += v4 * (v0 * cby - v1 * cbx);
  ux += hax;
  uy += hay;
  uz += haz;
  a.ux[p] = ux;
  a.uy[p] = uy;
  a.uz[p] = uz;
  // displacement in cell units, push.h:169-182
  v0 = one / sqrtf_narrow(one + (ux * ux + (uy * uy + uz * uz)));
  ux *= a.cdt_dx;
  uy *= a.cdt_dy;
  uz *= a.cdt_dz;
  ux *= v0;
  uy *= v0;
  uz *= v0;
  const R mx = dx + ux, my = dy + uy, mz = dz + uz;
  const R nx = mx + ux, ny = my + uy, nz = mz + uz;
  const R q = a.w[p] * a.qsp;
  int moved = 0;
  if (nx <= one && ny <= one && nz <= one && -nx <= one && -ny <= one && -nz <= one) {
    a.x[p] = nx;
    a.y[p] = ny;
    a.z[p] = nz;
    const R v5 = q * ux * uy * uz * one_third;
    deposit_streak(a.acc, ii, q, ux, uy, uz, mx, my, mz, v5);
  } else {
    R pos[3] = {dx, dy, dz};
    R rem[3] = {ux, uy, uz};
    move_p(a, p, pos, ii, q, rem);
    a.x[p] = pos[0];
    a.y[p] = pos[1];
    a.z[p] = pos[2];
    a.cell[p] = ii;
    moved = 1;
  }
  // one counter update per wave
  const unsigned long long active = __ballot(1);
  const unsigned long long movers = __ballot(moved);
  if ((int)(threadIdx.x & 63) == __ffsll((long long)active) - 1 && movers)
    atomicAdd(a.movers, (unsigned int)__popcll(movers));
}

// ---------------------------------------------------------------------------
// unload_accumulator_array (src/accumulator.cpp:45-122) over [1,n+1]^3.
template <class R>
__global__ void __launch_bounds__(kBlock)
k_unload(R* __restrict__ fields, const R* __restrict__ acc, Grid g, R cx, R cy, R cz) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int mx = g.nx + 1, my = g.ny + 1, mz = g.nz + 1;
  if (t >= mx * my * mz) return;
  const int x = 1 + t % mx;
  const int y = 1 + (t / mx) % my;
  const int z = 1 + t / (mx * my);
  const int i = voxel(g, x, y, z);
  const int dyo = g.sy, dzo = g.sz;
  const R* a0 = acc + (size_t)i * 12;
  const R* ay = acc + (size_t)(i - dyo) * 12;
  const R* az = acc + (size_t)(i - dzo) * 12;
  const R* ax = acc + (size_t)(i - 1) * 12;
  const R* ayz = acc + (size_t)(i - dyo - dzo) * 12;
  const R* azx = acc + (size_t)(i - 1 - dzo) * 12;
  const R* axy = acc + (size_t)(i - 1 - dyo) * 12;
  R* jx = fields + 6 * (size_t)g.ncells;
  R* jy = fields + 7 * (size_t)g.ncells;
  R* jz = fields + 8 * (size_t)g.ncells;
  jx[i] = cx * (a0[0] + ay[1] + az[2] + ayz[3]);
  jy[i] = cy * (a0[4] + az[5] + ax[6] + azx[7]);
  jz[i] = cz * (a0[8] + ax[9] + ay[10] + axy[11]);
}

// ---------------------------------------------------------------------------
// Periodic fold of the high ghost layer of J (serial_update_ghosts,
// src/fields.h:102-271).  The reference runs, for each component, a fold
// along a first transverse axis P and then along a second axis Q (jfx: y
// then z; jfy: z then x; jfz: x then y).  The resulting interior face
// values in closed form, with a = the values before the fold:
//   (1,1):        (a(1,1) + a(nP+1,1)) + (a(1,nQ+1) + a(nP+1,nQ+1))
//   (1,q), q>1:    a(1,q) + a(nP+1,q)
//   (p,1), p>1:    a(p,1) + a(p,nQ+1)
// Ghost cells the reference also touches are overwritten by the periodic
// copy that always follows, so only interior cells are written here, which
// keeps every read free of a race with a write.
template <class R>
__global__ void __launch_bounds__(kBlock) k_fold_j(R* __restrict__ fields, Grid g) {
  // work item = (component, coordinate along the component axis r,
  //              position on the two faces)
  const int n[3] = {g.nx, g.ny, g.nz};
  const int s[3] = {1, g.sy, g.sz};
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  for (int c = 0; c < 3; ++c) {
    const int ax_p = (c + 1) % 3, ax_q = (c + 2) % 3;
    // jfx folds y then z: P = y, Q = z; jfy: P = z, Q = x; jfz: P = x, Q = y
    const int nr = n[c], np_ = n[ax_p], nq = n[ax_q];
    // face positions: q in [1,nQ] with p = 1, and p in [2,nP] with q = 1
    const int per_r = nq + (np_ - 1);
    const int count = nr * per_r;
    if (t < count) {
      const int r = 1 + t / per_r;
      const int k = t % per_r;
      int p, q;
      if (k < nq) { p = 1; q = 1 + k; } else { q = 1; p = 2 + (k - nq); }
      R* J = fields + (size_t)(6 + c) * g.ncells;
      const int base = r * s[c];
      const int sp = s[ax_p], sq = s[ax_q];
      const int i = base + p * sp + q * sq;
      R v;
      if (p == 1 && q == 1) {
        v = (J[i] + J[base + (np_ + 1) * sp + sq]) +
            (J[base + sp + (nq + 1) * sq] + J[base + (np_ + 1) * sp + (nq + 1) * sq]);
      } else if (p == 1) {
        v = J[i] + J[base + (np_ + 1) * sp + q * sq];
      } else {
        v = J[i] + J[base + p * sp + (nq + 1) * sq];
      }
      J[i] = v;
      return;
    }
    t -= count;
  }
}

// ---------------------------------------------------------------------------
// Periodic ghost copy (serial_update_ghosts_B, src/fields.h:11-100) for three
// consecutive component arrays starting at comp0.  The reference's chained
// face copies give every ghost voxel the value of the interior voxel found
// by wrapping each coordinate (0 -> n, n+1 -> 1), so each ghost is written
// once from its wrapped source.
template <class R>
__global__ void __launch_bounds__(kBlock) k_copy_ghosts(R* __restrict__ fields, int comp0, Grid g) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int X = g.nx + 2, Y = g.ny + 2;
  const int zslab = X * Y;          // one z ghost plane
  const int yslab = X * g.nz;       // one y ghost plane, z interior
  const int xslab = g.ny * g.nz;    // one x ghost plane, y,z interior
  int x, y, z;
  int k = t;
  if (k < 2 * zslab) {
    z = (k < zslab) ? 0 : g.nz + 1;
    k %= zslab;
    x = k % X;
    y = k / X;
  } else if ((k -= 2 * zslab) < 2 * yslab) {
    y = (k < yslab) ? 0 : g.ny + 1;
    k %= yslab;
    x = k % X;
    z = 1 + k / X;
  } else if ((k -= 2 * yslab) < 2 * xslab) {
    x = (k < xslab) ? 0 : g.nx + 1;
    k %= xslab;
    y = 1 + k % g.ny;
    z = 1 + k / g.ny;
  } else {
    return;
  }
  const int sxw = (x == 0) ? g.nx : (x == g.nx + 1) ? 1 : x;
  const int syw = (y == 0) ? g.ny : (y == g.ny + 1) ? 1 : y;
  const int szw = (z == 0) ? g.nz : (z == g.nz + 1) ? 1 : z;
  const int dst = voxel(g, x, y, z), src = voxel(g, sxw, syw, szw);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    R* a = fields + (size_t)(comp0 + c) * g.ncells;
    a[dst] = a[src];
  }
}

// ---------------------------------------------------------------------------
// EM_Field_Solver::advance_b (src/fields.h:668-719), interior cells.
template <class R>
__global__ void __launch_bounds__(kBlock) k_advance_b(R* __restrict__ fields, Grid g, R px, R py, R pz) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.nx * g.ny * g.nz) return;
  const int x = 1 + t % g.nx;
  const int y = 1 + (t / g.nx) % g.ny;
  const int z = 1 + t / (g.nx * g.ny);
  const int i = voxel(g, x, y, z);
  const size_t N = g.ncells;
  const R* ex = fields;
  const R* ey = fields + N;
  const R* ez = fields + 2 * N;
  R* bx = fields + 3 * N;
  R* by = fields + 4 * N;
  R* bz = fields + 5 * N;
  const R e_x = ex[i], e_y = ey[i], e_z = ez[i];
  bx[i] -= (py * (ez[i + g.sy] - e_z) - pz * (ey[i + g.sz] - e_y));
  by[i] -= (pz * (ex[i + g.sz] - e_x) - px * (ez[i + 1] - e_z));
  bz[i] -= (px * (ey[i + 1] - e_y) - py * (ex[i + g.sy] - e_x));
}

// EM_Field_Solver::advance_e (src/fields.h:618-665), cells [1,n+1]^3.
template <class R>
__global__ void __launch_bounds__(kBlock)
k_advance_e(R* __restrict__ fields, Grid g, R px, R py, R pz, R cj) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int mx = g.nx + 1, my = g.ny + 1, mz = g.nz + 1;
  if (t >= mx * my * mz) return;
  const int x = 1 + t % mx;
  const int y = 1 + (t / mx) % my;
  const int z = 1 + t / (mx * my);
  const int i = voxel(g, x, y, z);
  const size_t N = g.ncells;
  R* ex = fields;
  R* ey = fields + N;
  R* ez = fields + 2 * N;
  const R* bx = fields + 3 * N;
  const R* by = fields + 4 * N;
  const R* bz = fields + 5 * N;
  const R* jx = fields + 6 * N;
  const R* jy = fields + 7 * N;
  const R* jz = fields + 8 * N;
  const R b_x = bx[i], b_y = by[i], b_z = bz[i];
  ex[i] = ex[i] + (-cj * jx[i]) + (py * (b_z - bz[i - g.sy]) - pz * (b_y - by[i - g.sz]));
  ey[i] = ey[i] + (-cj * jy[i]) + (pz * (b_x - bx[i - g.sz]) - px * (b_z - bz[i - 1]));
  ez[i] = ez[i] + (-cj * jz[i]) + (px * (b_y - by[i - 1]) - py * (b_x - bx[i - g.sy]));
}

// ---------------------------------------------------------------------------
// FP64 reductions.  Pass 1: each block reduces its share to one double per
// quantity; pass 2: one block sums the partials in a fixed order.  The result
// is deterministic for a given grid.

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

template <int NQ>
__device__ __forceinline__ void block_sum_store(double (&v)[NQ], double* out) {
  __shared__ double part[NQ][kBlock / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    v[q] = wave_sum(v[q]);
    if (lane == 0) part[q][wid] = v[q];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      double s = 0;
      for (int k = 0; k < kBlock / 64; ++k) s += part[q][k];
      out[q] = s;
    }
  }
}

// E^2 and B^2 over interior cells -> partials[block][2]
template <class R>
__global__ void __launch_bounds__(kBlock)
k_field_energy_partial(const R* __restrict__ fields, Grid g, double* __restrict__ partials) {
  const int nint = g.nx * g.ny * g.nz;
  const size_t N = g.ncells;
  double v[2] = {0, 0};
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nint; t += gridDim.x * blockDim.x) {
    const int x = 1 + t % g.nx;
    const int y = 1 + (t / g.nx) % g.ny;
    const int z = 1 + t / (g.nx * g.ny);
    const int i = voxel(g, x, y, z);
    const double e0 = fields[i], e1 = fields[N + i], e2 = fields[2 * N + i];
    const double b0 = fields[3 * N + i], b1 = fields[4 * N + i], b2 = fields[5 * N + i];
    v[0] += e0 * e0 + e1 * e1 + e2 * e2;
    v[1] += b0 * b0 + b1 * b1 + b2 * b2;
  }
  block_sum_store<2>(v, partials + 2 * blockIdx.x);
}

// sum w (gamma - 1) over particles -> partials[block][1]
template <class R>
__global__ void __launch_bounds__(kBlock)
k_kinetic_partial(const R* __restrict__ ux, const R* __restrict__ uy, const R* __restrict__ uz,
                  const R* __restrict__ w, int np, double* __restrict__ partials) {
  double v[1] = {0};
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < np; p += gridDim.x * blockDim.x) {
    const double a = ux[p], b = uy[p], c = uz[p];
    const double u2 = a * a + b * b + c * c;
    v[0] += (double)w[p] * (u2 / (sqrt(1.0 + u2) + 1.0));
  }
  block_sum_store<1>(v, partials + blockIdx.x);
}

// Final pass: out[q] = scale[q] * sum_b partials[b][q], in block order.
template <int NQ>
__global__ void __launch_bounds__(kBlock)
k_reduce_final(const double* __restrict__ partials, int nblocks, double s0, double s1,
               double* __restrict__ out) {
  double v[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) v[q] = 0;
  for (int b = threadIdx.x; b < nblocks; b += blockDim.x) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) v[q] += partials[b * NQ + q];
  }
  __shared__ double res[NQ];
  double tmp[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) tmp[q] = v[q];
  block_sum_store<NQ>(tmp, res);
  __syncthreads();
  if (threadIdx.x == 0) {
    out[0] = s0 * res[0];
    if (NQ > 1) out[1] = s1 * res[NQ > 1 ? 1 : 0];
  }
}

// ---------------------------------------------------------------------------
// Counting sort by cell: histogram, exclusive scan, scatter.

__global__ void __launch_bounds__(kBlock)
k_sort_histogram(const int* __restrict__ cell, int np, unsigned int* __restrict__ counts, int ncells,
                 int* err) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const int c = cell[p];
  if ((unsigned)c >= (unsigned)ncells) {
    atomicOr(err, kErrBadCell);
    return;
  }
  atomicAdd(counts + c, 1u);
}

// Block-local exclusive scan of 4 * kBlock counts; block totals to sums.
__global__ void __launch_bounds__(kBlock)
k_scan_blocks(const unsigned int* __restrict__ in, unsigned int* __restrict__ out, int n,
              unsigned int* __restrict__ sums) {
  __shared__ unsigned int s[kBlock];
  const int base = blockIdx.x * kBlock * 4 + threadIdx.x * 4;
  unsigned int v[4], tot = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    v[k] = (base + k < n) ? in[base + k] : 0u;
    tot += v[k];
  }
  s[threadIdx.x] = tot;
  __syncthreads();
  for (int off = 1; off < kBlock; off <<= 1) {
    unsigned int add = (threadIdx.x >= off) ? s[threadIdx.x - off] : 0u;
    __syncthreads();
    s[threadIdx.x] += add;
    __syncthreads();
  }
  unsigned int run = s[threadIdx.x] - tot;  // exclusive prefix of this thread
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (base + k < n) out[base + k] = run;
    run += v[k];
  }
  if (threadIdx.x == kBlock - 1) sums[blockIdx.x] = s[kBlock - 1];
}

// Serial-in-one-block exclusive scan of the block sums (n <= a few 10^4).
__global__ void __launch_bounds__(kBlock) k_scan_sums(unsigned int* __restrict__ sums, int n) {
  __shared__ unsigned int s[kBlock];
  unsigned int carry = 0;
  for (int base = 0; base < n; base += kBlock) {
    const int i = base + threadIdx.x;
    const unsigned int v = (i < n) ? sums[i] : 0u;
    s[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < kBlock; off <<= 1) {
      unsigned int add = (threadIdx.x >= off) ? s[threadIdx.x - off] : 0u;
      __syncthreads();
      s[threadIdx.x] += add;
      __syncthreads();
    }
    if (i < n) sums[i] = carry + s[threadIdx.x] - v;
    const unsigned int blocktot = s[kBlock - 1];
    __syncthreads();
    carry += blocktot;
  }
}

__global__ void __launch_bounds__(kBlock)
k_scan_add(unsigned int* __restrict__ out, int n, const unsigned int* __restrict__ sums) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] += sums[i / (kBlock * 4)];
}

template <class R> struct SortArgs {
  const R* src[7];
  R* dst[7];
  const int* src_cell;
  int* dst_cell;
  unsigned int* cursor;
  int np;
  int ncells;
};

template <class R>
__global__ void __launch_bounds__(kBlock) k_sort_scatter(SortArgs<R> a) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.np) return;
  const int c = a.src_cell[p];
  if ((unsigned)c >= (unsigned)a.ncells) return;
  const unsigned int d = atomicAdd(a.cursor + c, 1u);
  if (d >= (unsigned)a.np) return;
#pragma unroll
  for (int k = 0; k < 7; ++k) a.dst[k][d] = a.src[k][p];
  a.dst_cell[d] = c;
}

}  // namespace cpic
