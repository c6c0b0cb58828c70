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
//   fields         9 x Ng (R), component-major (as the host layout)
//   interpolators  Ng x 20 (R), cell-major: the 18 coefficients of
//                  src/types.h:64-84 padded to 20 so a cell is five 4-wide
//                  vector loads
//   accumulators   12 x Ng (R), slot-major: jx[0..3], jy[0..3], jz[0..3]
//   particles      per-cell columns, see cpic_push_kernels.h
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
  kErrBadCell = 1,       // particle cell outside the voxel array
  kErrStreakLimit = 2,   // move_p did not finish within kMaxStreaks
  kErrColumnFull = 4,    // a cell received more particles than its column holds
  kErrOverflowFull = 8,  // the spill list of crossing records is full
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
// load_interpolator_array (src/interpolator.cpp:4-124): the interpolator of
// voxel i from the fields (component stride ng, voxel strides sy and sz), as
// five 4-vectors (ex, dexdy, dexdz, d2exdydz; ey ...; ez ...; cbx, dcbxdx,
// cby, dcbydy; cbz, dcbzdz, 0, 0).
template <class R, class V4>
__device__ __forceinline__ void interpolator_at(const R* __restrict__ fields, size_t ng, int sy,
                                                int sz, int i, V4& o0, V4& o1, V4& o2, V4& o3,
                                                V4& o4) {
  const R* ex = fields;
  const R* ey = fields + ng;
  const R* ez = fields + 2 * ng;
  const R* bx = fields + 3 * ng;
  const R* by = fields + 4 * ng;
  const R* bz = fields + 5 * ng;
  const R fourth = 1.0 / 4.0;
  const R half = 1.0 / 2.0;
  {
    const R a = ex[i], b = ex[i + sy], c = ex[i + sz], d = ex[i + sy + sz];
    o0.x = fourth * ((d + a) + (b + c));
    o0.y = fourth * ((d - a) + (b - c));
    o0.z = fourth * ((d - a) - (b - c));
    o0.w = fourth * ((d + a) - (b + c));
  }
  {
    const R a = ey[i], b = ey[i + sz], c = ey[i + 1], d = ey[i + 1 + sz];
    o1.x = fourth * ((d + a) + (b + c));
    o1.y = fourth * ((d - a) + (b - c));
    o1.z = fourth * ((d - a) - (b - c));
    o1.w = fourth * ((d + a) - (b + c));
  }
  {
    const R a = ez[i], b = ez[i + 1], c = ez[i + sy], d = ez[i + 1 + sy];
    o2.x = fourth * ((d + a) + (b + c));
    o2.y = fourth * ((d - a) + (b - c));
    o2.z = fourth * ((d - a) - (b - c));
    o2.w = fourth * ((d + a) - (b + c));
  }
  {
    const R a = bx[i], b = bx[i + 1];
    o3.x = half * (b + a);
    o3.y = half * (b - a);
    const R c = by[i], d = by[i + sy];
    o3.z = half * (d + c);
    o3.w = half * (d - c);
    const R e = bz[i], f = bz[i + sz];
    o4.x = half * (f + e);
    o4.y = half * (f - e);
    o4.z = 0;
    o4.w = 0;
  }
}

// One thread per interior cell, x fastest so neighbouring lanes read
// neighbouring voxels.
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
  typedef typename Vec4T<R>::type V4;
  V4 o0, o1, o2, o3, o4;
  interpolator_at(fields, (size_t)g.ncells, g.sy, g.sz, i, o0, o1, o2, o3, o4);
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
// unload_accumulator_array (src/accumulator.cpp:45-122) over [1,n+1]^3.
// The device accumulators are slot-major, acc[k * Ng + voxel], k = 0..11
// for jx[0..3], jy[0..3], jz[0..3].
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
  const size_t N = g.ncells;
  const int sy = g.sy, sz = g.sz;
#define A(k, v) acc[(size_t)(k) * N + (size_t)(v)]
  R* jx = fields + 6 * N;
  R* jy = fields + 7 * N;
  R* jz = fields + 8 * N;
  jx[i] = cx * (A(0, i) + A(1, i - sy) + A(2, i - sz) + A(3, i - sy - sz));
  jy[i] = cy * (A(4, i) + A(5, i - sz) + A(6, i - 1) + A(7, i - 1 - sz));
  jz[i] = cz * (A(8, i) + A(9, i - 1) + A(10, i - sy) + A(11, i - 1 - sy));
#undef A
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
// The ghost voxels whose periodic source is interior voxel (x, y, z): each
// coordinate either its own or its ghost image (n + 1 for 1, 0 for n; both
// when n = 1), other than the voxel itself.  Every ghost voxel is the image
// of exactly one interior voxel (the one k_copy_ghosts copies it from), so a
// kernel that computes interior values can write the ghost copies itself.
template <class F>
__device__ __forceinline__ void for_each_ghost_image(const Grid& g, int x, int y, int z, F&& f) {
  const bool yz_inner = y != 1 && y != g.ny && z != 1 && z != g.nz;
  if (yz_inner) {
    // the common case: on the x faces only, one image (every wave holds
    // such a lane)
    if (g.nx > 1 && (x == 1 || x == g.nx)) f(voxel(g, x == 1 ? g.nx + 1 : 0, y, z));
    if (g.nx > 1 || x != 1) return;
  }
  int xs[3] = {x, x, x}, ys[3] = {y, y, y}, zs[3] = {z, z, z};
  int nxs = 1, nys = 1, nzs = 1;
  if (x == 1) xs[nxs++] = g.nx + 1;
  if (x == g.nx) xs[nxs++] = 0;
  if (y == 1) ys[nys++] = g.ny + 1;
  if (y == g.ny) ys[nys++] = 0;
  if (z == 1) zs[nzs++] = g.nz + 1;
  if (z == g.nz) zs[nzs++] = 0;
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
      for (int a = 0; a < 3; ++a)
        if (c < nzs && b < nys && a < nxs && (a | b | c)) f(voxel(g, xs[a], ys[b], zs[c]));
}

// EM_Field_Solver::advance_b (src/fields.h:668-719), interior cells, then
// the periodic ghost copy of B (serial_update_ghosts_B, src/fields.h:11-100):
// each thread also writes its voxel's new B into the ghosts it is the source
// of (no thread reads B ghosts here).
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
  const R nbx = bx[i] - (py * (ez[i + g.sy] - e_z) - pz * (ey[i + g.sz] - e_y));
  const R nby = by[i] - (pz * (ex[i + g.sz] - e_x) - px * (ez[i + 1] - e_z));
  const R nbz = bz[i] - (px * (ey[i + 1] - e_y) - py * (ex[i + g.sy] - e_x));
  bx[i] = nbx;
  by[i] = nby;
  bz[i] = nbz;
  for_each_ghost_image(g, x, y, z, [&](int v) {
    bx[v] = nbx;
    by[v] = nby;
    bz[v] = nbz;
  });
}

// EM_Field_Solver::advance_e (src/fields.h:618-665), cells [1,n+1]^3, after
// the J fold.  The periodic ghost copy of J (serial_update_ghosts,
// src/fields.h:11-100) happens here too: a voxel on the high ghost layer
// reads J from its periodic source (the value the copy gives it), and each
// interior thread writes its voxel's J into the ghosts it is the source of.
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
  R* jx = fields + 6 * N;
  R* jy = fields + 7 * N;
  R* jz = fields + 8 * N;
  const bool interior = x <= g.nx && y <= g.ny && z <= g.nz;
  const int js = interior ? i
                          : voxel(g, x > g.nx ? 1 : x, y > g.ny ? 1 : y, z > g.nz ? 1 : z);
  const R j_x = jx[js], j_y = jy[js], j_z = jz[js];
  const R b_x = bx[i], b_y = by[i], b_z = bz[i];
  ex[i] = ex[i] + (-cj * j_x) + (py * (b_z - bz[i - g.sy]) - pz * (b_y - by[i - g.sz]));
  ey[i] = ey[i] + (-cj * j_y) + (pz * (b_x - bx[i - g.sz]) - px * (b_z - bz[i - 1]));
  ez[i] = ez[i] + (-cj * j_z) + (px * (b_y - by[i - 1]) - py * (b_x - bx[i - g.sy]));
  if (interior)
    for_each_ghost_image(g, x, y, z, [&](int v) {
      jx[v] = j_x;
      jy[v] = j_y;
      jz[v] = j_z;
    });
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

// ---------------------------------------------------------------------------
// The electrostatic 1D solver (ES_Field_Solver_1D, src/fields.h:467-545).

// serial_update_ghosts (src/fields.h:102-271) as the reference runs it,
// including the ghost cells its passes leave behind, which the ES solver
// keeps (no periodic copy follows): one thread per x plane for jfx (its y
// fold then its z fold), per y plane for jfy (z then x), per z plane for jfz
// (x then y).  A thread touches only its own plane of its component.
template <class R>
__global__ void __launch_bounds__(kBlock) k_fold_j_ref(R* __restrict__ fields, Grid g) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const size_t N = g.ncells;
  if (t < g.nx) {
    R* s = fields + 6 * N;
    const int x = t + 1;
    for (int z = 1; z <= g.nz + 1; ++z) s[voxel(g, x, 1, z)] += s[voxel(g, x, g.ny + 1, z)];
    for (int y = 1; y <= g.ny + 1; ++y) s[voxel(g, x, y, 1)] += s[voxel(g, x, y, g.nz + 1)];
  } else if (t < g.nx + g.ny) {
    R* s = fields + 7 * N;
    const int y = t - g.nx + 1;
    for (int x = 1; x <= g.nx + 1; ++x) s[voxel(g, x, y, 1)] += s[voxel(g, x, y, g.nz + 1)];
    for (int z = 1; z <= g.nz + 1; ++z) s[voxel(g, 1, y, z)] += s[voxel(g, g.nx + 1, y, z)];
  } else if (t < g.nx + g.ny + g.nz) {
    R* s = fields + 8 * N;
    const int z = t - g.nx - g.ny + 1;
    for (int y = 1; y <= g.ny + 1; ++y) s[voxel(g, 1, y, z)] += s[voxel(g, g.nx + 1, y, z)];
    for (int x = 1; x <= g.nx + 1; ++x) s[voxel(g, x, 1, z)] += s[voxel(g, x, g.ny + 1, z)];
  }
}

// ES advance_e, src/fields.h:537-542: E += -dt_eps0 J on every voxel,
// ghosts included
template <class R>
__global__ void __launch_bounds__(kBlock) k_es_advance_e(R* __restrict__ fields, int ncells,
                                                         R dt_eps0) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ncells) return;
  const size_t N = ncells;
  const R cj = dt_eps0;
#pragma unroll
  for (int c = 0; c < 3; ++c) fields[c * N + i] = fields[c * N + i] + (-cj * fields[(6 + c) * N + i]);
}

// ES e_energy (src/fields.h:482-509): E^2 over every voxel, ghosts
// included; with B^2 over interior cells as for the EM solver (the ES
// solver leaves B as it is) -> partials[block][2]
template <class R>
__global__ void __launch_bounds__(kBlock)
k_es_energy_partial(const R* __restrict__ fields, Grid g, double* __restrict__ partials) {
  const int n = g.ncells;
  const size_t N = g.ncells;
  const int X = g.sy, Y = g.sz / g.sy;
  double v[2] = {0, 0};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double e0 = fields[i], e1 = fields[N + i], e2 = fields[2 * N + i];
    v[0] += e0 * e0 + e1 * e1 + e2 * e2;
    const int x = i % X, y = (i / X) % Y, z = i / g.sz;
    if (x >= 1 && x <= g.nx && y >= 1 && y <= g.ny && z >= 1 && z <= g.nz) {
      const double b0 = fields[3 * N + i], b1 = fields[4 * N + i], b2 = fields[5 * N + i];
      v[1] += b0 * b0 + b1 * b1 + b2 * b2;
    }
  }
  block_sum_store<2>(v, partials + 2 * blockIdx.x);
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

// Largest of a non-negative double over the wave, then one 64-bit atomic
// max per wave on its bit pattern (the order of non-negative doubles is
// the order of their bit patterns as unsigned integers).
__device__ __forceinline__ void wave_max_store(double v, unsigned long long* out) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(out, (unsigned long long)__double_as_longlong(v));
}

// Gauss residual at the interior nodes in FP64:
//   r(v) = sum_a (e_a(v) - e_a(v - a)) / d_a  -  q(v) / (dx dy dz eps0),
// where the x edge of voxel v (ex, src/types.h:138-172) joins node v to node
// v + x, and v - a in the low ghost layer is read at its periodic image
// (the reference never writes the low ghost layer of E).  q: node charge,
// interior order.  mode 0 stores r in base; mode 1 compares with base.
// out[0] = max |r| (mode 0) or max |r - base| (mode 1), out[1] = max
// |q| / (dx dy dz eps0); res (may be null) receives r.
template <class R>
__global__ void __launch_bounds__(kBlock)
k_gauss_residual(const R* __restrict__ fields, Grid g, const double* __restrict__ q,
                 double inv_dx, double inv_dy, double inv_dz, double qscale,
                 double* __restrict__ res, double* __restrict__ base, int mode,
                 unsigned long long* __restrict__ out) {
  const int nint = g.nx * g.ny * g.nz;
  const size_t N = g.ncells;
  double m0 = 0, m1 = 0;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nint; t += gridDim.x * blockDim.x) {
    const int x = 1 + t % g.nx;
    const int y = 1 + (t / g.nx) % g.ny;
    const int z = 1 + t / (g.nx * g.ny);
    const int i = voxel(g, x, y, z);
    const int ixm = voxel(g, x == 1 ? g.nx : x - 1, y, z);
    const int iym = voxel(g, x, y == 1 ? g.ny : y - 1, z);
    const int izm = voxel(g, x, y, z == 1 ? g.nz : z - 1);
    const double div = ((double)fields[i] - (double)fields[ixm]) * inv_dx +
                       ((double)fields[N + i] - (double)fields[N + iym]) * inv_dy +
                       ((double)fields[2 * N + i] - (double)fields[2 * N + izm]) * inv_dz;
    const double rho = q[t] * qscale;
    const double r = div - rho;
    if (res) res[t] = r;
    if (mode == 0) {
      base[t] = r;
      m0 = fmax(m0, fabs(r));
    } else {
      m0 = fmax(m0, fabs(r - base[t]));
    }
    m1 = fmax(m1, fabs(rho));
  }
  wave_max_store(m0, out);
  wave_max_store(m1, out + 1);
}

}  // namespace cpic
