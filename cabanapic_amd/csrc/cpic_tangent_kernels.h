// cpic_tangent_kernels.h -- forward-mode (tangent-linear) derivative of the
// push + move_p + deposit.
//
// The reference has no derivative of its step.  The tangent of the step
// with respect to one deck parameter theta is carried alongside the primal
// state: every particle carries d(x,y,z,ux,uy,uz)/dtheta, and a second set
// of field arrays carries dE/dtheta, dB/dtheta and dJ/dtheta.  Every stage
// but the push is linear in the fields (interpolator load, unload, ghost
// fold and copy, advance_b, advance_e), so its tangent is the same stage
// applied to the tangent arrays.  The push is not: its tangent is computed
// here by running the push's own device functions (boris, displacement,
// streak) on dual numbers, whose value part repeats the primal arithmetic
// operation for operation (so the primal results are those of the column
// push, bit for bit) and whose tangent part, in FP64, is the derivative of
// each operation.  The derivative is the one of the piece of the step the
// particle is on: crossing decisions (which face, how many streaks) are
// taken on the values and held fixed, and a position clamped to a face has
// zero tangent.
#pragma once

#include "cpic_push_kernels.h"

namespace cpic {

// A value of type R and its derivative, in FP64.  Operators are hidden
// friends, so integer and floating constants convert implicitly.
template <class R> struct Dual {
  R v;
  double t;
  __device__ Dual() : v(0), t(0) {}
  __device__ Dual(double c) : v((R)c), t(0) {}
  __device__ Dual(R value, double tangent) : v(value), t(tangent) {}

  friend __device__ __forceinline__ Dual operator+(const Dual& a, const Dual& b) {
    return Dual(a.v + b.v, a.t + b.t);
  }
  friend __device__ __forceinline__ Dual operator-(const Dual& a, const Dual& b) {
    return Dual(a.v - b.v, a.t - b.t);
  }
  friend __device__ __forceinline__ Dual operator-(const Dual& a) { return Dual(-a.v, -a.t); }
  friend __device__ __forceinline__ Dual operator*(const Dual& a, const Dual& b) {
    return Dual(a.v * b.v, a.t * (double)b.v + (double)a.v * b.t);
  }
  friend __device__ __forceinline__ Dual operator/(const Dual& a, const Dual& b) {
    const R q = a.v / b.v;
    return Dual(q, (a.t - (double)q * b.t) / (double)b.v);
  }
  __device__ __forceinline__ Dual& operator+=(const Dual& b) { return *this = *this + b; }
  __device__ __forceinline__ Dual& operator-=(const Dual& b) { return *this = *this - b; }
  __device__ __forceinline__ Dual& operator*=(const Dual& b) { return *this = *this * b; }
  friend __device__ __forceinline__ bool operator<(const Dual& a, const Dual& b) { return a.v < b.v; }
  friend __device__ __forceinline__ bool operator>(const Dual& a, const Dual& b) { return a.v > b.v; }
  friend __device__ __forceinline__ bool operator<=(const Dual& a, const Dual& b) { return a.v <= b.v; }
  friend __device__ __forceinline__ bool operator==(const Dual& a, const Dual& b) { return a.v == b.v; }
};

// the narrowed square root of the push and its derivative 1 / (2 sqrt)
template <class R> __device__ __forceinline__ Dual<R> sqrtf_narrow(const Dual<R>& s) {
  const R r = sqrtf_narrow(s.v);
  return Dual<R>(r, s.t * 0.5 / (double)r);
}

// the streak correction term q a b c / 3 and its derivative
template <class R>
__device__ __forceinline__ Dual<R> streak_v5(const Dual<R>& q, const Dual<R>& a,
                                             const Dual<R>& b, const Dual<R>& c) {
  const double Q = q.v, A = a.v, B = b.v, C = c.v;
  const double t = (q.t * A * B * C + Q * a.t * B * C + Q * A * b.t * C + Q * A * B * c.t) * (1. / 3.);
  return Dual<R>(streak_v5(q.v, a.v, b.v, c.v), t);
}

template <class R> struct DualV4 {
  Dual<R> x, y, z, w;
};

template <class R> struct TangentPush {
  // the particles before the push (sorted by cell, so a cell's particles
  // sit in consecutive lanes) and after it, at the same index
  const R* src_part;      // [7][stride]: x, y, z, ux, uy, uz, w
  const int* src_cell;    // voxel of each particle
  const double* src_tpart;  // [6][stride]: tangents of x, y, z, ux, uy, uz
  const int* src_tag;     // upload index of each particle
  R* part;
  int* cell;
  double* tpart;
  int* tag;
  long stride;
  int n;
  const R* interp;          // primal interpolators
  const double* tinterp;    // tangent interpolators
  R* acc;                   // primal accumulators, slot-major [12][Ng]
  double* tacc;             // tangent accumulators
  CellGrid g;
  int ncells;               // Ng
  R qdt_2mc, cdt_dx, cdt_dy, cdt_dz, qsp;
  int* err;
  unsigned* movers;
};

// A block's accumulator window in LDS: kTanWindow consecutive voxels from
// base (one before the voxel of the block's first particle; particles are
// sorted by voxel, so the block's own cells and their x neighbours fall in
// it), and four side segments of kTanSide voxels each, the same run shifted
// by +y, -y, +z and -z (the neighbours that crossing particles enter, and
// the cells of particles that left the block's cells since the last sort).
// Deposits into the window are LDS atomics, flushed to the accumulators
// once per block with one atomic per value that received any (the flush's
// lanes take consecutive voxels, so the L2 passes them on to memory a line
// at a time); others are global atomics.
constexpr int kTanWindow = 64;
constexpr int kTanSide = 32;
template <class R> struct TanWindow {
  int base;
  int sy, sz;                      // voxel strides in y and z
  R (*v)[kTanWindow];              // [12][kTanWindow] primal
  double (*t)[kTanWindow];         // [12][kTanWindow] tangent
  R (*sv)[4][kTanSide];            // [12][4][kTanSide] primal, side segments
  double (*st)[4][kTanSide];       // [12][4][kTanSide] tangent, side segments
};

// offset of side segment s (0..3: +y, -y, +z, -z) from the window's base
template <class R> __device__ __forceinline__ int side_offset(const TanWindow<R>& w, int s) {
  return s == 0 ? w.sy : s == 1 ? -w.sy : s == 2 ? w.sz : -w.sz;
}

template <class R>
__device__ __forceinline__ void deposit_dual(const TangentPush<R>& a, const TanWindow<R>& w, int v,
                                             const Dual<R> o[12]) {
  const int r = v - w.base;
  const unsigned c = (unsigned)r;
  // the first segment that holds v: the run, then the side segments
  int seg = -1;
  unsigned cs = 0;
  if (c >= (unsigned)kTanWindow) {
#pragma unroll
    for (int s2 = 3; s2 >= 0; --s2) {
      const unsigned d = (unsigned)(r - side_offset(w, s2));
      if (d < (unsigned)kTanSide) seg = s2, cs = d;
    }
    if (seg >= 0) {
#pragma unroll
      for (int k = 0; k < 12; ++k) {
        __hip_atomic_fetch_add(&w.sv[k][seg][cs], o[k].v, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_WORKGROUP);
        __hip_atomic_fetch_add(&w.st[k][seg][cs], o[k].t, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_WORKGROUP);
      }
      return;
    }
  }
  if (c < (unsigned)kTanWindow) {
    // workgroup scope: besides being all that LDS needs, it keeps these
    // atomics distinct from the global ones below, which the compiler would
    // otherwise merge with them into one flat atomic (counted by lgkmcnt, so
    // that every later LDS wait also waits for a global atomic)
#pragma unroll
    for (int k = 0; k < 12; ++k) {
      __hip_atomic_fetch_add(&w.v[k][c], o[k].v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      __hip_atomic_fetch_add(&w.t[k][c], o[k].t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    return;
  }
  const size_t N = a.ncells;
#pragma unroll
  for (int k = 0; k < 12; ++k) {
    unsafeAtomicAdd(a.acc + k * N + v, o[k].v);
    unsafeAtomicAdd(a.tacc + k * N + v, o[k].t);
  }
}

// Adds o (the deposits of the calling lane into voxel key, key < 0 for
// none) to the accumulators, summed first over each run of consecutive lanes
// of the wave with the same key (a segmented inclusive scan; the last lane
// of a run adds the run's total).  Particles of one cell that sit in
// consecutive lanes thus cost one atomic per value instead of one per lane.
// Every lane of the wave must call it.
template <class R>
__device__ __forceinline__ void deposit_runs(const TangentPush<R>& a, const TanWindow<R>& w,
                                             int key, Dual<R> o[12]) {
  const int lane = threadIdx.x & 63;
  const int before = __shfl_up(key, 1, 64);
  int head = (lane == 0 || before != key) ? 1 : 0;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int h = __shfl_up(head, off, 64);
    const bool add = lane >= off && !head;
#pragma unroll
    for (int k = 0; k < 12; ++k) {
      const R yv = __shfl_up(o[k].v, off, 64);
      const double yt = __shfl_up(o[k].t, off, 64);
      if (add) {
        o[k].v = yv + o[k].v;
        o[k].t = yt + o[k].t;
      }
    }
    if (lane >= off) head |= h;
  }
  const int after = __shfl_down(key, 1, 64);
  if (key >= 0 && (lane == 63 || after != key)) deposit_dual(a, w, key, o);
}

// The push of particle p (nothing for p >= n); returns the voxel of its own
// cell, whose deposits are left in own[], or -1.
template <class R>
__device__ __forceinline__ int push_tangent_particle(const TangentPush<R>& a,
                                                     const TanWindow<R>& win, int p,
                                                     Dual<R> own[12]) {
  if (p >= a.n) return -1;
  typedef Dual<R> D;
  typedef typename Vec4T<R>::type V4;
  typedef double T4 __attribute__((ext_vector_type(4)));
  const long S = a.stride;
  const int v0 = a.src_cell[p];
  a.tag[p] = a.src_tag[p];
  const int X = a.g.sy, Y = a.g.sz / a.g.sy;
  const int ix = v0 % X - 1, iy = (v0 / X) % Y - 1, iz = v0 / a.g.sz - 1;
  if (v0 < 0 || v0 >= a.ncells || ix < 0 || ix >= a.g.nx || iy < 0 || iy >= a.g.ny || iz < 0 ||
      iz >= a.g.nz) {
    atomicOr(a.err, kErrBadCell);
    // carried over unchanged
    for (int k = 0; k < 7; ++k) a.part[k * S + p] = a.src_part[k * S + p];
    for (int k = 0; k < 6; ++k) a.tpart[k * S + p] = a.src_tpart[k * S + p];
    a.cell[p] = v0;
    return -1;
  }
  int ci = ix + a.g.nx * (iy + a.g.ny * iz);

  const V4* f = reinterpret_cast<const V4*>(a.interp + (size_t)v0 * kInterpStride);
  const T4* tf = reinterpret_cast<const T4*>(a.tinterp + (size_t)v0 * kInterpStride);
  DualV4<R> e[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    const V4 fv = f[k];
    const T4 tv = tf[k];
    e[k].x = D(fv.x, tv.x);
    e[k].y = D(fv.y, tv.y);
    e[k].z = D(fv.z, tv.z);
    e[k].w = D(fv.w, tv.w);
  }
  const D dx(a.src_part[p], a.src_tpart[p]);
  const D dy(a.src_part[S + p], a.src_tpart[S + p]);
  const D dz(a.src_part[2 * S + p], a.src_tpart[2 * S + p]);
  D ux(a.src_part[3 * S + p], a.src_tpart[3 * S + p]);
  D uy(a.src_part[4 * S + p], a.src_tpart[4 * S + p]);
  D uz(a.src_part[5 * S + p], a.src_tpart[5 * S + p]);
  const R w = a.src_part[6 * S + p];

  boris(e[0], e[1], e[2], e[3], e[4], D(a.qdt_2mc, 0.), dx, dy, dz, ux, uy, uz);
  D ddx, ddy, ddz;
  displacement(ux, uy, uz, D(a.cdt_dx, 0.), D(a.cdt_dy, 0.), D(a.cdt_dz, 0.), ddx, ddy, ddz);
  const D one = 1.;
  const D one_third = 1. / 3.;
  const D mx = dx + ddx, my = dy + ddy, mz = dz + ddz;
  const D nx = mx + ddx, ny = my + ddy, nz = mz + ddz;
  const D q = D(w, 0.) * D(a.qsp, 0.);
  D pos[3];
  if (nx <= one && ny <= one && nz <= one && -nx <= one && -ny <= one && -nz <= one) {
    const D v5 = q * ddx * ddy * ddz * one_third;
    streak_values(q, ddx, ddy, ddz, mx, my, mz, v5, own);
    pos[0] = nx;
    pos[1] = ny;
    pos[2] = nz;
  } else {
    // one counter update per wave: the lowest active lane adds the number
    // of lanes that leave their cell
    const unsigned long long leaving = __ballot(1);
    if ((threadIdx.x & 63) == (unsigned)__ffsll((long long)leaving) - 1)
      atomicAdd(a.movers, (unsigned)__popcll(leaving));
    pos[0] = dx;
    pos[1] = dy;
    pos[2] = dz;
    D rem[3] = {ddx, ddy, ddz};
    int d = streak(pos, rem, q, own);  // the first streak, in the own cell
    for (int it = 1; it < kMaxStreaks && d >= 0; ++it) {
      ci = face_neighbor(a.g, ci, d);
      D o[12];
      d = streak(pos, rem, q, o);
      deposit_dual(a, win, interior_to_voxel(a.g, ci), o);
    }
    if (d >= 0) atomicOr(a.err, kErrStreakLimit);
  }
  a.part[p] = pos[0].v;
  a.part[S + p] = pos[1].v;
  a.part[2 * S + p] = pos[2].v;
  a.part[3 * S + p] = ux.v;
  a.part[4 * S + p] = uy.v;
  a.part[5 * S + p] = uz.v;
  a.tpart[p] = pos[0].t;
  a.tpart[S + p] = pos[1].t;
  a.tpart[2 * S + p] = pos[2].t;
  a.tpart[3 * S + p] = ux.t;
  a.tpart[4 * S + p] = uy.t;
  a.tpart[5 * S + p] = uz.t;
  a.part[6 * S + p] = w;
  a.cell[p] = interior_to_voxel(a.g, ci);
  return v0;
}

// Runs of equal keys in consecutive lanes of the wave: for a lane with
// key >= 0, the first lane of its run and the run's length.  The particles
// come mostly sorted by voxel (sorted a few steps ago), so a run stands for
// many particles of one voxel and takes one atomic instead of one per lane.
__device__ __forceinline__ void wave_key_run(int key, int& first, int& len) {
  const int lane = threadIdx.x & 63;
  const int prev = __shfl_up(key, 1, 64);
  const bool head = key >= 0 && (lane == 0 || prev != key);
  const unsigned long long heads = __ballot(head);
  const unsigned long long bounds = __ballot(head || key < 0);
  const unsigned long long upto = lane == 63 ? heads : heads & ((2ull << lane) - 1);
  first = upto ? 63 - __clzll(upto) : lane;
  const unsigned long long after = lane == 63 ? 0ull : bounds & ~((2ull << lane) - 1);
  const int next = after ? __ffsll((long long)after) - 1 : 64;
  len = next - first;
}

// Counting sort of the particles by voxel before a tangent push, so that
// the run sums of the push's own-cell deposits see a cell's particles in
// consecutive lanes (particles keep drifting out of the order of their
// upload).  Count: particles per voxel (an invalid voxel counts in bin 0
// and is flagged by the push), one atomic per run of a voxel in a wave,
// whose result gives every particle its rank within its voxel.
__global__ void __launch_bounds__(kBlock)
k_tangent_bin_count(const int* __restrict__ cell, int n, int ncells, int* __restrict__ bins,
                    int* __restrict__ rank) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  int key = -1;
  if (p < n) {
    const int v = cell[p];
    key = (v >= 0 && v < ncells) ? v : 0;
  }
  int first, len;
  wave_key_run(key, first, len);
  const int lane = threadIdx.x & 63;
  int base = 0;
  if (key >= 0 && lane == first) base = atomicAdd(bins + key, len);
  base = __shfl(base, first, 64);
  if (key >= 0) rank[p] = base + (lane - first);
}

// Scatter: every particle to its voxel's range (start: the exclusive scan
// of the counts) at its rank; the order within a voxel is the order in
// which the count's atomics arrived.
template <class R>
__global__ void __launch_bounds__(kBlock)
k_tangent_bin_scatter(const R* __restrict__ part, const int* __restrict__ cell,
                      const double* __restrict__ tpart, const int* __restrict__ tag, int n,
                      long stride, int ncells, const int* __restrict__ start,
                      const int* __restrict__ rank, R* __restrict__ part_out,
                      int* __restrict__ cell_out, double* __restrict__ tpart_out,
                      int* __restrict__ tag_out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int v = cell[p];
  const int key = (v >= 0 && v < ncells) ? v : 0;
  const int o = start[key] + rank[p];
#pragma unroll
  for (int k = 0; k < 7; ++k) part_out[k * stride + o] = part[k * stride + p];
#pragma unroll
  for (int k = 0; k < 6; ++k) tpart_out[k * stride + o] = tpart[k * stride + p];
  cell_out[o] = v;
  tag_out[o] = tag[p];
}

// One thread per particle: the push of src/push.h:65-295 with the move_p
// loop of src/move_p.h:93-371, on dual numbers, deposits added atomically
// (primal to acc, tangent to tacc).  The deposits into the particle's own
// cell (its whole track, or the first streak of a particle that leaves) are
// summed over the wave's runs of lanes in one cell first; the later streaks
// of a particle that leaves are added directly.  The primal positions,
// momenta and cells equal those of the column push; the primal accumulators
// equal them up to the order of the floating-point sums.  The FP32 kernel is
// held to 128 registers (4 waves per SIMD instead of 3), which pays for this
// latency-bound kernel; the FP64 one would spill too much.
template <class R>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(sizeof(R) == 4 ? 4 : 1)))
k_push_tangent(TangentPush<R> a) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  typedef Dual<R> D;
  __shared__ R wv[12][kTanWindow];
  __shared__ double wt[12][kTanWindow];
  __shared__ R sv[12][4][kTanSide];
  __shared__ double st[12][4][kTanSide];
  const int p0 = blockIdx.x * blockDim.x;
  TanWindow<R> win{(p0 < a.n ? a.src_cell[p0] : 0) - 1, a.g.sy, a.g.sz, wv, wt, sv, st};
  for (int i = threadIdx.x; i < 12 * kTanWindow; i += blockDim.x) {
    wv[i / kTanWindow][i % kTanWindow] = 0;
    wt[i / kTanWindow][i % kTanWindow] = 0;
  }
  for (int i = threadIdx.x; i < 12 * 4 * kTanSide; i += blockDim.x) {
    const int k = i / (4 * kTanSide), sg = (i / kTanSide) % 4, c = i % kTanSide;
    sv[k][sg][c] = 0;
    st[k][sg][c] = 0;
  }
  __syncthreads();
  D own[12];  // deposits into the particle's own cell
  // every lane takes part in the wave's run sums, so a lane without a
  // particle (or with an invalid cell) keeps key -1 and no deposit
  const int key = push_tangent_particle(a, win, p, own);
  deposit_runs(a, win, key, own);
  __syncthreads();
  // the window's sums, one global atomic per value that received any
  const size_t N = a.ncells;
  for (int i = threadIdx.x; i < 12 * kTanWindow; i += blockDim.x) {
    const int k = i / kTanWindow, c = i % kTanWindow;
    const int v = win.base + c;
    const R x = wv[k][c];
    const double y = wt[k][c];
    if ((x != R(0) || y != 0.) && v >= 0 && v < a.ncells) {
      unsafeAtomicAdd(a.acc + k * N + v, x);
      unsafeAtomicAdd(a.tacc + k * N + v, y);
    }
  }
  for (int i = threadIdx.x; i < 12 * 4 * kTanSide; i += blockDim.x) {
    const int k = i / (4 * kTanSide), sg = (i / kTanSide) % 4, c = i % kTanSide;
    const int v = win.base + side_offset(win, sg) + c;
    const R x = sv[k][sg][c];
    const double y = st[k][sg][c];
    if ((x != R(0) || y != 0.) && v >= 0 && v < a.ncells) {
      unsafeAtomicAdd(a.acc + k * N + v, x);
      unsafeAtomicAdd(a.tacc + k * N + v, y);
    }
  }
}

// Appends the step's row (E, B, dE, dB) to the history at the row a
// device counter names, then advances the counter: the history of graph
// replays, whose kernel arguments cannot change from step to step.
__global__ void k_append_row(const double* __restrict__ row, double* __restrict__ hist,
                             int* __restrict__ counter, int cap) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  const int r = *counter;
  if (r < cap)
    for (int k = 0; k < 4; ++k) hist[4 * (size_t)r + k] = row[k];
  *counter = r + 1;
}

// sum over interior cells of E . tE and B . tB -> partials[block][2] (FP64):
// the derivatives of the field energies 0.5 sum E^2 and 0.5 sum B^2
template <class R>
__global__ void __launch_bounds__(kBlock)
k_field_energy_tangent(const R* __restrict__ fields, const double* __restrict__ tfields, Grid g,
                       double* __restrict__ partials) {
  const int nint = g.nx * g.ny * g.nz;
  const size_t N = g.ncells;
  double v[2] = {0, 0};
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nint; t += gridDim.x * blockDim.x) {
    const int x = 1 + t % g.nx;
    const int y = 1 + (t / g.nx) % g.ny;
    const int z = 1 + t / (g.nx * g.ny);
    const int i = voxel(g, x, y, z);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      v[0] += (double)fields[k * N + i] * tfields[k * N + i];
      v[1] += (double)fields[(3 + k) * N + i] * tfields[(3 + k) * N + i];
    }
  }
  block_sum_store<2>(v, partials + 2 * blockIdx.x);
}

}  // namespace cpic
