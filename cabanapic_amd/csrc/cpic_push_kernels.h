// cpic_push_kernels.h -- the push + move_p + deposit on cell-binned particles.
//
// Particle storage ("columns").  Particles are kept binned by cell.  Interior
// cell ci (x fastest, 0 <= ci < Nc) owns a column of D rows; 64 consecutive
// cells form a group, and row r of a group is 64 consecutive slots, one per
// cell:
//     slot(ci, r) = ((ci / 64) * D + r) * 64 + ci % 64
// A slot holds one particle as 8 reals (x, y, z, ux, uy, uz, w, unused), two
// 4-wide vector accesses, so a scattered particle store touches one 32-byte
// (float) or 64-byte (double) piece of memory.
// A wave owns the 64 cells of one group, one cell per lane.  Looping over the
// rows, every load and store of a row is one contiguous 64-wide access, and
// each lane only ever deposits current into its own cell: the 12 accumulator
// values of the cell live in the lane's registers and are written once, with
// no atomics and in a fixed order (so the deposit is deterministic).
//
// Occupancy.  A column uses rows [0, len) and a bit mask per cell marks the
// occupied rows (words mask[w * Nc_pad + ci], bit r of word w = row 64w + r).
// The push normally keeps a particle that stays in its cell in its row, so
// the stores of a row stay one contiguous access; the rows of particles that
// left become holes, which the particles arriving from the neighbours fill
// before the column grows.  Every few steps a compacting push packs each
// column again (stayers to rows 0, 1, 2, ...).
//
// Cell crossings.  The reference's move_p loop (src/move_p.h:93-371) walks a
// particle through up to four streaks, each in one cell.  Here every streak
// is executed by the lane that owns the streak's cell:
//   pass 1   pushes every particle of the cell (src/push.h:65-295) and runs
//            its first streak.  A particle that ends in the cell is written
//            to the output column ("stayer").  A particle that crosses face
//            d is written as a record (position on the entry face of the
//            neighbour, remaining displacement, momentum, weight) into slot
//            k of the (cell, d) record list.  The list of one (cell, d) is
//            contiguous, so each lane writes and later reads its records
//            sequentially.
//   pass 2-4 the owner of the neighbour cell reads the records addressed to
//            it from its six neighbours, runs the next streak, and either
//            appends the particle to its own column or forwards a new record.
// The arithmetic of every streak is the reference's, so positions, momenta
// and cells are bit-identical to one uninterrupted move_p.  Record lists
// that fill up spill into an overflow list that a final fallback kernel
// finishes with atomic deposits (rare by construction of the capacities).
#pragma once

#include "cpic_kernels.h"

namespace cpic {

constexpr int kRecordLen = 12;  // pos[3], rem[3], u[3], w, 2 pad
constexpr int kSlotLen = 8;     // x, y, z, ux, uy, uz, w, pad

__device__ __forceinline__ size_t col_slot(int ci, int row, int D) {
  return ((size_t)(ci >> 6) * (size_t)D + (size_t)row) * 64 + (size_t)(ci & 63);
}

struct CellGrid {
  int nx, ny, nz;  // interior extents
  int nc;          // nx*ny*nz
  int sy, sz;      // voxel strides
};

__device__ __forceinline__ int interior_to_voxel(const CellGrid& g, int ci) {
  const int ix = ci % g.nx;
  const int rest = ci / g.nx;
  const int iy = rest % g.ny;
  const int iz = rest / g.ny;
  return (ix + 1) + g.sy * (iy + 1) + g.sz * (iz + 1);
}

// Interior cell across face d of cell ci (d = 0..5: -x,-y,-z,+x,+y,+z), with
// the periodic wrap of move_p.h:246-295.
__device__ __forceinline__ int face_neighbor(const CellGrid& g, int ci, int d) {
  int ix = ci % g.nx;
  const int rest = ci / g.nx;
  int iy = rest % g.ny;
  int iz = rest / g.ny;
  const int step = (d < 3) ? -1 : 1;
  const int ax = (d < 3) ? d : d - 3;
  if (ax == 0) ix = (ix + step + g.nx) % g.nx;
  else if (ax == 1) iy = (iy + step + g.ny) % g.ny;
  else iz = (iz + step + g.nz) % g.nz;
  return ix + g.nx * (iy + g.ny * iz);
}

__device__ __forceinline__ int wave_max_int(int v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
  return v;
}

__device__ __forceinline__ int wave_sum_int(int v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// correction term of a streak (move_p.h:154): real_t product scaled by the
// double 1/3, rounded back to real_t
__device__ __forceinline__ float streak_v5(float q, float a, float b, float c) {
  return (float)((double)(q * a * b * c) * (1. / 3.));
}
__device__ __forceinline__ double streak_v5(double q, double a, double b, double c) {
  return q * a * b * c * (1. / 3.);
}

// sqrtf of push.h:148,169 narrows a double argument to float
__device__ __forceinline__ float sqrtf_narrow(float v) { return sqrtf(v); }
__device__ __forceinline__ double sqrtf_narrow(double v) { return (double)sqrtf((float)v); }

// The four quadrant contributions of one current component for a streak
// with displacement dA along the component axis and midpoint (mB, mC) on
// the transverse axes (push.h:218-232, move_p.h:156-170).
template <class R>
__device__ __forceinline__ void quadrants(R q, R dA, R mB, R mC, R v5, R out[4]) {
  R t = q * dA;
  R s1 = t * mB;
  R s0 = t - s1;
  s1 += t;
  R c = R(1) + mC;
  R s2 = s0 * c;
  R s3 = s1 * c;
  c = R(1) - mC;
  s0 *= c;
  s1 *= c;
  out[0] = s0 + v5;
  out[1] = s1 - v5;
  out[2] = s2 - v5;
  out[3] = s3 + v5;
}

template <class R>
__device__ __forceinline__ void streak_values(R q, R dx, R dy, R dz, R mx, R my, R mz, R v5,
                                              R o[12]) {
  quadrants(q, dx, my, mz, v5, o + 0);  // jx: (x,y,z)
  quadrants(q, dy, mz, mx, v5, o + 4);  // jy: (y,z,x)
  quadrants(q, dz, mx, my, v5, o + 8);  // jz: (z,x,y)
}

// One iteration of the move_p loop body (move_p.h:101-209, 218-227,
// 368-370) for a particle in the calling lane's cell: the streak up to the
// first face or the end of the track.  Adds the streak's 12 contributions to
// acc, advances pos/rem, and returns the face crossed (0..5: -x,-y,-z,
// +x,+y,+z) or -1 when the track ends in the cell.  On a crossing pos is
// already expressed in the neighbour's coordinates (the crossed coordinate
// flipped to the entry face).
template <class R>
__device__ __forceinline__ int streak(R pos[3], R rem[3], R q, R acc[12]) {
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
  R o[12];
  streak_values(q, dsp[0], dsp[1], dsp[2], mid[0], mid[1], mid[2], v5, o);
#pragma unroll
  for (int k = 0; k < 12; ++k) acc[k] += o[k];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    rem[k] -= dsp[k];
    pos[k] += dsp[k] + dsp[k];
  }
  if (axis == 3) return -1;
  const R side = dir[axis];
  if (axis == 0) pos[0] = -side;
  else if (axis == 1) pos[1] = -side;
  else pos[2] = -side;
  return axis + ((side > 0) ? 3 : 0);
}

template <class R> struct ColState {
  const R* in;  // input columns, kSlotLen reals per slot
  R* out;       // output columns
  const int* cnt_in;   // rows in use per cell (len)
  int* cnt_out;
  const unsigned long long* mask_in;  // occupied rows, [w][Nc_pad]
  unsigned long long* mask_out;
  int D;
  int MW;      // mask words per cell = ceil(D / 64)
  int nc_pad;  // interior cells rounded up to whole groups
};

__device__ __forceinline__ unsigned long long low_bits(int n) {
  return n >= 64 ? ~0ull : (n <= 0 ? 0ull : ((1ull << n) - 1));
}

// Hands out free rows of one lane's output column: first the holes below
// len (in row order), then new rows at the end.  The mask word being edited
// is cached in a register and written back when the cursor leaves it.
template <class R> struct RowCursor {
  int len, hw;
  unsigned long long word, freeb;
  bool have, append;
  __device__ void start(int l) {
    len = l;
    hw = -1;
    word = freeb = 0;
    have = append = false;
  }
  __device__ int next(const ColState<R>& cs, int ci) {
    for (int guard = 0; guard < 64; ++guard) {
      if (!append) {
        if (have && freeb) {
          const int b = __ffsll((long long)freeb) - 1;
          freeb &= freeb - 1;
          word |= 1ull << b;
          return hw * 64 + b;
        }
        if (have) cs.mask_out[(size_t)hw * cs.nc_pad + ci] = word;
        have = false;
        ++hw;
        if (hw * 64 >= len) {
          append = true;
          continue;
        }
        word = cs.mask_out[(size_t)hw * cs.nc_pad + ci];
        freeb = ~word & low_bits(len - hw * 64);
        have = true;
      } else {
        if (len >= cs.D) return -1;
        const int row = len++;
        const int w = row >> 6;
        if (!have || w != hw) {
          if (have) cs.mask_out[(size_t)hw * cs.nc_pad + ci] = word;
          hw = w;
          word = cs.mask_out[(size_t)hw * cs.nc_pad + ci];
          have = true;
        }
        word |= 1ull << (row & 63);
        return row;
      }
    }
    return -1;
  }
  __device__ void finish(const ColState<R>& cs, int ci) {
    if (have) cs.mask_out[(size_t)hw * cs.nc_pad + ci] = word;
    have = false;
  }
};

template <class R> struct RecordList {
  R* rec;    // [Nc][6][M] records of kRecordLen
  int* cnt;  // [6][Nc]
  int M;
};

template <class R>
__device__ __forceinline__ R* record_at(const RecordList<R>& L, int ci, int d, int k) {
  return L.rec + ((((size_t)ci * 6 + d) * L.M) + k) * kRecordLen;
}

template <class R> struct Overflow {
  R* rec;           // [cap] records
  int* cell;        // [cap] interior cell the record is addressed to
  unsigned* count;  // number of records written (may exceed cap: error)
  int cap;
};

template <class R>
__device__ __forceinline__ void put_record(const RecordList<R>& L, const Overflow<R>& ovf,
                                           const CellGrid& g, int ci, int d, int& k, const R pos[3],
                                           const R rem[3], R ux, R uy, R uz, R w, int* err) {
  typedef typename Vec4T<R>::type V4;
  V4 a, b, c;
  a.x = pos[0]; a.y = pos[1]; a.z = pos[2]; a.w = rem[0];
  b.x = rem[1]; b.y = rem[2]; b.z = ux;     b.w = uy;
  c.x = uz;     c.y = w;      c.z = 0;      c.w = 0;
  V4* dst;
  if (k < L.M) {
    dst = reinterpret_cast<V4*>(record_at(L, ci, d, k));
  } else {
    const unsigned slot = atomicAdd(ovf.count, 1u);
    if (slot >= (unsigned)ovf.cap) {
      atomicOr(err, kErrOverflowFull);
      ++k;
      return;
    }
    ovf.cell[slot] = face_neighbor(g, ci, d);
    dst = reinterpret_cast<V4*>(ovf.rec + (size_t)slot * kRecordLen);
  }
  dst[0] = a;
  dst[1] = b;
  dst[2] = c;
  ++k;
}

template <class R>
__device__ __forceinline__ void store_particle(const ColState<R>& cs, int ci, int row, R x, R y, R z,
                                               R ux, R uy, R uz, R w) {
  typedef typename Vec4T<R>::type V4;
  V4 a, b;
  a.x = x;  a.y = y;  a.z = z;  a.w = ux;
  b.x = uy; b.y = uz; b.z = w;  b.w = 0;
  V4* dst = reinterpret_cast<V4*>(cs.out + col_slot(ci, row, cs.D) * kSlotLen);
  dst[0] = a;
  dst[1] = b;
}

template <class R> struct PushPass {
  ColState<R> cs;
  RecordList<R> out;  // records of first crossings
  Overflow<R> ovf;
  const R* interp;
  R* acc;  // slot-major [12][Ng]
  CellGrid g;
  int ncells;  // Ng
  R qdt_2mc, cdt_dx, cdt_dy, cdt_dz, qsp;
  int accumulate;  // 1: acc += (push semantics), 0: acc = (after a clear)
  int compact;     // 1: pack stayers to rows 0,1,..; 0: keep stayers in their row
  int* err;
  unsigned* movers;
  int* max_count;
};

// Pass 1: push + first streak, one lane per cell (src/push.h:65-295).
template <class R>
__global__ void __launch_bounds__(kBlock) k_push_pass(PushPass<R> a) {
  const int ci = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = ci < a.g.nc;
  const int len = valid ? a.cs.cnt_in[ci] : 0;
  const int lmax = wave_max_int(len);
  const int v = valid ? interior_to_voxel(a.g, ci) : 0;
  typedef typename Vec4T<R>::type V4;
  V4 e0 = {}, e1 = {}, e2 = {}, b0 = {}, b1 = {};
  if (valid) {
    const V4* f = reinterpret_cast<const V4*>(a.interp + (size_t)v * kInterpStride);
    e0 = f[0];
    e1 = f[1];
    e2 = f[2];
    b0 = f[3];
    b1 = f[4];
  }
  const R one = 1.;
  const R one_third = 1. / 3.;
  const R two_fifteenths = 2. / 15.;
  const R qdt_2mc = a.qdt_2mc;
  R acc[12];
#pragma unroll
  for (int k = 0; k < 12; ++k) acc[k] = 0;
  int ns = 0, hi = 0, nmov = 0;
  int kd[6] = {0, 0, 0, 0, 0, 0};
  unsigned long long m_in = 0, m_out = 0;
  const size_t MP = a.cs.nc_pad;

  for (int j = 0; j < lmax; ++j) {
    const int b = j & 63;
    if (b == 0) m_in = valid ? a.cs.mask_in[(size_t)(j >> 6) * MP + ci] : 0ull;
    if (j < len && ((m_in >> b) & 1ull)) {
      const V4* src = reinterpret_cast<const V4*>(a.cs.in + col_slot(ci, j, a.cs.D) * kSlotLen);
      const V4 pa = src[0], pb = src[1];
      const R dx = pa.x, dy = pa.y, dz = pa.z;
      R ux = pa.w, uy = pb.x, uz = pb.y;
      const R w = pb.z;
      // field gather, push.h:124-138
      const R hax = qdt_2mc * ((e0.x + dy * e0.y) + dz * (e0.z + dy * e0.w));
      const R hay = qdt_2mc * ((e1.x + dz * e1.y) + dx * (e1.z + dz * e1.w));
      const R haz = qdt_2mc * ((e2.x + dx * e2.y) + dy * (e2.z + dx * e2.w));
      const R cbx = b0.x + dx * b0.y;
      const R cby = b0.z + dy * b0.w;
      const R cbz = b1.x + dz * b1.y;
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
      // displacement in cell units, push.h:169-182
      v0 = one / sqrtf_narrow(one + (ux * ux + (uy * uy + uz * uz)));
      R ddx = ux * a.cdt_dx;
      R ddy = uy * a.cdt_dy;
      R ddz = uz * a.cdt_dz;
      ddx *= v0;
      ddy *= v0;
      ddz *= v0;
      const R mx = dx + ddx, my = dy + ddy, mz = dz + ddz;
      const R nx = mx + ddx, ny = my + ddy, nz = mz + ddz;
      const R q = w * a.qsp;
      R px, py, pz;
      bool stays;
      if (nx <= one && ny <= one && nz <= one && -nx <= one && -ny <= one && -nz <= one) {
        const R v5 = q * ddx * ddy * ddz * one_third;
        R o[12];
        streak_values(q, ddx, ddy, ddz, mx, my, mz, v5, o);
#pragma unroll
        for (int k = 0; k < 12; ++k) acc[k] += o[k];
        px = nx;
        py = ny;
        pz = nz;
        stays = true;
      } else {
        ++nmov;
        R pos[3] = {dx, dy, dz};
        R rem[3] = {ddx, ddy, ddz};
        const int d = streak(pos, rem, q, acc);
        px = pos[0];
        py = pos[1];
        pz = pos[2];
        stays = d < 0;
        if (!stays) {
          int k = kd[d];
          put_record(a.out, a.ovf, a.g, ci, d, k, pos, rem, ux, uy, uz, w, a.err);
          kd[d] = k;
        }
      }
      if (stays) {
        int row;
        if (a.compact) {
          row = ns++;
        } else {
          row = j;
          m_out |= 1ull << b;
          hi = j + 1;
        }
        store_particle(a.cs, ci, row, px, py, pz, ux, uy, uz, w);
      }
    }
    if (!a.compact && (b == 63 || j == lmax - 1)) {
      if (valid) a.cs.mask_out[(size_t)(j >> 6) * MP + ci] = m_out;
      m_out = 0;
    }
  }
  if (valid) {
    const size_t N = a.ncells;
    if (a.accumulate) {
#pragma unroll
      for (int k = 0; k < 12; ++k) a.acc[k * N + v] += acc[k];
    } else {
#pragma unroll
      for (int k = 0; k < 12; ++k) a.acc[k * N + v] = acc[k];
    }
    if (a.compact) {
      for (int w = 0; w < a.cs.MW; ++w) a.cs.mask_out[(size_t)w * MP + ci] = low_bits(ns - 64 * w);
      a.cs.cnt_out[ci] = ns;
    } else {
      for (int w = (lmax + 63) >> 6; w < a.cs.MW; ++w) a.cs.mask_out[(size_t)w * MP + ci] = 0ull;
      a.cs.cnt_out[ci] = hi;
    }
#pragma unroll
    for (int d = 0; d < 6; ++d) a.out.cnt[(size_t)d * a.g.nc + ci] = min(kd[d], a.out.M);
  }
  const int wm = wave_sum_int(nmov);
  if ((threadIdx.x & 63) == 0) {
    if (wm) atomicAdd(a.movers, (unsigned)wm);
    atomicMax(a.max_count, lmax);
  }
}

template <class R> struct FinishPass {
  ColState<R> cs;  // cs.out / cs.cnt_out: the columns being filled
  RecordList<R> in;
  RecordList<R> out;
  Overflow<R> ovf;
  R* acc;
  CellGrid g;
  int ncells;
  R qsp;
  int last;  // 1: no further pass; new crossings go to the overflow list
  int* err;
};

__device__ __forceinline__ int sel6(int d, int a0, int a1, int a2, int a3, int a4, int a5) {
  return d == 0 ? a0 : d == 1 ? a1 : d == 2 ? a2 : d == 3 ? a3 : d == 4 ? a4 : a5;
}

// Passes 2-4: the next streak of every particle that entered the lane's cell.
// Each lane walks its own arrivals (all directions, in direction then record
// order) in one loop, so a wave iterates max over lanes of the lane totals.
template <class R>
__global__ void __launch_bounds__(kBlock) k_finish_pass(FinishPass<R> a) {
  const int ci = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = ci < a.g.nc;
  const size_t NC = a.g.nc;
  int s0 = 0, s1 = 0, s2 = 0, s3 = 0, s4 = 0, s5 = 0;
  int c0 = 0, c1 = 0, c2 = 0, c3 = 0, c4 = 0, c5 = 0;
  if (valid) {
    // the source of a record of direction d sits across the opposite face
    s0 = face_neighbor(a.g, ci, 3);
    s1 = face_neighbor(a.g, ci, 4);
    s2 = face_neighbor(a.g, ci, 5);
    s3 = face_neighbor(a.g, ci, 0);
    s4 = face_neighbor(a.g, ci, 1);
    s5 = face_neighbor(a.g, ci, 2);
    c0 = a.in.cnt[0 * NC + s0];
    c1 = a.in.cnt[1 * NC + s1];
    c2 = a.in.cnt[2 * NC + s2];
    c3 = a.in.cnt[3 * NC + s3];
    c4 = a.in.cnt[4 * NC + s4];
    c5 = a.in.cnt[5 * NC + s5];
  }
  const int total = c0 + c1 + c2 + c3 + c4 + c5;
  const int tmax = wave_max_int(total);
  int kd[6] = {0, 0, 0, 0, 0, 0};
  if (tmax > 0) {
    const int v = valid ? interior_to_voxel(a.g, ci) : 0;
    RowCursor<R> rows;
    rows.start(valid ? a.cs.cnt_out[ci] : 0);
    R acc[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) acc[k] = 0;
    typedef typename Vec4T<R>::type V4;
    int d = 0, k = 0;
    for (int i = 0; i < tmax; ++i) {
      if (i < total) {
        while (d < 6 && k >= sel6(d, c0, c1, c2, c3, c4, c5)) {
          k = 0;
          ++d;
        }
        if (d >= 6) break;  // unreachable while the counts are consistent
        const int si = sel6(d, s0, s1, s2, s3, s4, s5);
        const V4* src = reinterpret_cast<const V4*>(record_at(a.in, si, d, k));
        ++k;
        const V4 r0 = src[0], r1 = src[1], r2 = src[2];
        R pos[3] = {r0.x, r0.y, r0.z};
        R rem[3] = {r0.w, r1.x, r1.y};
        const R ux = r1.z, uy = r1.w, uz = r2.x, w = r2.y;
        const R q = w * a.qsp;
        const int d2 = streak(pos, rem, q, acc);
        if (d2 < 0) {
          const int row = rows.next(a.cs, ci);
          if (row >= 0)
            store_particle(a.cs, ci, row, pos[0], pos[1], pos[2], ux, uy, uz, w);
          else
            atomicOr(a.err, kErrColumnFull);
        } else {
          int kk = a.last ? a.out.M : kd[d2];  // last pass: straight to overflow
          put_record(a.out, a.ovf, a.g, ci, d2, kk, pos, rem, ux, uy, uz, w, a.err);
          if (!a.last) kd[d2] = kk;
        }
      }
    }
    if (valid) {
      rows.finish(a.cs, ci);
      a.cs.cnt_out[ci] = rows.len;
      if (total > 0) {
        const size_t N = a.ncells;
#pragma unroll
        for (int kq = 0; kq < 12; ++kq) a.acc[kq * N + v] += acc[kq];
      }
    }
  }
  if (valid && !a.last) {
#pragma unroll
    for (int dd = 0; dd < 6; ++dd) a.out.cnt[(size_t)dd * NC + ci] = min(kd[dd], a.out.M);
  }
}

// Fallback for spilled records: one thread per record walks the particle to
// the end of its track with atomic deposits and appends it to its final
// cell's column.  Runs after all passes, so it never races with an owner.
template <class R> struct SpillPass {
  ColState<R> cs;
  Overflow<R> ovf;
  R* acc;
  CellGrid g;
  int ncells;
  R qsp;
  int* err;
};

template <class R>
__global__ void __launch_bounds__(kBlock) k_spill_pass(SpillPass<R> a) {
  const unsigned total = min(*a.ovf.count, (unsigned)a.ovf.cap);
  typedef typename Vec4T<R>::type V4;
  for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
    const V4* src = reinterpret_cast<const V4*>(a.ovf.rec + (size_t)i * kRecordLen);
    const V4 r0 = src[0], r1 = src[1], r2 = src[2];
    R pos[3] = {r0.x, r0.y, r0.z};
    R rem[3] = {r0.w, r1.x, r1.y};
    const R ux = r1.z, uy = r1.w, uz = r2.x, w = r2.y;
    const R q = w * a.qsp;
    int ci = a.ovf.cell[i];
    if ((unsigned)ci >= (unsigned)a.g.nc) {
      atomicOr(a.err, kErrBadCell);
      continue;
    }
    bool done = false;
    for (int it = 0; it < kMaxStreaks; ++it) {
      R acc[12];
#pragma unroll
      for (int k = 0; k < 12; ++k) acc[k] = 0;
      const int d = streak(pos, rem, q, acc);
      const int v = interior_to_voxel(a.g, ci);
      const size_t N = a.ncells;
#pragma unroll
      for (int k = 0; k < 12; ++k) unsafeAtomicAdd(a.acc + k * N + v, acc[k]);
      if (d < 0) {
        done = true;
        break;
      }
      ci = face_neighbor(a.g, ci, d);
    }
    if (!done) {
      atomicOr(a.err, kErrStreakLimit);
      continue;
    }
    const int row = atomicAdd(a.cs.cnt_out + ci, 1);
    if (row < a.cs.D) {
      store_particle(a.cs, ci, row, pos[0], pos[1], pos[2], ux, uy, uz, w);
      atomicOr(a.cs.mask_out + (size_t)(row >> 6) * a.cs.nc_pad + ci, 1ull << (row & 63));
    } else {
      atomicOr(a.err, kErrColumnFull);
    }
  }
}

// sum w (gamma - 1) over the particles of the columns -> partials[block]
template <class R>
__global__ void __launch_bounds__(kBlock)
k_kinetic_columns(ColState<R> cs, int nc, double* __restrict__ partials) {
  double v[1] = {0};
  for (int ci = blockIdx.x * blockDim.x + threadIdx.x; ci < nc; ci += gridDim.x * blockDim.x) {
    const int n = cs.cnt_in[ci];
    unsigned long long m = 0;
    for (int j = 0; j < n; ++j) {
      if ((j & 63) == 0) m = cs.mask_in[(size_t)(j >> 6) * cs.nc_pad + ci];
      if (!((m >> (j & 63)) & 1ull)) continue;
      const R* p = cs.in + col_slot(ci, j, cs.D) * kSlotLen;
      const double a = p[3], b = p[4], c = p[5];
      const double u2 = a * a + b * b + c * c;
      v[0] += (double)p[6] * (u2 / (sqrt(1.0 + u2) + 1.0));
    }
  }
  block_sum_store<1>(v, partials + blockIdx.x);
}

}  // namespace cpic
