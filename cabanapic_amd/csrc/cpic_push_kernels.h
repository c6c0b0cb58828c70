// cpic_push_kernels.h -- the push + move_p + deposit on cell-binned particles.
//
// Particle storage ("columns").  Particles are kept binned by cell.  Interior
// cell ci (x fastest, 0 <= ci < Nc) owns a column of D rows; 64 consecutive
// cells form a group, and row r of a group is 64 consecutive slots, one per
// cell:
//     slot(ci, r) = ((ci / 64) * D + r) * 64 + ci % 64
// A slot holds one particle as 8 reals (x, y, z, ux, uy, uz, w, tag), two
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
// The push normally keeps every particle in its row, so the stores of a row
// stay one contiguous access; the rows of particles that left become holes,
// which the particles arriving from the neighbours fill before the column
// grows.  Every few steps a compacting push packs each column again (rows
// 0, 1, 2, ...).
//
// Cell crossings.  The reference's move_p loop (src/move_p.h:93-371) walks a
// particle through up to four streaks, each in one cell.  Here every streak
// is executed by the lane that owns the streak's cell:
//   pass 1   pushes every particle of the cell (src/push.h:65-295) and runs
//            its first streak.  Every particle is stored in its own row (a
//            contiguous 64-wide access): one that stays with its new
//            position, one that leaves through face d with its position
//            before the push and its new momentum.  A departure keeps its
//            row reserved in the mask and sets a bit in the departure mask of
//            face d.
//   pass 2   the owner of the neighbour cell finds the departures through
//            its neighbours' departure masks, reads them from their rows,
//            recomputes displacement and first streak with the same
//            operations (bit-identical), and runs the next streak; the
//            particle is appended to its column or written as a record.
//            Each cell then releases the rows of its own departures.
//   pass 3-4 the streaks after the 2nd and 3rd crossings, from record lists
//            read by the owner of the cell each record entered.
// Writing departures in place instead of as scattered records keeps every
// write of pass 1 a full-row access.  The arithmetic of every streak is the
// reference's, so positions, momenta and cells are bit-identical to one
// uninterrupted move_p.  Record lists that fill up spill into an overflow
// list that a final fallback kernel finishes with atomic deposits (rare by
// construction of the capacities).
#pragma once

#include "cpic_kernels.h"

namespace cpic {

// A crossing record: pos[3], rem[3], u[3], w, tag, then padding to whole 32-byte
// sectors (64 bytes in float, 96 in double), so that a record write never
// leaves a partially written sector, which memory would have to read back.
template <class R> struct RecordLen { static constexpr int value = sizeof(R) == 4 ? 16 : 12; };
constexpr int record_bytes(int real_size) { return real_size == 4 ? 64 : 96; }
constexpr int kSlotLen = 8;     // x, y, z, ux, uy, uz, w, tag

__device__ __forceinline__ size_t col_slot(int ci, int row, int D) {
  return ((size_t)(ci >> 6) * (size_t)D + (size_t)row) * 64 + (size_t)(ci & 63);
}

// Third word of a gfx9 buffer descriptor for raw dword access (32-bit data
// format; no swizzle, no index stride).
constexpr int kBufferWord3 = 0x00020000;
typedef unsigned U32x4 __attribute__((ext_vector_type(4)));

// One slot (two 4-wide vectors) stored through buffer descriptor rs at byte
// offset off; an offset at or past the descriptor's size drops the store.
__device__ __forceinline__ void store_slot(__amdgpu_buffer_rsrc_t rs, int off,
                                           const Vec4T<float>::type& a,
                                           const Vec4T<float>::type& b) {
  __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(U32x4, a), rs, off, 0, 0);
  __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(U32x4, b), rs, off + 16, 0, 0);
}
__device__ __forceinline__ void store_slot(__amdgpu_buffer_rsrc_t rs, int off,
                                           const Vec4T<double>::type& a,
                                           const Vec4T<double>::type& b) {
  typedef double D2 __attribute__((ext_vector_type(2)));
  const D2 p[4] = {a.xy, a.zw, b.xy, b.zw};
#pragma unroll
  for (int k = 0; k < 4; ++k)
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(U32x4, p[k]), rs, off + 16 * k, 0, 0);
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

// All six face neighbours of interior cell ci, nb[d] = face_neighbor(g, ci,
// d), and its voxel, from one set of divisions.
__device__ __forceinline__ void face_neighbors6(const CellGrid& g, int ci, int nb[6],
                                                int& voxel) {
  const int ix = ci % g.nx;
  const int rest = ci / g.nx;
  const int iy = rest % g.ny;
  const int iz = rest / g.ny;
  const int sy = g.nx, sz = g.nx * g.ny;
  nb[0] = ix > 0 ? ci - 1 : ci + (g.nx - 1);
  nb[3] = ix < g.nx - 1 ? ci + 1 : ci - (g.nx - 1);
  nb[1] = iy > 0 ? ci - sy : ci + (g.ny - 1) * sy;
  nb[4] = iy < g.ny - 1 ? ci + sy : ci - (g.ny - 1) * sy;
  nb[2] = iz > 0 ? ci - sz : ci + (g.nz - 1) * sz;
  nb[5] = iz < g.nz - 1 ? ci + sz : ci - (g.nz - 1) * sz;
  voxel = (ix + 1) + g.sy * (iy + 1) + g.sz * (iz + 1);
}

// Wave-wide reductions; the result is read back as a scalar, so the
// compiler knows it is the same in every lane (branches and loops on it are
// then scalar, not exec-masked).
__device__ __forceinline__ int wave_max_int(int v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
  return __builtin_amdgcn_readfirstlane(v);
}

__device__ __forceinline__ int wave_sum_int(int v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return __builtin_amdgcn_readfirstlane(v);
}

// Marks a loaded vector as consumed at this point (an empty asm that reads
// and writes its registers): the wait for the load is placed here, and the
// load cannot be moved past it.
template <class V> __device__ __forceinline__ void settle(V& v) { asm volatile("" : "+v"(v)); }

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

// Field gather and relativistic Boris push with the tan-series rotation,
// push.h:124-163, for a particle at (dx, dy, dz) of a cell whose
// interpolator is (e0, e1, e2, b0, b1); updates (ux, uy, uz).
template <class V4, class R>
__device__ __forceinline__ void boris(const V4& e0, const V4& e1, const V4& e2, const V4& b0,
                                      const V4& b1, R qdt_2mc, R dx, R dy, R dz, R& ux, R& uy,
                                      R& uz) {
  const R one = 1.;
  const R one_third = 1. / 3.;
  const R two_fifteenths = 2. / 15.;
  const R hax = qdt_2mc * ((e0.x + dy * e0.y) + dz * (e0.z + dy * e0.w));
  const R hay = qdt_2mc * ((e1.x + dz * e1.y) + dx * (e1.z + dz * e1.w));
  const R haz = qdt_2mc * ((e2.x + dx * e2.y) + dy * (e2.z + dx * e2.w));
  const R cbx = b0.x + dx * b0.y;
  const R cby = b0.z + dy * b0.w;
  const R cbz = b1.x + dz * b1.y;
  ux += hax;
  uy += hay;
  uz += haz;
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
}

// Displacement in cell units of the pushed momentum, push.h:169-182; pass 2
// recomputes it from the stored momentum with the same operations.
template <class R>
__device__ __forceinline__ void displacement(R ux, R uy, R uz, R cdt_dx, R cdt_dy, R cdt_dz,
                                             R& ddx, R& ddy, R& ddz) {
  const R one = 1.;
  const R v0 = one / sqrtf_narrow(one + (ux * ux + (uy * uy + uz * uz)));
  ddx = ux * cdt_dx;
  ddy = uy * cdt_dy;
  ddz = uz * cdt_dz;
  ddx *= v0;
  ddy *= v0;
  ddz *= v0;
}

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

// The 12 accumulator values of a lane's cell.  Plain: one real_t sum each,
// added to in streak order.  Compensated (K = true): each sum carries the
// rounding error of its additions, recovered exactly by TwoSum (Knuth; six
// operations, no branch) and added back when the sum is read; the result is
// the sum rounded once or twice instead of once per addition.
template <class R, bool K> struct LaneAcc {
  R s[12];
  R c[K ? 12 : 1];
  __device__ __forceinline__ void zero() {
#pragma unroll
    for (int k = 0; k < 12; ++k) s[k] = 0;
    if (K) {
#pragma unroll
      for (int k = 0; k < 12; ++k) c[K ? k : 0] = 0;
    }
  }
  __device__ __forceinline__ void add(int k, R v) {
    if (K) {
      const R t = s[k] + v;
      const R vv = t - s[k];
      const R e = (s[k] - (t - vv)) + (v - vv);
      c[K ? k : 0] += e;
      s[k] = t;
    } else {
      s[k] += v;
    }
  }
  __device__ __forceinline__ R total(int k) const { return K ? s[k] + c[K ? k : 0] : s[k]; }
};

template <class R>
__device__ __forceinline__ void add12(R acc[12], const R o[12]) {
#pragma unroll
  for (int k = 0; k < 12; ++k) acc[k] += o[k];
}
template <class R, bool K>
__device__ __forceinline__ void add12(LaneAcc<R, K>& acc, const R o[12]) {
#pragma unroll
  for (int k = 0; k < 12; ++k) acc.add(k, o[k]);
}

// The geometry of one iteration of the move_p loop body (move_p.h:101-209,
// 218-227, 368-370) for a particle in the calling lane's cell: the streak
// up to the first face or the end of the track.  Returns the streak's
// displacement dsp and midpoint mid and its correction term v5 (the inputs
// of its 12 contributions, streak_values), advances pos/rem, and returns the
// face crossed (0..5: -x,-y,-z, +x,+y,+z) or -1 when the track ends in the
// cell.  On a crossing pos is already expressed in the neighbour's
// coordinates (the crossed coordinate flipped to the entry face).
template <class R>
__device__ __forceinline__ int streak_geometry(R pos[3], R rem[3], R q, R dsp[3], R mid[3],
                                               R& v5) {
  R len3[3];
  int up = 0;  // bit k: the track moves towards +k
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    dsp[k] = rem[k];
    mid[k] = pos[k];
    const R dir = (dsp[k] > 0) ? R(1) : R(-1);
    up |= (dsp[k] > 0) ? 1 << k : 0;
    len3[k] = (dsp[k] == 0) ? R(3.4e38) : (dir - mid[k]) / dsp[k];
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
  v5 = streak_v5(q, dsp[0], dsp[1], dsp[2]);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    rem[k] -= dsp[k];
    pos[k] += dsp[k] + dsp[k];
  }
  if (axis == 3) return -1;
  // the direction as a bit, not an array read at a run-time index, which
  // would put the array in scratch memory for the dual-number type
  const R side = ((up >> axis) & 1) ? R(1) : R(-1);
  if (axis == 0) pos[0] = -side;
  else if (axis == 1) pos[1] = -side;
  else pos[2] = -side;
  return axis + ((side > 0) ? 3 : 0);
}

// One whole iteration of the move_p loop body: streak_geometry, then the
// streak's 12 contributions added to acc.
template <class R, class A>
__device__ __forceinline__ int streak(R pos[3], R rem[3], R q, A& acc) {
  R dsp[3], mid[3], v5;
  const int d = streak_geometry(pos, rem, q, dsp, mid, v5);
  R o[12];
  streak_values(q, dsp[0], dsp[1], dsp[2], mid[0], mid[1], mid[2], v5, o);
  add12(acc, o);
  return d;
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
  // l: rows in use; w0: mask word 0 of the column (read by the caller
  // together with its other loads)
  __device__ void start(int l, unsigned long long w0) {
    len = l;
    append = false;
    if (l > 0) {
      hw = 0;
      word = w0;
      freeb = ~w0 & low_bits(l);
      have = true;
    } else {
      hw = -1;
      word = freeb = 0;
      have = false;
    }
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

// Record list k-th slot of (cell, d): [d][k][cell], so the records that
// neighbouring lanes write or read for the same (d, k) are adjacent.
constexpr int kMaxRecords = 248;  // record slots per (face, cell): counts fit a byte

template <class R> struct RecordList {
  R* rec;              // [6][M][Nc] records of RecordLen<R> reals
  unsigned char* cnt;  // [6][Nc] records per (face, cell), at most M <= kMaxRecords
  int M;
  int nc;
};

template <class R>
__device__ __forceinline__ R* record_at(const RecordList<R>& L, int ci, int d, int k) {
  return L.rec + ((((size_t)d * L.M) + k) * (size_t)L.nc + ci) * RecordLen<R>::value;
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
                                           const R rem[3], R ux, R uy, R uz, R w, R tag,
                                           int* err) {
  typedef typename Vec4T<R>::type V4;
  V4 a, b, c;
  a.x = pos[0]; a.y = pos[1]; a.z = pos[2]; a.w = rem[0];
  b.x = rem[1]; b.y = rem[2]; b.z = ux;     b.w = uy;
  c.x = uz;     c.y = w;      c.z = tag;    c.w = 0;
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
    dst = reinterpret_cast<V4*>(ovf.rec + (size_t)slot * RecordLen<R>::value);
  }
  dst[0] = a;
  dst[1] = b;
  dst[2] = c;
  if (RecordLen<R>::value == 16) dst[3] = V4{};
  ++k;
}

template <class R>
__device__ __forceinline__ void store_particle(const ColState<R>& cs, int ci, int row, R x, R y, R z,
                                               R ux, R uy, R uz, R w, R tag) {
  typedef typename Vec4T<R>::type V4;
  V4 a, b;
  a.x = x;  a.y = y;  a.z = z;  a.w = ux;
  b.x = uy; b.y = uz; b.z = w;  b.w = tag;
  V4* dst = reinterpret_cast<V4*>(cs.out + col_slot(ci, row, cs.D) * kSlotLen);
  dst[0] = a;
  dst[1] = b;
}

constexpr int kRowBatch = 2;  // rows loaded together by pass 1 (divides 64)

// Departure masks: bit r of word w of (face d, cell c) is set when the
// particle stored in row 64w + r of cell c's output column left the cell
// through face d in pass 1.  Layout [d][w][nc_pad]; entry d = 6 holds the
// union of the six (the cell's own departures), and dep_words[group] the
// number of words of the group's columns that pass 1 wrote.
constexpr int kDepMasks = 7;
__device__ __forceinline__ size_t dep_index(int d, int w, int MW, size_t MP, int ci) {
  return ((size_t)d * (size_t)MW + (size_t)w) * MP + (size_t)ci;
}

// Where pass 1 takes each cell's interpolator from.
enum Gather : int {
  kGatherRead = 0,   // the interpolator array (load_interpolator_array ran)
  kGatherFields = 1,  // computed from the fields, as load_interpolator_array would
  kGatherStore = 2,   // computed from the fields and also stored in the array
};

template <class R> struct PushPass {
  ColState<R> cs;
  unsigned long long* dep;  // departure masks written by this pass
  int* dep_words;           // [nc_pad / 64] words written per group
  R* interp;
  const R* fields;  // read when gather != kGatherRead
  int gather;
  R* acc;  // slot-major [12][Ng]
  CellGrid g;
  int ncells;  // Ng
  R qdt_2mc, cdt_dx, cdt_dy, cdt_dz, qsp;
  int accumulate;  // 1: acc += (push semantics), 0: acc = (after a clear)
  int* err;
  unsigned* movers;
  int* max_count;
};

// The interpolator of voxel v for pass 1: read from the array, or gathered
// from the fields with load_interpolator_array's arithmetic (bit-identical),
// which saves writing the array and reading it back.
template <class R, class V4>
__device__ __forceinline__ void cell_interpolator(const PushPass<R>& a, int v, V4& e0, V4& e1,
                                                  V4& e2, V4& b0, V4& b1) {
  V4* f = reinterpret_cast<V4*>(a.interp + (size_t)v * kInterpStride);
  if (a.gather == kGatherRead) {
    e0 = f[0];
    e1 = f[1];
    e2 = f[2];
    b0 = f[3];
    b1 = f[4];
    return;
  }
  interpolator_at(a.fields, (size_t)a.ncells, a.g.sy, a.g.sz, v, e0, e1, e2, b0, b1);
  if (a.gather == kGatherStore) {
    f[0] = e0;
    f[1] = e1;
    f[2] = e2;
    f[3] = b0;
    f[4] = b1;
  }
}

// Pass 1: push + first streak, one lane per cell (src/push.h:65-295).
// A particle that stays is stored with its new position.  A particle that
// leaves through face d (after depositing its first streak here) is stored
// in the same way but with its position before the push, its new momentum
// and its weight: the owner of the neighbour cell recomputes the
// displacement and the first streak from these in pass 2.  Its row is
// marked in the occupancy mask (reserved until pass 2 has read it) and in
// the departure mask of face d.
template <class R, bool K>
__global__ void __launch_bounds__(kBlock) k_push_pass(PushPass<R> a) {
  if (__builtin_amdgcn_readfirstlane(*a.err)) return;  // an earlier push failed
  const int ci = blockIdx.x * blockDim.x + threadIdx.x;  // < nc_pad
  const bool valid = ci < a.g.nc;
  const int len = valid ? min(a.cs.cnt_in[ci], a.cs.D) : 0;
  const int lmax = wave_max_int(len);
  const int v = valid ? interior_to_voxel(a.g, ci) : 0;
  typedef typename Vec4T<R>::type V4;
  V4 e0 = {}, e1 = {}, e2 = {}, b0 = {}, b1 = {};
  if (valid) cell_interpolator(a, v, e0, e1, e2, b0, b1);
  const R one = 1.;
  const R one_third = 1. / 3.;
  const R qdt_2mc = a.qdt_2mc;
  LaneAcc<R, K> acc;
  acc.zero();
  int ns = 0, nmov = 0;
  unsigned long long m_in = 0;
  const size_t MP = a.cs.nc_pad;
  const int MW = a.cs.MW;
  // departure bits of output word cw, one word per face
  int cw = 0;
  unsigned long long dw0 = 0, dw1 = 0, dw2 = 0, dw3 = 0, dw4 = 0, dw5 = 0;
  auto flush_dep = [&]() {
    a.dep[dep_index(0, cw, MW, MP, ci)] = dw0;
    a.dep[dep_index(1, cw, MW, MP, ci)] = dw1;
    a.dep[dep_index(2, cw, MW, MP, ci)] = dw2;
    a.dep[dep_index(3, cw, MW, MP, ci)] = dw3;
    a.dep[dep_index(4, cw, MW, MP, ci)] = dw4;
    a.dep[dep_index(5, cw, MW, MP, ci)] = dw5;
    a.dep[dep_index(6, cw, MW, MP, ci)] = dw0 | dw1 | dw2 | dw3 | dw4 | dw5;
    dw0 = dw1 = dw2 = dw3 = dw4 = dw5 = 0;
  };

  auto process = [&](bool have, const V4& pa, const V4& pb) {
    if (have) {
      const R dx = pa.x, dy = pa.y, dz = pa.z;
      R ux = pa.w, uy = pb.x, uz = pb.y;
      const R w = pb.z;
      boris(e0, e1, e2, b0, b1, qdt_2mc, dx, dy, dz, ux, uy, uz);
      // displacement in cell units, push.h:169-182
      R ddx, ddy, ddz;
      displacement(ux, uy, uz, a.cdt_dx, a.cdt_dy, a.cdt_dz, ddx, ddy, ddz);
      const R mx = dx + ddx, my = dy + ddy, mz = dz + ddz;
      const R nx = mx + ddx, ny = my + ddy, nz = mz + ddz;
      const R q = w * a.qsp;
      R px, py, pz;
      int d = -1;
      // the deposit's inputs: the whole move (it stays in the cell,
      // push.h:218-232) or the first streak of move_p
      R s0 = ddx, s1 = ddy, s2 = ddz, m0 = mx, m1 = my, m2 = mz, v5;
      if (nx <= one && ny <= one && nz <= one && -nx <= one && -ny <= one && -nz <= one) {
        v5 = q * ddx * ddy * ddz * one_third;
        px = nx;
        py = ny;
        pz = nz;
      } else {
        ++nmov;
        R pos[3] = {dx, dy, dz};
        R rem[3] = {ddx, ddy, ddz};
        R dsp[3], mid[3];
        d = streak_geometry(pos, rem, q, dsp, mid, v5);
        s0 = dsp[0];
        s1 = dsp[1];
        s2 = dsp[2];
        m0 = mid[0];
        m1 = mid[1];
        m2 = mid[2];
        // a particle that leaves is stored with its position before the push
        px = d < 0 ? pos[0] : dx;
        py = d < 0 ? pos[1] : dy;
        pz = d < 0 ? pos[2] : dz;
      }
      {
        R o[12];
        streak_values(q, s0, s1, s2, m0, m1, m2, v5, o);
#pragma unroll
        for (int k = 0; k < 12; ++k) acc.add(k, o[k]);
      }
      const int row = ns++;
      if ((row >> 6) != cw) {
        flush_dep();
        cw = row >> 6;
      }
      if (d >= 0) {
        const unsigned long long bit = 1ull << (row & 63);
        dw0 |= d == 0 ? bit : 0ull;
        dw1 |= d == 1 ? bit : 0ull;
        dw2 |= d == 2 ? bit : 0ull;
        dw3 |= d == 3 ? bit : 0ull;
        dw4 |= d == 4 ? bit : 0ull;
        dw5 |= d == 5 ? bit : 0ull;
      }
      store_particle(a.cs, ci, row, px, py, pz, ux, uy, uz, w, pb.w);
    }
  };

  // Rows are read kRowBatch at a time, all loads of a batch issued before
  // any store of it, so one wait for memory covers the whole batch (vector
  // loads and stores retire in order, so a wait for a load also waits for
  // the stores issued before it).  Every lane issues every load: lanes
  // without a particle in the row read one shared slot instead.
  const R* dummy = a.cs.in;
  for (int j0 = 0; j0 < lmax; j0 += kRowBatch) {
    if ((j0 & 63) == 0) {
      m_in = (valid && j0 < len) ? a.cs.mask_in[(size_t)(j0 >> 6) * MP + ci] : 0ull;
      settle(m_in);  // consume the word here, once per 64 rows
    }
    V4 pa[kRowBatch], pb[kRowBatch];
    bool have[kRowBatch];
#pragma unroll
    for (int u = 0; u < kRowBatch; ++u) {
      const int j = j0 + u;
      have[u] = valid && j < len && ((m_in >> (j & 63)) & 1ull);
      const V4* src = reinterpret_cast<const V4*>(
          have[u] ? a.cs.in + col_slot(ci, j, a.cs.D) * kSlotLen : dummy);
      pa[u] = src[0];
      pb[u] = src[1];
    }
    // one wait for the whole batch, here: the values are marked as used, so
    // the compiler neither moves the loads into the per-row branches nor
    // waits again inside them
#pragma unroll
    for (int u = 0; u < kRowBatch; ++u) {
      settle(pa[u]);
      settle(pb[u]);
    }
#pragma unroll
    for (int u = 0; u < kRowBatch; ++u)
      if (j0 + u < lmax) process(have[u], pa[u], pb[u]);
  }

  const size_t N = a.ncells;
  if (valid) {
    if (a.accumulate) {
#pragma unroll
      for (int k = 0; k < 12; ++k) a.acc[k * N + v] += acc.total(k);
    } else {
#pragma unroll
      for (int k = 0; k < 12; ++k) a.acc[k * N + v] = acc.total(k);
    }
  }
  // the remaining mask and departure words of the column: every lane writes
  // the group's departure words (those of its rows past its own end are
  // zero), the only ones pass 2 reads
  flush_dep();
  for (int w = 0; w < MW; ++w) a.cs.mask_out[(size_t)w * MP + ci] = low_bits(ns - 64 * w);
  const int gw = min((lmax + 63) >> 6, MW);
  for (int w = cw + 1; w < gw; ++w)
    for (int d = 0; d < kDepMasks; ++d) a.dep[dep_index(d, w, MW, MP, ci)] = 0ull;
  a.cs.cnt_out[ci] = ns;
  const int wm = wave_sum_int(nmov);
  if ((threadIdx.x & 63) == 0) {
    if (wm) atomicAdd(a.movers, (unsigned)wm);
    atomicMax(a.max_count, lmax);
    a.dep_words[ci >> 6] = (lmax + 63) >> 6;
  }
}

constexpr int kRecBatch = 2;  // records loaded together by the finish passes

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
  int* flag_set;   // if set: per group, marked when a record is addressed to it
  int* flag_skip;  // if set: per group; unmarked groups have no records (cleared here)
  int* err;
};

__device__ __forceinline__ int sel6(int d, int a0, int a1, int a2, int a3, int a4, int a5) {
  return d == 0 ? a0 : d == 1 ? a1 : d == 2 ? a2 : d == 3 ? a3 : d == 4 ? a4 : a5;
}

// Passes 2-4: the next streak of every particle that entered the lane's cell.
// Each lane walks its own arrivals (all directions, in direction then record
// order) in one loop, so a wave iterates max over lanes of the lane totals.
template <class R, bool K>
__global__ void __launch_bounds__(kBlock) k_finish_pass(FinishPass<R> a) {
  if (__builtin_amdgcn_readfirstlane(*a.err)) return;  // this push already failed
  const int ci = blockIdx.x * blockDim.x + threadIdx.x;
  if (a.flag_skip) {
    // no record of the previous pass is addressed to this wave's cells
    const int gi = __builtin_amdgcn_readfirstlane(ci >> 6);
    if (__builtin_amdgcn_readfirstlane(a.flag_skip[gi]) == 0) return;
    if ((threadIdx.x & 63) == 0) a.flag_skip[gi] = 0;
  }
  const bool valid = ci < a.g.nc;
  const size_t NC = a.g.nc;
  int s0 = 0, s1 = 0, s2 = 0, s3 = 0, s4 = 0, s5 = 0;
  int c0 = 0, c1 = 0, c2 = 0, c3 = 0, c4 = 0, c5 = 0;
  int v = 0;
  if (valid) {
    // the source of a record of direction d sits across the opposite face
    int nb[6];
    face_neighbors6(a.g, ci, nb, v);
    s0 = nb[3];
    s1 = nb[4];
    s2 = nb[5];
    s3 = nb[0];
    s4 = nb[1];
    s5 = nb[2];
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
    RowCursor<R> rows;
    {
      const int l = valid ? min(a.cs.cnt_out[ci], a.cs.D) : 0;
      unsigned long long w0 = valid ? a.cs.mask_out[ci] : 0ull;
      asm volatile("" : "+v"(w0));  // wait for the word here, before the loop
      rows.start(l, w0);
    }
    LaneAcc<R, K> acc;
    acc.zero();
    typedef typename Vec4T<R>::type V4;
    // (d, k): the next record, walked in direction then record order.  Every
    // lane issues every load of a batch (lanes past their last record read
    // one shared record), so one wait covers the batch.
    int d = 0, k = 0;
    const R* rdummy = a.in.rec;
    auto request = [&](int i, V4& r0, V4& r1, V4& r2) {
      const R* p = rdummy;
      if (i < total) {
        while (d < 5 && k >= sel6(d, c0, c1, c2, c3, c4, c5)) {
          k = 0;
          ++d;
        }
        p = record_at(a.in, sel6(d, s0, s1, s2, s3, s4, s5), d, k);
        ++k;
      }
      const V4* src = reinterpret_cast<const V4*>(p);
      r0 = src[0];
      r1 = src[1];
      r2 = src[2];
    };
    auto process = [&](const V4& r0, const V4& r1, const V4& r2) {
      R pos[3] = {r0.x, r0.y, r0.z};
      R rem[3] = {r0.w, r1.x, r1.y};
      const R ux = r1.z, uy = r1.w, uz = r2.x, w = r2.y, tag = r2.z;
      const R q = w * a.qsp;
      const int d2 = streak(pos, rem, q, acc);
      if (d2 < 0) {
        const int row = rows.next(a.cs, ci);
        if (row >= 0)
          store_particle(a.cs, ci, row, pos[0], pos[1], pos[2], ux, uy, uz, w, tag);
        else
          atomicOr(a.err, kErrColumnFull);
      } else {
        int kk = a.last ? a.out.M : kd[d2];  // last pass: straight to overflow
        put_record(a.out, a.ovf, a.g, ci, d2, kk, pos, rem, ux, uy, uz, w, tag, a.err);
        if (!a.last) kd[d2] = kk;
        if (a.flag_set) a.flag_set[face_neighbor(a.g, ci, d2) >> 6] = 1;
      }
    };
    for (int i0 = 0; i0 < tmax; i0 += kRecBatch) {
      V4 r0[kRecBatch], r1[kRecBatch], r2[kRecBatch];
#pragma unroll
      for (int u = 0; u < kRecBatch; ++u) request(i0 + u, r0[u], r1[u], r2[u]);
#pragma unroll
      for (int u = 0; u < kRecBatch; ++u) {
        settle(r0[u]);
        settle(r1[u]);
        settle(r2[u]);
      }
#pragma unroll
      for (int u = 0; u < kRecBatch; ++u)
        if (i0 + u < total) process(r0[u], r1[u], r2[u]);
    }
    if (valid) {
      rows.finish(a.cs, ci);
      a.cs.cnt_out[ci] = rows.len;
      if (total > 0) {
        const size_t N = a.ncells;
#pragma unroll
        for (int kq = 0; kq < 12; ++kq) a.acc[kq * N + v] += acc.total(kq);
      }
    }
  }
  if (valid && !a.last) {
#pragma unroll
    for (int dd = 0; dd < 6; ++dd)
      a.out.cnt[(size_t)dd * NC + ci] = (unsigned char)min(kd[dd], a.out.M);
  }
}

// Pass 1 in place (every step but the compacting ones).  Same arithmetic as
// k_push_pass; the memory pipeline differs.  Every row issues the same loads
// and stores on every lane: a lane without a particle in the row stores
// whatever it loaded when another lane of its 128-byte line holds one (the
// row is free in the output column, and arrivals overwrite it later), and
// otherwise its store is dropped by the buffer descriptor's range check.
// The loads of row j + 1 are issued before the stores of row j, and with a
// fixed number of instructions in between, the wait for them does not wait
// for those stores.
template <class R, bool K>
__global__ void __launch_bounds__(kBlock) k_push_rows(PushPass<R> a) {
  if (__builtin_amdgcn_readfirstlane(*a.err)) return;  // an earlier push failed
  const int ci = blockIdx.x * blockDim.x + threadIdx.x;  // < nc_pad
  const bool valid = ci < a.g.nc;
  const int len = valid ? min(a.cs.cnt_in[ci], a.cs.D) : 0;
  // rows walked: the longest column of the wave, rounded up to even so that
  // both rows of every loop iteration are processed (D is a multiple of 8,
  // so the extra row is a row past every column's end)
  const int lreal = wave_max_int(len);
  const int lmax = (lreal + 1) & ~1;
  const int v = valid ? interior_to_voxel(a.g, ci) : 0;
  typedef typename Vec4T<R>::type V4;
  V4 e0 = {}, e1 = {}, e2 = {}, b0 = {}, b1 = {};
  if (valid) cell_interpolator(a, v, e0, e1, e2, b0, b1);
  const R one = 1.;
  const R one_third = 1. / 3.;
  const R qdt_2mc = a.qdt_2mc;
  LaneAcc<R, K> acc;
  acc.zero();
  int hi = 0, nmov = 0;
  unsigned long long m_in = 0, m_out = 0;
  unsigned long long dw0 = 0, dw1 = 0, dw2 = 0, dw3 = 0, dw4 = 0, dw5 = 0;
  const size_t MP = a.cs.nc_pad;
  const int MW = a.cs.MW;
  const R* const dummy = a.cs.in;
  // Row stores go through a buffer descriptor over the wave's column group:
  // a store at an offset past its end is dropped without a branch.  A lane
  // stores whenever a lane of its 128-byte line of the row holds a particle
  // (a free row receives whatever was loaded), so memory only ever sees whole
  // lines (a partly written line costs a read-modify-write), and rows past
  // the ends of all the columns of a line cost nothing.
  const int lane = threadIdx.x & 63;
  const int grp = __builtin_amdgcn_readfirstlane(ci >> 6);
  const int gbytes = a.cs.D * 64 * kSlotLen * (int)sizeof(R);
  const __amdgpu_buffer_rsrc_t rows_out = __builtin_amdgcn_make_buffer_rsrc(
      a.cs.out + (size_t)grp * a.cs.D * 64 * kSlotLen, 0, gbytes, kBufferWord3);
  constexpr int kLine = 128 / (kSlotLen * (int)sizeof(R));  // lanes per line
  static_assert(kLine >= 1 && kLine < 64 && (kLine & (kLine - 1)) == 0, "line of slots");
  const int line0 = lane & ~(kLine - 1);

  // loads of row r into (x, y); returns whether the row holds a particle
  auto request = [&](int r, V4& x, V4& y) -> bool {
    if ((r & 63) == 0) {
      m_in = (valid && r < len) ? a.cs.mask_in[(size_t)(r >> 6) * MP + ci] : 0ull;
      settle(m_in);  // once per 64 rows
    }
    const bool ok = valid && r < len && ((m_in >> (r & 63)) & 1ull);
    const V4* src = reinterpret_cast<const V4*>(ok ? a.cs.in + col_slot(ci, r, a.cs.D) * kSlotLen
                                                   : dummy);
    x = src[0];
    y = src[1];
    return ok;
  };

  auto process = [&](int j, bool have, const V4& pa, const V4& pb) {
    const int b = j & 63;
    V4 o0 = pa, o1 = pb;  // what row j of the output receives
    if (have) {
      const R dx = pa.x, dy = pa.y, dz = pa.z;
      R ux = pa.w, uy = pb.x, uz = pb.y;
      const R w = pb.z;
      boris(e0, e1, e2, b0, b1, qdt_2mc, dx, dy, dz, ux, uy, uz);
      R ddx, ddy, ddz;
      displacement(ux, uy, uz, a.cdt_dx, a.cdt_dy, a.cdt_dz, ddx, ddy, ddz);
      const R mx = dx + ddx, my = dy + ddy, mz = dz + ddz;
      const R nx = mx + ddx, ny = my + ddy, nz = mz + ddz;
      const R q = w * a.qsp;
      R px, py, pz;
      // the deposit's inputs: the whole move (it stays in the cell,
      // push.h:218-232) or the first streak of move_p
      R s0 = ddx, s1 = ddy, s2 = ddz, m0 = mx, m1 = my, m2 = mz, v5;
      if (nx <= one && ny <= one && nz <= one && -nx <= one && -ny <= one && -nz <= one) {
        v5 = q * ddx * ddy * ddz * one_third;
        px = nx;
        py = ny;
        pz = nz;
      } else {
        ++nmov;
        R pos[3] = {dx, dy, dz};
        R rem[3] = {ddx, ddy, ddz};
        R dsp[3], mid[3];
        const int d = streak_geometry(pos, rem, q, dsp, mid, v5);
        s0 = dsp[0];
        s1 = dsp[1];
        s2 = dsp[2];
        m0 = mid[0];
        m1 = mid[1];
        m2 = mid[2];
        // a particle that leaves is stored with its position before the push
        px = d < 0 ? pos[0] : dx;
        py = d < 0 ? pos[1] : dy;
        pz = d < 0 ? pos[2] : dz;
        const unsigned long long bit = 1ull << b;
        dw0 |= d == 0 ? bit : 0ull;
        dw1 |= d == 1 ? bit : 0ull;
        dw2 |= d == 2 ? bit : 0ull;
        dw3 |= d == 3 ? bit : 0ull;
        dw4 |= d == 4 ? bit : 0ull;
        dw5 |= d == 5 ? bit : 0ull;
      }
      {
        R o[12];
        streak_values(q, s0, s1, s2, m0, m1, m2, v5, o);
#pragma unroll
        for (int k = 0; k < 12; ++k) acc.add(k, o[k]);
      }
      m_out |= 1ull << b;
      hi = j + 1;
      o0.x = px; o0.y = py; o0.z = pz; o0.w = ux;
      o1.x = uy; o1.y = uz; o1.z = w;  o1.w = pb.w;
    }
    const unsigned long long occ = __ballot(have);
    const bool st = (occ >> line0) & ((1ull << kLine) - 1);
    const int off = st ? (j * 64 + lane) * kSlotLen * (int)sizeof(R) : gbytes;
    store_slot(rows_out, off, o0, o1);
    if (b == 63 || j == lmax - 1) {
      const int wd = j >> 6;
      a.cs.mask_out[(size_t)wd * MP + ci] = m_out;
      a.dep[dep_index(0, wd, MW, MP, ci)] = dw0;
      a.dep[dep_index(1, wd, MW, MP, ci)] = dw1;
      a.dep[dep_index(2, wd, MW, MP, ci)] = dw2;
      a.dep[dep_index(3, wd, MW, MP, ci)] = dw3;
      a.dep[dep_index(4, wd, MW, MP, ci)] = dw4;
      a.dep[dep_index(5, wd, MW, MP, ci)] = dw5;
      a.dep[dep_index(6, wd, MW, MP, ci)] = dw0 | dw1 | dw2 | dw3 | dw4 | dw5;
      m_out = dw0 = dw1 = dw2 = dw3 = dw4 = dw5 = 0;
    }
  };

  // two row buffers in turn, one row ahead.  Each wait is preceded by the
  // same instructions on every path into it (two loads and two stores after
  // the awaited loads), so it waits for the awaited loads only; on entry,
  // two stores of values rewritten at the end stand in for the stores of
  // the previous row.
  V4 xa = {}, ya = {}, xb = {}, yb = {};
  bool ha = request(0, xa, ya);
  bool hb = request(1, xb, yb);
  a.cs.cnt_out[ci] = 0;
  a.cs.mask_out[ci] = 0ull;
  for (int j = 0; j < lmax; j += 2) {
    settle(xa);
    settle(ya);
    process(j, ha, xa, ya);
    ha = request(j + 2, xa, ya);
    settle(xb);
    settle(yb);
    process(j + 1, hb, xb, yb);
    hb = request(j + 3, xb, yb);
  }

  if (valid) {
    const size_t N = a.ncells;
    if (a.accumulate) {
#pragma unroll
      for (int k = 0; k < 12; ++k) a.acc[k * N + v] += acc.total(k);
    } else {
#pragma unroll
      for (int k = 0; k < 12; ++k) a.acc[k * N + v] = acc.total(k);
    }
  }
  for (int w = (lmax + 63) >> 6; w < MW; ++w) a.cs.mask_out[(size_t)w * MP + ci] = 0ull;
  a.cs.cnt_out[ci] = hi;
  const int wm = wave_sum_int(nmov);
  if ((threadIdx.x & 63) == 0) {
    if (wm) atomicAdd(a.movers, (unsigned)wm);
    atomicMax(a.max_count, lreal);
    a.dep_words[ci >> 6] = (lmax + 63) >> 6;
  }
}

template <class R> struct ArrivePass {
  ColState<R> cs;                        // cs.out: the columns of pass 1
  const unsigned long long* dep;         // departure masks of pass 1
  const int* dep_words;                  // words of them written per group
  RecordList<R> out;                     // records of the next crossings
  Overflow<R> ovf;
  R* acc;
  CellGrid g;
  int ncells;
  R cdt_dx, cdt_dy, cdt_dz, qsp;
  int* err;
};

// Pass 2: every particle that left a cell in pass 1 continues in the cell
// across the face it crossed.  The owner of that cell finds it through the
// departure masks of its six neighbours, reads it from the neighbour's
// column, recomputes its displacement and first streak (the same operations
// as pass 1, without the deposit, which pass 1 made), and runs the next
// streak.  Afterwards each cell releases the rows of its own departures.
template <class R, bool K>
__global__ void __launch_bounds__(kBlock) k_arrive_pass(ArrivePass<R> a) {
  if (__builtin_amdgcn_readfirstlane(*a.err)) return;  // this push already failed
  // groups in reverse order: the first blocks dispatched take the groups
  // pass 1 wrote last, whose lines the 256 MB Infinity Cache still holds
  // (14 us less than in pass 1's order), and the last ones end at the groups
  // the next pass 1 starts with
  const int ci = (gridDim.x - 1 - blockIdx.x) * blockDim.x + threadIdx.x;  // < nc_pad
  const bool valid = ci < a.g.nc;
  const size_t MP = a.cs.nc_pad;
  const int MW = a.cs.MW;
  const size_t NC = a.g.nc;
  typedef typename Vec4T<R>::type V4;
  // a particle that crossed face d came from the neighbour across the
  // opposite face
  int s0 = 0, s1 = 0, s2 = 0, s3 = 0, s4 = 0, s5 = 0;
  int v = 0;
  if (valid) {
    int nb[6];
    face_neighbors6(a.g, ci, nb, v);
    s0 = nb[3];
    s1 = nb[4];
    s2 = nb[5];
    s3 = nb[0];
    s4 = nb[1];
    s5 = nb[2];
  }
  RowCursor<R> rows;
  {
    const int l = valid ? min(a.cs.cnt_out[ci], a.cs.D) : 0;
    unsigned long long w0 = valid ? a.cs.mask_out[ci] : 0ull;
    settle(w0);
    rows.start(l, w0);
  }
  LaneAcc<R, K> acc;
  acc.zero();
  int kd[6] = {0, 0, 0, 0, 0, 0};
  int total_all = 0;
  const R* dummy = a.cs.out;

  // words of the departure masks to read: those pass 1 wrote for the
  // source groups (usually one).  Pass 1 writes only the words of its
  // group's rows, so a word at or past its group's count is no departure.
  int nw = 0;
  int ws0 = 0, ws1 = 0, ws2 = 0, ws3 = 0, ws4 = 0, ws5 = 0;
  if (valid) {
    ws0 = a.dep_words[s0 >> 6];
    ws1 = a.dep_words[s1 >> 6];
    ws2 = a.dep_words[s2 >> 6];
    ws3 = a.dep_words[s3 >> 6];
    ws4 = a.dep_words[s4 >> 6];
    ws5 = a.dep_words[s5 >> 6];
    nw = max(max(ws0, ws1), max(max(ws2, ws3), max(ws4, ws5)));
  }
  nw = min(wave_max_int(nw), MW);
  for (int w = 0; w < nw; ++w) {
    unsigned long long d0 = 0, d1 = 0, d2 = 0, d3 = 0, d4 = 0, d5 = 0;
    if (valid) {
      if (w < ws0) d0 = a.dep[dep_index(0, w, MW, MP, s0)];
      if (w < ws1) d1 = a.dep[dep_index(1, w, MW, MP, s1)];
      if (w < ws2) d2 = a.dep[dep_index(2, w, MW, MP, s2)];
      if (w < ws3) d3 = a.dep[dep_index(3, w, MW, MP, s3)];
      if (w < ws4) d4 = a.dep[dep_index(4, w, MW, MP, s4)];
      if (w < ws5) d5 = a.dep[dep_index(5, w, MW, MP, s5)];
    }
    const int total = __popcll(d0) + __popcll(d1) + __popcll(d2) + __popcll(d3) +
                      __popcll(d4) + __popcll(d5);
    total_all += total;
    const int tmax = wave_max_int(total);
    if (tmax == 0) continue;
    // the next departure to read: face d, bits left in the current word
    int d = 0;
    unsigned long long bits = d0;
    auto request = [&](int i, V4& x, V4& y) {
      const R* src = dummy;
      if (i < total) {
        while (bits == 0 && d < 5) {
          ++d;
          bits = d == 1 ? d1 : d == 2 ? d2 : d == 3 ? d3 : d == 4 ? d4 : d5;
        }
        if (bits != 0) {  // always, while the counts are consistent
          const int r = 64 * w + __ffsll((long long)bits) - 1;
          bits &= bits - 1;
          const int sc = sel6(d, s0, s1, s2, s3, s4, s5);
          src = a.cs.out + col_slot(sc, r, a.cs.D) * kSlotLen;
        } else {
          atomicOr(a.err, kErrBadCell);
        }
      }
      const V4* p = reinterpret_cast<const V4*>(src);
      x = p[0];
      y = p[1];
    };
    auto process = [&](const V4& pa, const V4& pb) {
      const R ux = pa.w, uy = pb.x, uz = pb.y, w8 = pb.z;
      R pos[3] = {pa.x, pa.y, pa.z};
      R rem[3];
      displacement(ux, uy, uz, a.cdt_dx, a.cdt_dy, a.cdt_dz, rem[0], rem[1], rem[2]);
      const R q = w8 * a.qsp;
      // the first streak, deposited by the neighbour in pass 1: its
      // geometry only
      {
        R dsp[3], mid[3], v5;
        if (streak_geometry(pos, rem, q, dsp, mid, v5) < 0) atomicOr(a.err, kErrBadCell);
      }
      const int d2n = streak(pos, rem, q, acc);
      if (d2n < 0) {
        const int row = rows.next(a.cs, ci);
        if (row >= 0)
          store_particle(a.cs, ci, row, pos[0], pos[1], pos[2], ux, uy, uz, w8, pb.w);
        else
          atomicOr(a.err, kErrColumnFull);
      } else {
        int kk = kd[d2n];
        put_record(a.out, a.ovf, a.g, ci, d2n, kk, pos, rem, ux, uy, uz, w8, pb.w, a.err);
        kd[d2n] = kk;
      }
    };
    for (int i0 = 0; i0 < tmax; i0 += kRecBatch) {
      V4 x[kRecBatch], y[kRecBatch];
#pragma unroll
      for (int u = 0; u < kRecBatch; ++u) request(i0 + u, x[u], y[u]);
#pragma unroll
      for (int u = 0; u < kRecBatch; ++u) {
        settle(x[u]);
        settle(y[u]);
      }
#pragma unroll
      for (int u = 0; u < kRecBatch; ++u)
        if (i0 + u < total) process(x[u], y[u]);
    }
  }

  if (valid) {
    rows.finish(a.cs, ci);
    a.cs.cnt_out[ci] = rows.len;
    if (total_all > 0) {
      const size_t N = a.ncells;
#pragma unroll
      for (int k = 0; k < 12; ++k) a.acc[k * N + v] += acc.total(k);
    }
    // release the rows of this cell's own departures: only their mask bits
    // are cleared here.  Other lanes may still be reading those rows in this
    // pass, but no lane writes them before the next pass, which is the first
    // that can hand them out again.
    const int own_words = min(a.dep_words[ci >> 6], MW);
    for (int w = 0; w < own_words; ++w) {
      const unsigned long long own = a.dep[dep_index(6, w, MW, MP, ci)];
      if (own) a.cs.mask_out[(size_t)w * MP + ci] &= ~own;
    }
#pragma unroll
    for (int dd = 0; dd < 6; ++dd)
      a.out.cnt[(size_t)dd * NC + ci] = (unsigned char)min(kd[dd], a.out.M);
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
  if (__builtin_amdgcn_readfirstlane(*a.err)) return;  // this push already failed
  const unsigned total = min(*a.ovf.count, (unsigned)a.ovf.cap);
  typedef typename Vec4T<R>::type V4;
  for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
    const V4* src = reinterpret_cast<const V4*>(a.ovf.rec + (size_t)i * RecordLen<R>::value);
    const V4 r0 = src[0], r1 = src[1], r2 = src[2];
    R pos[3] = {r0.x, r0.y, r0.z};
    R rem[3] = {r0.w, r1.x, r1.y};
    const R ux = r1.z, uy = r1.w, uz = r2.x, w = r2.y, tag = r2.z;
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
    // a row is taken only below D: a failed claim is handed back, so the
    // count never exceeds the column's capacity
    const int row = atomicAdd(a.cs.cnt_out + ci, 1);
    if (row < a.cs.D) {
      store_particle(a.cs, ci, row, pos[0], pos[1], pos[2], ux, uy, uz, w, tag);
      atomicOr(a.cs.mask_out + (size_t)(row >> 6) * a.cs.nc_pad + ci, 1ull << (row & 63));
    } else {
      atomicSub(a.cs.cnt_out + ci, 1);
      atomicOr(a.err, kErrColumnFull);
    }
  }
}

// uncenter_particles, src/uncenter_p.h:4-105: for every particle, a
// backward half rotation in B with qdt_4mc = -0.5 qdt_2mc and gamma of the
// momentum before it (a real_t sqrt, not the push's narrowed sqrtf), then
// u += hax with the forward qdt_2mc (uncenter_p.h:91-93 adds the forward
// half kick, where VPIC subtracts it).  One lane per cell, in place on the
// current columns; positions and cells are unchanged.
template <class R>
__global__ void __launch_bounds__(kBlock)
k_uncenter_columns(R* __restrict__ cols, const int* __restrict__ cnt,
                   const unsigned long long* __restrict__ mask, const R* __restrict__ interp,
                   CellGrid g, int D, int nc_pad, R qdt_2mc) {
  const int ci = blockIdx.x * blockDim.x + threadIdx.x;
  if (ci >= g.nc) return;
  typedef typename Vec4T<R>::type V4;
  const V4* f = reinterpret_cast<const V4*>(interp + (size_t)interior_to_voxel(g, ci) * kInterpStride);
  const V4 e0 = f[0], e1 = f[1], e2 = f[2], b0 = f[3], b1 = f[4];
  const R qdt_4mc = -0.5 * qdt_2mc;
  const R one = 1.;
  const R one_third = 1. / 3.;
  const R two_fifteenths = 2. / 15.;
  const int n = min(cnt[ci], D);
  unsigned long long m = 0;
  for (int j = 0; j < n; ++j) {
    if ((j & 63) == 0) m = mask[(size_t)(j >> 6) * nc_pad + ci];
    if (!((m >> (j & 63)) & 1ull)) continue;
    V4* slot = reinterpret_cast<V4*>(cols + col_slot(ci, j, D) * kSlotLen);
    V4 pa = slot[0], pb = slot[1];
    const R dx = pa.x, dy = pa.y, dz = pa.z;
    const R hax = qdt_2mc * ((e0.x + dy * e0.y) + dz * (e0.z + dy * e0.w));
    const R hay = qdt_2mc * ((e1.x + dz * e1.y) + dx * (e1.z + dz * e1.w));
    const R haz = qdt_2mc * ((e2.x + dx * e2.y) + dy * (e2.z + dx * e2.w));
    const R cbx = b0.x + dx * b0.y;
    const R cby = b0.z + dy * b0.w;
    const R cbz = b1.x + dz * b1.y;
    R ux = pa.w, uy = pb.x, uz = pb.y;
    R v0 = qdt_4mc / sqrt(one + (ux * ux + (uy * uy + uz * uz)));
    R v1 = cbx * cbx + (cby * cby + cbz * cbz);
    R v2 = (v0 * v0) * v1;
    const R v3 = v0 * (one + v2 * (one_third + v2 * two_fifteenths));
    R v4 = v3 / (one + v1 * (v3 * v3));
    v4 += v4;
    v0 = ux + v3 * (uy * cbz - uz * cby);
    v1 = uy + v3 * (uz * cbx - ux * cbz);
    v2 = uz + v3 * (ux * cby - uy * cbx);
    ux += v4 * (v1 * cbz - v2 * cby);
    uy += v4 * (v2 * cbx - v0 * cbz);
    uz += v4 * (v0 * cby - v1 * cbx);
    pa.w = ux + hax;
    pb.x = uy + hay;
    pb.y = uz + haz;
    slot[0] = pa;
    slot[1] = pb;
  }
}

// sum w (gamma - 1) over the particles of the columns -> partials[block]
template <class R>
__global__ void __launch_bounds__(kBlock)
k_kinetic_columns(ColState<R> cs, int nc, double* __restrict__ partials) {
  double v[1] = {0};
  for (int ci = blockIdx.x * blockDim.x + threadIdx.x; ci < nc; ci += gridDim.x * blockDim.x) {
    const int n = min(cs.cnt_in[ci], cs.D);
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

// Cloud-in-cell charge of the columns' particles at the nodes, FP64.  Node
// v is the low corner of voxel v.  A particle at offset o in [-1, 1] of
// cell i gives the fraction (1 - o)/2 of its charge w qsp to node i and
// (1 + o)/2 to node i + 1 along each axis (the weights of the charge whose
// flux the deposit of src/move_p.h:156-190 carries).  Nodes n+1 are the
// periodic images of nodes 1.  One lane per cell: the 8 corner sums in FP64
// registers, then one FP64 atomic per corner into q (interior node order,
// x fastest, 0-based).
template <class R>
__global__ void __launch_bounds__(kBlock)
k_node_charge_columns(ColState<R> cs, int nx, int ny, int nz, double qsp,
                      double* __restrict__ q) {
  const int nc = nx * ny * nz;
  for (int ci = blockIdx.x * blockDim.x + threadIdx.x; ci < nc; ci += gridDim.x * blockDim.x) {
    double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const int n = min(cs.cnt_in[ci], cs.D);
    unsigned long long m = 0;
    for (int j = 0; j < n; ++j) {
      if ((j & 63) == 0) m = cs.mask_in[(size_t)(j >> 6) * cs.nc_pad + ci];
      if (!((m >> (j & 63)) & 1ull)) continue;
      const R* p = cs.in + col_slot(ci, j, cs.D) * kSlotLen;
      const double ox = p[0], oy = p[1], oz = p[2];
      const double qq = (double)p[6] * qsp;
      const double wx[2] = {0.5 * (1.0 - ox), 0.5 * (1.0 + ox)};
      const double wy[2] = {0.5 * (1.0 - oy), 0.5 * (1.0 + oy)};
      const double wz[2] = {0.5 * (1.0 - oz), 0.5 * (1.0 + oz)};
#pragma unroll
      for (int k = 0; k < 8; ++k) s[k] += qq * wx[k & 1] * wy[(k >> 1) & 1] * wz[k >> 2];
    }
    const int ix = ci % nx, iy = (ci / nx) % ny, iz = ci / (nx * ny);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      if (s[k] == 0.0) continue;
      const int jx = ix + (k & 1) == nx ? 0 : ix + (k & 1);
      const int jy = iy + ((k >> 1) & 1) == ny ? 0 : iy + ((k >> 1) & 1);
      const int jz = iz + (k >> 2) == nz ? 0 : iz + (k >> 2);
      atomicAdd(q + jx + nx * (jy + ny * jz), s[k]);
    }
  }
}

}  // namespace cpic

namespace cpic {

// ---- column maintenance (upload, download, regrowth) -----------------------
// A slot's eighth real carries the particle's tag (its index in the upload,
// as the bits of a 32-bit or 64-bit integer), copied unchanged by every pass.

template <class R> struct TagBits;
template <> struct TagBits<float> {
  __device__ static float make(long long i) { return __int_as_float((int)i); }
  __device__ static long long get(float t) { return (long long)__float_as_int(t); }
};
template <> struct TagBits<double> {
  __device__ static double make(long long i) { return __longlong_as_double(i); }
  __device__ static long long get(double t) { return __double_as_longlong(t); }
};

// Upload: particle p (SoA src, 7 arrays of stride n) goes to column slot
// slot[p], which the host computed (cell and rank in the cell, input order).
template <class R>
__global__ void __launch_bounds__(kBlock)
k_bin_scatter(const R* __restrict__ src, const long long* __restrict__ slot, long n,
              R* __restrict__ cols) {
  typedef typename Vec4T<R>::type V4;
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < n;
       p += (long)gridDim.x * blockDim.x) {
    V4 a, b;
    a.x = src[p];
    a.y = src[n + p];
    a.z = src[2 * n + p];
    a.w = src[3 * n + p];
    b.x = src[4 * n + p];
    b.y = src[5 * n + p];
    b.z = src[6 * n + p];
    b.w = TagBits<R>::make(p);
    V4* dst = reinterpret_cast<V4*>(cols + (size_t)slot[p] * kSlotLen);
    dst[0] = a;
    dst[1] = b;
  }
}

// Occupancy masks of packed columns: rows [0, cnt) of every cell.
__global__ void __launch_bounds__(kBlock)
k_masks_packed(const int* __restrict__ cnt, unsigned long long* __restrict__ mask, int nc,
               int nc_pad, int MW) {
  const int ci = blockIdx.x * blockDim.x + threadIdx.x;
  if (ci >= nc_pad) return;
  const int n = ci < nc ? cnt[ci] : 0;
  for (int w = 0; w < MW; ++w) mask[(size_t)w * nc_pad + ci] = low_bits(n - 64 * w);
}

// Particles held by each cell (occupied rows below its count).
__global__ void __launch_bounds__(kBlock)
k_column_occupancy(const int* __restrict__ cnt, const unsigned long long* __restrict__ mask,
                   int nc, int nc_pad, int D, int* __restrict__ occ) {
  const int ci = blockIdx.x * blockDim.x + threadIdx.x;
  if (ci >= nc) return;
  const int n = min(cnt[ci], D);
  int k = 0;
  for (int w = 0; 64 * w < n; ++w)
    k += __popcll(mask[(size_t)w * nc_pad + ci] & low_bits(n - 64 * w));
  occ[ci] = k;
}

// Download: the particles of cell ci, occupied rows in row order, go to
// positions first[ci], first[ci] + 1, ... of the SoA dst (7 arrays of stride
// n), with their voxel in cell[] and their tag in tag[] (may be null).
template <class R>
__global__ void __launch_bounds__(kBlock)
k_gather_columns(const R* __restrict__ cols, const int* __restrict__ cnt,
                 const unsigned long long* __restrict__ mask, CellGrid g, int D, int nc_pad,
                 const long long* __restrict__ first, long n, R* __restrict__ dst,
                 int* __restrict__ cell, long long* __restrict__ tag) {
  typedef typename Vec4T<R>::type V4;
  const int ci = blockIdx.x * blockDim.x + threadIdx.x;
  if (ci >= g.nc) return;
  const int len = min(cnt[ci], D);
  const int v = interior_to_voxel(g, ci);
  long o = first[ci];
  unsigned long long m = 0;
  for (int j = 0; j < len; ++j) {
    if ((j & 63) == 0) m = mask[(size_t)(j >> 6) * nc_pad + ci];
    if (!((m >> (j & 63)) & 1ull)) continue;
    if (o >= n) break;  // more particles than the host counted: caught by the host
    const V4* s = reinterpret_cast<const V4*>(cols + col_slot(ci, j, D) * kSlotLen);
    const V4 a = s[0], b = s[1];
    dst[o] = a.x;
    dst[n + o] = a.y;
    dst[2 * n + o] = a.z;
    dst[3 * n + o] = a.w;
    dst[4 * n + o] = b.x;
    dst[5 * n + o] = b.y;
    dst[6 * n + o] = b.z;
    cell[o] = v;
    if (tag) tag[o] = TagBits<R>::get(b.w);
    ++o;
  }
}

// Snapshot for the per-step dumps: the in-cell offset x, the momentum ux and
// the voxel of every particle, each written at the particle's upload index
// (the order of the particle arrays the reference dumps, src/helpers.h:
// 31-70); lanes below nx also copy ex at voxel (lane + 1, 1, 1) (the line of
// src/fields.h:318-350), and lane 0 the last energies (3 doubles).  A tag
// outside [0, n) flags kErrBadCell.
template <class R>
__global__ void __launch_bounds__(kBlock)
k_snapshot(const R* __restrict__ cols, const int* __restrict__ cnt,
           const unsigned long long* __restrict__ mask, CellGrid g, int D, int nc_pad, long n,
           const R* __restrict__ ex, int sy, int sz, const double* __restrict__ energies,
           R* __restrict__ x_out, R* __restrict__ ux_out, int* __restrict__ cell_out,
           R* __restrict__ ex_out, double* __restrict__ energies_out, int* __restrict__ err) {
  typedef typename Vec4T<R>::type V4;
  const int ci = blockIdx.x * blockDim.x + threadIdx.x;
  if (ci < g.nx) ex_out[ci] = ex[(ci + 1) + sy + sz];
  if (ci < 3) energies_out[ci] = energies[ci];
  if (ci >= g.nc) return;
  const int len = min(cnt[ci], D);
  const int v = interior_to_voxel(g, ci);
  unsigned long long m = 0;
  for (int j = 0; j < len; ++j) {
    if ((j & 63) == 0) m = mask[(size_t)(j >> 6) * nc_pad + ci];
    if (!((m >> (j & 63)) & 1ull)) continue;
    const V4* s = reinterpret_cast<const V4*>(cols + col_slot(ci, j, D) * kSlotLen);
    const V4 a = s[0], b = s[1];
    const long long t = TagBits<R>::get(b.w);
    if (t < 0 || t >= n) {
      atomicOr(err, kErrBadCell);
      continue;
    }
    x_out[t] = a.x;
    ux_out[t] = a.w;
    cell_out[t] = v;
  }
}

// Regrowth: the occupied rows of every column, packed in row order, into
// columns of D_out rows.
template <class R>
__global__ void __launch_bounds__(kBlock)
k_relayout(const R* __restrict__ in, const int* __restrict__ cnt_in,
           const unsigned long long* __restrict__ mask_in, int D_in, R* __restrict__ out,
           int* __restrict__ cnt_out, int D_out, int nc, int nc_pad, int* __restrict__ err) {
  typedef typename Vec4T<R>::type V4;
  const int ci = blockIdx.x * blockDim.x + threadIdx.x;
  if (ci >= nc_pad) return;
  const int len = ci < nc ? min(cnt_in[ci], D_in) : 0;
  int k = 0;
  unsigned long long m = 0;
  for (int j = 0; j < len; ++j) {
    if ((j & 63) == 0) m = mask_in[(size_t)(j >> 6) * nc_pad + ci];
    if (!((m >> (j & 63)) & 1ull)) continue;
    if (k >= D_out) {
      atomicOr(err, kErrColumnFull);
      break;
    }
    const V4* s = reinterpret_cast<const V4*>(in + col_slot(ci, j, D_in) * kSlotLen);
    V4* d = reinterpret_cast<V4*>(out + col_slot(ci, k, D_out) * kSlotLen);
    d[0] = s[0];
    d[1] = s[1];
    ++k;
  }
  cnt_out[ci] = k;
}

}  // namespace cpic
