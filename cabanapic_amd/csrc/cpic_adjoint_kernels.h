// cpic_adjoint_kernels.h -- reverse-mode (adjoint) derivative of the step.
//
// The reference has no derivative of its step.  The adjoint runs the step
// backwards: given the derivative of an objective with respect to the state
// after a step (every particle's x, y, z, ux, uy, uz and the field arrays),
// it returns the derivative with respect to the state before the step.  It
// is the transpose of the linearization the tangent-linear step computes
// (cpic_tangent_kernels.h): the same primal values (recomputed operation for
// operation, so bit-identical to the forward push), the same derivative of
// every operation (the narrowed sqrtf differentiated as the exact square
// root), crossing decisions held as taken and a position clamped to a face
// given zero derivative.  Adjoints are FP64 whatever real_t is.
//
// The particle adjoint runs one thread per particle over a snapshot of the
// step's input (k_snapshot_columns: whole slots, cells in order, so a
// cell's particles sit in consecutive lanes).  Each thread recomputes its
// particle's push and streaks and takes the transpose in reverse order.  The
// deposit's transpose is a gather (each streak reads the accumulator adjoint
// of its cell); the transpose of the field gather (the interpolator adjoint)
// is summed over each run of lanes in one cell and added once per run.
// Particle adjoints are records indexed by the particle's tag (its upload
// index).
//
// The field stages are linear; their transposes are written as gathers: each
// output voxel sums what the forward stage's outputs read from it.
#pragma once

#include "cpic_push_kernels.h"

namespace cpic {

// ---------------------------------------------------------------------------
// Snapshots of the particle columns: cells in interior order, a cell's
// particles in row order, each as its whole column slot (x, y, z, ux, uy, uz,
// w, tag), with the interior cell of every slot; first[ci] = position of
// cell ci's first particle, first[nc] = n.

// Exclusive prefix sum of cnt[0..n) into out[0..n], out[n] = total, in three
// passes (block scans, a scan of the block totals, the block offsets added).
constexpr int kScanItems = 8;  // elements per thread of a block scan

__device__ __forceinline__ int block_exclusive_scan(int v, int* total) {
  __shared__ int wsum[kBlock / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  if (lane == 63) wsum[wid] = x;
  __syncthreads();
  int before = 0, all = 0;
#pragma unroll
  for (int k = 0; k < kBlock / 64; ++k) {
    if (k < wid) before += wsum[k];
    all += wsum[k];
  }
  __syncthreads();
  *total = all;
  return before + x - v;
}

__global__ void __launch_bounds__(kBlock)
k_scan_blocks(const int* __restrict__ in, int n, int* __restrict__ out, int* __restrict__ sums) {
  const int base = blockIdx.x * kBlock * kScanItems + threadIdx.x * kScanItems;
  int v[kScanItems], s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    v[k] = base + k < n ? in[base + k] : 0;
    s += v[k];
  }
  int total;
  int run = block_exclusive_scan(s, &total);
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    if (base + k < n) out[base + k] = run;
    run += v[k];
  }
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// one block: exclusive scan of the nb block sums in place, grand total to *total_out
__global__ void __launch_bounds__(kBlock) k_scan_sums(int* __restrict__ sums, int nb,
                                                      int* __restrict__ total_out) {
  int carry = 0;
  for (int b0 = 0; b0 < nb; b0 += kBlock) {
    const int i = b0 + threadIdx.x;
    const int v = i < nb ? sums[i] : 0;
    int total;
    const int ex = block_exclusive_scan(v, &total);
    if (i < nb) sums[i] = carry + ex;
    carry += total;
  }
  if (threadIdx.x == 0) *total_out = carry;
}

__global__ void __launch_bounds__(kBlock)
k_scan_add(int* __restrict__ out, int n, const int* __restrict__ sums) {
  const int base = blockIdx.x * kBlock * kScanItems + threadIdx.x * kScanItems;
  const int add = sums[blockIdx.x];
#pragma unroll
  for (int k = 0; k < kScanItems; ++k)
    if (base + k < n) out[base + k] += add;
}

// The context's columns into a snapshot: lane per cell, each occupied row
// (in row order) copied as its whole slot to slots[first[ci] + k]; cell[]
// receives the interior cell index of every snapshot slot.  Rows are taken
// kSnapRows at a time: the batch's loads are all issued before its stores,
// and the batch's slots land next to each other, so each lane completes
// its output lines in a few consecutive stores.
constexpr int kSnapRows = 4;
static_assert(64 % kSnapRows == 0, "a mask word holds whole batches of rows");

template <class R>
__global__ void __launch_bounds__(kBlock)
k_snapshot_columns(const R* __restrict__ cols, const int* __restrict__ cnt,
                   const unsigned long long* __restrict__ mask, int nc, int nc_pad, int D,
                   const int* __restrict__ first, R* __restrict__ slots, int* __restrict__ cell) {
  typedef typename Vec4T<R>::type V4;
  const int ci = blockIdx.x * blockDim.x + threadIdx.x;
  if (ci >= nc) return;
  const int len = min(cnt[ci], D);
  long o = first[ci];
  const long end = first[ci + 1];
  unsigned long long m = 0;
  for (int j0 = 0; j0 < len && o < end; j0 += kSnapRows) {
    // j0 is a multiple of kSnapRows, so a new mask word starts only at u = 0
    if ((j0 & 63) == 0) m = mask[(size_t)(j0 >> 6) * nc_pad + ci];
    V4 a[kSnapRows], b[kSnapRows];
    bool occ[kSnapRows];
#pragma unroll
    for (int u = 0; u < kSnapRows; ++u) {
      const int j = j0 + u;
      occ[u] = j < len && ((m >> (j & 63)) & 1ull);
      if (occ[u]) {
        const V4* s = reinterpret_cast<const V4*>(cols + col_slot(ci, j, D) * kSlotLen);
        a[u] = s[0];
        b[u] = s[1];
      }
    }
#pragma unroll
    for (int u = 0; u < kSnapRows; ++u) {
      if (occ[u] && o < end) {
        V4* d = reinterpret_cast<V4*>(slots + (size_t)o * kSlotLen);
        d[0] = a[u];
        d[1] = b[u];
        cell[o] = ci;
        ++o;
      }
    }
  }
}

// A snapshot back into packed columns: thread t < n moves snapshot slot t
// to row (t - first[cell]) of its cell; thread t < nc_pad sets the count of
// cell t.
template <class R>
__global__ void __launch_bounds__(kBlock)
k_restore_columns(const R* __restrict__ slots, const int* __restrict__ cell,
                  const int* __restrict__ first, long n, int nc, int nc_pad, int D,
                  R* __restrict__ cols, int* __restrict__ cnt, int* __restrict__ err) {
  typedef typename Vec4T<R>::type V4;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < nc_pad) cnt[t] = t < nc ? min(first[t + 1] - first[t], D) : 0;
  if (t >= n) return;
  const int ci = cell[t];
  const long rank = t - first[ci];
  if (rank >= D) {
    atomicOr(err, kErrColumnFull);
    return;
  }
  const V4* s = reinterpret_cast<const V4*>(slots + (size_t)t * kSlotLen);
  V4* d = reinterpret_cast<V4*>(cols + col_slot(ci, (int)rank, D) * kSlotLen);
  d[0] = s[0];
  d[1] = s[1];
}

// Particle adjoints are records of kAdjRec doubles per tag: x, y, z, ux,
// uy, uz and two zeros, so that a record is two whole 32-byte sectors.
constexpr int kAdjRec = 8;

// [6][n] (host layout, upload order) <-> [n][kAdjRec] records
__global__ void __launch_bounds__(kBlock)
k_adj_soa_to_rec(const double* __restrict__ soa, long n, double* __restrict__ rec) {
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < n;
       p += (long)gridDim.x * blockDim.x) {
    double* r = rec + (size_t)p * kAdjRec;
#pragma unroll
    for (int k = 0; k < 6; ++k) r[k] = soa[(size_t)k * n + p];
    r[6] = 0;
    r[7] = 0;
  }
}

__global__ void __launch_bounds__(kBlock)
k_adj_rec_to_soa(const double* __restrict__ rec, long n, double* __restrict__ soa) {
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < n;
       p += (long)gridDim.x * blockDim.x) {
    const double* r = rec + (size_t)p * kAdjRec;
#pragma unroll
    for (int k = 0; k < 6; ++k) soa[(size_t)k * n + p] = r[k];
  }
}

// ---------------------------------------------------------------------------
// Transposes of the push's pieces.  Each recomputes its forward values from
// its inputs and returns the adjoints of its inputs; adjoints already held
// for an input are added to.

// quadrants(q, dA, mB, mC, v5) of cpic_push_kernels.h, transposed: l[4] is
// the adjoint of the four outputs; adds to ldA, lmB, lmC, lv5.
template <class R>
__device__ __forceinline__ void quadrants_bwd(R q, R dA, R mB, R mC, const double l[4],
                                              double& ldA, double& lmB, double& lmC,
                                              double& lv5) {
  const R t = q * dA;
  const R s1a = t * mB;
  const R s0a = t - s1a;
  const R s1b = s1a + t;
  const R cp = R(1) + mC;
  const R cm = R(1) - mC;
  lv5 += l[0] - l[1] - l[2] + l[3];
  // s0b = s0a cm, s1c = s1b cm, s2 = s0a cp, s3 = s1b cp
  double ls0a = l[0] * (double)cm + l[2] * (double)cp;
  double ls1b = l[1] * (double)cm + l[3] * (double)cp;
  lmC += -(l[0] * (double)s0a + l[1] * (double)s1b) + (l[2] * (double)s0a + l[3] * (double)s1b);
  // s1b = s1a + t, s0a = t - s1a
  double ls1a = ls1b - ls0a;
  double lt = ls1b + ls0a;
  // s1a = t mB
  lt += ls1a * (double)mB;
  lmB += ls1a * (double)t;
  // t = q dA (q carries no derivative)
  ldA += lt * (double)q;
}

// streak_values transposed: l[12] adjoint of the 12 contributions.
template <class R>
__device__ __forceinline__ void streak_values_bwd(R q, R dx, R dy, R dz, R mx, R my, R mz,
                                                  const double l[12], double ld[3], double lm[3],
                                                  double& lv5) {
  quadrants_bwd(q, dx, my, mz, l + 0, ld[0], lm[1], lm[2], lv5);
  quadrants_bwd(q, dy, mz, mx, l + 4, ld[1], lm[2], lm[0], lv5);
  quadrants_bwd(q, dz, mx, my, l + 8, ld[2], lm[0], lm[1], lv5);
}

// the streak correction term q a b c / 3 (move_p.h:154), transposed as the
// tangent differentiates it (in FP64 from the primal operands)
template <class R>
__device__ __forceinline__ void streak_v5_bwd(R q, R a, R b, R c, double lv5, double& la,
                                              double& lb, double& lc) {
  const double Q = q, A = a, B = b, C = c;
  const double s = lv5 * (1. / 3.);
  la += s * Q * B * C;
  lb += s * Q * A * C;
  lc += s * Q * A * B;
}

// One streak (streak() of cpic_push_kernels.h) transposed.  pos, rem: the
// streak's inputs; lpos, lrem: on entry the adjoints of its outputs (pos and
// rem after the streak), on exit those of its inputs.  lacc: adjoint of the
// accumulators of the streak's cell.  Returns the face crossed (as streak).
template <class R>
__device__ __forceinline__ int streak_bwd(const R pos[3], const R rem[3], R q,
                                          const double lacc[12], double lpos[3],
                                          double lrem[3]) {
  R dir[3], len3[3], dsp0[3], dsp[3], mid[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    dsp0[k] = rem[k];
    mid[k] = pos[k];
    dir[k] = (dsp0[k] > 0) ? R(1) : R(-1);
    len3[k] = (dsp0[k] == 0) ? R(3.4e38) : (dir[k] - mid[k]) / dsp0[k];
  }
  R len = 2;
  int axis = 3;
  if (len3[0] < len) len = len3[0], axis = 0;
  if (len3[1] < len) len = len3[1], axis = 1;
  if (len3[2] < len) len = len3[2], axis = 2;
  const R lenh = len * R(0.5);
  R mid1[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    dsp[k] = dsp0[k] * lenh;
    mid1[k] = mid[k] + dsp[k];
  }
  const R v5 = streak_v5(q, dsp[0], dsp[1], dsp[2]);
  (void)v5;
  // the snap of the crossed coordinate: its output carries no derivative
  double lp[3] = {lpos[0], lpos[1], lpos[2]};
  if (axis < 3) lp[axis] = 0;
  double ldsp[3], lmid1[3] = {0, 0, 0}, lv5 = 0;
#pragma unroll
  for (int k = 0; k < 3; ++k) ldsp[k] = -lrem[k] + 2.0 * lp[k];  // rem -= dsp; pos += 2 dsp
  streak_values_bwd(q, dsp[0], dsp[1], dsp[2], mid1[0], mid1[1], mid1[2], lacc, ldsp, lmid1, lv5);
  streak_v5_bwd(q, dsp[0], dsp[1], dsp[2], lv5, ldsp[0], ldsp[1], ldsp[2]);
  // mid1 = mid + dsp
  double lmid[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    lmid[k] = lmid1[k];
    ldsp[k] += lmid1[k];
  }
  // dsp = dsp0 lenh
  double llenh = 0, ldsp0[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    ldsp0[k] = ldsp[k] * (double)lenh;
    llenh += ldsp[k] * (double)dsp0[k];
  }
  if (axis < 3) {
    // lenh = 0.5 len, len = (dir - mid) / dsp0 on the crossed axis
    const double llen = 0.5 * llenh;
    const double b = dsp0[axis];
    lmid[axis] -= llen / b;
    ldsp0[axis] -= llen * (double)len / b;
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    lpos[k] = lp[k] + lmid[k];
    lrem[k] += ldsp0[k];
  }
  if (axis == 3) return -1;
  return axis + ((dir[axis] > 0) ? 3 : 0);
}

// displacement() transposed: ld adjoint of (ddx, ddy, ddz); adds to lu.
template <class R>
__device__ __forceinline__ void displacement_bwd(R ux, R uy, R uz, R cdt_dx, R cdt_dy, R cdt_dz,
                                                 const double ld[3], double lu[3]) {
  const R one = 1.;
  const R s = one + (ux * ux + (uy * uy + uz * uz));
  const R r = sqrtf_narrow(s);
  const R v0 = one / r;
  const R a[3] = {ux * cdt_dx, uy * cdt_dy, uz * cdt_dz};
  const R c[3] = {cdt_dx, cdt_dy, cdt_dz};
  const R u[3] = {ux, uy, uz};
  double lv0 = 0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    lv0 += ld[k] * (double)a[k];
    lu[k] += ld[k] * (double)v0 * (double)c[k];
  }
  const double lr = -lv0 * (double)v0 / (double)r;
  const double ls = lr * 0.5 / (double)r;
#pragma unroll
  for (int k = 0; k < 3; ++k) lu[k] += ls * 2.0 * (double)u[k];
}

// boris() transposed.  Inputs: the interpolator (e0..b1) and the particle
// (d, u before the push); lu: on entry the adjoint of u after the push, on
// exit that of u before it.  Adds to ld (offsets) and lf[18] (interpolator
// coefficients in the order of src/types.h:64-84).
template <class V4, class R>
__device__ __forceinline__ void boris_bwd(const V4& e0, const V4& e1, const V4& e2, const V4& b0,
                                          const V4& b1, R qdt_2mc, R dx, R dy, R dz, R ux0,
                                          R uy0, R uz0, double lu[3], double ld[3],
                                          double lf[18]) {
  const R one = 1.;
  const R one_third = 1. / 3.;
  const R two_fifteenths = 2. / 15.;
  // forward, keeping the intermediates
  const R zx = e0.z + dy * e0.w, zy = e1.z + dz * e1.w, zz = e2.z + dx * e2.w;
  const R hax = qdt_2mc * ((e0.x + dy * e0.y) + dz * zx);
  const R hay = qdt_2mc * ((e1.x + dz * e1.y) + dx * zy);
  const R haz = qdt_2mc * ((e2.x + dx * e2.y) + dy * zz);
  const R cbx = b0.x + dx * b0.y;
  const R cby = b0.z + dy * b0.w;
  const R cbz = b1.x + dz * b1.y;
  const R ux = ux0 + hax, uy = uy0 + hay, uz = uz0 + haz;
  const R s = one + (ux * ux + (uy * uy + uz * uz));
  const R r = sqrtf_narrow(s);
  const R v0 = qdt_2mc / r;
  const R v1 = cbx * cbx + (cby * cby + cbz * cbz);
  const R v0sq = v0 * v0;
  const R v2 = v0sq * v1;
  const R inner = one_third + v2 * two_fifteenths;
  const R poly = one + v2 * inner;
  const R v3 = v0 * poly;
  const R v3sq = v3 * v3;
  const R den = one + v1 * v3sq;
  const R v4a = v3 / den;
  const R v4 = v4a + v4a;
  const R X0 = uy * cbz - uz * cby, X1 = uz * cbx - ux * cbz, X2 = ux * cby - uy * cbx;
  const R w0 = ux + v3 * X0, w1 = uy + v3 * X1, w2 = uz + v3 * X2;
  const R Y0 = w1 * cbz - w2 * cby, Y1 = w2 * cbx - w0 * cbz, Y2 = w0 * cby - w1 * cbx;
  // backward: u3 = u2 + ha, u2 = u1 + v4 Y
  const double lha[3] = {lu[0], lu[1], lu[2]};
  double lu1[3] = {lu[0], lu[1], lu[2]};
  double lv4 = lu[0] * (double)Y0 + lu[1] * (double)Y1 + lu[2] * (double)Y2;
  const double lY0 = lu[0] * (double)v4, lY1 = lu[1] * (double)v4, lY2 = lu[2] * (double)v4;
  double lw0 = 0, lw1 = 0, lw2 = 0, lcbx = 0, lcby = 0, lcbz = 0;
  // Y0 = w1 cbz - w2 cby
  lw1 += lY0 * (double)cbz; lcbz += lY0 * (double)w1; lw2 -= lY0 * (double)cby; lcby -= lY0 * (double)w2;
  // Y1 = w2 cbx - w0 cbz
  lw2 += lY1 * (double)cbx; lcbx += lY1 * (double)w2; lw0 -= lY1 * (double)cbz; lcbz -= lY1 * (double)w0;
  // Y2 = w0 cby - w1 cbx
  lw0 += lY2 * (double)cby; lcby += lY2 * (double)w0; lw1 -= lY2 * (double)cbx; lcbx -= lY2 * (double)w1;
  // w = u1 + v3 X
  lu1[0] += lw0; lu1[1] += lw1; lu1[2] += lw2;
  double lv3 = lw0 * (double)X0 + lw1 * (double)X1 + lw2 * (double)X2;
  const double lX0 = lw0 * (double)v3, lX1 = lw1 * (double)v3, lX2 = lw2 * (double)v3;
  // X0 = uy cbz - uz cby
  lu1[1] += lX0 * (double)cbz; lcbz += lX0 * (double)uy; lu1[2] -= lX0 * (double)cby; lcby -= lX0 * (double)uz;
  // X1 = uz cbx - ux cbz
  lu1[2] += lX1 * (double)cbx; lcbx += lX1 * (double)uz; lu1[0] -= lX1 * (double)cbz; lcbz -= lX1 * (double)ux;
  // X2 = ux cby - uy cbx
  lu1[0] += lX2 * (double)cby; lcby += lX2 * (double)ux; lu1[1] -= lX2 * (double)cbx; lcbx -= lX2 * (double)uy;
  // v4 = 2 v4a, v4a = v3 / den, den = 1 + v1 v3sq
  const double lv4a = 2.0 * lv4;
  lv3 += lv4a / (double)den;
  const double lden = -lv4a * (double)v4a / (double)den;
  double lv1 = lden * (double)v3sq;
  lv3 += lden * (double)v1 * 2.0 * (double)v3;
  // v3 = v0 poly, poly = 1 + v2 inner, inner = 1/3 + v2 2/15
  double lv0 = lv3 * (double)poly;
  const double lpoly = lv3 * (double)v0;
  double lv2 = lpoly * (double)inner;
  lv2 += lpoly * (double)v2 * (double)two_fifteenths;
  // v2 = v0sq v1
  lv1 += lv2 * (double)v0sq;
  lv0 += lv2 * (double)v1 * 2.0 * (double)v0;
  // v1 = |cb|^2
  lcbx += lv1 * 2.0 * (double)cbx;
  lcby += lv1 * 2.0 * (double)cby;
  lcbz += lv1 * 2.0 * (double)cbz;
  // v0 = qdt_2mc / r, r = sqrt(s), s = 1 + |u1|^2
  const double lr = -lv0 * (double)v0 / (double)r;
  const double ls = lr * 0.5 / (double)r;
  lu1[0] += ls * 2.0 * (double)ux;
  lu1[1] += ls * 2.0 * (double)uy;
  lu1[2] += ls * 2.0 * (double)uz;
  // u1 = u0 + ha
  lu[0] = lu1[0];
  lu[1] = lu1[1];
  lu[2] = lu1[2];
  const double lhx = lha[0] + lu1[0], lhy = lha[1] + lu1[1], lhz = lha[2] + lu1[2];
  // cbx = b0.x + dx b0.y, cby = b0.z + dy b0.w, cbz = b1.x + dz b1.y
  lf[12] += lcbx; ld[0] += lcbx * (double)b0.y; lf[13] += lcbx * (double)dx;
  lf[14] += lcby; ld[1] += lcby * (double)b0.w; lf[15] += lcby * (double)dy;
  lf[16] += lcbz; ld[2] += lcbz * (double)b1.y; lf[17] += lcbz * (double)dz;
  // hax = q ((e0.x + dy e0.y) + dz (e0.z + dy e0.w))
  {
    const double li = lhx * (double)qdt_2mc;
    lf[0] += li; ld[1] += li * (double)e0.y; lf[1] += li * (double)dy;
    ld[2] += li * (double)zx;
    const double lz = li * (double)dz;
    lf[2] += lz; ld[1] += lz * (double)e0.w; lf[3] += lz * (double)dy;
  }
  // hay = q ((e1.x + dz e1.y) + dx (e1.z + dz e1.w))
  {
    const double li = lhy * (double)qdt_2mc;
    lf[4] += li; ld[2] += li * (double)e1.y; lf[5] += li * (double)dz;
    ld[0] += li * (double)zy;
    const double lz = li * (double)dx;
    lf[6] += lz; ld[2] += lz * (double)e1.w; lf[7] += lz * (double)dz;
  }
  // haz = q ((e2.x + dx e2.y) + dy (e2.z + dx e2.w))
  {
    const double li = lhz * (double)qdt_2mc;
    lf[8] += li; ld[0] += li * (double)e2.y; lf[9] += li * (double)dx;
    ld[1] += li * (double)zz;
    const double lz = li * (double)dy;
    lf[10] += lz; ld[0] += lz * (double)e2.w; lf[11] += lz * (double)dx;
  }
}

template <class R> struct AdjointPush {
  const R* slots;      // [n][kSlotLen] the step's input particles (snapshot)
  const int* cell;     // [n] interior cell of each snapshot slot
  long n;
  const R* interp;     // the step's interpolators (cell-major, kInterpStride)
  const double* lacc;  // adjoint of the accumulators, slot-major [12][Ng]
  const double* lout;  // [n][kAdjRec]: adjoint of the step's output particles, by tag
  double* lin;         // [n][kAdjRec]: adjoint of its input particles, by tag
  double* linterp;     // [Ng][kInterpStride]: adjoint of the interpolators (added to)
  CellGrid g;
  int ncells;  // Ng
  R qdt_2mc, cdt_dx, cdt_dy, cdt_dz, qsp;
  int* err;
};

template <class R>
__device__ __forceinline__ void load_lacc(const AdjointPush<R>& a, int v, double l[12]) {
  const size_t N = a.ncells;
#pragma unroll
  for (int k = 0; k < 12; ++k) l[k] = a.lacc[k * N + v];
}

// Sums l[18] over each run of consecutive lanes of the wave with the same
// key (a segmented inclusive scan; the last lane of a run then adds the
// run's total to linterp of voxel key).  Every lane of the wave calls it;
// key < 0 for none.
__device__ __forceinline__ void add_interp_runs(double* linterp, int key, double l[18]) {
  const int lane = threadIdx.x & 63;
  const int before = __shfl_up(key, 1, 64);
  int head = (lane == 0 || before != key) ? 1 : 0;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int h = __shfl_up(head, off, 64);
    const bool add = lane >= off && !head;
#pragma unroll
    for (int k = 0; k < 18; ++k) {
      const double y = __shfl_up(l[k], off, 64);
      if (add) l[k] += y;
    }
    if (lane >= off) head |= h;
  }
  const int after = __shfl_down(key, 1, 64);
  if (key >= 0 && (lane == 63 || after != key)) {
    double* li = linterp + (size_t)key * kInterpStride;
#pragma unroll
    for (int k = 0; k < 18; ++k)
      if (l[k] != 0.0) unsafeAtomicAdd(li + k, l[k]);
  }
}

// One thread per particle of the snapshot (cell-sorted, so the particles of
// a cell sit in consecutive lanes): the particle's push and streaks
// recomputed and transposed (src/push.h:124-256, src/move_p.h:93-209 in
// reverse).  Writes the input-particle adjoint by tag; the interpolator
// adjoints of the cell's particles are summed over the wave's runs of lanes
// in one cell and added atomically.
template <class R>
__global__ void __launch_bounds__(kBlock) k_push_adjoint(AdjointPush<R> a) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double lf[18];
#pragma unroll
  for (int k = 0; k < 18; ++k) lf[k] = 0;
  int key = -1;
  if (p < a.n) {
    typedef typename Vec4T<R>::type V4;
    const int ci = a.cell[p];
    const int v = interior_to_voxel(a.g, ci);
    const V4* f = reinterpret_cast<const V4*>(a.interp + (size_t)v * kInterpStride);
    const V4 e0 = f[0], e1 = f[1], e2 = f[2], b0 = f[3], b1 = f[4];
    const V4* sl = reinterpret_cast<const V4*>(a.slots + (size_t)p * kSlotLen);
    const V4 pa = sl[0], pb = sl[1];
    const R one = 1.;
    const R one_third = 1. / 3.;
    const R dx = pa.x, dy = pa.y, dz = pa.z;
    const R ux0 = pa.w, uy0 = pb.x, uz0 = pb.y, w = pb.z;
    const long t = TagBits<R>::get(pb.w);
    if (t < 0 || t >= a.n) {
      atomicOr(a.err, kErrBadCell);
    } else {
      key = v;
      // the forward push of this particle
      R ux = ux0, uy = uy0, uz = uz0;
      boris(e0, e1, e2, b0, b1, a.qdt_2mc, dx, dy, dz, ux, uy, uz);
      R dd[3];
      displacement(ux, uy, uz, a.cdt_dx, a.cdt_dy, a.cdt_dz, dd[0], dd[1], dd[2]);
      const R mx = dx + dd[0], my = dy + dd[1], mz = dz + dd[2];
      const R nx = mx + dd[0], ny = my + dd[1], nz = mz + dd[2];
      const R q = w * a.qsp;
      // adjoints of the outputs
      const double* lo = a.lout + (size_t)t * kAdjRec;
      double lpos[3] = {lo[0], lo[1], lo[2]};
      double lu[3] = {lo[3], lo[4], lo[5]};
      double ld[3] = {0, 0, 0}, ldd[3] = {0, 0, 0};
      double lacc[12];
      if (nx <= one && ny <= one && nz <= one && -nx <= one && -ny <= one && -nz <= one) {
        // in-cell: pos' = (d + dd) + dd, deposit with v5 = q ddx ddy ddz / 3
        load_lacc(a, v, lacc);
        const R t1 = q * dd[0], t2 = t1 * dd[1];
        double lm[3] = {0, 0, 0}, lv5 = 0;
        streak_values_bwd(q, dd[0], dd[1], dd[2], mx, my, mz, lacc, ldd, lm, lv5);
        const double lt3 = lv5 * (double)one_third;
        const double lt2 = lt3 * (double)dd[2];
        ldd[2] += lt3 * (double)t2;
        const double lt1 = lt2 * (double)dd[1];
        ldd[1] += lt2 * (double)t1;
        ldd[0] += lt1 * (double)q;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          // m = d + dd, n = m + dd
          const double lmk = lm[k] + lpos[k];
          ldd[k] += lpos[k] + lmk;
          ld[k] += lmk;
        }
      } else {
        // the streaks: recount them forward, then transpose the last first
        int K = 0;
        {
          R pos[3] = {dx, dy, dz}, rem[3] = {dd[0], dd[1], dd[2]};
          R dsp[3], mid[3], v5;
          for (; K < kMaxStreaks; ++K) {
            if (streak_geometry(pos, rem, q, dsp, mid, v5) < 0) break;
          }
          ++K;
          if (K > kMaxStreaks) {
            atomicOr(a.err, kErrStreakLimit);
            K = kMaxStreaks;
          }
        }
        double lrem[3] = {0, 0, 0};  // the remainder after the last streak is no output
        for (int s = K - 1; s >= 0; --s) {
          // inputs and cell of streak s
          R pos[3] = {dx, dy, dz}, rem[3] = {dd[0], dd[1], dd[2]};
          int cs = ci;
          R dsp[3], mid[3], v5;
          for (int k = 0; k < s; ++k)
            cs = face_neighbor(a.g, cs, streak_geometry(pos, rem, q, dsp, mid, v5));
          load_lacc(a, interior_to_voxel(a.g, cs), lacc);
          streak_bwd(pos, rem, q, lacc, lpos, lrem);
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          ld[k] = lpos[k];
          ldd[k] = lrem[k];
        }
      }
      displacement_bwd(ux, uy, uz, a.cdt_dx, a.cdt_dy, a.cdt_dz, ldd, lu);
      boris_bwd(e0, e1, e2, b0, b1, a.qdt_2mc, dx, dy, dz, ux0, uy0, uz0, lu, ld, lf);
      double* li = a.lin + (size_t)t * kAdjRec;
      typedef double D4 __attribute__((ext_vector_type(4)));
      D4 r0, r1;
      r0.x = ld[0]; r0.y = ld[1]; r0.z = ld[2]; r0.w = lu[0];
      r1.x = lu[1]; r1.y = lu[2]; r1.z = 0;     r1.w = 0;
      reinterpret_cast<D4*>(li)[0] = r0;
      reinterpret_cast<D4*>(li)[1] = r1;
    }
  }
  add_interp_runs(a.linterp, key, lf);
}

// ---------------------------------------------------------------------------
// Field stages transposed (FP64, arrays of the field layout 9 x Ng).

__device__ __forceinline__ bool in_box(int x, int y, int z, int lo, int hx, int hy, int hz) {
  return x >= lo && x <= hx && y >= lo && y <= hy && z >= lo && z <= hz;
}

// load_interpolator_array transposed: adds to the E and B adjoints what
// the interior cells' interpolator adjoints (cell-major, kInterpStride) read
// from every voxel.  One thread per voxel of [1, n+1]^3.
__global__ void __launch_bounds__(kBlock)
k_load_interpolator_bwd(double* __restrict__ lf, const double* __restrict__ li, Grid g) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int mx = g.nx + 1, my = g.ny + 1, mz = g.nz + 1;
  if (t >= mx * my * mz) return;
  const int x = 1 + t % mx, y = 1 + (t / mx) % my, z = 1 + t / (mx * my);
  const int j = voxel(g, x, y, z);
  const size_t N = g.ncells;
  auto cell_ok = [&](int dx_, int dy_, int dz_) {
    return in_box(x - dx_, y - dy_, z - dz_, 1, g.nx, g.ny, g.nz);
  };
  auto L = [&](int i, int k) { return li[(size_t)i * kInterpStride + k]; };
  // the bilinear E coefficients: corners a (i), b, c, d of cell i
  //   o0 = (d + a) + (b + c), o1 = (d - a) + (b - c), o2 = (d - a) - (b - c),
  //   o3 = (d + a) - (b + c), all times 1/4
  auto corner = [&](int i, int k0, int role) {
    const double l0 = L(i, k0), l1 = L(i, k0 + 1), l2 = L(i, k0 + 2), l3 = L(i, k0 + 3);
    switch (role) {
      case 0: return 0.25 * (l0 - l1 - l2 + l3);   // a
      case 1: return 0.25 * (l0 + l1 - l2 - l3);   // b
      case 2: return 0.25 * (l0 - l1 + l2 - l3);   // c
      default: return 0.25 * (l0 + l1 + l2 + l3);  // d
    }
  };
  double s = 0;
  // ex: a = ex[i], b = ex[i + y], c = ex[i + z], d = ex[i + y + z]
  if (cell_ok(0, 0, 0)) s += corner(j, 0, 0);
  if (cell_ok(0, 1, 0)) s += corner(j - g.sy, 0, 1);
  if (cell_ok(0, 0, 1)) s += corner(j - g.sz, 0, 2);
  if (cell_ok(0, 1, 1)) s += corner(j - g.sy - g.sz, 0, 3);
  lf[j] += s;
  // ey: a = ey[i], b = ey[i + z], c = ey[i + x], d = ey[i + x + z]
  s = 0;
  if (cell_ok(0, 0, 0)) s += corner(j, 4, 0);
  if (cell_ok(0, 0, 1)) s += corner(j - g.sz, 4, 1);
  if (cell_ok(1, 0, 0)) s += corner(j - 1, 4, 2);
  if (cell_ok(1, 0, 1)) s += corner(j - 1 - g.sz, 4, 3);
  lf[N + j] += s;
  // ez: a = ez[i], b = ez[i + x], c = ez[i + y], d = ez[i + x + y]
  s = 0;
  if (cell_ok(0, 0, 0)) s += corner(j, 8, 0);
  if (cell_ok(1, 0, 0)) s += corner(j - 1, 8, 1);
  if (cell_ok(0, 1, 0)) s += corner(j - g.sy, 8, 2);
  if (cell_ok(1, 1, 0)) s += corner(j - 1 - g.sy, 8, 3);
  lf[2 * N + j] += s;
  // cbx: o = (b + a) / 2, (b - a) / 2 with a = cbx[i], b = cbx[i + x]
  s = 0;
  if (cell_ok(0, 0, 0)) s += 0.5 * (L(j, 12) - L(j, 13));
  if (cell_ok(1, 0, 0)) s += 0.5 * (L(j - 1, 12) + L(j - 1, 13));
  lf[3 * N + j] += s;
  s = 0;
  if (cell_ok(0, 0, 0)) s += 0.5 * (L(j, 14) - L(j, 15));
  if (cell_ok(0, 1, 0)) s += 0.5 * (L(j - g.sy, 14) + L(j - g.sy, 15));
  lf[4 * N + j] += s;
  s = 0;
  if (cell_ok(0, 0, 0)) s += 0.5 * (L(j, 16) - L(j, 17));
  if (cell_ok(0, 0, 1)) s += 0.5 * (L(j - g.sz, 16) + L(j - g.sz, 17));
  lf[5 * N + j] += s;
}

// unload_accumulator_array transposed: the accumulator adjoint of every
// interior cell from the J adjoint over [1, n+1]^3, which the forward
// stage overwrites (so its adjoint is cleared afterwards, k_clear_j_bwd).
__global__ void __launch_bounds__(kBlock)
k_unload_bwd(const double* __restrict__ lf, double* __restrict__ lacc, Grid g, double cx,
             double cy, double cz) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.nx * g.ny * g.nz) return;
  const int x = 1 + t % g.nx, y = 1 + (t / g.nx) % g.ny, z = 1 + t / (g.nx * g.ny);
  const int v = voxel(g, x, y, z);
  const size_t N = g.ncells;
  const double* jx = lf + 6 * N;
  const double* jy = lf + 7 * N;
  const double* jz = lf + 8 * N;
  const int sy = g.sy, sz = g.sz;
  double* A = lacc;
  A[0 * N + v] = cx * jx[v];
  A[1 * N + v] = cx * jx[v + sy];
  A[2 * N + v] = cx * jx[v + sz];
  A[3 * N + v] = cx * jx[v + sy + sz];
  A[4 * N + v] = cy * jy[v];
  A[5 * N + v] = cy * jy[v + sz];
  A[6 * N + v] = cy * jy[v + 1];
  A[7 * N + v] = cy * jy[v + 1 + sz];
  A[8 * N + v] = cz * jz[v];
  A[9 * N + v] = cz * jz[v + 1];
  A[10 * N + v] = cz * jz[v + sy];
  A[11 * N + v] = cz * jz[v + 1 + sy];
}

// the J adjoint after the transposed unload: J is overwritten by every step,
// so nothing before the step depends on it
__global__ void __launch_bounds__(kBlock) k_clear_j_bwd(double* __restrict__ lf, Grid g) {
  const size_t N = g.ncells;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < 3 * N;
       i += (size_t)gridDim.x * blockDim.x)
    lf[6 * N + i] = 0;
}

// advance_b transposed (B_a -= p (E_b(i+s) - E_b(i)) on interior i): B's
// adjoint is unchanged; E's adjoint at voxel j of [1, n+1]^3 gains what the
// interior cells that read E(j) contribute.
__global__ void __launch_bounds__(kBlock)
k_advance_b_bwd(double* __restrict__ lf, Grid g, double px, double py, double pz) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int mx = g.nx + 1, my = g.ny + 1, mz = g.nz + 1;
  if (t >= mx * my * mz) return;
  const int x = 1 + t % mx, y = 1 + (t / mx) % my, z = 1 + t / (mx * my);
  const int j = voxel(g, x, y, z);
  const size_t N = g.ncells;
  const double* lbx = lf + 3 * N;
  const double* lby = lf + 4 * N;
  const double* lbz = lf + 5 * N;
  auto in = [&](int dx_, int dy_, int dz_) {
    return in_box(x - dx_, y - dy_, z - dz_, 1, g.nx, g.ny, g.nz);
  };
  const bool c0 = in(0, 0, 0);
  // bx[i] -= py (ez[i+y] - ez[i]) - pz (ey[i+z] - ey[i])
  // by[i] -= pz (ex[i+z] - ex[i]) - px (ez[i+x] - ez[i])
  // bz[i] -= px (ey[i+x] - ey[i]) - py (ex[i+y] - ex[i])
  double sx = 0, sy_ = 0, sz_ = 0;
  if (c0) {
    sx += pz * lby[j] - py * lbz[j];
    sy_ += -pz * lbx[j] + px * lbz[j];
    sz_ += py * lbx[j] - px * lby[j];
  }
  if (in(0, 0, 1)) sx -= pz * lby[j - g.sz];
  if (in(0, 1, 0)) sx += py * lbz[j - g.sy];
  if (in(0, 0, 1)) sy_ += pz * lbx[j - g.sz];
  if (in(1, 0, 0)) sy_ -= px * lbz[j - 1];
  if (in(0, 1, 0)) sz_ -= py * lbx[j - g.sy];
  if (in(1, 0, 0)) sz_ += px * lby[j - 1];
  lf[j] += sx;
  lf[N + j] += sy_;
  lf[2 * N + j] += sz_;
}

// advance_e transposed (E(i) += -cj J(i) + p (B(i) - B(i-s)) on i of
// [1, n+1]^3): E's adjoint is unchanged; J's gains -cj E's on that range;
// B's at voxel j of [0, n+1]^3 gains what reads B(j).
__global__ void __launch_bounds__(kBlock)
k_advance_e_bwd(double* __restrict__ lf, Grid g, double px, double py, double pz, double cj) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int X = g.nx + 2, Y = g.ny + 2, Z = g.nz + 2;
  if (t >= X * Y * Z) return;
  const int x = t % X, y = (t / X) % Y, z = t / (X * Y);
  const int j = voxel(g, x, y, z);
  const size_t N = g.ncells;
  const double* lex = lf;
  const double* ley = lf + N;
  const double* lez = lf + 2 * N;
  auto in = [&](int dx_, int dy_, int dz_) {
    return in_box(x + dx_, y + dy_, z + dz_, 1, g.nx + 1, g.ny + 1, g.nz + 1);
  };
  const bool c0 = in(0, 0, 0);
  // ex[i] += py (bz[i] - bz[i-y]) - pz (by[i] - by[i-z])
  // ey[i] += pz (bx[i] - bx[i-z]) - px (bz[i] - bz[i-x])
  // ez[i] += px (by[i] - by[i-x]) - py (bx[i] - bx[i-y])
  double sbx = 0, sby = 0, sbz = 0;
  if (c0) {
    sbz += py * lex[j] - px * ley[j];
    sby += -pz * lex[j] + px * lez[j];
    sbx += pz * ley[j] - py * lez[j];
    lf[6 * N + j] += -cj * lex[j];
    lf[7 * N + j] += -cj * ley[j];
    lf[8 * N + j] += -cj * lez[j];
  }
  if (in(0, 1, 0)) sbz -= py * lex[j + g.sy];
  if (in(0, 0, 1)) sby += pz * lex[j + g.sz];
  if (in(0, 0, 1)) sbx -= pz * ley[j + g.sz];
  if (in(1, 0, 0)) sbz += px * ley[j + 1];
  if (in(1, 0, 0)) sby -= px * lez[j + 1];
  if (in(0, 1, 0)) sbx += py * lez[j + g.sy];
  lf[3 * N + j] += sbx;
  lf[4 * N + j] += sby;
  lf[5 * N + j] += sbz;
}

// The periodic ghost copy transposed, for three components from comp0:
// every interior voxel gains the adjoints of the ghosts that are its images
// (each coordinate c also appears as ghost 0 when c = n and as ghost n+1
// when c = 1), then (k_zero_ghosts_bwd) the ghost adjoints are cleared.
__global__ void __launch_bounds__(kBlock)
k_copy_ghosts_bwd(double* __restrict__ lf, int comp0, Grid g) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.nx * g.ny * g.nz) return;
  const int x = 1 + t % g.nx, y = 1 + (t / g.nx) % g.ny, z = 1 + t / (g.nx * g.ny);
  int ax[3], ay[3], az[3], nax = 0, nay = 0, naz = 0;
  ax[nax++] = x;
  if (x == g.nx) ax[nax++] = 0;
  if (x == 1) ax[nax++] = g.nx + 1;
  ay[nay++] = y;
  if (y == g.ny) ay[nay++] = 0;
  if (y == 1) ay[nay++] = g.ny + 1;
  az[naz++] = z;
  if (z == g.nz) az[naz++] = 0;
  if (z == 1) az[naz++] = g.nz + 1;
  if (nax * nay * naz == 1) return;
  const int j = voxel(g, x, y, z);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    double* a = lf + (size_t)(comp0 + c) * g.ncells;
    double s = 0;
    for (int i = 0; i < nax; ++i)
      for (int k = 0; k < nay; ++k)
        for (int m = 0; m < naz; ++m)
          if (i || k || m) s += a[voxel(g, ax[i], ay[k], az[m])];
    a[j] += s;
  }
}

__global__ void __launch_bounds__(kBlock)
k_zero_ghosts_bwd(double* __restrict__ lf, int comp0, Grid g) {
  const int X = g.nx + 2, Y = g.ny + 2, Z = g.nz + 2;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= X * Y * Z) return;
  const int x = t % X, y = (t / X) % Y, z = t / (X * Y);
  if (x >= 1 && x <= g.nx && y >= 1 && y <= g.ny && z >= 1 && z <= g.nz) return;
#pragma unroll
  for (int c = 0; c < 3; ++c) lf[(size_t)(comp0 + c) * g.ncells + t] = 0;
}

// k_fold_j transposed: every interior face value the fold wrote is a sum of
// itself and high-ghost values; the ghosts gain the face's adjoint.  One
// thread per ghost-layer source, gathering from the faces that read it.
__global__ void __launch_bounds__(kBlock) k_fold_j_bwd(double* __restrict__ lf, Grid g) {
  const int n[3] = {g.nx, g.ny, g.nz};
  const int s[3] = {1, g.sy, g.sz};
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  for (int c = 0; c < 3; ++c) {
    const int ax_p = (c + 1) % 3, ax_q = (c + 2) % 3;
    const int nr = n[c], np_ = n[ax_p], nq = n[ax_q];
    // sources: (p, nq+1) for p in [1, np+1] and (np+1, q) for q in [1, nq]
    const int per_r = (np_ + 1) + nq;
    const int count = nr * per_r;
    if (t < count) {
      const int r = 1 + t / per_r;
      const int k = t % per_r;
      double* J = lf + (size_t)(6 + c) * g.ncells;
      const int base = r * s[c];
      const int sp = s[ax_p], sq = s[ax_q];
      int p, q;
      if (k < np_ + 1) {
        p = 1 + k;
        q = nq + 1;
      } else {
        p = np_ + 1;
        q = 1 + (k - (np_ + 1));
      }
      // faces: (1,1) reads (np+1,1), (1,nq+1), (np+1,nq+1); (1,q>1) reads
      // (np+1,q); (p>1,1) reads (p,nq+1)
      double add = 0;
      if (p == np_ + 1 && q == nq + 1) {
        add = J[base + sp + sq];  // the corner feeds only (1,1)
      } else if (q == nq + 1) {
        add = J[base + p * sp + sq];  // (p, 1), p in [1, np]: (1,1) or (p>1,1)
      } else {
        add = J[base + sp + q * sq];  // (np+1, q): (1,1) when q = 1, else (1,q)
      }
      J[base + p * sp + q * sq] += add;
      return;
    }
    t -= count;
  }
}

// E energy adjoint seed and field-energy derivative helpers are host-side.

}  // namespace cpic
