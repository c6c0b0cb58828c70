// cpic_api.hip -- the C-ABI of the MI355X-native CabanaPIC hot path.
//
// A context owns the device arrays of one simulation and one HIP stream.
// Every entry point validates its arguments on the host before a kernel is
// launched, so no kernel is ever given a grid or an index range that does
// not match its buffers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/cabanapic_amd.h"
#include "cpic_push_kernels.h"
#include "cpic_tangent_kernels.h"
#include "cpic_adjoint_kernels.h"

using namespace cpic;

namespace {

thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

#define HIP_TRY(expr)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      return fail(CPIC_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                  __FILE__, __LINE__);                                                  \
  } while (0)

#define CHECK_LAUNCH() HIP_TRY(hipGetLastError())

inline int cdiv(long a, long b) { return (int)((a + b - 1) / b); }

// Runs an entry point's body; a host allocation failure (or any other C++
// exception) becomes an error code instead of crossing the C ABI.
template <class F>
int guarded(F&& f) {
  try {
    return f();
  } catch (const std::bad_alloc&) {
    return fail(CPIC_ERR_INVALID, "out of host memory");
  } catch (const std::exception& e) {
    return fail(CPIC_ERR_INVALID, "host error: %s", e.what());
  } catch (...) {
    return fail(CPIC_ERR_INVALID, "host error");
  }
}

// Pushes between column compactions when the caller names none.  The in-place
// pass fills the holes departures leave before it lengthens a column, and
// rows past the ends of all the columns of a line cost it nothing, so packing
// rarely pays: 2.70 / 2.67 / 2.65 ms per step for intervals 8 / 16 / 32 on
// 128^3 cells x 32 particles per cell over 200 steps.
constexpr int kCompactInterval = 32;

}  // namespace

namespace cpic_detail {
// the error path of the library's other translation units
int fail_msg(int code, const char* msg) { return fail(code, "%s", msg); }
}  // namespace cpic_detail

// The particle columns and everything sized by them (see
// cpic_push_kernels.h): replaced as a whole when the capacity changes.
struct ColumnSet {
  int D = 0;               // rows per column
  int MW = 0;              // mask words per cell
  void* cols[2] = {};      // kSlotLen reals per slot: x,y,z,ux,uy,uz,w,tag
  int* cnt[2] = {};        // rows in use per interior cell
  unsigned long long* mask[2] = {};  // occupied rows, [MW][nc_pad]
  // crossing records: list 1 (first and third crossings), list 2 (second)
  int M1 = 0, M2 = 0;
  void* rec1 = nullptr;
  void* rec2 = nullptr;
  unsigned char* rc1 = nullptr;  // record counts, [6][nc] bytes
  unsigned char* rc2 = nullptr;
  void* ovf_rec = nullptr;
  int* ovf_cell = nullptr;
  int ovf_cap = 0;
  unsigned long long* dep = nullptr;  // departure masks of pass 1, [7][MW][nc_pad]
  int* dep_words = nullptr;           // mask words written by pass 1, per group
  // per group, in dep_words' allocation: set by the first finish pass when
  // it writes a record addressed to a cell of the group, cleared by the
  // second, whose waves without a mark have nothing to do
  int* rec_flags = nullptr;
};

constexpr int kStatusStride = 64;  // ints per status line (256 bytes)

struct cpic_context : ColumnSet {
  int real_size = 4;
  long nx = 0, ny = 0, nz = 0;
  long ncells = 0;  // Ng
  long nc = 0;      // interior cells
  long nc_pad = 0;  // interior cells rounded up to whole blocks (kBlock lanes)
  long cap = 0;     // particle capacity declared at creation
  long column_rows = 0;  // requested rows per column (0 = automatic)
  int* occ_host = nullptr;        // pinned copies of the last push's error flags and max_count
  hipEvent_t occ_event = nullptr;  // recorded after those copies
  bool occ_pending = false;        // a copy is in flight
  long np = 0;      // particles resident (conserved by the push)
  Grid grid{};
  CellGrid cg{};

  // particle columns (double-buffered): the ColumnSet base
  int cur = 0;             // buffer holding the current particles
  long pushes = 0;         // pushes since the last upload (the first of every
                           // compaction interval packs the columns)
  // device status words, one allocation (status), each on its own 256-byte
  // line (every wave reads the error flags while pass 1's waves update the
  // counters atomically; sharing a line slowed the step by 20 us):
  // [0] error flags (err), [kStatusStride] largest column occupancy seen by
  // the last push (max_count), [2 kStatusStride] particles that left their
  // cell (movers), [3 kStatusStride] overflow records (ovf_count).  A push
  // clears the three counters with one launch and the host reads err and
  // max_count back with one copy.
  int* status = nullptr;
  unsigned int* ovf_count = nullptr;
  int* max_count = nullptr;

  void* fields = nullptr;
  void* interp = nullptr;
  void* acc = nullptr;             // slot-major [12][Ng]
  bool acc_ghosts_zero = true;     // ghost accumulators known to be zero
  double* partials = nullptr;  // reduction partials
  double* red_out = nullptr;   // [E, B, KE]
  int* err = nullptr;
  unsigned int* movers = nullptr;
  double* hist_dev = nullptr;  // per-step energy rows of cpic_step
  long hist_cap = 0;           // rows allocated in hist_dev
  int n_partials = 0;
  double* gauss_q = nullptr;     // node charge, then the residual (cpic_gauss_residual)
  double* gauss_base = nullptr;  // baseline residual
  unsigned long long* gauss_out = nullptr;  // [max |r| or |r - base|, max |rho / eps0|]
  bool gauss_have_base = false;

  // dump snapshots (cpic_snapshot_begin / _wait): two slots, each a device
  // staging area and its pinned host copy, filled on the copy stream
  struct Snapshot {
    void* dev = nullptr;
    void* host = nullptr;
    size_t bytes = 0;
    hipEvent_t staged = nullptr;  // on the step stream, after the gather
    hipEvent_t ready = nullptr;   // on the copy stream, after the download
    bool pending = false;         // begun, not yet waited for
    // what the pending snapshot holds, fixed at its begin: particle count,
    // line length, byte offsets of x, ux, cell, ex line and energies
    int64_t n = 0;
    int nx = 0;
    size_t off[5] = {};
  } snap[2];
  hipStream_t copy_stream = nullptr;

  hipStream_t stream = nullptr;
  int device = 0;
  int solver = CPIC_FIELD_SOLVER_EM;  // field solver of advance_b/e and energies
  bool compensated = false;           // compensated lane accumulators in the push

  // stage timing: bit s set = stage s bracketed by events
  unsigned timing = 0;
  struct Pending {
    int stage;
    hipEvent_t a, b;
  };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> free_events;
  long stage_count[CPIC_NUM_STAGES] = {};
  double stage_ms[CPIC_NUM_STAGES] = {};

  size_t rsz() const { return (size_t)real_size; }
  size_t col_len() const { return (size_t)nc_pad * (size_t)D; }
};

namespace {

int alloc(void** p, size_t bytes) {
  HIP_TRY(hipMalloc(p, bytes == 0 ? 16 : bytes));
  return CPIC_OK;
}

template <class F>
int with_real(cpic_context* c, F&& f) {
  if (c->real_size == 4) return f(float{});
  return f(double{});
}

int get_event(cpic_context* c, hipEvent_t* e) {
  if (!c->free_events.empty()) {
    *e = c->free_events.back();
    c->free_events.pop_back();
    return CPIC_OK;
  }
  // timing-only events: no system-scope fence when they complete, whose
  // cache write-back and invalidation would stall the next kernel by about
  // 10 us (the times are read only after a stream synchronisation)
  HIP_TRY(hipEventCreateWithFlags(e, hipEventDisableSystemFence));
  return CPIC_OK;
}

// Brackets the launches of one stage with events when its timing is on.
struct StageTimer {
  cpic_context* c;
  int stage;
  bool on;
  hipEvent_t a = nullptr, b = nullptr;
  int status = CPIC_OK;
  StageTimer(cpic_context* ctx, int s) : c(ctx), stage(s), on((ctx->timing >> s) & 1u) {
    if (!on) return;
    if ((status = get_event(c, &a)) != CPIC_OK) return;
    if ((status = get_event(c, &b)) != CPIC_OK) return;
    if (hipEventRecord(a, c->stream) != hipSuccess)
      status = fail(CPIC_ERR_HIP, "hipEventRecord failed");
  }
  int finish() {
    if (!on || status != CPIC_OK) return status;
    HIP_TRY(hipEventRecord(b, c->stream));
    c->pending.push_back({stage, a, b});
    return CPIC_OK;
  }
};

// Folds finished event pairs into the per-stage totals (after a sync).
int collect_timing(cpic_context* c) {
  for (auto& p : c->pending) {
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, p.a, p.b));
    c->stage_count[p.stage] += 1;
    c->stage_ms[p.stage] += ms;
    c->free_events.push_back(p.a);
    c->free_events.push_back(p.b);
  }
  c->pending.clear();
  return CPIC_OK;
}

int check_ctx(cpic_context* c) {
  if (!c) return fail(CPIC_ERR_INVALID, "null context");
  return CPIC_OK;
}

template <class R>
int fill_zero(cpic_context* c, void* p, size_t n) {
  if (n == 0) return CPIC_OK;
  const int blocks = (int)std::min<size_t>((n + kBlock - 1) / kBlock, 65536);
  k_fill_zero<R><<<blocks, kBlock, 0, c->stream>>>((R*)p, n);
  CHECK_LAUNCH();
  return CPIC_OK;
}

// ---- stages ---------------------------------------------------------------

template <class R>
int stage_load_interpolator(cpic_context* c) {
  StageTimer t(c, CPIC_STAGE_INTERP);
  const long n = c->nx * c->ny * c->nz;
  k_load_interpolator<R><<<cdiv(n, kBlock), kBlock, 0, c->stream>>>((const R*)c->fields,
                                                                      (R*)c->interp, c->grid);
  CHECK_LAUNCH();
  return t.finish();
}

template <class R>
int stage_clear(cpic_context* c) {
  StageTimer t(c, CPIC_STAGE_CLEAR);
  int s = fill_zero<R>(c, c->acc, (size_t)c->ncells * 12);
  if (s) return s;
  c->acc_ghosts_zero = true;
  return t.finish();
}

template <class R>
ColState<R> col_state(cpic_context* c, int from) {
  ColState<R> cs;
  cs.in = (const R*)c->cols[from];
  cs.out = (R*)c->cols[1 - from];
  cs.cnt_in = c->cnt[from];
  cs.cnt_out = c->cnt[1 - from];
  cs.mask_in = c->mask[from];
  cs.mask_out = c->mask[1 - from];
  cs.D = c->D;
  cs.MW = c->MW;
  cs.nc_pad = (int)c->nc_pad;
  return cs;
}

// push + move_p + deposit: pass 1, three finishing passes, spill fallback.
// accumulate = 1 adds to the accumulators as the reference's push does;
// accumulate = 0 overwrites them (valid right after a clear).  gather: where
// pass 1 takes the interpolators from (the array, or the fields).
template <class R>
int stage_push(cpic_context* c, R qdt_2mc, R cdt_dx, R cdt_dy, R cdt_dz, R qsp, int accumulate,
               int compact, int gather = kGatherRead) {
  StageTimer t(c, CPIC_STAGE_PUSH);
  int s;
  // max_count, movers and ovf_count (and the unused words between them)
  if ((s = fill_zero<int>(c, c->max_count, 2 * kStatusStride + 1))) return s;
  if (c->np == 0) return t.finish();
  if (c->D == 0) return fail(CPIC_ERR_INVALID, "particles resident without columns");
  const int blocks = cdiv(c->nc, kBlock);
  Overflow<R> ovf{(R*)c->ovf_rec, c->ovf_cell, c->ovf_count, c->ovf_cap};
  RecordList<R> L1{(R*)c->rec1, c->rc1, c->M1, (int)c->nc};
  RecordList<R> L2{(R*)c->rec2, c->rc2, c->M2, (int)c->nc};
  const ColState<R> cs = col_state<R>(c, c->cur);

  PushPass<R> p;
  p.cs = cs;
  p.dep = c->dep;
  p.dep_words = c->dep_words;
  p.interp = (R*)c->interp;
  p.fields = (const R*)c->fields;
  p.gather = gather;
  p.acc = (R*)c->acc;
  p.g = c->cg;
  p.ncells = (int)c->ncells;
  p.qdt_2mc = qdt_2mc;
  p.cdt_dx = cdt_dx;
  p.cdt_dy = cdt_dy;
  p.cdt_dz = cdt_dz;
  p.qsp = qsp;
  p.accumulate = accumulate;
  p.err = c->err;
  p.movers = c->movers;
  p.max_count = c->max_count;
  // k_push_rows keeps every particle in its row; k_push_pass packs the
  // columns.  The row stores of k_push_rows address one column group with
  // 32-bit byte offsets; columns too deep for that are packed at every push.
  const bool group_fits = (int64_t)c->D * 64 * kSlotLen * (int64_t)sizeof(R) < INT32_MAX;
  if (compact || !group_fits)
    c->compensated ? k_push_pass<R, true><<<blocks, kBlock, 0, c->stream>>>(p)
                   : k_push_pass<R, false><<<blocks, kBlock, 0, c->stream>>>(p);
  else
    c->compensated ? k_push_rows<R, true><<<blocks, kBlock, 0, c->stream>>>(p)
                   : k_push_rows<R, false><<<blocks, kBlock, 0, c->stream>>>(p);
  CHECK_LAUNCH();

  // the streak after the 1st crossing, found through the departure masks
  ArrivePass<R> ar;
  ar.cs = cs;
  ar.dep = c->dep;
  ar.dep_words = c->dep_words;
  ar.out = L2;
  ar.ovf = ovf;
  ar.acc = (R*)c->acc;
  ar.g = c->cg;
  ar.ncells = (int)c->ncells;
  ar.cdt_dx = cdt_dx;
  ar.cdt_dy = cdt_dy;
  ar.cdt_dz = cdt_dz;
  ar.qsp = qsp;
  ar.err = c->err;
  c->compensated ? k_arrive_pass<R, true><<<blocks, kBlock, 0, c->stream>>>(ar)
                 : k_arrive_pass<R, false><<<blocks, kBlock, 0, c->stream>>>(ar);
  CHECK_LAUNCH();

  // the streaks after the 2nd and 3rd crossings, from the record lists
  const RecordList<R> ins[2] = {L2, L1};
  const RecordList<R> outs[2] = {L1, L2};
  for (int pass = 0; pass < 2; ++pass) {
    FinishPass<R> f;
    f.cs = cs;
    f.in = ins[pass];
    f.out = outs[pass];
    f.ovf = ovf;
    f.acc = (R*)c->acc;
    f.g = c->cg;
    f.ncells = (int)c->ncells;
    f.qsp = qsp;
    f.last = (pass == 1) ? 1 : 0;
    f.flag_set = (pass == 0) ? c->rec_flags : nullptr;
    f.flag_skip = (pass == 1) ? c->rec_flags : nullptr;
    f.err = c->err;
    c->compensated ? k_finish_pass<R, true><<<blocks, kBlock, 0, c->stream>>>(f)
                   : k_finish_pass<R, false><<<blocks, kBlock, 0, c->stream>>>(f);
    CHECK_LAUNCH();
  }
  SpillPass<R> sp{cs, ovf, (R*)c->acc, c->cg, (int)c->ncells, qsp, c->err};
  k_spill_pass<R><<<64, kBlock, 0, c->stream>>>(sp);
  CHECK_LAUNCH();
  c->cur = 1 - c->cur;
  return t.finish();
}

template <class R>
int stage_unload(cpic_context* c, R dx, R dy, R dz, R dt) {
  StageTimer t(c, CPIC_STAGE_UNLOAD);
  // real_t products, double quotient, stored as real_t (accumulator.cpp:66-68)
  const R cx = 0.25 / (dy * dz * dt);
  const R cy = 0.25 / (dz * dx * dt);
  const R cz = 0.25 / (dx * dy * dt);
  const long n = (c->nx + 1) * (c->ny + 1) * (c->nz + 1);
  k_unload<R><<<cdiv(n, kBlock), kBlock, 0, c->stream>>>((R*)c->fields, (const R*)c->acc, c->grid,
                                                         cx, cy, cz);
  CHECK_LAUNCH();
  return t.finish();
}

template <class R>
int launch_copy_ghosts(cpic_context* c, int comp0) {
  const long X = c->nx + 2, Y = c->ny + 2;
  const long n = 2 * (X * Y + X * c->nz + c->ny * c->nz);
  k_copy_ghosts<R><<<cdiv(n, kBlock), kBlock, 0, c->stream>>>((R*)c->fields, comp0, c->grid);
  CHECK_LAUNCH();
  return CPIC_OK;
}

template <class R>
int launch_fold_j(cpic_context* c) {
  const long n = c->nx * (c->nz + c->ny - 1) + c->ny * (c->nx + c->nz - 1) +
                 c->nz * (c->ny + c->nx - 1);
  k_fold_j<R><<<cdiv(n, kBlock), kBlock, 0, c->stream>>>((R*)c->fields, c->grid);
  CHECK_LAUNCH();
  return CPIC_OK;
}

template <class R>
int stage_advance_b(cpic_context* c, R px, R py, R pz) {
  StageTimer t(c, CPIC_STAGE_ADVANCE_B);
  // ES_Field_Solver_1D::advance_b does nothing (src/fields.h:470-480)
  if (c->solver == CPIC_FIELD_SOLVER_ES_1D) return t.finish();
  const long n = c->nx * c->ny * c->nz;
  // the kernel also writes the B ghosts (the periodic copy)
  k_advance_b<R><<<cdiv(n, kBlock), kBlock, 0, c->stream>>>((R*)c->fields, c->grid, px, py, pz);
  CHECK_LAUNCH();
  return t.finish();
}

template <class R>
int stage_ghosts_j(cpic_context* c) {
  int s = launch_fold_j<R>(c);
  if (s) return s;
  return launch_copy_ghosts<R>(c, 6);
}

template <class R>
int stage_advance_e(cpic_context* c, R px, R py, R pz, R dt_eps0) {
  StageTimer t(c, CPIC_STAGE_ADVANCE_E);
  if (c->solver == CPIC_FIELD_SOLVER_ES_1D) {
    // ES_Field_Solver_1D::advance_e, src/fields.h:511-545
    k_fold_j_ref<R><<<cdiv(c->nx + c->ny + c->nz, kBlock), kBlock, 0, c->stream>>>((R*)c->fields,
                                                                                 c->grid);
    CHECK_LAUNCH();
    k_es_advance_e<R><<<cdiv(c->ncells, kBlock), kBlock, 0, c->stream>>>(
        (R*)c->fields, (int)c->ncells, dt_eps0);
    CHECK_LAUNCH();
    return t.finish();
  }
  // the J fold, then advance_e, which also writes the J ghosts (the
  // periodic copy that follows the fold in stage_ghosts_j)
  int s = launch_fold_j<R>(c);
  if (s) return s;
  const long n = (c->nx + 1) * (c->ny + 1) * (c->nz + 1);
  k_advance_e<R><<<cdiv(n, kBlock), kBlock, 0, c->stream>>>((R*)c->fields, c->grid, px, py, pz,
                                                            dt_eps0);
  CHECK_LAUNCH();
  return t.finish();
}

// E and B energies into out[0..1] (device), FP64.
template <class R>
int launch_field_energy(cpic_context* c, double* out) {
  const bool es = c->solver == CPIC_FIELD_SOLVER_ES_1D;
  const long n = es ? c->ncells : c->nx * c->ny * c->nz;
  const int blocks = std::max(1, std::min(cdiv(n, kBlock), c->n_partials));
  if (es)
    k_es_energy_partial<R><<<blocks, kBlock, 0, c->stream>>>((const R*)c->fields, c->grid,
                                                             c->partials);
  else
    k_field_energy_partial<R><<<blocks, kBlock, 0, c->stream>>>((const R*)c->fields, c->grid,
                                                                c->partials);
  CHECK_LAUNCH();
  k_reduce_final<2><<<1, kBlock, 0, c->stream>>>(c->partials, blocks, 0.5, 0.5, out);
  CHECK_LAUNCH();
  return CPIC_OK;
}

template <class R>
int launch_kinetic(cpic_context* c, double scale, double* out) {
  double* part = c->partials + 2 * (size_t)c->n_partials;
  int blocks = 1;
  if (c->D > 0 && c->np > 0) {
    blocks = std::max(1, std::min(cdiv(c->nc, kBlock), c->n_partials));
    const ColState<R> cs = col_state<R>(c, c->cur);
    k_kinetic_columns<R><<<blocks, kBlock, 0, c->stream>>>(cs, (int)c->nc, part);
  } else {
    k_fill_zero<double><<<1, kBlock, 0, c->stream>>>(part, 1);
  }
  CHECK_LAUNCH();
  k_reduce_final<1><<<1, kBlock, 0, c->stream>>>(part, blocks, scale, 0.0, out);
  CHECK_LAUNCH();
  return CPIC_OK;
}

// ---- particle columns --------------------------------------------------------
// Capacities: D rows per column (see upload_particles); M1 record slots per
// (cell, face) for the common first crossings, M2 for the rarer later ones;
// spills beyond them are finished by the fallback kernel.

void free_column_set(ColumnSet& s) {
  // best effort: a free that fails leaves nothing to act upon
  for (int b = 0; b < 2; ++b) {
    if (s.cols[b]) (void)hipFree(s.cols[b]);
    if (s.cnt[b]) (void)hipFree(s.cnt[b]);
    if (s.mask[b]) (void)hipFree(s.mask[b]);
  }
  void* ptrs[] = {s.rec1, s.rec2, s.rc1, s.rc2, s.ovf_rec, s.ovf_cell, s.dep, s.dep_words};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  s = ColumnSet{};
}

// Allocates and clears a complete column set for D rows per column into
// out; on failure frees what it allocated and leaves out empty.  The
// context's current set is not touched.
int alloc_column_set(cpic_context* c, int D, int M1, int M2, ColumnSet& out) {
  ColumnSet s;
  s.D = D;
  s.MW = (D + 63) / 64;
  s.M1 = M1;
  s.M2 = M2;
  s.ovf_cap = (int)std::max(65536L, c->nc / 4);
  const size_t len = (size_t)c->nc_pad * (size_t)D;
  const size_t words = (size_t)s.MW * c->nc_pad;
  const size_t per = (size_t)record_bytes(c->real_size) * 6 * (size_t)c->nc;
  auto run = [&]() -> int {
    int t;
    for (int b = 0; b < 2; ++b) {
      if ((t = alloc(&s.cols[b], c->rsz() * kSlotLen * len))) return t;
      if ((t = alloc((void**)&s.cnt[b], sizeof(int) * c->nc_pad))) return t;
      if ((t = alloc((void**)&s.mask[b], sizeof(unsigned long long) * words))) return t;
    }
    if ((t = alloc(&s.rec1, per * M1))) return t;
    if ((t = alloc(&s.rec2, per * M2))) return t;
    if ((t = alloc((void**)&s.rc1, 6 * (size_t)c->nc))) return t;
    if ((t = alloc((void**)&s.rc2, 6 * (size_t)c->nc))) return t;
    if ((t = alloc(&s.ovf_rec, (size_t)record_bytes(c->real_size) * s.ovf_cap))) return t;
    if ((t = alloc((void**)&s.ovf_cell, sizeof(int) * s.ovf_cap))) return t;
    if ((t = alloc((void**)&s.dep, sizeof(unsigned long long) * kDepMasks * words))) return t;
    if ((t = alloc((void**)&s.dep_words, sizeof(int) * 2 * (c->nc_pad / 64)))) return t;
    s.rec_flags = s.dep_words + c->nc_pad / 64;
    for (int b = 0; b < 2; ++b) {
      if ((t = fill_zero<int>(c, s.cnt[b], c->nc_pad))) return t;
      if ((t = fill_zero<unsigned long long>(c, s.mask[b], words))) return t;
    }
    if ((t = fill_zero<unsigned long long>(c, s.dep, kDepMasks * words))) return t;
    if ((t = fill_zero<int>(c, s.dep_words, 2 * (c->nc_pad / 64)))) return t;
    return CPIC_OK;
  };
  if (int t = run()) {
    (void)hipStreamSynchronize(c->stream);
    free_column_set(s);
    return t;
  }
  out = s;
  return CPIC_OK;
}

// Replaces the context's column set by s (after the stream has drained).
// If the stream reports an error, s is freed and the current set is kept.
int install_column_set(cpic_context* c, ColumnSet& s) {
  const hipError_t e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) {
    free_column_set(s);
    return fail(CPIC_ERR_HIP, "column set installation: %s", hipGetErrorString(e));
  }
  ColumnSet& mine = *c;
  free_column_set(mine);
  mine = s;
  s = ColumnSet{};
  return CPIC_OK;
}

// Device scratch of a synchronous host transfer, freed on every path.
struct Scratch {
  void* p = nullptr;
  ~Scratch() {
    if (p) (void)hipFree(p);
  }
};

int check_periodic(int boundary) {
  if (boundary != CPIC_BOUNDARY_PERIODIC)
    return fail(CPIC_ERR_UNSUPPORTED,
                "only periodic boundaries are supported (the reference exits for Reflect)");
  return CPIC_OK;
}

int check_device_errors(cpic_context* c) {
  int flags = 0;
  HIP_TRY(hipMemcpyAsync(&flags, c->err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (flags)
    return fail(CPIC_ERR_DEVICE, "device kernel flagged invalid particles (flags=%d)", flags);
  return CPIC_OK;
}

// ---- column headroom -------------------------------------------------------
// Each push copies its largest input column occupancy (rows in use) and the
// device error flags to pinned host memory.  Before the next push the host
// reads the copies (one step behind, so the GPU does not idle).  A set flag
// fails the push: the kernels of a push also return at once when a flag is
// set, so pushes already queued behind a failed one change nothing.  A
// column within max(8, D/4) rows of its capacity makes the host re-lay the
// columns out on the device with 1.5x the rows, so that particles bunching
// in one cell (a nonlinear two-stream phase reaches 4x the mean count) do
// not overflow it.  A push that overflows anyway fails with kErrColumnFull.

int note_occupancy(cpic_context* c) {
  if (c->D == 0 || c->np == 0) return CPIC_OK;
  if (!c->occ_host) {
    HIP_TRY(hipHostMalloc((void**)&c->occ_host, (kStatusStride + 1) * sizeof(int),
                          hipHostMallocDefault));
    HIP_TRY(hipEventCreateWithFlags(&c->occ_event, hipEventDisableTiming));
  }
  // error flags and max_count, the first and second status lines
  HIP_TRY(hipMemcpyAsync(c->occ_host, c->status, (kStatusStride + 1) * sizeof(int),
                         hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipEventRecord(c->occ_event, c->stream));
  c->occ_pending = true;
  return CPIC_OK;
}

// The particles packed into columns of D rows (a multiple of 64, more than
// the current rows), on the device; row order within a column is kept.
template <class R>
int regrow_columns(cpic_context* c, int D) {
  ColumnSet s;
  if (int t = alloc_column_set(c, D, c->M1, c->M2, s)) return t;
  const int blocks = cdiv(c->nc_pad, kBlock);
  k_relayout<R><<<blocks, kBlock, 0, c->stream>>>(
      (const R*)c->cols[c->cur], c->cnt[c->cur], c->mask[c->cur], c->D, (R*)s.cols[0], s.cnt[0],
      D, (int)c->nc, (int)c->nc_pad, c->err);
  k_masks_packed<<<blocks, kBlock, 0, c->stream>>>(s.cnt[0], s.mask[0], (int)c->nc,
                                                   (int)c->nc_pad, s.MW);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    (void)hipStreamSynchronize(c->stream);
    free_column_set(s);
    return fail(CPIC_ERR_HIP, "column regrowth launch failed: %s", hipGetErrorString(e));
  }
  if (int t = install_column_set(c, s)) return t;
  c->cur = 0;
  c->pushes = 0;
  return CPIC_OK;
}

int check_headroom(cpic_context* c) {
  if (!c->occ_pending) return CPIC_OK;
  HIP_TRY(hipEventSynchronize(c->occ_event));
  c->occ_pending = false;
  if (c->occ_host[0])
    return fail(CPIC_ERR_DEVICE, "an earlier push failed on the device (flags=%d)",
                c->occ_host[0]);
  const int max_count = c->occ_host[kStatusStride];
  if (c->D == 0 || max_count < c->D - std::max(8, c->D / 4)) return CPIC_OK;
  const int D = (c->D * 3 / 2 + 63) / 64 * 64;  // 1.5x the rows
  return with_real(c, [&](auto r) { return regrow_columns<decltype(r)>(c, D); });
}

}  // namespace

namespace {

// ---- dump snapshots ----------------------------------------------------------
// Layout of a slot (device staging and pinned host copy alike): x[n], ux[n]
// (reals), cell[n] (int32), ex[nx] (reals), energies[3] (double), each part
// starting on a 256-byte boundary.
struct SnapLayout {
  size_t x, ux, cell, ex, energies, bytes;
};

SnapLayout snap_layout(const cpic_context* c) {
  auto up = [](size_t b) { return (b + 255) / 256 * 256; };
  const size_t n = (size_t)c->np, rs = c->rsz();
  SnapLayout L;
  L.x = 0;
  L.ux = L.x + up(rs * n);
  L.cell = L.ux + up(rs * n);
  L.ex = L.cell + up(sizeof(int) * n);
  L.energies = L.ex + up(rs * (size_t)c->nx);
  L.bytes = L.energies + 3 * sizeof(double);
  return L;
}

void free_snapshot(cpic_context::Snapshot& sn) {
  if (sn.dev) (void)hipFree(sn.dev);
  if (sn.host) (void)hipHostFree(sn.host);
  if (sn.staged) (void)hipEventDestroy(sn.staged);
  if (sn.ready) (void)hipEventDestroy(sn.ready);
  sn = cpic_context::Snapshot{};
}

int snapshot_begin(cpic_context* c, int slot) {
  if (int s = check_ctx(c)) return s;
  if (slot != 0 && slot != 1) return fail(CPIC_ERR_INVALID, "snapshot slot must be 0 or 1");
  auto& sn = c->snap[slot];
  if (sn.pending)
    return fail(CPIC_ERR_INVALID, "snapshot slot %d was begun and not waited for", slot);
  const SnapLayout L = snap_layout(c);
  // first use, or the particle count changed: (re)allocate this slot and the
  // other one too unless it holds a snapshot not yet waited for
  for (int k : {slot, 1 - slot}) {
    auto& q = c->snap[k];
    if (q.bytes == L.bytes || q.pending) continue;
    HIP_TRY(hipStreamSynchronize(c->stream));
    free_snapshot(q);
    if (int s = alloc(&q.dev, L.bytes)) return s;
    if (hipHostMalloc(&q.host, L.bytes, hipHostMallocDefault) != hipSuccess) {
      q.host = nullptr;
      free_snapshot(q);
      return fail(CPIC_ERR_HIP, "pinned allocation of %zu bytes failed", L.bytes);
    }
    HIP_TRY(hipEventCreateWithFlags(&q.staged, hipEventDisableTiming));
    HIP_TRY(hipEventCreateWithFlags(&q.ready, hipEventDisableTiming));
    q.bytes = L.bytes;
  }
  if (!c->copy_stream)
    HIP_TRY(hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking));
  char* d = (char*)sn.dev;
  const long nthreads = std::max<long>(std::max<long>(c->nc, c->nx), 3);
  const bool have = c->np > 0 && c->D > 0;
  int s = with_real(c, [&](auto r) -> int {
    using R = decltype(r);
    k_snapshot<R><<<cdiv(nthreads, kBlock), kBlock, 0, c->stream>>>(
        have ? (const R*)c->cols[c->cur] : nullptr, have ? c->cnt[c->cur] : nullptr,
        have ? c->mask[c->cur] : nullptr, c->cg, c->D, (int)c->nc_pad, have ? c->np : 0,
        (const R*)c->fields, c->grid.sy, c->grid.sz, c->red_out, (R*)(d + L.x),
        (R*)(d + L.ux), (int*)(d + L.cell), (R*)(d + L.ex), (double*)(d + L.energies), c->err);
    CHECK_LAUNCH();
    return CPIC_OK;
  });
  if (s) return s;
  HIP_TRY(hipEventRecord(sn.staged, c->stream));
  HIP_TRY(hipStreamWaitEvent(c->copy_stream, sn.staged, 0));
  HIP_TRY(hipMemcpyAsync(sn.host, sn.dev, L.bytes, hipMemcpyDeviceToHost, c->copy_stream));
  HIP_TRY(hipEventRecord(sn.ready, c->copy_stream));
  sn.n = c->np;
  sn.nx = (int)c->nx;
  const size_t off[5] = {L.x, L.ux, L.cell, L.ex, L.energies};
  std::copy(off, off + 5, sn.off);
  sn.pending = true;
  return CPIC_OK;
}

int snapshot_wait(cpic_context* c, int slot, cpic_snapshot* out) {
  if (int s = check_ctx(c)) return s;
  if (slot != 0 && slot != 1) return fail(CPIC_ERR_INVALID, "snapshot slot must be 0 or 1");
  if (!out) return fail(CPIC_ERR_INVALID, "null snapshot");
  auto& sn = c->snap[slot];
  if (!sn.pending) return fail(CPIC_ERR_INVALID, "no snapshot begun on slot %d", slot);
  HIP_TRY(hipEventSynchronize(sn.ready));
  sn.pending = false;
  // the layout of the begin: an upload since then may have changed the
  // particle count
  const char* h = (const char*)sn.host;
  out->n = sn.n;
  out->nx = sn.nx;
  out->real_size = c->real_size;
  out->x = h + sn.off[0];
  out->ux = h + sn.off[1];
  out->cell = (const int32_t*)(h + sn.off[2]);
  out->ex_line = h + sn.off[3];
  const double* e = (const double*)(h + sn.off[4]);
  out->e_energy = e[0];
  out->b_energy = e[1];
  out->kinetic_energy = e[2];
  return CPIC_OK;
}

}  // namespace

namespace {

// nsteps iterations of the loop body of example/example.cpp:216-279 on the
// context's stream, no host synchronisation; energies of step s (1-based)
// into dhist row s - 1 (3 doubles: E, B, kinetic) when dhist is not null,
// as sp's energy and kinetic intervals ask.
template <class R>
int run_steps(cpic_context* c, const cpic_step_params* sp, int64_t nsteps, double* dhist) {
  const int ei = sp->energy_interval;
  const int ki = sp->kinetic_interval;
  const R px = (R)sp->px, py = (R)sp->py, pz = (R)sp->pz;
  const R half = (R)0.5;
  const R hpx = half * px, hpy = half * py, hpz = half * pz;
  for (int64_t step = 1; step <= nsteps; ++step) {
    int t;
    // With particles resident, pass 1 gathers each cell's interpolator from
    // the fields itself; the array is written on the last step only, so that
    // it holds the last push's interpolators when the call returns.
    const bool fused = c->np > 0 && c->D > 0;
    if (!fused && (t = stage_load_interpolator<R>(c))) return t;
    // the push overwrites every interior accumulator, so the clear is only
    // needed while ghost accumulators may hold data
    if (!c->acc_ghosts_zero && (t = stage_clear<R>(c))) return t;
    if ((t = check_headroom(c))) return t;
    const int ci = sp->compact_interval > 0 ? sp->compact_interval : kCompactInterval;
    const int compact = (c->pushes++ % ci == 0) ? 1 : 0;
    const int gather = !fused ? kGatherRead : step == nsteps ? kGatherStore : kGatherFields;
    if ((t = stage_push<R>(c, (R)sp->qdt_2mc, (R)sp->cdt_dx, (R)sp->cdt_dy, (R)sp->cdt_dz,
                           (R)sp->qsp, 0, compact, gather)))
      return t;
    if ((t = note_occupancy(c))) return t;
    if ((t = stage_unload<R>(c, (R)sp->dx, (R)sp->dy, (R)sp->dz, (R)sp->dt))) return t;
    if ((t = stage_advance_b<R>(c, hpx, hpy, hpz))) return t;
    if ((t = stage_advance_e<R>(c, px, py, pz, (R)sp->dt_eps0))) return t;
    if ((t = stage_advance_b<R>(c, hpx, hpy, hpz))) return t;
    const bool do_e = ei > 0 && step % ei == 0, do_k = ki > 0 && step % ki == 0;
    if (do_e || do_k) {
      StageTimer tm(c, CPIC_STAGE_ENERGY);
      double* row = dhist ? dhist + 3 * (step - 1) : c->red_out;
      if (do_e && (t = launch_field_energy<R>(c, row))) return t;
      if (do_k && (t = launch_kinetic<R>(c, (double)sp->me * sp->c * sp->c, row + 2))) return t;
      if ((t = tm.finish())) return t;
    }
  }
  return CPIC_OK;
}

}  // namespace

// ===========================================================================

extern "C" {

const char* cpic_last_error(void) { return g_last_error.c_str(); }

int cpic_create(const cpic_config* cfg, cpic_context** out) {
  return guarded([&]() -> int {
    if (!cfg || !out) return fail(CPIC_ERR_INVALID, "null argument");
    *out = nullptr;
    if (cfg->real_size != 4 && cfg->real_size != 8)
      return fail(CPIC_ERR_INVALID, "real_size must be 4 or 8 (got %d)", cfg->real_size);
    if (cfg->num_ghosts != 1)
      return fail(CPIC_ERR_UNSUPPORTED, "num_ghosts must be 1 (got %d)", cfg->num_ghosts);
    if (cfg->nx < 1 || cfg->ny < 1 || cfg->nz < 1)
      return fail(CPIC_ERR_INVALID, "grid extents must be >= 1");
    const long X = cfg->nx + 2, Y = cfg->ny + 2, Z = cfg->nz + 2;
    if ((double)X * Y * Z > 2147483647.0 / 20)
      return fail(CPIC_ERR_INVALID, "grid too large for 32-bit voxel indexing");
    if (cfg->max_particles < 0 || cfg->max_particles > 2147483647L)
      return fail(CPIC_ERR_INVALID, "max_particles out of range");
    if (cfg->column_rows < 0 || cfg->column_rows > (1L << 20))
      return fail(CPIC_ERR_INVALID, "column_rows out of range [0, 2^20]");
  
    cpic_context* c = new (std::nothrow) cpic_context();
    if (!c) return fail(CPIC_ERR_INVALID, "out of host memory");
    c->real_size = cfg->real_size;
    c->nx = cfg->nx;
    c->ny = cfg->ny;
    c->nz = cfg->nz;
    c->ncells = X * Y * Z;
    c->cap = cfg->max_particles;
    c->column_rows = (long)cfg->column_rows;
    c->grid.nx = (int)c->nx;
    c->grid.ny = (int)c->ny;
    c->grid.nz = (int)c->nz;
    c->grid.sy = (int)X;
    c->grid.sz = (int)(X * Y);
    c->grid.ncells = (int)c->ncells;
    c->nc = cfg->nx * cfg->ny * cfg->nz;
    c->nc_pad = (c->nc + kBlock - 1) / kBlock * kBlock;
    c->cg.nx = (int)c->nx;
    c->cg.ny = (int)c->ny;
    c->cg.nz = (int)c->nz;
    c->cg.nc = (int)c->nc;
    c->cg.sy = (int)X;
    c->cg.sz = (int)(X * Y);
    c->n_partials = 4096;
  
    auto bail = [&](int s) {
      cpic_destroy(c);
      return s;
    };
    int s;
    if (hipGetDevice(&c->device) != hipSuccess) return bail(fail(CPIC_ERR_HIP, "no HIP device"));
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
      return bail(fail(CPIC_ERR_HIP, "hipStreamCreate failed"));
    if ((s = alloc(&c->fields, c->rsz() * 9 * c->ncells))) return bail(s);
    if ((s = alloc(&c->interp, c->rsz() * kInterpStride * c->ncells))) return bail(s);
    if ((s = alloc(&c->acc, c->rsz() * 12 * c->ncells))) return bail(s);
    if ((s = alloc((void**)&c->partials, sizeof(double) * 3 * c->n_partials))) return bail(s);
    if ((s = alloc((void**)&c->red_out, sizeof(double) * 4))) return bail(s);
    if ((s = alloc((void**)&c->status, 4 * kStatusStride * sizeof(int)))) return bail(s);
    c->err = c->status;
    c->max_count = c->status + kStatusStride;
    c->movers = reinterpret_cast<unsigned int*>(c->status + 2 * kStatusStride);
    c->ovf_count = reinterpret_cast<unsigned int*>(c->status + 3 * kStatusStride);
    s = with_real(c, [&](auto r) {
      using R = decltype(r);
      int t;
      if ((t = fill_zero<R>(c, c->fields, 9 * (size_t)c->ncells))) return t;
      if ((t = fill_zero<R>(c, c->interp, kInterpStride * (size_t)c->ncells))) return t;
      if ((t = fill_zero<R>(c, c->acc, 12 * (size_t)c->ncells))) return t;
      return CPIC_OK;
    });
    if (s) return bail(s);
    if (hipMemsetAsync(c->status, 0, 4 * kStatusStride * sizeof(int), c->stream) != hipSuccess ||
        hipMemsetAsync(c->red_out, 0, sizeof(double) * 4, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
      return bail(fail(CPIC_ERR_HIP, "context initialisation failed"));
    *out = c;
    return CPIC_OK;
  });
}

int cpic_destroy(cpic_context* c) {
  return guarded([&]() -> int {
    if (!c) return CPIC_OK;
    // teardown is best effort: errors cannot be acted upon here
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto& p : c->pending) {
      (void)hipEventDestroy(p.a);
      (void)hipEventDestroy(p.b);
    }
    for (auto e : c->free_events) (void)hipEventDestroy(e);
    free_column_set(*c);
    void* ptrs[] = {c->status,  c->fields,     c->interp,   c->acc,     c->partials,
                    c->red_out, c->hist_dev,   c->gauss_q,  c->gauss_base, c->gauss_out};
    for (void* p : ptrs)
      if (p) (void)hipFree(p);
    if (c->occ_event) (void)hipEventDestroy(c->occ_event);
    if (c->occ_host) (void)hipHostFree(c->occ_host);
    if (c->copy_stream) (void)hipStreamSynchronize(c->copy_stream);
    for (auto& sn : c->snap) free_snapshot(sn);
    if (c->copy_stream) (void)hipStreamDestroy(c->copy_stream);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return CPIC_OK;
  });
}

int cpic_synchronize(cpic_context* c) {
  return guarded([&]() -> int {
    if (int s = check_ctx(c)) return s;
    HIP_TRY(hipStreamSynchronize(c->stream));
    return collect_timing(c);
  });
}

int cpic_get_stream(cpic_context* c, void** stream) {
  return guarded([&]() -> int {
    if (int s = check_ctx(c)) return s;
    if (!stream) return fail(CPIC_ERR_INVALID, "null stream pointer");
    *stream = (void*)c->stream;
    return CPIC_OK;
  });
}

namespace {

// Column rows for a largest cell count maxc at upload: 2.5x (at least +48),
// rounded up to whole 64-row mask words (the masks and departure words are
// allocated per word, so the rows up to the next word cost only slots); every
// particle in one cell where that costs at most 64 MB of columns; at least
// the configured column_rows.
int column_rows_for(const cpic_context* c, int maxc, int64_t n) {
  auto up8 = [](long v) { return (int)((v + 7) / 8 * 8); };
  int D = (std::max({maxc * 5 / 2, maxc + 48, 16}) + 63) / 64 * 64;
  const double worst_bytes = (double)c->nc_pad * (double)up8(n) * kSlotLen * (double)c->rsz();
  if (worst_bytes <= 64.0 * 1024 * 1024) D = std::max(D, up8(n));
  if (c->column_rows > 0) D = std::max(D, up8(c->column_rows));
  return D;
}

// Binning on the device: the host validates the cells and computes each
// particle's slot (its cell's column, rows in input order); the particles
// and slots go to a device staging area, and a kernel scatters them into
// freshly cleared columns.  A new column set is allocated before the old one
// is released, so a failed allocation leaves the resident particles as they
// were; a failure after that leaves the context without particles.
int upload_particles(cpic_context* c, int64_t n, const void* x, const void* y, const void* z,
                     const void* ux, const void* uy, const void* uz, const void* w,
                     const int32_t* cell) {
  if (int s = check_ctx(c)) return s;
  if (n < 0 || n > c->cap)
    return fail(CPIC_ERR_INVALID, "particle count %lld exceeds capacity %ld", (long long)n, c->cap);
  const void* src[7] = {x, y, z, ux, uy, uz, w};
  if (n > 0) {
    for (int k = 0; k < 7; ++k)
      if (!src[k]) return fail(CPIC_ERR_INVALID, "null particle array");
    if (!cell) return fail(CPIC_ERR_INVALID, "null cell array");
  }
  // every particle must sit in an interior voxel (src/types.h:40)
  const long X = c->nx + 2, Y = c->ny + 2;
  std::vector<int> count(c->nc_pad, 0);
  std::vector<long long> slot((size_t)n);  // (cell << 32 | rank), then the slot
  int maxc = 0;
  for (int64_t p = 0; p < n; ++p) {
    const long v = cell[p];
    const long ix = v % X, iy = (v / X) % Y, iz = v / (X * Y);
    if (v < 0 || v >= c->ncells || ix < 1 || ix > c->nx || iy < 1 || iy > c->ny || iz < 1 ||
        iz > c->nz)
      return fail(CPIC_ERR_INVALID, "particle %lld has non-interior cell %ld", (long long)p, v);
    const long ci = (ix - 1) + c->nx * ((iy - 1) + c->ny * (iz - 1));
    slot[p] = ((long long)ci << 32) | (long long)count[ci];
    maxc = std::max(maxc, ++count[ci]);
  }
  const int D = column_rows_for(c, maxc, n);
  auto up8 = [](long v) { return (int)((v + 7) / 8 * 8); };
  const int M1 = std::min({D, std::max(16, up8(maxc / 2)), kMaxRecords});
  const int M2 = std::min(D, 8);
  for (int64_t p = 0; p < n; ++p) {
    const long long ci = slot[p] >> 32, r = slot[p] & 0xffffffffll;
    slot[p] = ((ci >> 6) * D + r) * 64 + (ci & 63);
  }
  const size_t rs = c->rsz();
  Scratch stage;
  if (n > 0)
    if (int s = alloc(&stage.p, 7 * rs * (size_t)n + sizeof(long long) * (size_t)n)) return s;
  ColumnSet fresh;
  const bool realloc = !(c->D == D && c->M1 == M1 && c->M2 == M2 && c->cnt[0]);
  if (realloc)
    if (int s = alloc_column_set(c, D, M1, M2, fresh)) return s;
  char* sp = (char*)stage.p;
  long long* slot_dev = (long long*)(sp + 7 * rs * (size_t)n);
  auto stage_in = [&]() -> int {
    for (int k = 0; k < 7 && n > 0; ++k)
      HIP_TRY(hipMemcpyAsync(sp + k * rs * (size_t)n, src[k], rs * (size_t)n,
                             hipMemcpyHostToDevice, c->stream));
    if (n > 0)
      HIP_TRY(hipMemcpyAsync(slot_dev, slot.data(), sizeof(long long) * (size_t)n,
                             hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return CPIC_OK;
  };
  if (int s = stage_in()) {
    free_column_set(fresh);
    return s;
  }
  // from here on the resident particles are replaced
  c->np = 0;
  if (realloc)
    if (int s = install_column_set(c, fresh)) return s;
  int s = with_real(c, [&](auto r) -> int {
    using R = decltype(r);
    int t;
    if ((t = fill_zero<R>(c, c->cols[0], kSlotLen * c->col_len()))) return t;
    if (n > 0) {
      const int blocks = (int)std::min<int64_t>((n + kBlock - 1) / kBlock, 65536);
      k_bin_scatter<R><<<blocks, kBlock, 0, c->stream>>>((const R*)sp, slot_dev, (long)n,
                                                         (R*)c->cols[0]);
      CHECK_LAUNCH();
    }
    HIP_TRY(hipMemcpyAsync(c->cnt[0], count.data(), sizeof(int) * c->nc_pad,
                           hipMemcpyHostToDevice, c->stream));
    k_masks_packed<<<cdiv(c->nc_pad, kBlock), kBlock, 0, c->stream>>>(
        c->cnt[0], c->mask[0], (int)c->nc, (int)c->nc_pad, c->MW);
    CHECK_LAUNCH();
    // a fresh particle state: earlier device errors no longer apply
    if ((t = fill_zero<int>(c, c->err, 1))) return t;
    HIP_TRY(hipStreamSynchronize(c->stream));
    return CPIC_OK;
  });
  if (s) return s;
  c->cur = 0;
  c->pushes = 0;
  c->occ_pending = false;
  c->np = n;
  return CPIC_OK;
}

// Gathering on the device: per-cell particle counts come to the host, which
// sets each cell's first output position; a kernel writes the particles
// (cells in interior order, rows in order) to a staging area that is then
// downloaded.  tag (may be null) receives each particle's upload index.
int download_particles(cpic_context* c, int64_t n, void* x, void* y, void* z, void* ux,
                       void* uy, void* uz, void* w, int32_t* cell, int64_t* tag) {
  if (int s = check_ctx(c)) return s;
  if (n != c->np)
    return fail(CPIC_ERR_INVALID, "download of %lld particles, %ld resident", (long long)n, c->np);
  if (n == 0) return CPIC_OK;
  void* dst[7] = {x, y, z, ux, uy, uz, w};
  const size_t rs = c->rsz();
  const size_t nc = (size_t)c->nc;
  Scratch stage;
  const size_t bytes = sizeof(long long) * nc + (7 * rs + sizeof(int) + sizeof(long long)) * n;
  if (int s = alloc(&stage.p, bytes)) return s;
  char* sp = (char*)stage.p;
  long long* first_dev = (long long*)sp;
  char* soa = sp + sizeof(long long) * nc;
  int* cell_dev = (int*)(soa + 7 * rs * n);
  long long* tag_dev = (long long*)(cell_dev + n);
  int* occ_dev = (int*)soa;  // the counts use the SoA area before the gather
  std::vector<int> occ(nc);
  k_column_occupancy<<<cdiv(c->nc, kBlock), kBlock, 0, c->stream>>>(
      c->cnt[c->cur], c->mask[c->cur], (int)c->nc, (int)c->nc_pad, c->D, occ_dev);
  CHECK_LAUNCH();
  HIP_TRY(hipMemcpyAsync(occ.data(), occ_dev, sizeof(int) * nc, hipMemcpyDeviceToHost,
                         c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  std::vector<long long> first(nc);
  long long total = 0;
  for (size_t i = 0; i < nc; ++i) {
    first[i] = total;
    total += occ[i];
  }
  if (total != n)
    return fail(CPIC_ERR_DEVICE, "columns hold %lld particles, %lld expected", total,
                (long long)n);
  HIP_TRY(hipMemcpyAsync(first_dev, first.data(), sizeof(long long) * nc, hipMemcpyHostToDevice,
                         c->stream));
  int s = with_real(c, [&](auto r) -> int {
    using R = decltype(r);
    k_gather_columns<R><<<cdiv(c->nc, kBlock), kBlock, 0, c->stream>>>(
        (const R*)c->cols[c->cur], c->cnt[c->cur], c->mask[c->cur], c->cg, c->D,
        (int)c->nc_pad, first_dev, (long)n, (R*)soa, cell_dev, tag ? tag_dev : nullptr);
    CHECK_LAUNCH();
    return CPIC_OK;
  });
  if (s) return s;
  for (int k = 0; k < 7; ++k)
    if (dst[k])
      HIP_TRY(hipMemcpyAsync(dst[k], soa + k * rs * n, rs * n, hipMemcpyDeviceToHost,
                             c->stream));
  if (cell)
    HIP_TRY(hipMemcpyAsync(cell, cell_dev, sizeof(int) * n, hipMemcpyDeviceToHost, c->stream));
  if (tag)
    HIP_TRY(hipMemcpyAsync(tag, tag_dev, sizeof(long long) * n, hipMemcpyDeviceToHost,
                           c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return CPIC_OK;
}

}  // namespace

// Host allocation failures are returned as errors, never thrown across the ABI.
int cpic_upload_particles(cpic_context* c, int64_t n, const void* x, const void* y, const void* z,
                          const void* ux, const void* uy, const void* uz, const void* w,
                          const int32_t* cell) {
  return guarded([&] { return upload_particles(c, n, x, y, z, ux, uy, uz, w, cell); });
}

int cpic_download_particles(cpic_context* c, int64_t n, void* x, void* y, void* z, void* ux,
                            void* uy, void* uz, void* w, int32_t* cell) {
  return guarded([&] { return download_particles(c, n, x, y, z, ux, uy, uz, w, cell, nullptr); });
}

int cpic_download_particle_tags(cpic_context* c, int64_t n, int64_t* tag) {
  if (!tag) return fail(CPIC_ERR_INVALID, "null tag array");
  return guarded([&] {
    return download_particles(c, n, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                              nullptr, nullptr, tag);
  });
}

int cpic_reserve_column_rows(cpic_context* c, int64_t rows) {
  return guarded([&]() -> int {
    if (int s = check_ctx(c)) return s;
    if (rows < 0 || rows > (1L << 20))
      return fail(CPIC_ERR_INVALID, "rows out of range [0, 2^20]");
    if (c->D == 0 || rows <= c->D) return CPIC_OK;
    const int D = (int)((rows + 63) / 64 * 64);
    int s = with_real(c, [&](auto r) { return regrow_columns<decltype(r)>(c, D); });
    if (s) return s;
    HIP_TRY(hipStreamSynchronize(c->stream));
    return check_device_errors(c);
  });
}

int cpic_column_rows(cpic_context* c, int64_t* rows) {
  if (int s = check_ctx(c)) return s;
  if (!rows) return fail(CPIC_ERR_INVALID, "null pointer");
  *rows = c->D;
  return CPIC_OK;
}

int cpic_num_particles(cpic_context* c, int64_t* n) {
  return guarded([&]() -> int {
    if (int s = check_ctx(c)) return s;
    if (!n) return fail(CPIC_ERR_INVALID, "null pointer");
    *n = c->np;
    return CPIC_OK;
  });
}

int cpic_upload_fields(cpic_context* c, const void* f) {
  return guarded([&]() -> int {
    if (int s = check_ctx(c)) return s;
    if (!f) return fail(CPIC_ERR_INVALID, "null fields");
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipMemcpy(c->fields, f, c->rsz() * 9 * c->ncells, hipMemcpyHostToDevice));
    return CPIC_OK;
  });
}

int cpic_download_fields(cpic_context* c, void* f) {
  return guarded([&]() -> int {
    if (int s = check_ctx(c)) return s;
    if (!f) return fail(CPIC_ERR_INVALID, "null fields");
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipMemcpy(f, c->fields, c->rsz() * 9 * c->ncells, hipMemcpyDeviceToHost));
    return CPIC_OK;
  });
}

int cpic_upload_interpolators(cpic_context* c, const void* in) {
  return guarded([&]() -> int {
    if (int s = check_ctx(c)) return s;
    if (!in) return fail(CPIC_ERR_INVALID, "null interpolators");
    // component-major host layout -> cell-major padded device layout
    const size_t rs = c->rsz();
    std::vector<unsigned char> buf(rs * kInterpStride * c->ncells, 0);
    const unsigned char* src = (const unsigned char*)in;
    for (long i = 0; i < c->ncells; ++i)
      for (int k = 0; k < 18; ++k)
        memcpy(&buf[(i * kInterpStride + k) * rs], &src[((size_t)k * c->ncells + i) * rs], rs);
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipMemcpy(c->interp, buf.data(), buf.size(), hipMemcpyHostToDevice));
    return CPIC_OK;
  });
}

int cpic_download_interpolators(cpic_context* c, void* out) {
  return guarded([&]() -> int {
    if (int s = check_ctx(c)) return s;
    if (!out) return fail(CPIC_ERR_INVALID, "null interpolators");
    const size_t rs = c->rsz();
    std::vector<unsigned char> buf(rs * kInterpStride * c->ncells);
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipMemcpy(buf.data(), c->interp, buf.size(), hipMemcpyDeviceToHost));
    unsigned char* dst = (unsigned char*)out;
    for (long i = 0; i < c->ncells; ++i)
      for (int k = 0; k < 18; ++k)
        memcpy(&dst[((size_t)k * c->ncells + i) * rs], &buf[(i * kInterpStride + k) * rs], rs);
    return CPIC_OK;
  });
}

int cpic_upload_accumulators(cpic_context* c, const void* in) {
  return guarded([&]() -> int {
    if (int s = check_ctx(c)) return s;
    if (!in) return fail(CPIC_ERR_INVALID, "null accumulators");
    const size_t rs = c->rsz();
    std::vector<unsigned char> buf(rs * 12 * c->ncells);
    const unsigned char* src = (const unsigned char*)in;
    for (long i = 0; i < c->ncells; ++i)
      for (int k = 0; k < 12; ++k)
        memcpy(&buf[((size_t)k * c->ncells + i) * rs], &src[((size_t)i * 12 + k) * rs], rs);
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipMemcpy(c->acc, buf.data(), buf.size(), hipMemcpyHostToDevice));
    c->acc_ghosts_zero = false;
    return CPIC_OK;
  });
}

int cpic_download_accumulators(cpic_context* c, void* out) {
  return guarded([&]() -> int {
    if (int s = check_ctx(c)) return s;
    if (!out) return fail(CPIC_ERR_INVALID, "null accumulators");
    const size_t rs = c->rsz();
    std::vector<unsigned char> buf(rs * 12 * c->ncells);
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipMemcpy(buf.data(), c->acc, buf.size(), hipMemcpyDeviceToHost));
    unsigned char* dst = (unsigned char*)out;
    for (long i = 0; i < c->ncells; ++i)
      for (int k = 0; k < 12; ++k)
        memcpy(&dst[((size_t)i * 12 + k) * rs], &buf[((size_t)k * c->ncells + i) * rs], rs);
    return CPIC_OK;
  });
}

int cpic_load_interpolator_array(cpic_context* c) {
  return guarded([&]() -> int {
    if (int s = check_ctx(c)) return s;
    return with_real(c, [&](auto r) { return stage_load_interpolator<decltype(r)>(c); });
  });
}

int cpic_clear_accumulator_array(cpic_context* c) {
  return guarded([&]() -> int {
    if (int s = check_ctx(c)) return s;
    return with_real(c, [&](auto r) { return stage_clear<decltype(r)>(c); });
  });
}

int cpic_push(cpic_context* c, double qdt_2mc, double cdt_dx, double cdt_dy, double cdt_dz,
              double qsp, int32_t boundary) {
  return guarded([&]() -> int {
    if (int s = check_ctx(c)) return s;
    if (int s = check_periodic(boundary)) return s;
    return with_real(c, [&](auto r) {
      using R = decltype(r);
      if (int t = check_headroom(c)) return t;
      // the first push after an upload and then every kCompactInterval-th one
      // pack the columns
      const int compact = (c->pushes++ % kCompactInterval == 0) ? 1 : 0;
      if (int t = stage_push<R>(c, (R)qdt_2mc, (R)cdt_dx, (R)cdt_dy, (R)cdt_dz, (R)qsp, 1, compact))
        return t;
      return note_occupancy(c);
    });
  });
}

int cpic_uncenter_particles(cpic_context* c, double qdt_2mc) {
  return guarded([&]() -> int {
    if (int s = check_ctx(c)) return s;
    if (c->np == 0 || c->D == 0) return CPIC_OK;
    return with_real(c, [&](auto r) {
      using R = decltype(r);
      k_uncenter_columns<R><<<cdiv(c->nc, kBlock), kBlock, 0, c->stream>>>(
          (R*)c->cols[c->cur], c->cnt[c->cur], c->mask[c->cur], (const R*)c->interp, c->cg, c->D,
          (int)c->nc_pad, (R)qdt_2mc);
      CHECK_LAUNCH();
      return CPIC_OK;
    });
  });
}

int cpic_unload_accumulator_array(cpic_context* c, double dx, double dy, double dz, double dt) {
  return guarded([&]() -> int {
    if (int s = check_ctx(c)) return s;
    return with_real(c, [&](auto r) {
      using R = decltype(r);
      return stage_unload<R>(c, (R)dx, (R)dy, (R)dz, (R)dt);
    });
  });
}

int cpic_advance_b(cpic_context* c, double px, double py, double pz) {
  return guarded([&]() -> int {
    if (int s = check_ctx(c)) return s;
    return with_real(c, [&](auto r) {
      using R = decltype(r);
      return stage_advance_b<R>(c, (R)px, (R)py, (R)pz);
    });
  });
}

int cpic_advance_e(cpic_context* c, double px, double py, double pz, double dt_eps0) {
  return guarded([&]() -> int {
    if (int s = check_ctx(c)) return s;
    return with_real(c, [&](auto r) {
      using R = decltype(r);
      return stage_advance_e<R>(c, (R)px, (R)py, (R)pz, (R)dt_eps0);
    });
  });
}

int cpic_set_field_solver(cpic_context* c, int32_t solver) {
  return guarded([&]() -> int {
    if (int s = check_ctx(c)) return s;
    if (solver != CPIC_FIELD_SOLVER_EM && solver != CPIC_FIELD_SOLVER_ES_1D)
      return fail(CPIC_ERR_INVALID, "unknown field solver %d", solver);
    // example/example.cpp:69-74
    if (solver == CPIC_FIELD_SOLVER_ES_1D && (c->ny > 1 || c->nz > 1))
      return fail(CPIC_ERR_UNSUPPORTED, "the ES field solver supports 1D (ny = nz = 1) only");
    c->solver = solver;
    return CPIC_OK;
  });
}

int cpic_set_compensated_deposit(cpic_context* c, int32_t enabled) {
  return guarded([&]() -> int {
    if (int s = check_ctx(c)) return s;
    if (enabled != 0 && enabled != 1) return fail(CPIC_ERR_INVALID, "enabled must be 0 or 1");
    c->compensated = enabled == 1;
    return CPIC_OK;
  });
}

int cpic_update_ghosts_j(cpic_context* c) {
  return guarded([&]() -> int {
    if (int s = check_ctx(c)) return s;
    return with_real(c, [&](auto r) { return stage_ghosts_j<decltype(r)>(c); });
  });
}

int cpic_energies(cpic_context* c, double* e, double* b) {
  return guarded([&]() -> int {
    if (int s = check_ctx(c)) return s;
    int s = with_real(c, [&](auto r) { return launch_field_energy<decltype(r)>(c, c->red_out); });
    if (s) return s;
    double h[2];
    HIP_TRY(hipMemcpyAsync(h, c->red_out, sizeof h, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (e) *e = h[0];
    if (b) *b = h[1];
    return collect_timing(c);
  });
}

int cpic_kinetic_energy(cpic_context* c, double me, double cc, double* ke) {
  return guarded([&]() -> int {
    if (int s = check_ctx(c)) return s;
    int s = with_real(c, [&](auto r) {
      return launch_kinetic<decltype(r)>(c, me * cc * cc, c->red_out + 2);
    });
    if (s) return s;
    double h = 0;
    HIP_TRY(hipMemcpyAsync(&h, c->red_out + 2, sizeof h, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (ke) *ke = h;
    return collect_timing(c);
  });
}

int cpic_gauss_residual(cpic_context* c, double dx, double dy, double dz, double eps0,
                        double qsp, int32_t mode, double* out, double* residual) {
  return guarded([&]() -> int {
    if (int s = check_ctx(c)) return s;
    if (mode != CPIC_GAUSS_BASELINE && mode != CPIC_GAUSS_CHANGE)
      return fail(CPIC_ERR_INVALID, "Gauss check mode %d", (int)mode);
    if (!(dx > 0 && dy > 0 && dz > 0 && eps0 > 0))
      return fail(CPIC_ERR_INVALID, "cell sizes and eps0 must be positive");
    if (mode == CPIC_GAUSS_CHANGE && !c->gauss_have_base)
      return fail(CPIC_ERR_INVALID, "no Gauss baseline: call with CPIC_GAUSS_BASELINE first");
    const size_t nc = (size_t)c->nc;
    if (!c->gauss_q) {
      if (int s = alloc((void**)&c->gauss_q, sizeof(double) * nc)) return s;
      if (int s = alloc((void**)&c->gauss_base, sizeof(double) * nc)) return s;
      if (int s = alloc((void**)&c->gauss_out, sizeof(unsigned long long) * 2)) return s;
    }
    int s = with_real(c, [&](auto r) {
      using R = decltype(r);
      if (int t = fill_zero<double>(c, c->gauss_q, nc)) return t;
      if (int t = fill_zero<unsigned long long>(c, c->gauss_out, 2)) return t;
      const int blocks = std::max(1, std::min(cdiv((long)nc, kBlock), 8192));
      if (c->D > 0 && c->np > 0) {
        const ColState<R> cs = col_state<R>(c, c->cur);
        k_node_charge_columns<R><<<blocks, kBlock, 0, c->stream>>>(
            cs, (int)c->nx, (int)c->ny, (int)c->nz, qsp, c->gauss_q);
        CHECK_LAUNCH();
      }
      // the residual overwrites the node charge in place (same index)
      k_gauss_residual<R><<<blocks, kBlock, 0, c->stream>>>(
          (const R*)c->fields, c->grid, c->gauss_q, 1.0 / dx, 1.0 / dy, 1.0 / dz,
          1.0 / (dx * dy * dz * eps0), c->gauss_q, c->gauss_base,
          mode == CPIC_GAUSS_BASELINE ? 0 : 1, c->gauss_out);
      CHECK_LAUNCH();
      return CPIC_OK;
    });
    if (s) return s;
    unsigned long long h[2];
    HIP_TRY(hipMemcpyAsync(h, c->gauss_out, sizeof h, hipMemcpyDeviceToHost, c->stream));
    if (residual)
      HIP_TRY(hipMemcpyAsync(residual, c->gauss_q, sizeof(double) * nc, hipMemcpyDeviceToHost,
                             c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (mode == CPIC_GAUSS_BASELINE) c->gauss_have_base = true;
    if (out) {
      std::memcpy(&out[0], &h[0], sizeof(double));
      std::memcpy(&out[1], &h[1], sizeof(double));
    }
    return collect_timing(c);
  });
}

int cpic_sort_particles(cpic_context* c) {
  return guarded([&]() -> int {
    // The particles are binned by cell at upload and stay binned: every push
    // writes each particle into the column of the cell it ends in.
    return check_ctx(c);
  });
}

int cpic_step_async(cpic_context* c, const cpic_step_params* sp, int64_t nsteps) {
  return guarded([&]() -> int {
    if (int s = check_ctx(c)) return s;
    if (!sp) return fail(CPIC_ERR_INVALID, "null step parameters");
    if (nsteps < 0) return fail(CPIC_ERR_INVALID, "negative step count");
    if (int s = check_periodic(sp->boundary)) return s;
    return with_real(c, [&](auto r) { return run_steps<decltype(r)>(c, sp, nsteps, nullptr); });
  });
}

int cpic_snapshot_begin(cpic_context* c, int32_t slot) {
  return guarded([&] { return snapshot_begin(c, slot); });
}

int cpic_snapshot_wait(cpic_context* c, int32_t slot, cpic_snapshot* out) {
  return guarded([&] { return snapshot_wait(c, slot, out); });
}

int cpic_step(cpic_context* c, const cpic_step_params* sp, int64_t nsteps, double* hist) {
  return guarded([&]() -> int {
    if (int s = check_ctx(c)) return s;
    if (!sp) return fail(CPIC_ERR_INVALID, "null step parameters");
    if (nsteps < 0) return fail(CPIC_ERR_INVALID, "negative step count");
    if (int s = check_periodic(sp->boundary)) return s;
    double* dhist = nullptr;
    const int ei = sp->energy_interval;
    const int ki = sp->kinetic_interval;
    if (hist && (ei > 0 || ki > 0) && nsteps > 0) {
      if (c->hist_cap < nsteps) {
        HIP_TRY(hipStreamSynchronize(c->stream));
        if (c->hist_dev) HIP_TRY(hipFree(c->hist_dev));
        c->hist_dev = nullptr;
        c->hist_cap = 0;
        HIP_TRY(hipMalloc(&c->hist_dev, sizeof(double) * 3 * nsteps));
        c->hist_cap = nsteps;
      }
      dhist = c->hist_dev;
    }
    int s = with_real(c, [&](auto r) { return run_steps<decltype(r)>(c, sp, nsteps, dhist); });
    if (s == CPIC_OK && dhist) {
      std::vector<double> h(3 * nsteps);
      hipError_t e = hipMemcpyAsync(h.data(), dhist, sizeof(double) * 3 * nsteps,
                                    hipMemcpyDeviceToHost, c->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
      if (e != hipSuccess) {
        s = fail(CPIC_ERR_HIP, "history download failed: %s", hipGetErrorString(e));
      } else {
        for (int64_t step = 1; step <= nsteps; ++step) {
          if (ei > 0 && step % ei == 0)
            for (int k = 0; k < 2; ++k) hist[3 * (step - 1) + k] = h[3 * (step - 1) + k];
          if (ki > 0 && step % ki == 0) hist[3 * (step - 1) + 2] = h[3 * (step - 1) + 2];
        }
      }
    }
    if (s) return s;
    if ((s = check_device_errors(c))) return s;
    return collect_timing(c);
  });
}

int cpic_set_timing(cpic_context* c, int32_t enabled) {
  return guarded([&]() -> int {
    return cpic_set_timing_stages(c, enabled ? (1u << CPIC_NUM_STAGES) - 1 : 0u);
  });
}

int cpic_set_timing_stages(cpic_context* c, uint32_t stages) {
  return guarded([&]() -> int {
    if (int s = check_ctx(c)) return s;
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (int s = collect_timing(c)) return s;
    c->timing = stages & ((1u << CPIC_NUM_STAGES) - 1);
    for (int k = 0; k < CPIC_NUM_STAGES; ++k) {
      c->stage_count[k] = 0;
      c->stage_ms[k] = 0;
    }
    return CPIC_OK;
  });
}

int cpic_stage_times(cpic_context* c, int64_t* counts, double* ms) {
  return guarded([&]() -> int {
    if (int s = check_ctx(c)) return s;
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (int s = collect_timing(c)) return s;
    for (int k = 0; k < CPIC_NUM_STAGES; ++k) {
      if (counts) counts[k] = c->stage_count[k];
      if (ms) ms[k] = c->stage_ms[k];
    }
    return CPIC_OK;
  });
}

int cpic_last_push_movers(cpic_context* c, int64_t* movers) {
  return guarded([&]() -> int {
    if (int s = check_ctx(c)) return s;
    unsigned int m = 0;
    HIP_TRY(hipMemcpyAsync(&m, c->movers, sizeof m, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (movers) *movers = m;
    return CPIC_OK;
  });
}

int cpic_device_errors(cpic_context* c, int32_t* flags) {
  return guarded([&]() -> int {
    if (int s = check_ctx(c)) return s;
    int f = 0;
    HIP_TRY(hipMemcpyAsync(&f, c->err, sizeof f, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (flags) *flags = f;
    return CPIC_OK;
  });
}

}  // extern "C"

#include "cpic_tangent.inc"
#include "cpic_adjoint.inc"
