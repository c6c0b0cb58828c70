// cpic_api.hip -- the C-ABI of the MI355X-native CabanaPIC hot path.
//
// A context owns the device arrays of one simulation and one HIP stream.
// Every entry point validates its arguments on the host before a kernel is
// launched, so no kernel is ever given a grid or an index range that does
// not match its buffers.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/cabanapic_amd.h"
#include "cpic_kernels.h"

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

}  // namespace

struct cpic_context {
  int real_size = 4;
  long nx = 0, ny = 0, nz = 0;
  long ncells = 0;  // Ng
  long cap = 0;     // particle capacity
  long np = 0;      // particles resident
  Grid grid{};

  // device arrays
  void* part[7] = {};     // x,y,z,ux,uy,uz,w
  int* cell = nullptr;
  void* alt[7] = {};      // sort destination (allocated on first sort)
  int* alt_cell = nullptr;
  void* fields = nullptr;
  void* interp = nullptr;
  void* acc = nullptr;
  double* partials = nullptr;  // reduction partials
  double* red_out = nullptr;   // [E, B, KE]
  int* err = nullptr;
  unsigned int* movers = nullptr;
  unsigned int* sort_counts = nullptr;
  unsigned int* sort_sums = nullptr;
  double* hist_dev = nullptr;  // per-step energy rows of cpic_step
  long hist_cap = 0;           // rows allocated in hist_dev
  int n_partials = 0;

  hipStream_t stream = nullptr;
  int device = 0;

  // stage timing
  bool timing = false;
  struct Pending {
    int stage;
    hipEvent_t a, b;
  };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> free_events;
  long stage_count[CPIC_NUM_STAGES] = {};
  double stage_ms[CPIC_NUM_STAGES] = {};

  size_t rsz() const { return (size_t)real_size; }
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
  HIP_TRY(hipEventCreate(e));
  return CPIC_OK;
}

// Brackets the launches of one stage with events when timing is on.
struct StageTimer {
  cpic_context* c;
  int stage;
  hipEvent_t a = nullptr, b = nullptr;
  int status = CPIC_OK;
  StageTimer(cpic_context* ctx, int s) : c(ctx), stage(s) {
    if (!c->timing) return;
    if ((status = get_event(c, &a)) != CPIC_OK) return;
    if ((status = get_event(c, &b)) != CPIC_OK) return;
    if (hipEventRecord(a, c->stream) != hipSuccess)
      status = fail(CPIC_ERR_HIP, "hipEventRecord failed");
  }
  int finish() {
    if (!c->timing || status != CPIC_OK) return status;
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
  return t.finish();
}

template <class R>
int stage_push(cpic_context* c, R qdt_2mc, R cdt_dx, R cdt_dy, R cdt_dz, R qsp) {
  StageTimer t(c, CPIC_STAGE_PUSH);
  if (int z = fill_zero<unsigned int>(c, c->movers, 1)) return z;
  if (c->np > 0) {
    PushArgs<R> a;
    a.x = (R*)c->part[0];
    a.y = (R*)c->part[1];
    a.z = (R*)c->part[2];
    a.ux = (R*)c->part[3];
    a.uy = (R*)c->part[4];
    a.uz = (R*)c->part[5];
    a.w = (const R*)c->part[6];
    a.cell = c->cell;
    a.interp = (const R*)c->interp;
    a.acc = (R*)c->acc;
    a.np = (int)c->np;
    a.g = c->grid;
    a.qdt_2mc = qdt_2mc;
    a.cdt_dx = cdt_dx;
    a.cdt_dy = cdt_dy;
    a.cdt_dz = cdt_dz;
    a.qsp = qsp;
    a.err = c->err;
    a.movers = c->movers;
    k_push<R><<<cdiv(c->np, kBlock), kBlock, 0, c->stream>>>(a);
    CHECK_LAUNCH();
  }
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
  const long n = c->nx * c->ny * c->nz;
  k_advance_b<R><<<cdiv(n, kBlock), kBlock, 0, c->stream>>>((R*)c->fields, c->grid, px, py, pz);
  CHECK_LAUNCH();
  int s = launch_copy_ghosts<R>(c, 3);
  if (s) return s;
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
  int s = stage_ghosts_j<R>(c);
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
  const long n = c->nx * c->ny * c->nz;
  const int blocks = std::max(1, std::min(cdiv(n, kBlock), c->n_partials));
  k_field_energy_partial<R><<<blocks, kBlock, 0, c->stream>>>((const R*)c->fields, c->grid,
                                                              c->partials);
  CHECK_LAUNCH();
  k_reduce_final<2><<<1, kBlock, 0, c->stream>>>(c->partials, blocks, 0.5, 0.5, out);
  CHECK_LAUNCH();
  return CPIC_OK;
}

template <class R>
int launch_kinetic(cpic_context* c, double scale, double* out) {
  const int blocks = std::max(1, std::min(cdiv(std::max(c->np, 1L), kBlock), c->n_partials));
  k_kinetic_partial<R><<<blocks, kBlock, 0, c->stream>>>(
      (const R*)c->part[3], (const R*)c->part[4], (const R*)c->part[5], (const R*)c->part[6],
      (int)c->np, c->partials + 2 * (size_t)c->n_partials);
  CHECK_LAUNCH();
  k_reduce_final<1><<<1, kBlock, 0, c->stream>>>(c->partials + 2 * (size_t)c->n_partials, blocks,
                                                 scale, 0.0, out);
  CHECK_LAUNCH();
  return CPIC_OK;
}

template <class R>
int stage_sort(cpic_context* c) {
  StageTimer t(c, CPIC_STAGE_SORT);
  if (c->np == 0) return t.finish();
  if (!c->alt_cell) {
    for (int k = 0; k < 7; ++k) {
      int s = alloc(&c->alt[k], c->rsz() * c->cap);
      if (s) return s;
    }
    int s = alloc((void**)&c->alt_cell, sizeof(int) * c->cap);
    if (s) return s;
  }
  const long ng = c->ncells;
  const int nsum = cdiv(ng, kBlock * 4);
  int s = fill_zero<unsigned int>(c, c->sort_counts, (size_t)ng);
  if (s) return s;
  k_sort_histogram<<<cdiv(c->np, kBlock), kBlock, 0, c->stream>>>(c->cell, (int)c->np,
                                                                   c->sort_counts, (int)ng, c->err);
  CHECK_LAUNCH();
  // in-place exclusive scan of the counts: counts -> offsets
  k_scan_blocks<<<nsum, kBlock, 0, c->stream>>>(c->sort_counts, c->sort_counts, (int)ng,
                                                c->sort_sums);
  CHECK_LAUNCH();
  k_scan_sums<<<1, kBlock, 0, c->stream>>>(c->sort_sums, nsum);
  CHECK_LAUNCH();
  k_scan_add<<<cdiv(ng, kBlock), kBlock, 0, c->stream>>>(c->sort_counts, (int)ng, c->sort_sums);
  CHECK_LAUNCH();
  SortArgs<R> a;
  for (int k = 0; k < 7; ++k) {
    a.src[k] = (const R*)c->part[k];
    a.dst[k] = (R*)c->alt[k];
  }
  a.src_cell = c->cell;
  a.dst_cell = c->alt_cell;
  a.cursor = c->sort_counts;
  a.np = (int)c->np;
  a.ncells = (int)ng;
  k_sort_scatter<R><<<cdiv(c->np, kBlock), kBlock, 0, c->stream>>>(a);
  CHECK_LAUNCH();
  for (int k = 0; k < 7; ++k) std::swap(c->part[k], c->alt[k]);
  std::swap(c->cell, c->alt_cell);
  return t.finish();
}

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

}  // namespace

// ===========================================================================

extern "C" {

const char* cpic_last_error(void) { return g_last_error.c_str(); }

int cpic_create(const cpic_config* cfg, cpic_context** out) {
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

  cpic_context* c = new (std::nothrow) cpic_context();
  if (!c) return fail(CPIC_ERR_INVALID, "out of host memory");
  c->real_size = cfg->real_size;
  c->nx = cfg->nx;
  c->ny = cfg->ny;
  c->nz = cfg->nz;
  c->ncells = X * Y * Z;
  c->cap = cfg->max_particles;
  c->grid.nx = (int)c->nx;
  c->grid.ny = (int)c->ny;
  c->grid.nz = (int)c->nz;
  c->grid.sy = (int)X;
  c->grid.sz = (int)(X * Y);
  c->grid.ncells = (int)c->ncells;
  c->n_partials = 4096;

  auto bail = [&](int s) {
    cpic_destroy(c);
    return s;
  };
  int s;
  if (hipGetDevice(&c->device) != hipSuccess) return bail(fail(CPIC_ERR_HIP, "no HIP device"));
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(CPIC_ERR_HIP, "hipStreamCreate failed"));
  for (int k = 0; k < 7; ++k)
    if ((s = alloc(&c->part[k], c->rsz() * c->cap))) return bail(s);
  if ((s = alloc((void**)&c->cell, sizeof(int) * c->cap))) return bail(s);
  if ((s = alloc(&c->fields, c->rsz() * 9 * c->ncells))) return bail(s);
  if ((s = alloc(&c->interp, c->rsz() * kInterpStride * c->ncells))) return bail(s);
  if ((s = alloc(&c->acc, c->rsz() * 12 * c->ncells))) return bail(s);
  if ((s = alloc((void**)&c->partials, sizeof(double) * 3 * c->n_partials))) return bail(s);
  if ((s = alloc((void**)&c->red_out, sizeof(double) * 4))) return bail(s);
  if ((s = alloc((void**)&c->err, sizeof(int)))) return bail(s);
  if ((s = alloc((void**)&c->movers, sizeof(unsigned int)))) return bail(s);
  if ((s = alloc((void**)&c->sort_counts, sizeof(unsigned int) * c->ncells))) return bail(s);
  if ((s = alloc((void**)&c->sort_sums, sizeof(unsigned int) * (cdiv(c->ncells, kBlock * 4) + 1))))
    return bail(s);
  s = with_real(c, [&](auto r) {
    using R = decltype(r);
    int t;
    if ((t = fill_zero<R>(c, c->fields, 9 * (size_t)c->ncells))) return t;
    if ((t = fill_zero<R>(c, c->interp, kInterpStride * (size_t)c->ncells))) return t;
    if ((t = fill_zero<R>(c, c->acc, 12 * (size_t)c->ncells))) return t;
    return CPIC_OK;
  });
  if (s) return bail(s);
  if (hipMemsetAsync(c->err, 0, sizeof(int), c->stream) != hipSuccess ||
      hipMemsetAsync(c->movers, 0, sizeof(unsigned int), c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess)
    return bail(fail(CPIC_ERR_HIP, "context initialisation failed"));
  *out = c;
  return CPIC_OK;
}

int cpic_destroy(cpic_context* c) {
  if (!c) return CPIC_OK;
  // teardown is best effort: errors cannot be acted upon here
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto& p : c->pending) {
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  for (auto e : c->free_events) (void)hipEventDestroy(e);
  for (int k = 0; k < 7; ++k) {
    if (c->part[k]) (void)hipFree(c->part[k]);
    if (c->alt[k]) (void)hipFree(c->alt[k]);
  }
  void* ptrs[] = {c->cell,   c->alt_cell, c->fields,      c->interp,    c->acc,     c->partials,
                  c->red_out, c->err,      c->movers,      c->sort_counts, c->sort_sums, c->hist_dev};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return CPIC_OK;
}

int cpic_synchronize(cpic_context* c) {
  if (int s = check_ctx(c)) return s;
  HIP_TRY(hipStreamSynchronize(c->stream));
  return collect_timing(c);
}

int cpic_get_stream(cpic_context* c, void** stream) {
  if (int s = check_ctx(c)) return s;
  if (!stream) return fail(CPIC_ERR_INVALID, "null stream pointer");
  *stream = (void*)c->stream;
  return CPIC_OK;
}

int cpic_upload_particles(cpic_context* c, int64_t n, const void* x, const void* y, const void* z,
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
  for (int64_t p = 0; p < n; ++p) {
    const long v = cell[p];
    const long ix = v % X, iy = (v / X) % Y, iz = v / (X * Y);
    if (v < 0 || v >= c->ncells || ix < 1 || ix > c->nx || iy < 1 || iy > c->ny || iz < 1 ||
        iz > c->nz)
      return fail(CPIC_ERR_INVALID, "particle %lld has non-interior cell %ld", (long long)p, v);
  }
  HIP_TRY(hipStreamSynchronize(c->stream));
  for (int k = 0; k < 7; ++k)
    if (n > 0) HIP_TRY(hipMemcpy(c->part[k], src[k], c->rsz() * n, hipMemcpyHostToDevice));
  if (n > 0) HIP_TRY(hipMemcpy(c->cell, cell, sizeof(int) * n, hipMemcpyHostToDevice));
  c->np = n;
  return CPIC_OK;
}

int cpic_download_particles(cpic_context* c, int64_t n, void* x, void* y, void* z, void* ux,
                            void* uy, void* uz, void* w, int32_t* cell) {
  if (int s = check_ctx(c)) return s;
  if (n != c->np)
    return fail(CPIC_ERR_INVALID, "download of %lld particles, %ld resident", (long long)n, c->np);
  void* dst[7] = {x, y, z, ux, uy, uz, w};
  HIP_TRY(hipStreamSynchronize(c->stream));
  for (int k = 0; k < 7; ++k)
    if (dst[k] && n > 0) HIP_TRY(hipMemcpy(dst[k], c->part[k], c->rsz() * n, hipMemcpyDeviceToHost));
  if (cell && n > 0) HIP_TRY(hipMemcpy(cell, c->cell, sizeof(int) * n, hipMemcpyDeviceToHost));
  return CPIC_OK;
}

int cpic_num_particles(cpic_context* c, int64_t* n) {
  if (int s = check_ctx(c)) return s;
  if (!n) return fail(CPIC_ERR_INVALID, "null pointer");
  *n = c->np;
  return CPIC_OK;
}

int cpic_upload_fields(cpic_context* c, const void* f) {
  if (int s = check_ctx(c)) return s;
  if (!f) return fail(CPIC_ERR_INVALID, "null fields");
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipMemcpy(c->fields, f, c->rsz() * 9 * c->ncells, hipMemcpyHostToDevice));
  return CPIC_OK;
}

int cpic_download_fields(cpic_context* c, void* f) {
  if (int s = check_ctx(c)) return s;
  if (!f) return fail(CPIC_ERR_INVALID, "null fields");
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipMemcpy(f, c->fields, c->rsz() * 9 * c->ncells, hipMemcpyDeviceToHost));
  return CPIC_OK;
}

int cpic_upload_interpolators(cpic_context* c, const void* in) {
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
}

int cpic_download_interpolators(cpic_context* c, void* out) {
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
}

int cpic_upload_accumulators(cpic_context* c, const void* in) {
  if (int s = check_ctx(c)) return s;
  if (!in) return fail(CPIC_ERR_INVALID, "null accumulators");
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipMemcpy(c->acc, in, c->rsz() * 12 * c->ncells, hipMemcpyHostToDevice));
  return CPIC_OK;
}

int cpic_download_accumulators(cpic_context* c, void* out) {
  if (int s = check_ctx(c)) return s;
  if (!out) return fail(CPIC_ERR_INVALID, "null accumulators");
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipMemcpy(out, c->acc, c->rsz() * 12 * c->ncells, hipMemcpyDeviceToHost));
  return CPIC_OK;
}

int cpic_load_interpolator_array(cpic_context* c) {
  if (int s = check_ctx(c)) return s;
  return with_real(c, [&](auto r) { return stage_load_interpolator<decltype(r)>(c); });
}

int cpic_clear_accumulator_array(cpic_context* c) {
  if (int s = check_ctx(c)) return s;
  return with_real(c, [&](auto r) { return stage_clear<decltype(r)>(c); });
}

int cpic_push(cpic_context* c, double qdt_2mc, double cdt_dx, double cdt_dy, double cdt_dz,
              double qsp, int32_t boundary) {
  if (int s = check_ctx(c)) return s;
  if (int s = check_periodic(boundary)) return s;
  return with_real(c, [&](auto r) {
    using R = decltype(r);
    return stage_push<R>(c, (R)qdt_2mc, (R)cdt_dx, (R)cdt_dy, (R)cdt_dz, (R)qsp);
  });
}

int cpic_unload_accumulator_array(cpic_context* c, double dx, double dy, double dz, double dt) {
  if (int s = check_ctx(c)) return s;
  return with_real(c, [&](auto r) {
    using R = decltype(r);
    return stage_unload<R>(c, (R)dx, (R)dy, (R)dz, (R)dt);
  });
}

int cpic_advance_b(cpic_context* c, double px, double py, double pz) {
  if (int s = check_ctx(c)) return s;
  return with_real(c, [&](auto r) {
    using R = decltype(r);
    return stage_advance_b<R>(c, (R)px, (R)py, (R)pz);
  });
}

int cpic_advance_e(cpic_context* c, double px, double py, double pz, double dt_eps0) {
  if (int s = check_ctx(c)) return s;
  return with_real(c, [&](auto r) {
    using R = decltype(r);
    return stage_advance_e<R>(c, (R)px, (R)py, (R)pz, (R)dt_eps0);
  });
}

int cpic_update_ghosts_j(cpic_context* c) {
  if (int s = check_ctx(c)) return s;
  return with_real(c, [&](auto r) { return stage_ghosts_j<decltype(r)>(c); });
}

int cpic_energies(cpic_context* c, double* e, double* b) {
  if (int s = check_ctx(c)) return s;
  int s = with_real(c, [&](auto r) { return launch_field_energy<decltype(r)>(c, c->red_out); });
  if (s) return s;
  double h[2];
  HIP_TRY(hipMemcpyAsync(h, c->red_out, sizeof h, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (e) *e = h[0];
  if (b) *b = h[1];
  return collect_timing(c);
}

int cpic_kinetic_energy(cpic_context* c, double me, double cc, double* ke) {
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
}

int cpic_sort_particles(cpic_context* c) {
  if (int s = check_ctx(c)) return s;
  return with_real(c, [&](auto r) { return stage_sort<decltype(r)>(c); });
}

int cpic_step(cpic_context* c, const cpic_step_params* sp, int64_t nsteps, double* hist) {
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
  int s = with_real(c, [&](auto r) {
    using R = decltype(r);
    const R px = (R)sp->px, py = (R)sp->py, pz = (R)sp->pz;
    const R half = (R)0.5;
    const R hpx = half * px, hpy = half * py, hpz = half * pz;
    for (int64_t step = 1; step <= nsteps; ++step) {
      int t;
      if (sp->sort_interval > 0 && (step - 1) % sp->sort_interval == 0)
        if ((t = stage_sort<R>(c))) return t;
      if ((t = stage_load_interpolator<R>(c))) return t;
      if ((t = stage_clear<R>(c))) return t;
      if ((t = stage_push<R>(c, (R)sp->qdt_2mc, (R)sp->cdt_dx, (R)sp->cdt_dy, (R)sp->cdt_dz,
                             (R)sp->qsp)))
        return t;
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
  });
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
}

int cpic_set_timing(cpic_context* c, int32_t enabled) {
  if (int s = check_ctx(c)) return s;
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (int s = collect_timing(c)) return s;
  c->timing = enabled != 0;
  for (int k = 0; k < CPIC_NUM_STAGES; ++k) {
    c->stage_count[k] = 0;
    c->stage_ms[k] = 0;
  }
  return CPIC_OK;
}

int cpic_stage_times(cpic_context* c, int64_t* counts, double* ms) {
  if (int s = check_ctx(c)) return s;
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (int s = collect_timing(c)) return s;
  for (int k = 0; k < CPIC_NUM_STAGES; ++k) {
    if (counts) counts[k] = c->stage_count[k];
    if (ms) ms[k] = c->stage_ms[k];
  }
  return CPIC_OK;
}

int cpic_last_push_movers(cpic_context* c, int64_t* movers) {
  if (int s = check_ctx(c)) return s;
  unsigned int m = 0;
  HIP_TRY(hipMemcpyAsync(&m, c->movers, sizeof m, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (movers) *movers = m;
  return CPIC_OK;
}

int cpic_device_errors(cpic_context* c, int32_t* flags) {
  if (int s = check_ctx(c)) return s;
  int f = 0;
  HIP_TRY(hipMemcpyAsync(&f, c->err, sizeof f, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (flags) *flags = f;
  return CPIC_OK;
}

}  // extern "C"
