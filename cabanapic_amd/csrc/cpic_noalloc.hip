// cpic_noalloc.hip -- the field energies of caller-owned device fields with
// no device allocation (cpic_field_energies_noalloc).
//
// The reduction's block partials, which cpic_energies keeps in its context,
// live here in a file-scope device array.  Every GPU holds its own copy of the
// array until the process ends; its device address is looked up once per
// device and kept.  Calls that overlap on one GPU share the array and race.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "../../include/cabanapic_amd.h"
#include "cpic_kernels.h"

using namespace cpic;

namespace cpic_detail {
int fail_msg(int code, const char* msg);  // sets cpic_last_error (cpic_api.hip)
}
using cpic_detail::fail_msg;

namespace {

// At most this many blocks write partials: the grid-stride kernel covers any
// grid with them, so the array does not grow with the grid.
constexpr int kScratchBlocks = 4096;
constexpr int kScratchDoubles = 2 * kScratchBlocks;  // [block][E, B]: 64 KiB

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
double* g_scratch_ptr[kMaxDevices] = {};  // device address of `scratch` per device

}  // namespace

static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {

// The device address of `scratch` on the current device (looked up on the
// first call there).
hipError_t scratch_on_current_device(double** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_mutex);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[dev] = (double*)p;
  }
  *out = g_scratch_ptr[dev];
  return hipSuccess;
}

template <class R>
hipError_t launch(const void* fields, const Grid& g, double* part, double* out, hipStream_t st) {
  const long n = (long)g.nx * g.ny * g.nz;
  const int blocks = (int)std::max(1L, std::min((n + kBlock - 1) / kBlock, (long)kScratchBlocks));
  k_field_energy_partial<R><<<blocks, kBlock, 0, st>>>((const R*)fields, g, part);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  k_reduce_final<2><<<1, kBlock, 0, st>>>(part, blocks, 0.5, 0.5, out);
  return hipGetLastError();
}

}  // namespace

extern "C" int cpic_field_energies_noalloc(const void* fields, int32_t real_size, int64_t nx,
                                           int64_t ny, int64_t nz, double* out, void* stream) {
  if (!fields || !out) return fail_msg(CPIC_ERR_INVALID, "null fields or output pointer");
  if (real_size != 4 && real_size != 8) return fail_msg(CPIC_ERR_INVALID, "real_size must be 4 or 8");
  if (nx < 1 || ny < 1 || nz < 1) return fail_msg(CPIC_ERR_INVALID, "grid extents must be >= 1");
  // the voxel indexing of the kernels is 32-bit, as for a context; the
  // partials do not grow with the grid (at most kScratchBlocks blocks)
  const double ng = (double)(nx + 2) * (double)(ny + 2) * (double)(nz + 2);
  if (ng * 9 > 2147483647.0)
    return fail_msg(CPIC_ERR_INVALID, "grid too large for 32-bit voxel indexing");
  double* part = nullptr;
  if (scratch_on_current_device(&part) != hipSuccess)
    return fail_msg(CPIC_ERR_HIP, "device address of the reduction scratch not found");
  Grid g;
  g.nx = (int)nx;
  g.ny = (int)ny;
  g.nz = (int)nz;
  g.sy = (int)(nx + 2);
  g.sz = (int)((nx + 2) * (ny + 2));
  g.ncells = (int)ng;
  const hipStream_t st = (hipStream_t)stream;
  const hipError_t e = real_size == 4 ? launch<float>(fields, g, part, out, st)
                                      : launch<double>(fields, g, part, out, st);
  return e == hipSuccess ? CPIC_OK : fail_msg(CPIC_ERR_HIP, hipGetErrorString(e));
}
