// sis_noalloc.hip -- per-agent MaxCut cuts without any device allocation.
//
// sis_maxcut_cuts_noalloc runs the kernels of sis_solver_maxcut_cuts
// (k_cut_rows, k_cut_reduce, k_best) on caller-owned device arrays.  The one
// buffer beyond the inputs and outputs, the per-block partial sums, is the
// file-scope device array `scratch` below: every GPU the library runs on holds
// its own copy of it until the process ends.  Its device address is resolved
// once per device with hipGetSymbolAddress (in host code the array's name is
// not a device address) and only that address is passed to the kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <mutex>
#include <string>

#include "sis_hip_util.hpp"
#include "sis_internal.hpp"
#include "sis_kernels.hpp"

using namespace sis;

namespace {

constexpr size_t kScratchDoubles = SIS_NOALLOC_SCRATCH_BYTES / sizeof(double);
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
std::array<double *, kMaxDevices> g_scratch{};  // per device, resolved on first use

int scratch_for(int dev, double **out) {
  if (dev < 0 || dev >= kMaxDevices) return fail(SIS_ERR_ARG, "device index out of range");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch[(size_t)dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(SIS_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[(size_t)dev] = static_cast<double *>(p);
  }
  *out = g_scratch[(size_t)dev];
  return SIS_OK;
}

}  // namespace

extern "C" int sis_maxcut_cuts_noalloc(const int64_t *rp, const int32_t *col, const double *val,
                                       int64_t n, const uint64_t *spins, int64_t agents, double S,
                                       double *cut, int64_t *best_agent, void *stream) {
  if (!rp || !col || !val || !spins || !cut)
    return fail(SIS_ERR_ARG, "sis_maxcut_cuts_noalloc: null argument");
  if (n < 1 || agents < 1 || agents > (1 << 18))
    return fail(SIS_ERR_ARG, "sis_maxcut_cuts_noalloc: need n >= 1 and 1 <= agents <= 262144");
  // the block partition of sis_solver_maxcut_cuts, so the sums are identical
  const int64_t W = (agents + 63) / 64, lda = 64 * W;
  const int64_t rows = std::max<int64_t>(4, std::min<int64_t>(256, (n + 1023) / 1024));
  const int64_t nblk = (n + rows - 1) / rows;
  if ((size_t)nblk * (size_t)lda > kScratchDoubles)
    return fail(SIS_ERR_ARG, "sis_maxcut_cuts_noalloc: " + std::to_string(nblk) + " blocks x " +
                                 std::to_string(lda) + " agents of partial sums exceed the " +
                                 std::to_string(SIS_NOALLOC_SCRATCH_BYTES) + "-byte scratch");
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  double *part = nullptr;
  if (int st = scratch_for(dev, &part)) return st;
  hipStream_t st = static_cast<hipStream_t>(stream);
  launch_cut_rows(st, rp, col, val, nullptr, spins, n, W, lda, nblk, rows, part);
  launch_cut_reduce(st, part, nblk, agents, lda, 0, S, cut);
  if (best_agent) launch_best(st, cut, agents, 0, best_agent);
  HIPCHK(hipGetLastError());
  return SIS_OK;
}
