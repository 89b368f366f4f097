// noalloc.hip -- the factorisation entry point that allocates nothing.
//
// hamsx_factor_noalloc factors a caller-owned device matrix with the kernels
// of hamsx_factor (tournament-pivoted blocked LU, lu.hip).  The workspace
// those kernels need besides the matrix and the pivots -- the factored
// diagonal block, the row permutation and the tournament's candidate lists
// -- lives in `scratch`, a file-scope device array of fixed size that each
// GPU holds for the life of the process.  Its device address is looked up
// once per device (hipGetSymbolAddress, with that device current) and kept;
// host code never passes the array's own name to a kernel.  Every call on
// one GPU shares that GPU's array: calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <mutex>

#include "hamsx.h"
#include "hamsx_internal.hpp"

namespace {

constexpr int kMaxN = HAMSX_NOALLOC_MAX_N;
constexpr int kMaxSets = (kMaxN + hamsx::kSelRows - 1) / hamsx::kSelRows;
// layout (bytes): diagonal block (kPanelNB^2 complex128), permutation (kMaxN
// int), candidate lists (2 kMaxSets kPanelNB int), candidate counts (2 kMaxSets int)
constexpr size_t kDiagOff = 0;
constexpr size_t kPermOff = kDiagOff + sizeof(double2) * hamsx::kPanelNB * hamsx::kPanelNB;
constexpr size_t kIdxOff = kPermOff + sizeof(int) * (size_t)kMaxN;
constexpr size_t kCntOff = kIdxOff + sizeof(int) * 2 * (size_t)kMaxSets * hamsx::kPanelNB;
constexpr size_t kScratchBytes = kCntOff + sizeof(int) * 2 * (size_t)kMaxSets;
static_assert(kPermOff % 256 == 0 && kIdxOff % 256 == 0 && kCntOff % 256 == 0, "aligned sections");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
unsigned char* g_scratch[kMaxDevices] = {};  // device address of `scratch` per device

int scratch_on(int dev, unsigned char** out) {
  std::lock_guard<std::mutex> lock(g_mutex);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      char msg[160];
      std::snprintf(msg, sizeof msg, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
      return hamsx::set_error(HAMSX_EDEVICE, msg);
    }
    g_scratch[dev] = static_cast<unsigned char*>(p);
  }
  *out = g_scratch[dev];
  return 0;
}

}  // namespace

extern "C" int hamsx_factor_noalloc(int n, double* a, int* ipiv, int* info, void* stream) {
  if (n < 1 || n > kMaxN) {
    char msg[160];
    std::snprintf(msg, sizeof msg, "hamsx_factor_noalloc: n = %d outside [1, %d] (HAMSX_NOALLOC_MAX_N)", n, kMaxN);
    return hamsx::set_error(HAMSX_EARG, msg);
  }
  if (!a || !ipiv || !info) return hamsx::set_error(HAMSX_EARG, "hamsx_factor_noalloc: a, ipiv and info are required");
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hamsx::set_error(HAMSX_EDEVICE, hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices) return hamsx::set_error(HAMSX_EDEVICE, "hamsx_factor_noalloc: device index too large");
  unsigned char* base = nullptr;
  int rc = scratch_on(dev, &base);
  if (rc) return rc;
  double2* diag = reinterpret_cast<double2*>(base + kDiagOff);
  int* perm = reinterpret_cast<int*>(base + kPermOff);
  int* widx = reinterpret_cast<int*>(base + kIdxOff);
  int* wcnt = reinterpret_cast<int*>(base + kCntOff);
  // one stream, no look-ahead (the same factors bit for bit); the kernels set
  // info to zero and the permutation to the identity before they use them,
  // and every candidate list is written before it is read
  hamsx::launch::lu_factor(static_cast<hipStream_t>(stream), reinterpret_cast<double2*>(a), n, ipiv, perm, widx,
                           wcnt, info, diag, nullptr);
  e = hipGetLastError();
  if (e != hipSuccess) return hamsx::set_error(HAMSX_EDEVICE, hipGetErrorString(e));
  return 0;
}
