// noalloc.hip -- an_bicgstab_solve_noalloc: the one-shot BiCGSTAB + Jacobi
// solve of an_bicgstab_solve for callers that cannot allocate device memory.
//
// Every buffer the solve needs beyond the caller's inputs and output (row
// blocks of the CSR-stream SpMV, the nine Krylov work vectors, the Jacobi
// diagonal, the reduction partials and the device scalars) is carved out of
// one file-scope device array, `scratch`, by an arena DevPool; the host
// reads the Krylov scalars through a plain host array.  No hipMalloc,
// hipFree or hipHostMalloc is called.  The kernels are the ones
// an_bicgstab_solve launches, so the result is the same bit for bit.
//
// Each GPU holds its own copy of `scratch` (kScratchBytes) from module load
// until the process ends.  Its device address is looked up once per device
// with hipGetSymbolAddress on the first call made with that device current,
// and only that address is handed to kernels and copies.  All calls on one
// GPU share the array: calls that overlap on one GPU (from several host
// threads, or a call still running on the device) race and are not
// supported.  The arena starts empty on every call and no buffer is read
// before it is written in the same call, so nothing needs clearing.
#include <mutex>

#include "../../include/anabasis_hip.h"
#include "kernels.hpp"
#include "solver.hpp"

constexpr size_t kScratchBytes = AN_NOALLOC_SCRATCH_BYTES;  // per GPU

// Host code sees `scratch` only as a stand-in symbol: its device address
// comes from hipGetSymbolAddress (scratch_address below).
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;

std::mutex g_scratchMutex;
void* g_scratchAddr[kMaxDevices] = {};

// Device address of `scratch` on the current device (looked up once).
void* scratch_address() {
  int dev = 0;
  AN_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) throw an::Error(AN_EINVAL, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratchMutex);
  if (!g_scratchAddr[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw an::Error(AN_EHIP, std::string("hipGetSymbolAddress(scratch) failed: ") +
                                   hipGetErrorString(e));
    g_scratchAddr[dev] = p;
  }
  return g_scratchAddr[dev];
}

// Bytes of scratch for n rows when every row block holds 64 rows.
constexpr size_t scratch_need_bound(long long n);

// Bytes of scratch the solve takes for n rows and nRowBlocks CSR row blocks,
// mirroring the buffers build_row_blocks, exclusive_scan and BicgWork
// request (each rounded up to 256 bytes by the arena).
constexpr size_t scratch_need(long long n, long long nRowBlocks) {
  auto r = [](long long b) { return (size_t)((b + 255) & ~255LL); };
  const long long nsl = (n + 63) / 64;
  const long long nbSpmv = (nRowBlocks + 3) / 4;  // 4 waves per SpMV block
  const long long redBlocks = nbSpmv > 2048 ? nbSpmv + 16 : 2048;
  size_t t = 2 * r(4 * (nsl + 1));          // row-block counts and offsets
  t += r(8 * ((nsl + 2047) / 2048 + 1));     // scan tile sums (2048 per tile)
  t += r(8 * (nRowBlocks + 1));               // row-block starts
  t += r(8 * 4 * redBlocks);                  // reduction partials
  t += r(8 * 40);                             // device scalars
  t += 9 * r(8 * n);                          // dinv, r, r0, p, v, ph, s, sh, t
  return t;
}
constexpr size_t scratch_need_bound(long long n) { return scratch_need(n, (n + 63) / 64); }

static_assert(scratch_need_bound(AN_NOALLOC_MAX_ROWS) <= kScratchBytes,
              "AN_NOALLOC_MAX_ROWS rows must fit the scratch array");
}  // namespace

using namespace an;

extern "C" an_status an_bicgstab_solve_noalloc(int n, const int* rowOffsets, const int* cols,
                                               int64_t nnz, const double* values,
                                               const double* b, double* x, int maxIter,
                                               double relTol, double absTol,
                                               an_solve_info* info) {
  AN_TRY
  if (n <= 0 || nnz < n || !rowOffsets || !cols || !values || !b || !x)
    throw Error(AN_EINVAL, "NULL argument or empty system");
  if (nnz > 0x7fffffffLL) throw Error(AN_EINVAL, "nnz exceeds int32 range");
  // checked before any device work; the row-block count (long rows make
  // more blocks) is checked again once it is known
  if (n > AN_NOALLOC_MAX_ROWS)
    throw Error(AN_EINVAL, "an_bicgstab_solve_noalloc: " + std::to_string(n) +
                               " rows exceed the maximum of " +
                               std::to_string(AN_NOALLOC_MAX_ROWS) + " (" +
                               std::to_string(kScratchBytes >> 20) + " MiB of scratch)");
  DevPool arena(scratch_address(), kScratchBytes);
  PoolScope scope(&arena);
  DevCSR<double> A;
  A.n = n;
  A.nnz = nnz;
  // borrowed caller arrays: owned by the arena pool, whose put() is a no-op
  A.rowOff.p = const_cast<int*>(rowOffsets);
  A.rowOff.n = (size_t)n + 1;
  A.rowOff.pool = &arena;
  A.cols.p = const_cast<int*>(cols);
  A.cols.n = (size_t)nnz;
  A.cols.pool = &arena;
  A.vals.p = const_cast<double*>(values);
  A.vals.n = (size_t)nnz;
  A.vals.pool = &arena;
  A.finish_pattern(0, false);  // no 16-bit column offsets: the scratch is sized by rows
  if (scratch_need(n, A.rb.count) > kScratchBytes)
    throw Error(AN_EINVAL, "an_bicgstab_solve_noalloc: the matrix's " +
                               std::to_string(A.rb.count) + " row blocks need more than the " +
                               std::to_string(kScratchBytes >> 20) + " MiB of scratch");
  BicgWork w(false);
  an_status st = bicgstab_device(A, b, x, maxIter, relTol, absTol, w, 0, info);
  if (st == AN_ENOCONV) set_last_error("BiCGSTAB did not converge");
  return st;
  AN_CATCH
}
