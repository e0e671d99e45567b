// an_internal.hpp -- shared internals of the anabasis MI355X library.
//
// Device data layout (HBM):
//   cells           double4 cgeo[nC] = (ccx, ccy, ccz, vol)
//   internal faces  int2 fPN[nIF] (owner, neighbour); SoA doubles
//                   fLam (interp. weight), fDpn (n.d), fA (area),
//                   fN{x,y,z} (unit normal), fD{x,y,z} (cc_N - cc_P)
//   boundary faces  int bP[nBF] (owner), int bPatch[nBF] (0-based),
//                   SoA doubles bDpn, bA, bN{x,y,z}, bD{x,y,z} (xf - cc_P)
//   addressing      CSR rowOff[nC+1], cols[nnz], diagPos[nC],
//                   facePP/PN/NP/NN[nIF];
//                   cell->internal-face incidence in face order
//                   cfOff[nC+1], cfEnt[2nIF] = face<<1 | side (0 owner,
//                   1 neighbour), cfSlot[2nIF] = nnzSlot<<1 | firstUse;
//                   cell->boundary-face incidence cbOff[nC+1], cbFace[nBF]
//                   (boundary-face index bf = f - nIF, in face order).
// Every derived per-face quantity is computed on the host with the
// reference's expressions (-ffp-contract=off), so kernels that read it see
// the reference's bits.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "host_mesh.hpp"

namespace an {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string& msg);

#define AN_HIP(stmt)                                                             \
  do {                                                                           \
    hipError_t _e = (stmt);                                                      \
    if (_e != hipSuccess)                                                        \
      throw ::an::Error(3, std::string("HIP error ") + hipGetErrorString(_e) +  \
                               " at " + __FILE__ + ":" + std::to_string(__LINE__)); \
  } while (0)

#define AN_CHECK_LAUNCH() AN_HIP(hipGetLastError())

// Entry-point wrappers of the C ABI: every C++ exception becomes an
// an_status plus the thread-local message of an_last_error().
#define AN_TRY try {
#define AN_CATCH                                                     \
  }                                                                  \
  catch (const an::Error& e) {                                       \
    ::an::set_last_error(e.what());                                        \
    return (an_status)e.code;                                        \
  }                                                                  \
  catch (const std::bad_alloc&) {                                    \
    ::an::set_last_error("out of host memory");                            \
    return AN_ENOMEM;                                                \
  }                                                                  \
  catch (const std::exception& e) {                                  \
    ::an::set_last_error(e.what());                                        \
    return AN_EINVAL;                                                \
  }                                                                  \
  catch (...) {                                                      \
    ::an::set_last_error("unknown error");                                 \
    return AN_EINTERNAL;                                             \
  }

// Cache of device blocks (sizes rounded up to powers of two) for code that
// allocates and frees many buffers, all used on one stream: a freed block
// goes back to the cache instead of hipFree (which waits for the device)
// and is handed out again in stream order.  A DevBuf allocated while a
// PoolScope is active on its thread takes its block from that pool and
// returns it there; the pool must outlive such buffers.
//
// Arena mode (constructed over a caller-owned device range): blocks are
// consecutive 256-byte-aligned pieces of that range, never returned; the
// pool neither allocates nor frees device memory, and a request beyond the
// range throws AN_ENOMEM.
class DevPool {
 public:
  DevPool() = default;
  DevPool(void* arenaBase, size_t arenaBytes) : base_((char*)arenaBase), cap_(arenaBytes) {}
  DevPool(const DevPool&) = delete;
  DevPool& operator=(const DevPool&) = delete;
  ~DevPool() {
    for (auto& f : free_) (void)hipFree(f.second);
  }
  static size_t bucket(size_t bytes) {
    size_t b = 256;
    while (b < bytes) b <<= 1;
    return b;
  }
  // block size handed out for a request of `bytes`
  size_t round(size_t bytes) const { return base_ ? ((bytes + 255) & ~size_t(255)) : bucket(bytes); }
  size_t used() const { return used_; }
  void* get(size_t bytes) {
    if (base_) {
      if (bytes > cap_ - used_)
        throw Error(5, "scratch of " + std::to_string(cap_) + " bytes exhausted (" +
                           std::to_string(used_) + " in use, " + std::to_string(bytes) +
                           " more requested)");
      void* p = base_ + used_;
      used_ += bytes;
      return p;
    }
    auto it = free_.find(bytes);
    if (it != free_.end()) {
      void* p = it->second;
      free_.erase(it);
      return p;
    }
    void* p = nullptr;
    AN_HIP(hipMalloc(&p, bytes));
    return p;
  }
  void put(void* p, size_t bytes) {
    if (!base_) free_.emplace(bytes, p);
  }

 private:
  std::multimap<size_t, void*> free_;
  char* base_ = nullptr;
  size_t cap_ = 0, used_ = 0;
};
inline thread_local DevPool* tlPool = nullptr;
struct PoolScope {
  DevPool* prev;
  explicit PoolScope(DevPool* p) : prev(tlPool) { tlPool = p; }
  ~PoolScope() { tlPool = prev; }
  PoolScope(const PoolScope&) = delete;
  PoolScope& operator=(const PoolScope&) = delete;
};

template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevPool* pool = nullptr;  // owner of p when it came from a pool
  size_t bytes = 0;         // pool block size
  bool borrowed = false;    // p belongs to another buffer (borrow())
  DevBuf() = default;
  explicit DevBuf(size_t count) { alloc(count); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept
      : p(o.p), n(o.n), pool(o.pool), bytes(o.bytes), borrowed(o.borrowed) {
    o.p = nullptr;
    o.n = 0;
    o.pool = nullptr;
    o.bytes = 0;
    o.borrowed = false;
  }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      n = o.n;
      pool = o.pool;
      bytes = o.bytes;
      borrowed = o.borrowed;
      o.p = nullptr;
      o.n = 0;
      o.pool = nullptr;
      o.bytes = 0;
      o.borrowed = false;
    }
    return *this;
  }
  // a view of memory another buffer owns (released without a free); the
  // owner must outlive it
  void borrow(T* q, size_t count) {
    release();
    p = q;
    n = count;
    borrowed = true;
  }
  ~DevBuf() { release(); }
  void alloc(size_t count) {
    release();
    n = count;
    if (!count) return;
    if (tlPool) {
      pool = tlPool;
      bytes = pool->round(count * sizeof(T));
      p = (T*)pool->get(bytes);
    } else {
      AN_HIP(hipMalloc((void**)&p, count * sizeof(T)));
    }
  }
  void ensure(size_t count) { if (count > n || !p) alloc(count); }
  void release() {
    if (p && !borrowed) {
      if (pool)
        pool->put(p, bytes);
      else
        (void)hipFree(p);
    }
    p = nullptr;
    n = 0;
    pool = nullptr;
    bytes = 0;
    borrowed = false;
  }
  void upload(const T* h, size_t count, hipStream_t s = 0) {
    ensure(count);
    if (count) AN_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
  }
  void upload(const std::vector<T>& h, hipStream_t s = 0) { upload(h.data(), h.size(), s); }
  void download(T* h, size_t count, hipStream_t s = 0) const {
    if (count) AN_HIP(hipMemcpyAsync(h, p, count * sizeof(T), hipMemcpyDeviceToHost, s));
  }
  void zero(hipStream_t s = 0) {
    if (n) AN_HIP(hipMemsetAsync(p, 0, n * sizeof(T), s));
  }
};

// CSR addressing + incidence built on the GPU (addressing.hip).
struct DevAddressing {
  int nC = 0, nF = 0, nIF = 0, nBF = 0;
  long long nnz = 0;
  DevBuf<int> rowOff, cols, diagPos;
  DevBuf<int> facePP, facePN, faceNP, faceNN;
  DevBuf<int> cfOff, cfEnt, cfSlot;
  DevBuf<int> cbOff, cbFace;
  DevBuf<int> owner, neigh;  // raw face addressing on device
};

// Builds addressing from owner[nF], neigh[nIF] (device pointers).
void build_addressing(DevAddressing& A, int nC, int nF, int nIF, const int* d_owner,
                      const int* d_neigh, hipStream_t s);

// The GSV1 least-squares gradient stored as its per-term operator (a source
// and a coefficient vector per term), laid out for thread-per-cell reads
// that coalesce: cells in slices of 64 (one wave), term j of the slice's
// cell i at off[slice] + 64 j + i, a slice as wide as its longest row.
// src >= 0 is a neighbour cell, src < 0 the boundary face -src - 1.
struct LsqOperator {
  DevBuf<int> off, len, src;
  DevBuf<double> cx, cy, cz;
  long long slots = 0;
};

struct DevMesh {
  HostMesh host;
  bool onDevice = false;
  int nC = 0, nF = 0, nIF = 0, nBF = 0, nPatches = 0;
  // internal faces with !(|n.d| > 1e-30), which the non-orthogonal flux and
  // its divergence skip; none on a valid mesh, and the kernels then drop the
  // per-face test
  int nTinyDpn = 0;
  DevAddressing addr;
  DevBuf<double4> cgeo;
  DevBuf<int2> fPN;
  DevBuf<double> fLam, fDpn, fA, fNx, fNy, fNz, fDx, fDy, fDz;
  DevBuf<int> bP, bPatch;
  DevBuf<double> bDpn, bA, bNx, bNy, bNz, bDx, bDy, bDz;
  // non-orthogonal face vector T = A n - (A / n.d) d, internal and boundary
  DevBuf<double> fTx, fTy, fTz, bTx, bTy, bTz;
  DevBuf<double> gsvInv;  // GSV1 LSQ normal-matrix inverse per cell (9 x nC)
  LsqOperator lsqOp;      // the GSV1 gradient's per-term operator, sliced
  void to_device(hipStream_t s);
};

// exclusive scan of n ints into out[0..n]; returns out[n] (synchronises).
long long exclusive_scan(const int* d_in, int* d_out, int n, hipStream_t s);

inline int grid_for(long long n, int block) {
  long long g = (n + block - 1) / block;
  return (int)(g > 0 ? g : 1);
}

// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X; remap
// the hardware block id so each XCD works on one contiguous chunk of rows and
// keeps its neighbour gathers inside its own L2.
__device__ __forceinline__ int xcd_block(int nblocks) {
  const int b = blockIdx.x;
  const int per = (nblocks + 7) >> 3;
  const int xcd = b & 7, k = b >> 3;
  const int full = nblocks & 7 ? (nblocks & 7) : 8;  // XCDs with 'per' blocks
  int base = xcd < full ? xcd * per : full * per + (xcd - full) * (per - 1);
  return base + k;
}

}  // namespace an
