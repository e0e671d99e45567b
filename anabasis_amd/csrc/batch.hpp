// batch.hpp -- many independent PCG+AMG solves in one block-diagonal system,
// iterated by one persistent kernel launch (captured in a HIP graph) or by
// a captured graph of a few iterations replayed until every problem is done.
#pragma once

#include <vector>

#include "solver.hpp"

namespace an {

// Row blocks that never cross a problem boundary, with the problem of each
// block (one 64-lane wave per block; per-wave partial sums are reduced per
// problem in block order, so every problem's scalars are deterministic).
struct BatchLayout {
  int rows = 64;  // rows per block (<= 64); nonzero cap = 12 per lane
  RowBlocks rb;
  DevBuf<int> wPid, wOff;  // problem of each block; first block of each problem
  int nW = 0;
};

class BatchSolver {
 public:
  BatchSolver(int nProblems, const std::vector<int>& rowsPerProblem,
              const std::vector<const int*>& rowOffsets, const std::vector<const int*>& cols,
              bool deviceInput, const an_pcg_options& o);
  ~BatchSolver();
  void set_values(const std::vector<const double*>& values, bool deviceInput);
  // b, x per problem (device pointers inside the batch vectors are managed
  // here); x holds x0 on entry.
  an_status solve(const std::vector<const double*>& b, const std::vector<double*>& x,
                  bool deviceIO, std::vector<an_solve_info>& info);

  int nP = 0, n = 0;
  long long nnz = 0;
  an_pcg_options opt;
  std::vector<int> pOff;          // host: first row of each problem
  std::vector<long long> pNnzOff; // host: first nonzero of each problem
  DevCSR<double> A;               // block-diagonal matrix
  AmgHierarchy amg;
  BatchLayout lay;
  // autotuner: rows per block tried, their SpMV+dot time (s), choice
  std::vector<int> tuneRows;
  std::vector<double> tuneSeconds;
  // launch: AN_BATCH_EAGER, AN_BATCH_GRAPH or AN_BATCH_PERSISTENT, or
  // AN_BATCH_AUTO (the tuner's choice, `chosen`); the tuner times one PCG
  // iteration of graph replay and of the persistent kernel at
  // tuneBlocksPerCU workgroups per CU (launchSeconds, same order)
  int launch = 0;
  int chosen = 2;
  int persistBlocksPerCU = 0;     // chosen grid of the persistent kernel
  std::vector<int> tuneBlocksPerCU;
  std::vector<double> launchSeconds;
  int graphIters = 0;             // iterations per captured graph launch
  int lastLaunches = 0;           // kernel-sequence launches of the last solve
  double lastSolveSeconds = 0.0;

 private:
  hipStream_t s = nullptr;
  DevBuf<double> scal, tol2, partial, bB, xB, r, z, p, q;
  DevBuf<float> r32, z32, p32, q32, e32;
  DevBuf<int> act, its, nAct;
  int* hAct = nullptr;  // pinned
  double* hScal = nullptr;  // pinned, nP * S_COUNT
  hipGraphExec_t g64 = nullptr, g32 = nullptr;
  // the persistent kernel: barrier words, its arguments and level views in
  // device memory (per precision), the coarsest dense solve's chunk sums,
  // and the captured {barrier reset, kernel} graph per precision
  struct PersistGraph {
    hipGraphExec_t exec = nullptr;
    int grid = 0, maxIter = 0;
    DevBuf<char> args, levels;
    void drop();
  };
  DevBuf<unsigned> bar;
  DevBuf<double> dpart;
  PersistGraph gp64, gp32;
  int cus = 0, maxPerCU64 = 0, maxPerCU32 = 0;
  long long waitTicks = 0;
  bool valuesSet = false;

  int mode() const;  // the launch a solve uses
  template <typename T>
  void build_persist(const DevCSR<T>& M, T* xv, T* rv, T* zv, T* pv, T* qv, int tolOff,
                     int maxIter, PersistGraph& g);
  // (re)captures g for this grid and iteration cap when needed, launches it
  template <typename T>
  void launch_persist(const DevCSR<T>& M, T* xv, T* rv, T* zv, T* pv, T* qv, int tolOff,
                      PersistGraph& g, int blocksPerCU, int maxIter);
  void check_persist();  // synchronises; throws when a barrier wait gave up
  void tune_launch();

  void build_layout(int rows, BatchLayout& L);
  double time_layout(const BatchLayout& L);
  template <typename T>
  void pcg_prologue(const DevCSR<T>& M, T* xv, T* rv, T* zv, T* pv);
  template <typename T>
  void pcg_iteration(const DevCSR<T>& M, int it, T* xv, T* rv, T* zv, T* pv, T* qv);
  template <typename T>
  hipGraphExec_t capture(const DevCSR<T>& M, T* xv, T* rv, T* zv, T* pv, T* qv);
  template <typename T>
  int run_pcg(const DevCSR<T>& M, T* xv, T* rv, T* zv, T* pv, T* qv, hipGraphExec_t& g,
              PersistGraph& gp, int maxIter);
  void residual64();  // r = b - A x, S_RR per problem
  void finalize(int slot);
  void read_scalars();
};

}  // namespace an
