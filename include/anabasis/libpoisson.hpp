// libpoisson.hpp -- the libpoisson C++ API over the anabasis_hip C ABI.
//
// Same type and function names, signatures and error behaviour as the
// reference headers, so an application written against libpoisson compiles
// against this header and links libanabasis_hip.so instead of the CUDA/HYPRE
// build:
//   mesh.h              Mesh, PatchInfo, read_openfoam_polymesh
//                       (libpoisson/include/mesh.h:5-26)
//   bc.h                ScalarBCType, ScalarPatchBC, ScalarBCSet,
//                       BoundaryFaceData, make_*_bc, build_boundary_face_data
//                       (libpoisson/include/bc.h:5-37)
//   gradient.h          compute_lsq_gradient (libpoisson/include/gradient.h:4-8)
//   scalar_elliptic.h   CSRPattern, build_scalar_pattern,
//                       assemble_scalar_elliptic_system
//                       (libpoisson/include/scalar_elliptic.h:4-31)
//   hypre_backend.h     HypreOptions, HypreSolveInfo, solve_system_hypre_gpu,
//                       HypreReusableSystem + init/update/solve/destroy
//                       (libpoisson/include/hypre_backend.h:4-82)
//   poisson_library.h   EllipticOptions, EllipticResult,
//                       solve_scalar_elliptic, solve_poisson
//                       (libpoisson/include/poisson_library.h:4-35)
//
// Differences a caller can see:
//   - no MPI/HYPRE initialisation is needed (init_hypre_gpu_runtime has no
//     counterpart); failures throw std::runtime_error with the library's
//     message instead of calling MPI_Abort;
//   - the pressure/Poisson solver is PCG + aggregation AMG: HypreOptions'
//     amg* fields are accepted and ignored, maxIter/relTol/absTol/tol keep
//     their meaning (stop at ||r|| <= max(relTol ||b||, absTol));
//   - Mesh additionally carries patchStartFace/patchNFaces (as
//     generic_simple_v1's mesh does) and a handle to the device copy.
//
// Header-only; C++17.
#pragma once

#include <anabasis_hip.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

// HYPRE scalar types of the reference's contract (int32 indices, FP64).
typedef int HYPRE_Int;
typedef int HYPRE_BigInt;
typedef double HYPRE_Complex;

namespace anabasis {

inline void check(an_status st) {
  if (st != AN_OK) throw std::runtime_error(an_last_error());
}

struct MeshHandle {
  an_mesh_t h = nullptr;
  explicit MeshHandle(an_mesh_t m) : h(m) {}
  ~MeshHandle() {
    if (h) an_mesh_destroy(h);
  }
  MeshHandle(const MeshHandle&) = delete;
  MeshHandle& operator=(const MeshHandle&) = delete;
};

}  // namespace anabasis

struct PatchInfo {
  std::string name;
  int nFaces = 0;
  int startFace = 0;
  std::string type;
};

struct Mesh {
  std::vector<std::array<double, 3>> P;
  std::vector<std::vector<int>> faces;
  std::vector<int> owner, neigh, bPatch;
  std::vector<std::string> patchNames;
  int nFaces = 0;
  int nInternalFaces = 0;
  int nCells = 0;
  std::vector<std::array<double, 3>> cc, xf, nf, Sf;
  std::vector<double> vol, Af;
  std::vector<std::vector<int>> cellNbrs, cellBFace, cellFaces, cellOrient;
  double maxNonOrthDeg = 0.0;
  std::vector<int> patchStartFace, patchNFaces;
  // device-resident copy used by every library call
  std::shared_ptr<anabasis::MeshHandle> handle;
};

namespace anabasis {

template <typename T>
inline void get_array(an_mesh_t m, const char* name, std::vector<T>& v, size_t n) {
  v.resize(n);
  check(an_mesh_get_array(m, name, v.data()));
}

inline std::vector<std::array<double, 3>> to_vec3(const std::vector<double>& a) {
  std::vector<std::array<double, 3>> out(a.size() / 3);
  for (size_t i = 0; i < out.size(); ++i) out[i] = {a[3 * i], a[3 * i + 1], a[3 * i + 2]};
  return out;
}

inline std::vector<double> flat3(const std::vector<std::array<double, 3>>& a) {
  std::vector<double> out(3 * a.size());
  for (size_t i = 0; i < a.size(); ++i)
    for (int d = 0; d < 3; ++d) out[3 * i + d] = a[i][d];
  return out;
}

// Fills the host Mesh (reference layout, mesh.cu:160-288) from a handle.
inline Mesh mesh_from_handle(an_mesh_t h) {
  Mesh m;
  m.handle = std::make_shared<MeshHandle>(h);
  an_mesh_info info{};
  check(an_mesh_get_info(h, &info));
  m.nFaces = info.nFaces;
  m.nInternalFaces = info.nInternalFaces;
  m.nCells = info.nCells;
  m.maxNonOrthDeg = info.maxNonOrthDeg;
  std::vector<double> pts, tmp;
  get_array(h, "points", pts, 3 * (size_t)info.nPoints);
  m.P = to_vec3(pts);
  std::vector<int> off, fv;
  get_array(h, "faceOffsets", off, (size_t)info.nFaces + 1);
  get_array(h, "faceVerts", fv, (size_t)off.back());
  m.faces.resize(info.nFaces);
  for (int f = 0; f < info.nFaces; ++f) m.faces[f].assign(fv.begin() + off[f], fv.begin() + off[f + 1]);
  get_array(h, "owner", m.owner, (size_t)info.nFaces);
  get_array(h, "neigh", m.neigh, (size_t)info.nInternalFaces);
  get_array(h, "bPatch", m.bPatch, (size_t)info.nFaces);
  for (int k = 0; k < info.nPatches; ++k) {
    char name[256];
    int s = 0, n = 0;
    check(an_mesh_get_patch(h, k, name, (int)sizeof(name), &s, &n));
    m.patchNames.emplace_back(name);
    m.patchStartFace.push_back(s);
    m.patchNFaces.push_back(n);
  }
  get_array(h, "cc", tmp, 3 * (size_t)m.nCells);
  m.cc = to_vec3(tmp);
  get_array(h, "xf", tmp, 3 * (size_t)m.nFaces);
  m.xf = to_vec3(tmp);
  get_array(h, "nf", tmp, 3 * (size_t)m.nFaces);
  m.nf = to_vec3(tmp);
  get_array(h, "Sf", tmp, 3 * (size_t)m.nFaces);
  m.Sf = to_vec3(tmp);
  get_array(h, "vol", m.vol, (size_t)m.nCells);
  get_array(h, "Af", m.Af, (size_t)m.nFaces);
  // derived connectivity exactly as mesh.cu:188-199, 262-277
  m.cellFaces.assign(m.nCells, {});
  m.cellOrient.assign(m.nCells, {});
  for (int f = 0; f < m.nFaces; ++f) {
    m.cellFaces[m.owner[f]].push_back(f);
    m.cellOrient[m.owner[f]].push_back(+1);
  }
  for (int f = 0; f < m.nInternalFaces; ++f) {
    m.cellFaces[m.neigh[f]].push_back(f);
    m.cellOrient[m.neigh[f]].push_back(-1);
  }
  m.cellNbrs.assign(m.nCells, {});
  m.cellBFace.assign(m.nCells, {});
  for (int f = 0; f < m.nFaces; ++f) {
    const int P = m.owner[f];
    if (f < m.nInternalFaces) {
      m.cellNbrs[P].push_back(m.neigh[f]);
      m.cellNbrs[m.neigh[f]].push_back(P);
    } else {
      m.cellBFace[P].push_back(f);
    }
  }
  for (auto& nb : m.cellNbrs) {
    std::sort(nb.begin(), nb.end());
    nb.erase(std::unique(nb.begin(), nb.end()), nb.end());
  }
  return m;
}

// Handle of a Mesh built by the caller (not by read_openfoam_polymesh).
inline an_mesh_t handle_of(const Mesh& m) {
  if (m.handle) return m.handle->h;
  std::vector<double> pts = flat3(m.P);
  std::vector<int> off(1, 0), fv;
  for (const auto& f : m.faces) {
    fv.insert(fv.end(), f.begin(), f.end());
    off.push_back((int)fv.size());
  }
  std::vector<const char*> names;
  std::vector<int> start, count;
  for (size_t k = 0; k < m.patchNames.size(); ++k) {
    names.push_back(m.patchNames[k].c_str());
    int s = m.nFaces, e = m.nInternalFaces - 1;
    for (int f = m.nInternalFaces; f < m.nFaces; ++f)
      if (m.bPatch[f] == (int)k + 1) {
        s = std::min(s, f);
        e = std::max(e, f);
      }
    start.push_back(s <= e ? s : m.nFaces);
    count.push_back(s <= e ? e - s + 1 : 0);
  }
  an_mesh_t h = nullptr;
  check(an_mesh_from_arrays((int)m.P.size(), pts.data(), m.nFaces, off.data(), fv.data(),
                            m.owner.data(), m.nInternalFaces, m.neigh.data(),
                            (int)m.patchNames.size(), names.data(), start.data(), count.data(),
                            &h));
  const_cast<Mesh&>(m).handle = std::make_shared<MeshHandle>(h);
  return h;
}

}  // namespace anabasis

inline Mesh read_openfoam_polymesh(const std::string& polyMeshDir) {
  an_mesh_t h = nullptr;
  anabasis::check(an_mesh_read_polymesh(polyMeshDir.c_str(), &h));
  return anabasis::mesh_from_handle(h);
}

// ------------------------------------------------------------------- bc.h
enum class ScalarBCType { Dirichlet, Neumann };

using ScalarBCFunc = std::function<double(const std::array<double, 3>& x,
                                          const std::array<double, 3>& outwardNormal)>;

struct ScalarPatchBC {
  std::string patchName;
  ScalarBCType type = ScalarBCType::Dirichlet;
  ScalarBCFunc evaluator;
};

struct ScalarBCSet {
  std::vector<ScalarPatchBC> patches;
};

struct BoundaryFaceData {
  std::vector<ScalarBCType> type;  // [nFaces], boundary entries used
  std::vector<double> value;
};

inline ScalarPatchBC make_dirichlet_patch_bc(const std::string& patchName, ScalarBCFunc fn) {
  return ScalarPatchBC{patchName, ScalarBCType::Dirichlet, std::move(fn)};
}
inline ScalarPatchBC make_neumann_patch_bc(const std::string& patchName, ScalarBCFunc fn) {
  return ScalarPatchBC{patchName, ScalarBCType::Neumann, std::move(fn)};
}
inline ScalarPatchBC make_dirichlet_constant_bc(const std::string& patchName, double value) {
  return make_dirichlet_patch_bc(patchName, [value](const auto&, const auto&) { return value; });
}
inline ScalarPatchBC make_neumann_constant_bc(const std::string& patchName, double value) {
  return make_neumann_patch_bc(patchName, [value](const auto&, const auto&) { return value; });
}

// bc.cu:3-35: evaluates every boundary face's patch BC at (xf, nf).
inline BoundaryFaceData build_boundary_face_data(const Mesh& mesh, const ScalarBCSet& bcSet) {
  BoundaryFaceData out;
  out.type.assign(mesh.nFaces, ScalarBCType::Dirichlet);
  out.value.assign(mesh.nFaces, 0.0);
  for (int f = mesh.nInternalFaces; f < mesh.nFaces; ++f) {
    const int p1 = mesh.bPatch[f];
    if (p1 <= 0) throw std::runtime_error("Boundary face missing patch index at face " + std::to_string(f));
    if (p1 - 1 >= (int)mesh.patchNames.size())
      throw std::runtime_error("Invalid patch index at face " + std::to_string(f));
    const std::string& name = mesh.patchNames[p1 - 1];
    const ScalarPatchBC* bc = nullptr;
    for (const auto& b : bcSet.patches)
      if (b.patchName == name) bc = &b;  // the last spec of a name wins, as a std::map insert
    if (!bc) throw std::runtime_error("No BC supplied for patch '" + name + "'");
    if (!bc->evaluator) throw std::runtime_error("BC evaluator missing for patch '" + name + "'");
    out.type[f] = bc->type;
    out.value[f] = bc->evaluator(mesh.xf[f], mesh.nf[f]);
  }
  return out;
}

namespace anabasis {
// per-boundary-face arrays of the C ABI (index f - nInternalFaces)
inline void bf_arrays(const Mesh& mesh, const BoundaryFaceData& bc, std::vector<int>& t,
                      std::vector<double>& v) {
  const int nB = mesh.nFaces - mesh.nInternalFaces;
  t.resize(nB);
  v.resize(nB);
  for (int i = 0; i < nB; ++i) {
    const int f = mesh.nInternalFaces + i;
    t[i] = bc.type[f] == ScalarBCType::Dirichlet ? AN_BC_DIRICHLET : AN_BC_NEUMANN;
    v[i] = bc.value[f];
  }
}
}  // namespace anabasis

// ------------------------------------------------------------- gradient.h
// gradient.cu:21-83 (LP boundary rule)
inline void compute_lsq_gradient(const Mesh& mesh, const std::vector<double>& phi,
                                 const BoundaryFaceData& bcFaceData,
                                 std::vector<std::array<double, 3>>& grad) {
  std::vector<int> t;
  std::vector<double> v, g(3 * (size_t)mesh.nCells);
  anabasis::bf_arrays(mesh, bcFaceData, t, v);
  anabasis::check(an_lsq_gradient(anabasis::handle_of(mesh), AN_GRAD_LP, t.data(), v.data(),
                                  phi.data(), g.data(), AN_HOST));
  grad = anabasis::to_vec3(g);
}

// ------------------------------------------------------ scalar_elliptic.h
struct CSRPattern {
  int nRows = 0;
  int nnz = 0;
  std::vector<HYPRE_Int> ncols;
  std::vector<int> rowOffsets;
  std::vector<int> diagPos;
  std::vector<HYPRE_BigInt> rows;
  std::vector<HYPRE_BigInt> cols;
  std::vector<int> facePP, facePN, faceNP, faceNN;
};

// scalar_elliptic.cu:16-58 (built on the GPU, bit-identical)
inline CSRPattern build_scalar_pattern(const Mesh& mesh) {
  an_mesh_t h = anabasis::handle_of(mesh);
  an_mesh_info info{};
  anabasis::check(an_mesh_get_info(h, &info));
  CSRPattern p;
  p.nRows = mesh.nCells;
  p.nnz = (int)info.nnz;
  p.rowOffsets.resize(mesh.nCells + 1);
  p.cols.resize(p.nnz);
  p.diagPos.resize(mesh.nCells);
  const size_t nIF = (size_t)mesh.nInternalFaces;
  p.facePP.resize(nIF);
  p.facePN.resize(nIF);
  p.faceNP.resize(nIF);
  p.faceNN.resize(nIF);
  anabasis::check(an_mesh_csr(h, p.rowOffsets.data(), p.cols.data(), p.diagPos.data(),
                              p.facePP.data(), p.facePN.data(), p.faceNP.data(), p.faceNN.data(),
                              AN_HOST));
  p.rows.resize(mesh.nCells);
  p.ncols.resize(mesh.nCells);
  for (int c = 0; c < mesh.nCells; ++c) {
    p.rows[c] = c;
    p.ncols[c] = p.rowOffsets[c + 1] - p.rowOffsets[c];
  }
  return p;
}

// scalar_elliptic.cu:60-136
inline void assemble_scalar_elliptic_system(const Mesh& mesh, const CSRPattern& pat,
                                            const std::vector<double>& gammaFace,
                                            const std::vector<double>& cellSource,
                                            const BoundaryFaceData& bcFaceData,
                                            const std::vector<std::array<double, 3>>& grad,
                                            std::vector<HYPRE_Complex>& values,
                                            std::vector<HYPRE_Complex>& rhs, bool includeNonOrth,
                                            bool useReferenceCell, int refCell,
                                            double refValue) {
  std::vector<int> t;
  std::vector<double> v;
  anabasis::bf_arrays(mesh, bcFaceData, t, v);
  std::vector<double> g = anabasis::flat3(grad);
  values.assign(pat.nnz, 0.0);
  rhs.assign(mesh.nCells, 0.0);
  anabasis::check(an_elliptic_assemble(anabasis::handle_of(mesh), gammaFace.data(),
                                       cellSource.data(), t.data(), v.data(), g.data(),
                                       includeNonOrth ? 1 : 0, useReferenceCell ? 1 : 0, refCell,
                                       refValue, values.data(), rhs.data(), AN_HOST));
}

// -------------------------------------------------------- hypre_backend.h
struct HypreOptions {
  int maxIter = 4000;
  double relTol = 0.0;
  double absTol = -1.0;
  double tol = -1.0;  // if absTol < 0, tol is the absolute tolerance
  int monitor = 1;
  // BoomerAMG settings of the reference: accepted, not used (the
  // preconditioner is the aggregation AMG)
  int amgMaxIter = 1;
  int amgRelaxType = 18;
  int amgCoarsenType = 8;
  int amgInterpType = 6;
  int amgAggLevels = 1;
  int amgAggInterpType = 4;
  int amgNumSweeps = 1;
  int amgRAP2 = 0;
  int amgPmax = 4;
  int amgKeepTranspose = 1;
  double amgTruncFactor = 0.0;
  double amgStrongThreshold = -1.0;
};

struct HypreSolveInfo {
  int iterations = 0;
  double finalRelResNorm = 0.0;
};

namespace anabasis {
inline an_pcg_options pcg_options_of(const HypreOptions& o) {
  an_pcg_options p;
  an_pcg_default_options(&p);
  p.maxIter = o.maxIter;
  p.relTol = o.relTol;
  p.absTol = o.absTol >= 0.0 ? o.absTol : (o.tol >= 0.0 ? o.tol : 0.0);
  return p;
}

struct Solver {
  an_solver_t h = nullptr;
  ~Solver() {
    if (h) an_solver_destroy(h);
  }
};
}  // namespace anabasis

namespace anabasis {
// opt.monitor: the residual norm of every iteration of the last solve, one
// line each (the reference's monitor sets HYPRE's PCG print level 2,
// hypre_backend.cu:63-64, 129-130)
inline void print_residual_history(an_solver_t s, const std::vector<double>& rhs) {
  int n = 0;
  check(an_pcg_residual_history(s, nullptr, 0, &n));
  std::vector<double> h((size_t)n);
  check(an_pcg_residual_history(s, h.data(), n, &n));
  double bb = 0.0;
  for (double v : rhs) bb += v * v;
  const double bn = std::sqrt(bb);
  std::printf("  PCG iteration   ||r||_2        rate       ||r||_2/||b||_2\n");
  for (int i = 0; i < n; ++i)
    std::printf("  %13d   %.6e   %.6f   %.6e\n", i, h[(size_t)i],
                i > 0 && h[(size_t)i - 1] > 0.0 ? h[(size_t)i] / h[(size_t)i - 1] : 0.0,
                bn > 0.0 ? h[(size_t)i] / bn : 0.0);
}
}  // namespace anabasis

// hypre_backend.cu:3-114: one-shot PCG solve; x is the initial guess and
// the result.  Non-convergence is reported in the info, not thrown.
inline HypreSolveInfo solve_system_hypre_gpu(const CSRPattern& pat,
                                             const std::vector<HYPRE_Complex>& values,
                                             const std::vector<HYPRE_Complex>& rhs,
                                             std::vector<double>& x, const HypreOptions& opt) {
  an_pcg_options po = anabasis::pcg_options_of(opt);
  anabasis::Solver s;
  anabasis::check(an_pcg_create(pat.nRows, pat.rowOffsets.data(), pat.cols.data(), pat.nnz, &po,
                                AN_HOST, &s.h));
  anabasis::check(an_pcg_set_values(s.h, values.data(), 1, AN_HOST));
  x.resize(pat.nRows, 0.0);
  an_solve_info info{};
  an_status st = an_pcg_solve(s.h, rhs.data(), x.data(), AN_HOST, &info);
  if (st != AN_OK && st != AN_ENOCONV) anabasis::check(st);
  if (opt.monitor) anabasis::print_residual_history(s.h, rhs);
  return HypreSolveInfo{info.iterations, info.finalRelResNorm};
}

struct HypreReusableSystem {
  bool initialized = false;
  bool isSetup = false;
  int nRows = 0;
  int nnz = 0;
  std::vector<HYPRE_BigInt> rows;
  std::vector<HYPRE_Int> ncols;
  std::vector<HYPRE_BigInt> cols;
  HypreOptions opt;
  std::shared_ptr<anabasis::Solver> solver;
  std::vector<double> rhs;
};

// hypre_backend.h:64-82: rows/ncols/cols describe the CSR pattern row by row
inline void init_reusable_hypre_system_gpu(HypreReusableSystem& sys,
                                           const std::vector<HYPRE_BigInt>& rows,
                                           const std::vector<HYPRE_Int>& ncols,
                                           const std::vector<HYPRE_BigInt>& cols,
                                           const HypreOptions& opt) {
  sys.rows = rows;
  sys.ncols = ncols;
  sys.cols = cols;
  sys.nRows = (int)rows.size();
  sys.nnz = (int)cols.size();
  sys.opt = opt;
  std::vector<int> ro(sys.nRows + 1, 0);
  for (int i = 0; i < sys.nRows; ++i) ro[i + 1] = ro[i] + ncols[i];
  if (ro.back() != sys.nnz) throw std::runtime_error("ncols does not match cols");
  an_pcg_options po = anabasis::pcg_options_of(opt);
  sys.solver = std::make_shared<anabasis::Solver>();
  anabasis::check(an_pcg_create(sys.nRows, ro.data(), cols.data(), sys.nnz, &po, AN_HOST,
                                &sys.solver->h));
  sys.initialized = true;
  sys.isSetup = false;
}

inline void update_reusable_hypre_system_gpu(HypreReusableSystem& sys,
                                             const std::vector<HYPRE_Complex>& values,
                                             const std::vector<HYPRE_Complex>& rhs,
                                             const HypreOptions& opt, bool doSetup) {
  if (!sys.initialized) throw std::runtime_error("reusable system not initialized");
  sys.opt = opt;
  anabasis::check(
      an_pcg_set_values(sys.solver->h, values.data(), (doSetup || !sys.isSetup) ? 1 : 0, AN_HOST));
  sys.rhs = rhs;
  sys.isSetup = true;
}

inline HypreSolveInfo solve_reusable_hypre_system_gpu(HypreReusableSystem& sys,
                                                      std::vector<double>& x) {
  if (!sys.isSetup) throw std::runtime_error("reusable system not set up");
  x.resize(sys.nRows, 0.0);
  an_solve_info info{};
  an_status st = an_pcg_solve(sys.solver->h, sys.rhs.data(), x.data(), AN_HOST, &info);
  if (st != AN_OK && st != AN_ENOCONV) anabasis::check(st);
  if (sys.opt.monitor) anabasis::print_residual_history(sys.solver->h, sys.rhs);
  return HypreSolveInfo{info.iterations, info.finalRelResNorm};
}

inline void destroy_reusable_hypre_system_gpu(HypreReusableSystem& sys) {
  sys.solver.reset();
  sys.initialized = false;
  sys.isSetup = false;
}

// ------------------------------------------------------ poisson_library.h
struct EllipticOptions {
  int nNonOrthCorr = 2;
  std::string gradScheme = "lsq";
  std::string laplacianScheme = "nonorth";
  bool useReferenceCell = false;
  int referenceCell = 0;
  double referenceValue = 0.0;
  HypreOptions hypre;
};

struct EllipticResult {
  std::vector<double> phi;
  HypreSolveInfo lastSolveInfo;
};

// poisson_library.cu:3-47 (the whole non-orthogonal loop runs on the GPU)
inline EllipticResult solve_scalar_elliptic(const Mesh& mesh, const std::vector<double>& gammaFace,
                                            const std::vector<double>& cellSource,
                                            const ScalarBCSet& bcSet,
                                            const EllipticOptions& opts) {
  if (opts.gradScheme != "lsq")
    throw std::runtime_error("Only gradScheme=lsq is implemented in this build.");
  if (opts.laplacianScheme != "orth" && opts.laplacianScheme != "nonorth")
    throw std::runtime_error("Use laplacianScheme=orth or laplacianScheme=nonorth.");
  const BoundaryFaceData bc = build_boundary_face_data(mesh, bcSet);
  std::vector<int> t;
  std::vector<double> v;
  anabasis::bf_arrays(mesh, bc, t, v);
  an_elliptic_options o;
  an_elliptic_default_options(&o);
  o.nNonOrthCorr = opts.nNonOrthCorr;
  o.nonOrth = opts.laplacianScheme == "nonorth" ? 1 : 0;
  o.useReferenceCell = opts.useReferenceCell ? 1 : 0;
  o.referenceCell = opts.referenceCell;
  o.referenceValue = opts.referenceValue;
  o.solver = anabasis::pcg_options_of(opts.hypre);
  EllipticResult out;
  out.phi.assign(mesh.nCells, 0.0);
  an_solve_info info{};
  an_status st = an_solve_scalar_elliptic(anabasis::handle_of(mesh), gammaFace.data(),
                                          cellSource.data(), t.data(), v.data(), &o,
                                          out.phi.data(), AN_HOST, &info);
  if (st != AN_OK && st != AN_ENOCONV) anabasis::check(st);
  out.lastSolveInfo = HypreSolveInfo{info.iterations, info.finalRelResNorm};
  return out;
}

inline EllipticResult solve_poisson(const Mesh& mesh, const std::vector<double>& cellSource,
                                    const ScalarBCSet& bcSet, const EllipticOptions& opts) {
  std::vector<double> gammaFace(mesh.nFaces, 1.0);
  return solve_scalar_elliptic(mesh, gammaFace, cellSource, bcSet, opts);
}
