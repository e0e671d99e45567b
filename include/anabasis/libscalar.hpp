// libscalar.hpp -- the libscalar C++ API over the anabasis_hip C ABI.
//
// Same names, types and behaviour as
// libscalar/include/scalar_transport_library.h:1-111 (namespace libscalar):
// ScalarBCType {Dirichlet, ZeroGradient, NeumannGradient}, ScalarPatchBC,
// ScalarBCSet, make_*_bc, ConvectionScheme, DiffusionScheme,
// LinearSolverType, PreconditionerType, ScalarTransportInputs/Options/
// Result and solve_steady_scalar_transport
// (LS/src/scalar_transport_library.cu:475-512).  The mesh type is the
// libpoisson Mesh of <anabasis/libpoisson.hpp>.  Failures throw
// std::runtime_error.  Header-only; C++17.
#pragma once

#include <anabasis/libpoisson.hpp>

namespace libscalar {

enum class ScalarBCType { Dirichlet, ZeroGradient, NeumannGradient };

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

inline ScalarPatchBC make_dirichlet_patch_bc(const std::string& patchName, ScalarBCFunc fn) {
  return ScalarPatchBC{patchName, ScalarBCType::Dirichlet, std::move(fn)};
}
inline ScalarPatchBC make_zero_gradient_patch_bc(const std::string& patchName) {
  return ScalarPatchBC{patchName, ScalarBCType::ZeroGradient,
                       [](const std::array<double, 3>&, const std::array<double, 3>&) {
                         return 0.0;
                       }};
}
inline ScalarPatchBC make_neumann_gradient_patch_bc(const std::string& patchName,
                                                    ScalarBCFunc fn) {
  return ScalarPatchBC{patchName, ScalarBCType::NeumannGradient, std::move(fn)};
}
inline ScalarPatchBC make_dirichlet_constant_bc(const std::string& patchName, double value) {
  return make_dirichlet_patch_bc(
      patchName,
      [value](const std::array<double, 3>&, const std::array<double, 3>&) { return value; });
}
inline ScalarPatchBC make_neumann_gradient_constant_bc(const std::string& patchName,
                                                       double value) {
  return make_neumann_gradient_patch_bc(
      patchName,
      [value](const std::array<double, 3>&, const std::array<double, 3>&) { return value; });
}

enum class ConvectionScheme { Central, Upwind };
enum class DiffusionScheme { Orth, NonOrth };
enum class LinearSolverType { BiCGSTAB };
enum class PreconditionerType { Jacobi };

struct ScalarTransportInputs {
  std::vector<double> faceFlux;   // size = mesh.nFaces, oriented owner->neigh/outward
  std::vector<double> gammaFace;  // size = mesh.nFaces
  std::vector<double> Su;         // size = mesh.nCells
  std::vector<double> Sp;         // size = mesh.nCells, implicit linearized source
};

struct ScalarTransportOptions {
  ConvectionScheme convectionScheme = ConvectionScheme::Central;
  DiffusionScheme diffusionScheme = DiffusionScheme::NonOrth;
  int nNonOrthCorr = 2;
  LinearSolverType linearSolver = LinearSolverType::BiCGSTAB;
  PreconditionerType preconditioner = PreconditionerType::Jacobi;
  int maxIter = 4000;
  double relTol = 0.0;
  double absTol = 1.0e-10;
  int monitor = 0;
};

struct ScalarTransportResult {
  std::vector<double> phi;
  int iterations = 0;
  double finalRelRes = 0.0;
};

// LS build_boundary_face_data (scalar_transport_library.cu:70-126): every
// boundary face gets its patch's type and the evaluator at (xf, nf); the
// patch index base (0 or 1) is detected from bPatch as the reference does.
inline void boundary_face_arrays(const Mesh& mesh, const ScalarBCSet& bcSet, std::vector<int>& t,
                                 std::vector<double>& v) {
  const int nP = (int)mesh.patchNames.size();
  int lo = 999999999, hi = -999999999;
  for (int f = mesh.nInternalFaces; f < mesh.nFaces; ++f) {
    lo = std::min(lo, mesh.bPatch[f]);
    hi = std::max(hi, mesh.bPatch[f]);
  }
  int base = 0;
  if (nP > 0) {
    if (lo == 0 && hi == nP - 1)
      base = 0;
    else if (lo == 1 && hi == nP)
      base = 1;
    else if (lo >= 1 && hi <= nP)
      base = 1;
  }
  const int nB = mesh.nFaces - mesh.nInternalFaces;
  t.assign(nB, AN_BC_ZERO_GRADIENT);
  v.assign(nB, 0.0);
  for (int f = mesh.nInternalFaces; f < mesh.nFaces; ++f) {
    const int raw = mesh.bPatch[f];
    const int patch = raw - base;
    if (patch < 0 || patch >= nP)
      throw std::runtime_error("Boundary face has invalid patch index: raw=" +
                               std::to_string(raw) + ", adjusted=" + std::to_string(patch) +
                               ", nPatches=" + std::to_string(nP));
    const std::string& name = mesh.patchNames[patch];
    const ScalarPatchBC* bc = nullptr;
    for (const auto& b : bcSet.patches)
      if (b.patchName == name) bc = &b;
    if (!bc) throw std::runtime_error("No scalar BC provided for patch '" + name + "'");
    const int i = f - mesh.nInternalFaces;
    t[i] = bc->type == ScalarBCType::Dirichlet      ? AN_BC_DIRICHLET
           : bc->type == ScalarBCType::ZeroGradient ? AN_BC_ZERO_GRADIENT
                                                    : AN_BC_NEUMANN;
    v[i] = bc->evaluator(mesh.xf[f], mesh.nf[f]);
  }
}

inline ScalarTransportResult solve_steady_scalar_transport(const Mesh& mesh,
                                                           const ScalarTransportInputs& in,
                                                           const ScalarBCSet& bcSet,
                                                           const ScalarTransportOptions& opt,
                                                           const std::vector<double>& x0 = {}) {
  if (!x0.empty() && (int)x0.size() != mesh.nCells)
    throw std::runtime_error("x0 must have size mesh.nCells");
  std::vector<int> t;
  std::vector<double> v;
  boundary_face_arrays(mesh, bcSet, t, v);
  an_transport_options o;
  an_transport_default_options(&o);
  o.upwind = opt.convectionScheme == ConvectionScheme::Upwind ? 1 : 0;
  o.nonOrth = opt.diffusionScheme == DiffusionScheme::NonOrth ? 1 : 0;
  o.nNonOrthCorr = opt.nNonOrthCorr;
  o.maxIter = opt.maxIter;
  o.relTol = opt.relTol;
  o.absTol = opt.absTol;
  ScalarTransportResult out;
  out.phi.assign(mesh.nCells, 0.0);
  an_solve_info info{};
  an_status st = an_solve_steady_scalar_transport(
      anabasis::handle_of(mesh), in.faceFlux.data(), in.gammaFace.data(), in.Su.data(),
      in.Sp.data(), t.data(), v.data(), &o, x0.empty() ? nullptr : x0.data(), out.phi.data(),
      AN_HOST, &info);
  if (st != AN_OK && st != AN_ENOCONV) anabasis::check(st);
  out.iterations = info.iterations;
  out.finalRelRes = info.finalRelResNorm;
  return out;
}

}  // namespace libscalar
