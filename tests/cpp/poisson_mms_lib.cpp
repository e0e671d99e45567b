// poisson_mms_lib.cpp -- the reference's library-driven Poisson MMS driver
// (apps/poisson_mms_lib/main.cu:1-214) written against the libpoisson C++
// API of <anabasis/libpoisson.hpp>: same options, same BC construction
// (Dirichlet phi_exact unless listed in -neumann-patches, then the exact
// outward normal derivative), same summary lines.  No MPI/HYPRE runtime.
#include <anabasis/libpoisson.hpp>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <set>
#include <sstream>

namespace {

constexpr double kPi = 3.141592653589793238462643383279502884;

// manufactured solution of libpoisson/include/mms.h:5-35
double phi_exact(const std::array<double, 3>& x) {
  return std::sin(kPi * x[0]) * std::sin(kPi * x[1]) * std::sin(kPi * x[2]);
}
std::array<double, 3> grad_phi_exact(const std::array<double, 3>& x) {
  return {kPi * std::cos(kPi * x[0]) * std::sin(kPi * x[1]) * std::sin(kPi * x[2]),
          kPi * std::sin(kPi * x[0]) * std::cos(kPi * x[1]) * std::sin(kPi * x[2]),
          kPi * std::sin(kPi * x[0]) * std::sin(kPi * x[1]) * std::cos(kPi * x[2])};
}
double rhs_exact(const std::array<double, 3>& x) { return 3.0 * kPi * kPi * phi_exact(x); }
double normal_grad_exact(const std::array<double, 3>& x, const std::array<double, 3>& n) {
  const auto g = grad_phi_exact(x);
  return g[0] * n[0] + g[1] * n[1] + g[2] * n[2];
}

struct Params {
  std::string polyMeshDir = "./mesh/constant/polyMesh";
  int device = 0;
  int nNonOrthCorr = 10;
  int maxIter = 4000;
  double tol = 1e-12;
  std::string gradScheme = "lsq";
  std::string laplacianScheme = "nonorth";
  std::string dirichletPatchesCsv, neumannPatchesCsv;
  int useReferenceCell = 0, referenceCell = 0;
  double referenceValue = 0.0;
};

std::vector<std::string> split_csv(const std::string& s) {
  std::vector<std::string> out;
  std::stringstream ss(s);
  std::string item;
  while (std::getline(ss, item, ',')) {
    std::string t;
    for (char ch : item)
      if (!std::isspace((unsigned char)ch)) t.push_back(ch);
    if (!t.empty()) out.push_back(t);
  }
  return out;
}

void parse_args(int argc, char** argv, Params& p) {
  for (int i = 1; i < argc; ++i) {
    auto val = [&](const char* opt) -> const char* {
      if (i + 1 >= argc) {
        std::fprintf(stderr, "Missing value after %s\n", opt);
        std::exit(1);
      }
      return argv[++i];
    };
    const char* a = argv[i];
    if (!std::strcmp(a, "-polyMeshDir")) p.polyMeshDir = val(a);
    else if (!std::strcmp(a, "-device")) p.device = std::atoi(val(a));
    else if (!std::strcmp(a, "-nNonOrthCorr")) p.nNonOrthCorr = std::atoi(val(a));
    else if (!std::strcmp(a, "-maxit")) p.maxIter = std::atoi(val(a));
    else if (!std::strcmp(a, "-tol")) p.tol = std::atof(val(a));
    else if (!std::strcmp(a, "-grad-scheme")) p.gradScheme = val(a);
    else if (!std::strcmp(a, "-laplacian-scheme")) p.laplacianScheme = val(a);
    else if (!std::strcmp(a, "-dirichlet-patches")) p.dirichletPatchesCsv = val(a);
    else if (!std::strcmp(a, "-neumann-patches")) p.neumannPatchesCsv = val(a);
    else if (!std::strcmp(a, "-use-reference-cell")) p.useReferenceCell = std::atoi(val(a));
    else if (!std::strcmp(a, "-reference-cell")) p.referenceCell = std::atoi(val(a));
    else if (!std::strcmp(a, "-reference-value")) p.referenceValue = std::atof(val(a));
    else if (!std::strncmp(a, "-amg-", 5) || !std::strcmp(a, "-monitor")) (void)val(a);  // HYPRE knobs
    else {
      std::fprintf(stderr, "Unknown option: %s\n", a);
      std::exit(1);
    }
  }
}

}  // namespace

int main(int argc, char** argv) {
  Params par;
  parse_args(argc, argv, par);
  try {
    anabasis::check(an_set_device(par.device));
    std::printf("Reading mesh from %s\n", par.polyMeshDir.c_str());
    Mesh mesh = read_openfoam_polymesh(par.polyMeshDir);
    std::printf("Mesh stats: cells=%d, faces=%d, internalFaces=%d, boundaryPatches=%zu, "
                "maxNonOrthDeg=%.3f\n",
                mesh.nCells, mesh.nFaces, mesh.nInternalFaces, mesh.patchNames.size(),
                mesh.maxNonOrthDeg);
    for (size_t i = 0; i < mesh.patchNames.size(); ++i)
      std::printf("  patch[%zu] = %s\n", i, mesh.patchNames[i].c_str());
    const auto nl = split_csv(par.neumannPatchesCsv);
    const std::set<std::string> neumannSet(nl.begin(), nl.end());
    ScalarBCSet bcSet;
    for (const auto& name : mesh.patchNames) {
      if (neumannSet.count(name))
        bcSet.patches.push_back(make_neumann_patch_bc(
            name, [](const std::array<double, 3>& x, const std::array<double, 3>& n) {
              return normal_grad_exact(x, n);
            }));
      else
        bcSet.patches.push_back(make_dirichlet_patch_bc(
            name, [](const std::array<double, 3>& x, const std::array<double, 3>&) {
              return phi_exact(x);
            }));
    }
    std::vector<double> cellSource(mesh.nCells);
    for (int c = 0; c < mesh.nCells; ++c) cellSource[c] = rhs_exact(mesh.cc[c]);
    EllipticOptions opts;
    opts.nNonOrthCorr = par.nNonOrthCorr;
    opts.gradScheme = par.gradScheme;
    opts.laplacianScheme = par.laplacianScheme;
    opts.useReferenceCell = par.useReferenceCell != 0;
    opts.referenceCell = par.referenceCell;
    opts.referenceValue = par.referenceValue;
    opts.hypre.maxIter = par.maxIter;
    opts.hypre.relTol = 0.0;
    opts.hypre.absTol = par.tol;
    opts.hypre.tol = -1.0;
    EllipticResult result = solve_poisson(mesh, cellSource, bcSet, opts);
    const std::vector<double>& phi = result.phi;
    double volSum = 0, l1 = 0, l2 = 0, linf = 0;
    double phiMin = std::numeric_limits<double>::max(), phiMax = -phiMin;
    for (int c = 0; c < mesh.nCells; ++c) {
      const double err = std::fabs(phi[c] - phi_exact(mesh.cc[c]));
      const double V = mesh.vol[c];
      volSum += V;
      l1 += V * err;
      l2 += V * err * err;
      linf = std::max(linf, err);
      phiMin = std::min(phiMin, phi[c]);
      phiMax = std::max(phiMax, phi[c]);
    }
    l1 /= std::max(volSum, 1e-30);
    l2 = std::sqrt(l2 / std::max(volSum, 1e-30));
    std::printf("\n=== Poisson MMS summary ===\n");
    std::printf("gradScheme      : %s\n", par.gradScheme.c_str());
    std::printf("laplacianScheme : %s\n", par.laplacianScheme.c_str());
    std::printf("nNonOrthCorr    : %d\n", par.nNonOrthCorr);
    std::printf("last hypreIts   : %d\n", result.lastSolveInfo.iterations);
    std::printf("last relRes     : %.12e\n", result.lastSolveInfo.finalRelResNorm);
    std::printf("solution min/max: %.12e  %.12e\n", phiMin, phiMax);
    std::printf("L1 error        : %.12e\n", l1);
    std::printf("L2 error        : %.12e\n", l2);
    std::printf("Linf error      : %.12e\n", linf);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  return 0;
}
