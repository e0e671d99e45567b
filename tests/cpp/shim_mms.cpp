// shim_mms.cpp -- exercises the libpoisson / libscalar C++ headers
// (<anabasis/libpoisson.hpp>, <anabasis/libscalar.hpp>) the way the
// reference's library callers use them, on four manufactured-solution cases:
//
//   poisson <polyMesh>                -lap u = 3 pi^2 u, u = sin sin sin, Dirichlet on
//                                     every patch (apps/poisson_mms_lib/main.cu:115-171)
//   mixed   <polyMesh> <dirichletCsv> u = cos cos cos, Dirichlet on the listed patches and
//                                     the exact normal derivative on the others
//                                     (apps/poisson_mms_mixed_lib/main.cu:150-224)
//   neumann <polyMesh> <refCell>      u = cos cos cos, pure Neumann, source mean removed,
//                                     reference cell pinned to the exact value
//                                     (apps/poisson_mms_neumann_lib/main.cu:157-200)
//   scalar  <polyMesh> <conv> <diff>  div(U phi) - 0.1 lap phi = S, U = (1+y, 0.5-x, 0.25),
//                                     phi = sin sin sin, Dirichlet everywhere; conv is
//                                     central|upwind, diff orth|nonorth
//                                     (apps/scalar_mms_lib/main.cu:16-30, 86-125, 200-260)
//
// Output: one "key value" pair per line (cells, l2, linf, its, relres, plus
// the source mean for the Neumann case), read by tests/test_cpp_shim.py.
#include <anabasis/libpoisson.hpp>
#include <anabasis/libscalar.hpp>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <set>
#include <string>
#include <vector>

namespace {

using Vec3 = std::array<double, 3>;
constexpr double kPi = 3.141592653589793238462643383279502884;

double dot(const Vec3& a, const Vec3& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// u = prod sin(pi x_i) and its gradient
double sin3(const Vec3& x) {
  return std::sin(kPi * x[0]) * std::sin(kPi * x[1]) * std::sin(kPi * x[2]);
}
Vec3 grad_sin3(const Vec3& x) {
  const double s[3] = {std::sin(kPi * x[0]), std::sin(kPi * x[1]), std::sin(kPi * x[2])};
  const double c[3] = {std::cos(kPi * x[0]), std::cos(kPi * x[1]), std::cos(kPi * x[2])};
  return {kPi * c[0] * s[1] * s[2], kPi * s[0] * c[1] * s[2], kPi * s[0] * s[1] * c[2]};
}
// u = prod cos(pi x_i) and its gradient
double cos3(const Vec3& x) {
  return std::cos(kPi * x[0]) * std::cos(kPi * x[1]) * std::cos(kPi * x[2]);
}
Vec3 grad_cos3(const Vec3& x) {
  const double s[3] = {std::sin(kPi * x[0]), std::sin(kPi * x[1]), std::sin(kPi * x[2])};
  const double c[3] = {std::cos(kPi * x[0]), std::cos(kPi * x[1]), std::cos(kPi * x[2])};
  return {-kPi * s[0] * c[1] * c[2], -kPi * c[0] * s[1] * c[2], -kPi * c[0] * c[1] * s[2]};
}

struct Norms {
  double l2 = 0.0, linf = 0.0;
};

// volume-weighted L2 and max-norm of phi - exact over the cells
Norms error_norms(const Mesh& m, const std::vector<double>& phi,
                  const std::function<double(const Vec3&)>& exact) {
  Norms n;
  double vsum = 0.0, e2 = 0.0;
  for (int c = 0; c < m.nCells; ++c) {
    const double e = std::fabs(phi[(size_t)c] - exact(m.cc[(size_t)c]));
    vsum += m.vol[(size_t)c];
    e2 += m.vol[(size_t)c] * e * e;
    n.linf = std::max(n.linf, e);
  }
  n.l2 = std::sqrt(e2 / std::max(vsum, 1e-30));
  return n;
}

std::set<std::string> csv_set(const char* s) {
  std::set<std::string> out;
  std::string cur;
  for (const char* p = s; ; ++p) {
    if (*p == ',' || *p == '\0') {
      if (!cur.empty()) out.insert(cur);
      cur.clear();
      if (*p == '\0') break;
    } else if (*p != ' ') {
      cur.push_back(*p);
    }
  }
  return out;
}

void report(const Mesh& m, const Norms& n, int its, double relres) {
  std::printf("cells %d\nl2 %.17g\nlinf %.17g\nits %d\nrelres %.17g\n", m.nCells, n.l2, n.linf,
              its, relres);
}

// -lap u = 3 pi^2 u for u = exact (sin3 or cos3): Dirichlet data on the
// patches in `dirichlet` (every patch when it is empty), the exact outward
// normal derivative on the others
int run_poisson(const Mesh& m, double (*exact)(const Vec3&), Vec3 (*grad)(const Vec3&),
                const std::set<std::string>& dirichlet) {
  ScalarBCSet bcs;
  for (const auto& name : m.patchNames) {
    if (dirichlet.empty() || dirichlet.count(name))
      bcs.patches.push_back(
          make_dirichlet_patch_bc(name, [exact](const Vec3& x, const Vec3&) { return exact(x); }));
    else
      bcs.patches.push_back(make_neumann_patch_bc(
          name, [grad](const Vec3& x, const Vec3& nrm) { return dot(grad(x), nrm); }));
  }
  std::vector<double> f((size_t)m.nCells);
  for (int c = 0; c < m.nCells; ++c) f[(size_t)c] = 3.0 * kPi * kPi * exact(m.cc[(size_t)c]);
  EllipticOptions opt;
  opt.nNonOrthCorr = 10;
  opt.hypre.relTol = 0.0;
  opt.hypre.absTol = 1e-12;
  opt.hypre.monitor = 0;
  const EllipticResult r = solve_poisson(m, f, bcs, opt);
  report(m, error_norms(m, r.phi, exact), r.lastSolveInfo.iterations,
         r.lastSolveInfo.finalRelResNorm);
  return 0;
}

int run_neumann(const Mesh& m, int refCell) {
  if (refCell < 0 || refCell >= m.nCells) {
    std::fprintf(stderr, "reference cell %d outside [0,%d)\n", refCell, m.nCells);
    return 2;
  }
  ScalarBCSet bcs;
  for (const auto& name : m.patchNames)
    bcs.patches.push_back(make_neumann_patch_bc(
        name, [](const Vec3& x, const Vec3& nrm) { return dot(grad_cos3(x), nrm); }));
  // a pure-Neumann problem is solvable only for a zero-mean source: remove
  // the discrete (volume-weighted) mean of 3 pi^2 u
  std::vector<double> f((size_t)m.nCells);
  double fv = 0.0, v = 0.0;
  for (int c = 0; c < m.nCells; ++c) {
    f[(size_t)c] = 3.0 * kPi * kPi * cos3(m.cc[(size_t)c]);
    fv += f[(size_t)c] * m.vol[(size_t)c];
    v += m.vol[(size_t)c];
  }
  const double mean = fv / std::max(v, 1e-30);
  for (double& x : f) x -= mean;
  EllipticOptions opt;
  opt.nNonOrthCorr = 10;
  opt.useReferenceCell = true;
  opt.referenceCell = refCell;
  opt.referenceValue = cos3(m.cc[(size_t)refCell]);
  opt.hypre.relTol = 0.0;
  opt.hypre.absTol = 1e-12;
  opt.hypre.monitor = 0;
  const EllipticResult r = solve_poisson(m, f, bcs, opt);
  report(m, error_norms(m, r.phi, cos3), r.lastSolveInfo.iterations,
         r.lastSolveInfo.finalRelResNorm);
  std::printf("srcmean %.17g\n", mean);
  return 0;
}

int run_scalar(const Mesh& m, const std::string& conv, const std::string& diff) {
  const double gamma = 0.1;
  auto U = [](const Vec3& x) { return Vec3{1.0 + x[1], 0.5 - x[0], 0.25}; };
  libscalar::ScalarTransportInputs in;
  in.faceFlux.resize((size_t)m.nFaces);
  in.gammaFace.assign((size_t)m.nFaces, gamma);
  for (int f = 0; f < m.nFaces; ++f) in.faceFlux[(size_t)f] = dot(U(m.xf[(size_t)f]), m.Sf[(size_t)f]);
  in.Su.resize((size_t)m.nCells);
  in.Sp.assign((size_t)m.nCells, 0.0);
  for (int c = 0; c < m.nCells; ++c) {
    const Vec3& x = m.cc[(size_t)c];
    // U is divergence-free, so div(U phi) = U . grad phi
    in.Su[(size_t)c] = dot(U(x), grad_sin3(x)) + 3.0 * gamma * kPi * kPi * sin3(x);
  }
  libscalar::ScalarBCSet bcs;
  for (const auto& name : m.patchNames)
    bcs.patches.push_back(libscalar::make_dirichlet_patch_bc(
        name, [](const Vec3& x, const Vec3&) { return sin3(x); }));
  libscalar::ScalarTransportOptions opt;
  if (conv != "central" && conv != "upwind") {
    std::fprintf(stderr, "convection scheme must be central or upwind\n");
    return 2;
  }
  if (diff != "orth" && diff != "nonorth") {
    std::fprintf(stderr, "diffusion scheme must be orth or nonorth\n");
    return 2;
  }
  opt.convectionScheme =
      conv == "upwind" ? libscalar::ConvectionScheme::Upwind : libscalar::ConvectionScheme::Central;
  opt.diffusionScheme =
      diff == "nonorth" ? libscalar::DiffusionScheme::NonOrth : libscalar::DiffusionScheme::Orth;
  opt.nNonOrthCorr = diff == "nonorth" ? 2 : 0;
  opt.relTol = 0.0;
  opt.absTol = 1e-12;
  const libscalar::ScalarTransportResult r =
      libscalar::solve_steady_scalar_transport(m, in, bcs, opt);
  report(m, error_norms(m, r.phi, sin3), r.iterations, r.finalRelRes);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: %s poisson|mixed|neumann|scalar <polyMeshDir> [args]\n", argv[0]);
    return 2;
  }
  const std::string mode = argv[1];
  try {
    const Mesh m = read_openfoam_polymesh(argv[2]);
    if (mode == "poisson") return run_poisson(m, sin3, grad_sin3, {});
    if (mode == "mixed") {
      const auto dir = csv_set(argc > 3 ? argv[3] : "");
      if (dir.empty()) {
        std::fprintf(stderr, "mixed needs at least one Dirichlet patch\n");
        return 2;
      }
      return run_poisson(m, cos3, grad_cos3, dir);
    }
    if (mode == "neumann") return run_neumann(m, argc > 3 ? std::atoi(argv[3]) : 0);
    if (mode == "scalar")
      return run_scalar(m, argc > 3 ? argv[3] : "central", argc > 4 ? argv[4] : "orth");
    std::fprintf(stderr, "unknown mode '%s'\n", mode.c_str());
    return 2;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
}
