// generic_simple_v1 -- the steady SIMPLE solver's command line over
// libanabasis_hip.so: the reference driver's external interface (GSV1
// main.cu:2936-3904) with every solve on the MI355X.
//
//   -case-config <file>   key/value case file + inline velocity/pressure BC
//                         lines (main.cu:162-331); flags given on the
//                         command line override the file's keys
//   -<flag> <value>       the driver's flags (main.cu:333-401)
//
// Output: the start-up report, the patch geometry summaries and the runtime
// BC coverage table, one log line per printed iteration, the final summary
// with the force reports, <outPrefix>_final.vtu (and _iterNNNNNN.vtu
// snapshots) and <outPrefix>_summary.txt, in the reference's formats.
//
// Input errors (case grammar, flags, mesh, BC grammar and coverage) are
// reported before the GPU is touched.
#include <anabasis_hip.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <map>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

using V3 = std::array<double, 3>;

struct Fail : std::runtime_error {
  using std::runtime_error::runtime_error;
};

void check(an_status st, const char* what) {
  if (st != AN_OK) throw Fail(std::string(what) + ": " + an_last_error());
}

// ------------------------------------------------------------- params ---
// Defaults of the reference driver (main.cu:73-101).
struct Params {
  std::string polyMeshDir = "/tmp/meshCase/constant/polyMesh";
  std::string outPrefix = "pipe_poiseuille_gpu";
  double rho = 1.0, mu = 0.05, Re = 1.0, Umean = 1.0, CFL = 1.0;
  bool muExplicit = false;
  double pipeDiameter = 0.05, pipeLength = 0.50;
  std::string wallPatch = "patch_0_0", inletPatch = "patch_1_0", outletPatch = "patch_2_0";
  std::string bcConfigPath;
  int device = 0, velRestart = 40, velMaxit = 100, monitor = 1, writeVtu = 1;
  int nVelNonOrthCorr = 0, nNonOrthCorr = 2, nPressureCorr = 0, nsteps = 1000, printEvery = 25,
      writeEvery = 0;
  double velTol = 1e-10, velRelTol = 0.0, corrTol = 1e-12, tolMass = 1e-10, tolVel = 1e-10;
  double uRelax = 0.7, pRelax = 1.0;
  int pUseAmg = 1, pMaxit = 4000, pAmgMaxit = 1, pAmgRelaxType = 18, pAmgCoarsenType = 8,
      pAmgInterpType = 3, pAmgAggLevels = 1, pAmgAggInterpType = 4, pAmgPmax = 4,
      pAmgKeepTranspose = 0;
  double pTol = 1e-10, pRelTol = 0.0, pAmgTruncFactor = 0.2;
  int profileSteps = 0, pAmgRebuildEvery = 1, pAmgSetupScope = 0;
  int forceEnable = 0;
  std::string forcePatch;
  int forceNormalSign = -1;
  V3 dragDir{{1.0, 0.0, 0.0}}, liftDir{{0.0, 1.0, 0.0}}, spanDir{{0.0, 0.0, 1.0}};
  double forceUref = -1.0, forceAreaRef = -1.0;
};

// case key -> command-line flag (main.cu:184-250)
const std::map<std::string, std::string>& case_keys() {
  static const std::map<std::string, std::string> k = {
      {"polyMeshDir", "-polyMeshDir"}, {"outPrefix", "-out-prefix"}, {"rho", "-rho"},
      {"mu", "-mu"}, {"re", "-re"}, {"Re", "-re"}, {"uMean", "-uMean"}, {"Umean", "-uMean"},
      {"pipeD", "-pipeD"}, {"pipeL", "-pipeL"}, {"wallPatch", "-wall-patch"},
      {"inletPatch", "-inlet-patch"}, {"outletPatch", "-outlet-patch"},
      {"bcConfig", "-bc-config"}, {"device", "-device"}, {"cfl", "-cfl"}, {"CFL", "-cfl"},
      {"nsteps", "-nsteps"}, {"printEvery", "-print-every"}, {"writeVtu", "-write-vtu"},
      {"writeEvery", "-write-every"}, {"monitor", "-monitor"}, {"velRestart", "-vel-restart"},
      {"velMaxit", "-vel-maxit"}, {"velTol", "-vel-tol"}, {"velRelTol", "-vel-reltol"},
      {"nVelNonOrthCorr", "-nVelNonOrthCorr"}, {"nNonOrthCorr", "-nNonOrthCorr"},
      {"nPressureCorr", "-nPressureCorr"}, {"uRelax", "-u-relax"}, {"pRelax", "-p-relax"},
      {"corrTol", "-corrTol"}, {"tolMass", "-tolMass"}, {"tolVel", "-tolVel"},
      {"pUseAmg", "-p-use-amg"}, {"pMaxit", "-p-maxit"}, {"pTol", "-p-tol"},
      {"pRelTol", "-p-reltol"}, {"pAmgSetupScope", "-p-amg-setup-scope"},
      {"pAmgMaxit", "-p-amg-maxit"}, {"pAmgCoarsenType", "-p-amg-coarsen-type"},
      {"pAmgInterpType", "-p-amg-interp-type"}, {"pAmgAggLevels", "-p-amg-agg-levels"},
      {"pAmgAggInterpType", "-p-amg-agg-interp-type"}, {"pAmgRelaxType", "-p-amg-relax-type"},
      {"pAmgPmax", "-p-amg-pmax"}, {"pAmgTruncFactor", "-p-amg-trunc-factor"},
      {"pAmgKeepTranspose", "-p-amg-keep-transpose"},
      {"pAmgRebuildEvery", "-p-amg-rebuild-every"}, {"forceEnable", "-force-enable"},
      {"forcePatch", "-force-patch"}, {"forceNormalSign", "-force-normal-sign"},
      {"forceDragDir", "-force-drag-dir"}, {"forceLiftDir", "-force-lift-dir"},
      {"forceSpanDir", "-force-span-dir"}, {"forceUref", "-force-uref"},
      {"forceAreaRef", "-force-area-ref"}, {"profileSteps", "-profile-steps"},
      {"assemblyBackend", "-assembly-backend"}};
  return k;
}

std::string case_line(std::string s) {
  s = s.substr(0, s.find('#'));
  const auto a = s.find_first_not_of(" \t\r\n\f\v");
  if (a == std::string::npos) return "";
  const auto b = s.find_last_not_of(" \t\r\n\f\v");
  return s.substr(a, b - a + 1);
}

std::vector<std::string> tokens(const std::string& s) {
  std::istringstream is(s);
  std::vector<std::string> t;
  for (std::string w; is >> w;) t.push_back(w);
  return t;
}

// -case-config: the file's keys become flags placed before the command
// line's own (so the command line wins); inline BC lines go to
// <case>.generated.bc, passed as -bc-config (main.cu:162-331).
std::vector<std::string> expand_case(int argc, char** argv) {
  std::string casePath;
  for (int i = 1; i < argc; ++i)
    if (!std::strcmp(argv[i], "-case-config")) {
      if (i + 1 >= argc) throw Fail("Missing value after -case-config");
      casePath = argv[i + 1];
      break;
    }
  std::vector<std::string> out{argv[0]};
  if (casePath.empty()) {
    for (int i = 1; i < argc; ++i) out.emplace_back(argv[i]);
    return out;
  }
  std::ifstream in(casePath);
  if (!in) throw Fail("Could not open case config file: " + casePath);
  std::vector<std::string> bcLines;
  bool bcKey = false;
  std::string raw;
  for (int lineNo = 1; std::getline(in, raw); ++lineNo) {
    const std::string line = case_line(raw);
    if (line.empty()) continue;
    const auto t = tokens(line);
    if (t[0] == "velocity" || t[0] == "pressure") {
      bcLines.push_back(line);
      continue;
    }
    const std::string where =
        "Case config parse error in '" + casePath + "' at line " + std::to_string(lineNo) + ": ";
    if (t.size() != 2) throw Fail(where + "expected '<key> <value>' or a velocity/pressure BC line");
    const auto it = case_keys().find(t[0]);
    if (it == case_keys().end()) throw Fail(where + "unknown key '" + t[0] + "'");
    bcKey = bcKey || t[0] == "bcConfig";
    out.push_back(it->second);
    out.push_back(t[1]);
  }
  if (!bcLines.empty() && bcKey)
    throw Fail("Case config cannot contain both bcConfig and inline velocity/pressure BC lines");
  if (!bcLines.empty()) {
    const std::string gen = casePath + ".generated.bc";
    std::ofstream bc(gen);
    if (!bc) throw Fail("Could not write generated BC file: " + gen);
    bc << "# Auto-generated from " << casePath << "\n";
    for (const auto& l : bcLines) bc << l << "\n";
    out.push_back("-bc-config");
    out.push_back(gen);
  }
  for (int i = 1; i < argc; ++i) {
    if (!std::strcmp(argv[i], "-case-config")) {
      ++i;
      continue;
    }
    out.emplace_back(argv[i]);
  }
  return out;
}

V3 vec3_arg(const std::string& raw, const char* opt) {
  std::string t = raw;
  for (char& c : t)
    if (c == ',' || c == ';' || c == ':') c = ' ';
  std::istringstream is(t);
  V3 v{{0.0, 0.0, 0.0}};
  if (!(is >> v[0] >> v[1] >> v[2]))
    throw Fail(std::string("Could not parse ") + opt + " value '" + raw + "'. Expected x,y,z");
  return v;
}

// The driver's flags (main.cu:333-401); unknown flags are ignored, as there.
void parse_flags(const std::vector<std::string>& a, Params& p) {
  const size_t n = a.size();
  for (size_t i = 1; i < n; ++i) {
    const std::string& f = a[i];
    auto val = [&]() -> const std::string& {
      if (i + 1 >= n) throw Fail("Missing value after " + f);
      return a[++i];
    };
    auto dbl = [&](double& d) { d = std::atof(val().c_str()); };
    auto in = [&](int& k) { k = std::atoi(val().c_str()); };
    if (f == "-polyMeshDir") p.polyMeshDir = val();
    else if (f == "-out-prefix") p.outPrefix = val();
    else if (f == "-rho") dbl(p.rho);
    else if (f == "-mu") { dbl(p.mu); p.muExplicit = true; }
    else if (f == "-re") dbl(p.Re);
    else if (f == "-uMean" || f == "-lidU") dbl(p.Umean);
    else if (f == "-pipeD") dbl(p.pipeDiameter);
    else if (f == "-pipeL") dbl(p.pipeLength);
    else if (f == "-wall-patch") p.wallPatch = val();
    else if (f == "-inlet-patch" || f == "-lid-patch") p.inletPatch = val();
    else if (f == "-outlet-patch") p.outletPatch = val();
    else if (f == "-bc-config") p.bcConfigPath = val();
    else if (f == "-force-enable") in(p.forceEnable);
    else if (f == "-force-patch") p.forcePatch = val();
    else if (f == "-force-normal-sign") in(p.forceNormalSign);
    else if (f == "-force-drag-dir") p.dragDir = vec3_arg(val(), "-force-drag-dir");
    else if (f == "-force-lift-dir") p.liftDir = vec3_arg(val(), "-force-lift-dir");
    else if (f == "-force-span-dir") p.spanDir = vec3_arg(val(), "-force-span-dir");
    else if (f == "-force-uref") dbl(p.forceUref);
    else if (f == "-force-area-ref") dbl(p.forceAreaRef);
    else if (f == "-cfl") dbl(p.CFL);
    else if (f == "-device") in(p.device);
    else if (f == "-vel-restart") in(p.velRestart);
    else if (f == "-vel-maxit") in(p.velMaxit);
    else if (f == "-vel-tol") dbl(p.velTol);
    else if (f == "-vel-reltol") dbl(p.velRelTol);
    else if (f == "-nVelNonOrthCorr") in(p.nVelNonOrthCorr);
    else if (f == "-nNonOrthCorr") in(p.nNonOrthCorr);
    else if (f == "-nPressureCorr") in(p.nPressureCorr);
    else if (f == "-u-relax") dbl(p.uRelax);
    else if (f == "-p-relax") dbl(p.pRelax);
    else if (f == "-corrTol") dbl(p.corrTol);
    else if (f == "-monitor") in(p.monitor);
    else if (f == "-write-vtu") in(p.writeVtu);
    else if (f == "-nsteps") in(p.nsteps);
    else if (f == "-print-every") in(p.printEvery);
    else if (f == "-write-every") in(p.writeEvery);
    else if (f == "-tolMass") dbl(p.tolMass);
    else if (f == "-tolVel") dbl(p.tolVel);
    else if (f == "-p-use-amg") in(p.pUseAmg);
    else if (f == "-p-maxit") in(p.pMaxit);
    else if (f == "-p-tol") dbl(p.pTol);
    else if (f == "-p-reltol") dbl(p.pRelTol);
    else if (f == "-p-amg-maxit") in(p.pAmgMaxit);
    else if (f == "-p-amg-relax-type") in(p.pAmgRelaxType);
    else if (f == "-p-amg-coarsen-type") in(p.pAmgCoarsenType);
    else if (f == "-p-amg-interp-type") in(p.pAmgInterpType);
    else if (f == "-p-amg-agg-levels") in(p.pAmgAggLevels);
    else if (f == "-p-amg-agg-interp-type") in(p.pAmgAggInterpType);
    else if (f == "-p-amg-pmax") in(p.pAmgPmax);
    else if (f == "-p-amg-trunc-factor") dbl(p.pAmgTruncFactor);
    else if (f == "-p-amg-keep-transpose") in(p.pAmgKeepTranspose);
    else if (f == "-profile-steps") in(p.profileSteps);
    else if (f == "-p-amg-rebuild-every") in(p.pAmgRebuildEvery);
    else if (f == "-p-amg-setup-scope") {
      const std::string v = val();
      if (v == "outer") p.pAmgSetupScope = 0;
      else if (v == "pressure" || v == "solve") p.pAmgSetupScope = 1;
      else throw Fail("Unknown -p-amg-setup-scope '" + v + "'. Use outer or pressure.");
    }
  }
}

// --------------------------------------------------------------- mesh ---
struct Mesh {
  an_mesh_t h = nullptr;
  an_mesh_info info{};
  std::vector<double> P, cc, vol, xf, nf, Af;
  std::vector<int> faceOff, faceVerts, owner, neigh;
  std::vector<std::string> patchNames;
  std::vector<int> patchStart, patchN;
  ~Mesh() {
    if (h) an_mesh_destroy(h);
  }
  V3 at(const std::vector<double>& a, int i) const {
    return {a[3 * (size_t)i], a[3 * (size_t)i + 1], a[3 * (size_t)i + 2]};
  }
  int patch(const std::string& name) const {
    for (size_t k = 0; k < patchNames.size(); ++k)
      if (patchNames[k] == name) return (int)k;
    return -1;
  }
};

template <typename T>
void get_array(const Mesh& m, const char* name, std::vector<T>& v, size_t n) {
  v.resize(n);
  check(an_mesh_get_array(m.h, name, v.data()), name);
}

void load_mesh(const std::string& dir, Mesh& m) {
  check(an_mesh_read_polymesh(dir.c_str(), &m.h), "read_openfoam_polymesh");
  check(an_mesh_get_info(m.h, &m.info), "mesh info");
  const an_mesh_info& i = m.info;
  get_array(m, "points", m.P, 3 * (size_t)i.nPoints);
  get_array(m, "faceOffsets", m.faceOff, (size_t)i.nFaces + 1);
  get_array(m, "faceVerts", m.faceVerts, (size_t)m.faceOff.back());
  get_array(m, "owner", m.owner, (size_t)i.nFaces);
  get_array(m, "neigh", m.neigh, (size_t)i.nInternalFaces);
  get_array(m, "cc", m.cc, 3 * (size_t)i.nCells);
  get_array(m, "vol", m.vol, (size_t)i.nCells);
  get_array(m, "xf", m.xf, 3 * (size_t)i.nFaces);
  get_array(m, "nf", m.nf, 3 * (size_t)i.nFaces);
  get_array(m, "Af", m.Af, (size_t)i.nFaces);
  for (int k = 0; k < i.nPatches; ++k) {
    char name[256];
    int s = 0, n = 0;
    check(an_mesh_get_patch(m.h, k, name, sizeof name, &s, &n), "patch");
    m.patchNames.emplace_back(name);
    m.patchStart.push_back(s);
    m.patchN.push_back(n);
  }
}

// ----------------------------------------------------------- geometry ---
double dot(const V3& a, const V3& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
double norm(const V3& a) { return std::sqrt(dot(a, a)); }
V3 sub(const V3& a, const V3& b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
V3 scale(double s, const V3& a) { return {s * a[0], s * a[1], s * a[2]}; }
V3 unit_or_zero(const V3& v) {
  const double m = norm(v);
  return m <= 1.0e-300 ? V3{{0.0, 0.0, 0.0}} : scale(1.0 / m, v);
}

void print_patch_summary(const Mesh& m, int k) {
  double g[18];
  check(an_patch_geometry(m.h, k, g), "patch geometry");
  std::printf("Patch summary [%s]\n", m.patchNames[(size_t)k].c_str());
  std::printf("  faces      : %d\n", (int)g[0]);
  std::printf("  area       : %.12e\n", g[1]);
  std::printf("  centroid   : [%.6e, %.6e, %.6e]\n", g[2], g[3], g[4]);
  std::printf("  avg normal : [%.6e, %.6e, %.6e]\n", g[8], g[9], g[10]);
  std::printf("  planarity  : %.6f\n", g[11]);
}

// ------------------------------------------------------------- forces ---
// Reports of main.cu:843-973 and 2687-2933 (the same per-face formulas):
// the DFG paper-formula cylinder force (host, per face), the vector
// wall-shear traction on the cylinder and on a chosen patch (reduced on
// the GPU by an_simple_patch_forces), with projections and coefficients.
struct PaperForce {
  bool valid = false;
  std::string patch;
  int nFaces = 0;
  double area = 0, FDp = 0, FDv = 0, FD = 0, FLp = 0, FLv = 0, FL = 0, CD = 0, CL = 0;
  double rho = 0, mu = 0, Ubar = 0, D = 0, H = 0, denom = 0;
  double dnMin = 1e300, dnMax = 0, maxDvtDn = 0;
};

PaperForce paper_force(const Mesh& m, int k, const std::vector<double>& u,
                       const std::vector<double>& v, const std::vector<double>& w,
                       const std::vector<double>& p, double rho, double mu, double D, double H,
                       double Ubar) {
  PaperForce r;
  if (k < 0) return r;
  r.valid = true;
  r.patch = m.patchNames[(size_t)k];
  r.rho = rho;
  r.mu = mu;
  r.D = D;
  r.H = H;
  r.Ubar = Ubar;
  r.denom = rho * Ubar * Ubar * D * H;
  for (int f = m.patchStart[(size_t)k]; f < m.patchStart[(size_t)k] + m.patchN[(size_t)k]; ++f) {
    const int P = m.owner[(size_t)f];
    const double A = m.Af[(size_t)f];
    if (A <= 1.0e-300) continue;
    V3 n = scale(-1.0, m.at(m.nf, f));  // solid -> fluid
    const double nm = norm(n);
    if (nm <= 1.0e-300) continue;
    n = scale(1.0 / nm, n);
    V3 t{{n[1], -n[0], 0.0}};
    const double tm = norm(t);
    if (tm <= 1.0e-14) continue;
    t = scale(1.0 / tm, t);
    const double vt = u[(size_t)P] * t[0] + v[(size_t)P] * t[1] + w[(size_t)P] * t[2];
    const V3 dx = sub(m.at(m.cc, P), m.at(m.xf, f));
    double dn = dot(dx, n);
    if (dn <= 1.0e-14) dn = std::fabs(dn);
    if (dn <= 1.0e-14) dn = norm(dx);
    if (dn <= 1.0e-14) continue;
    const double dvt = vt / dn, pf = p[(size_t)P];
    r.FDp += (-pf * n[0]) * A;
    r.FDv += (mu * dvt * n[1]) * A;
    r.FLp += (-pf * n[1]) * A;
    r.FLv += (-mu * dvt * n[0]) * A;
    r.area += A;
    r.nFaces += 1;
    r.dnMin = std::min(r.dnMin, dn);
    r.dnMax = std::max(r.dnMax, dn);
    r.maxDvtDn = std::max(r.maxDvtDn, std::fabs(dvt));
  }
  r.FD = r.FDp + r.FDv;
  r.FL = r.FLp + r.FLv;
  if (r.denom > 1.0e-300) {
    r.CD = 2.0 * r.FD / r.denom;
    r.CL = 2.0 * r.FL / r.denom;
  }
  if (r.dnMin == 1.0e300) r.dnMin = 0.0;
  return r;
}

// an_simple_patch_forces out[15]: Fp[3], Fv[3], F[3], area, nFaces, dnMin,
// dnMax, maxUt, maxShear
struct Traction {
  bool valid = false;
  std::string patch;
  V3 Fp{}, Fv{}, F{};
  double area = 0, dnMin = 0, dnMax = 0, maxUt = 0, maxShear = 0;
  int nFaces = 0;
};

Traction traction(an_simple_t s, const Mesh& m, int k, int sign) {
  Traction t;
  if (k < 0) return t;
  double o[15];
  check(an_simple_patch_forces(s, k, sign, o), "patch forces");
  t.valid = true;
  t.patch = m.patchNames[(size_t)k];
  t.Fp = {o[0], o[1], o[2]};
  t.Fv = {o[3], o[4], o[5]};
  t.F = {o[6], o[7], o[8]};
  t.area = o[9];
  t.nFaces = (int)o[10];
  t.dnMin = o[11];
  t.dnMax = o[12];
  t.maxUt = o[13];
  t.maxShear = o[14];
  return t;
}

// ---------------------------------------------------------------- VTU ---
// One polyhedron (VTK type 42) per cell: its faces in the order owner
// faces then neighbour faces, a neighbour's face reversed; points listed
// once per cell in first-seen order (main.cu:975-983).
void write_vtu(const std::string& file, const Mesh& m, const std::vector<std::string>& names,
               const std::vector<const std::vector<double>*>& data,
               const std::vector<double>& u, const std::vector<double>& v,
               const std::vector<double>& w) {
  const int nC = m.info.nCells, nF = m.info.nFaces, nIF = m.info.nInternalFaces;
  std::vector<std::vector<int>> cellFaces((size_t)nC);
  std::vector<std::vector<char>> cellRev((size_t)nC);
  for (int f = 0; f < nF; ++f) {
    cellFaces[(size_t)m.owner[(size_t)f]].push_back(f);
    cellRev[(size_t)m.owner[(size_t)f]].push_back(0);
  }
  for (int f = 0; f < nIF; ++f) {
    cellFaces[(size_t)m.neigh[(size_t)f]].push_back(f);
    cellRev[(size_t)m.neigh[(size_t)f]].push_back(1);
  }
  std::vector<long long> conn, offsets((size_t)nC), faces, faceOffsets((size_t)nC);
  long long nConn = 0, nFaceWords = 0;
  for (int c = 0; c < nC; ++c) {
    std::set<int> seen;
    faces.push_back((long long)cellFaces[(size_t)c].size());
    ++nFaceWords;
    for (size_t j = 0; j < cellFaces[(size_t)c].size(); ++j) {
      const int f = cellFaces[(size_t)c][j];
      std::vector<int> fv(m.faceVerts.begin() + m.faceOff[(size_t)f],
                          m.faceVerts.begin() + m.faceOff[(size_t)f + 1]);
      if (cellRev[(size_t)c][j]) std::reverse(fv.begin(), fv.end());
      faces.push_back((long long)fv.size());
      for (int pt : fv) {
        faces.push_back(pt);
        if (seen.insert(pt).second) {
          conn.push_back(pt);
          ++nConn;
        }
      }
      nFaceWords += 1 + (long long)fv.size();
    }
    offsets[(size_t)c] = nConn;
    faceOffsets[(size_t)c] = nFaceWords;
  }
  std::ofstream o(file);
  if (!o) throw Fail("Could not open " + file + " for writing");
  o << std::setprecision(15);
  o << "<?xml version=\"1.0\"?>\n<VTKFile type=\"UnstructuredGrid\" version=\"1.0\" "
       "byte_order=\"LittleEndian\">\n  <UnstructuredGrid>\n    <Piece NumberOfPoints=\""
    << m.info.nPoints << "\" NumberOfCells=\"" << nC << "\">\n      <Points>\n"
    << "        <DataArray type=\"Float64\" NumberOfComponents=\"3\" format=\"ascii\">\n";
  for (int i = 0; i < m.info.nPoints; ++i)
    o << "          " << m.P[3 * (size_t)i] << " " << m.P[3 * (size_t)i + 1] << " "
      << m.P[3 * (size_t)i + 2] << "\n";
  auto ints = [&](const char* type, const char* name, const std::vector<long long>& a) {
    o << "        <DataArray type=\"" << type << "\" Name=\"" << name
      << "\" format=\"ascii\">\n          ";
    for (long long x : a) o << x << " ";
    o << "\n        </DataArray>\n";
  };
  o << "        </DataArray>\n      </Points>\n      <Cells>\n";
  ints("Int64", "connectivity", conn);
  ints("Int64", "offsets", offsets);
  ints("UInt8", "types", std::vector<long long>((size_t)nC, 42));
  ints("Int64", "faces", faces);
  ints("Int64", "faceoffsets", faceOffsets);
  o << "      </Cells>\n      <CellData>\n";
  o << "        <DataArray type=\"Float64\" Name=\"U\" NumberOfComponents=\"3\" "
       "format=\"ascii\">\n";
  for (int c = 0; c < nC; ++c)
    o << "          " << u[(size_t)c] << " " << v[(size_t)c] << " " << w[(size_t)c] << "\n";
  o << "        </DataArray>\n";
  for (size_t k = 0; k < names.size(); ++k) {
    o << "        <DataArray type=\"Float64\" Name=\"" << names[k]
      << "\" NumberOfComponents=\"1\" format=\"ascii\">\n";
    for (double x : *data[k]) o << "          " << x << "\n";
    o << "        </DataArray>\n";
  }
  o << "      </CellData>\n    </Piece>\n  </UnstructuredGrid>\n</VTKFile>\n";
}

std::vector<double> field(an_simple_t s, const char* name, size_t n) {
  std::vector<double> v(n);
  check(an_simple_get_field(s, name, v.data(), AN_HOST), name);
  return v;
}

// --------------------------------------------------------------- main ---
int run(int argc, char** argv) {
  Params par;
  parse_flags(expand_case(argc, argv), par);
  if (par.nPressureCorr > 0)
    throw Fail("nPressureCorr > 0 (extra pressure corrections after the non-orthogonal "
               "passes) is not supported; use nPressureCorr 0");

  Mesh mesh;
  load_mesh(par.polyMeshDir, mesh);
  const an_mesh_info& mi = mesh.info;
  double bb[6] = {1e300, -1e300, 1e300, -1e300, 1e300, -1e300}, vmin = 1e300, vmax = -1e300;
  for (int i = 0; i < mi.nPoints; ++i)
    for (int d = 0; d < 3; ++d) {
      bb[2 * d] = std::min(bb[2 * d], mesh.P[3 * (size_t)i + (size_t)d]);
      bb[2 * d + 1] = std::max(bb[2 * d + 1], mesh.P[3 * (size_t)i + (size_t)d]);
    }
  for (double v : mesh.vol) {
    vmin = std::min(vmin, v);
    vmax = std::max(vmax, v);
  }
  const int wallPatch = mesh.patch(par.wallPatch), inletPatch = mesh.patch(par.inletPatch),
            outletPatch = mesh.patch(par.outletPatch), cylPatch = mesh.patch("patch_3_0");
  if (wallPatch < 0 || inletPatch < 0 || outletPatch < 0)
    throw Fail("Could not find wall/inlet/outlet patch.");
  const double mu = par.muExplicit ? par.mu : par.rho * par.Umean * par.pipeDiameter / par.Re;

  std::printf("============================================================\n");
  std::printf("generic_simple_v1: OpenFOAM polyMesh SIMPLE solver\n");
  std::printf("MI355X: assembly, gradients, fluxes, BiCGSTAB momentum and PCG + aggregation "
              "AMG pressure solves on the GPU\n");
  std::printf("============================================================\n");
  std::printf("polyMeshDir : %s\n", par.polyMeshDir.c_str());
  std::printf("outPrefix   : %s\n", par.outPrefix.c_str());
  std::printf("------------------------------------------------------------\n");
  std::printf("Points         : %d\n", mi.nPoints);
  std::printf("Faces          : %d\n", mi.nFaces);
  std::printf("Internal faces : %d\n", mi.nInternalFaces);
  std::printf("Cells          : %d\n", mi.nCells);
  std::printf("BBox           : [%g, %g] x [%g, %g] x [%g, %g]\n", bb[0], bb[1], bb[2], bb[3],
              bb[4], bb[5]);
  std::printf("Volume min/max : %.8e / %.8e\n", vmin, vmax);
  std::printf("maxNonOrthDeg  : %.6f\n", mi.maxNonOrthDeg);
  std::printf("Momentum solve : BiCGSTAB + Jacobi (u, v, w fused, one matrix)\n");
  std::printf("Pressure solve : PCG + aggregation AMG%s\n",
              par.pUseAmg ? "" : " (p-use-amg 0 accepted; the AMG stays on)");
  std::printf("rho            : %.8g\n", par.rho);
  std::printf("Re             : %.8g\n", par.Re);
  std::printf("Umean          : %.8g\n", par.Umean);
  std::printf("pipeDiameter   : %.8g\n", par.pipeDiameter);
  std::printf("pipeLength     : %.8g\n", par.pipeLength);
  std::printf("wall/in/out    : %s / %s / %s\n", par.wallPatch.c_str(), par.inletPatch.c_str(),
              par.outletPatch.c_str());
  std::printf("mu             : %.8e\n", mu);
  std::printf("mu source      : %s\n",
              par.muExplicit ? "direct -mu / case mu" : "legacy rho*Umean*pipeD/Re");
  std::printf("steady mode    : ON (no pseudo-time mass term)\n");
  std::printf("nsteps         : %d\n", par.nsteps);
  std::printf("nVelNonOrthCorr: %d\n", par.nVelNonOrthCorr);
  std::printf("nNonOrthCorr   : %d\n", par.nNonOrthCorr);
  std::printf("nPressureCorr  : %d\n", par.nPressureCorr);
  if (par.profileSteps > 0) std::printf("profileSteps   : %d\n", par.profileSteps);
  std::printf("velTol / velRelTol : %.3e / %.3e\n", par.velTol, par.velRelTol);
  std::printf("pTol   / pRelTol   : %.3e / %.3e\n", par.pTol, par.pRelTol);
  std::printf("pAmgRebuildEvery   : %d\n", par.pAmgRebuildEvery);
  std::printf("------------------------------------------------------------\n");

  if (par.bcConfigPath.empty())
    throw Fail("generic_simple_v1 requires explicit runtime BCs.\n"
               "Add velocity/pressure lines to the case file, or set bcConfig.");
  an_bc_t bc = nullptr;
  check(an_bc_load(par.bcConfigPath.c_str(), &bc), "BC config");
  struct BcGuard {
    an_bc_t b;
    ~BcGuard() { an_bc_destroy(b); }
  } bcGuard{bc};
  std::printf("------------------------------------------------------------\n");
  std::printf("All mesh patches detected in polyMesh:\n");
  for (int k = 0; k < mi.nPatches; ++k) print_patch_summary(mesh, k);
  std::printf("------------------------------------------------------------\n");
  std::printf("Runtime BC patch coverage:\n");
  int len = 0, complete = 0;
  check(an_bc_coverage(mesh.h, bc, 0, nullptr, 0, &len, &complete), "BC coverage");
  std::vector<char> text((size_t)len + 1);
  check(an_bc_coverage(mesh.h, bc, 0, text.data(), len + 1, &len, &complete), "BC coverage");
  std::fputs(text.data(), stdout);
  if (!complete) {
    check(an_bc_coverage(mesh.h, bc, 1, nullptr, 0, &len, &complete), "BC coverage");
    text.assign((size_t)len + 1, 0);
    check(an_bc_coverage(mesh.h, bc, 1, text.data(), len + 1, &len, &complete), "BC coverage");
    std::fflush(stdout);
    std::fputs(text.data(), stderr);
    return 1;
  }
  const int nP = mi.nPatches, nF = mi.nFaces, nC = mi.nCells;
  std::vector<int> tU((size_t)nP), tP((size_t)nP);
  std::vector<double> ub((size_t)nF), vb((size_t)nF), wb((size_t)nF), pb((size_t)nF);
  check(an_bc_apply(mesh.h, bc, tU.data(), tP.data(), ub.data(), vb.data(), wb.data(),
                    pb.data()),
        "BC evaluation");
  bool pinned = false;  // a fixed pressure somewhere (bc_specs.cu:143-149)
  for (int k = 0; k < nP; ++k) pinned = pinned || tP[(size_t)k] == AN_BC_DIRICHLET;
  print_patch_summary(mesh, wallPatch);
  if (cylPatch >= 0) print_patch_summary(mesh, cylPatch);
  print_patch_summary(mesh, inletPatch);
  print_patch_summary(mesh, outletPatch);
  std::printf("Pressure reference required: %s\n", pinned ? "no" : "yes");
  std::printf("BC config source         : %s\n", par.bcConfigPath.c_str());
  std::fflush(stdout);

  // ------------------------------------------------------------- GPU ---
  check(an_set_device(par.device), "device");
  std::printf("Running on HIP device %d\n", par.device);
  an_simple_params sp;
  an_simple_default_params(&sp);
  sp.rho = par.rho;
  sp.mu = mu;
  sp.nVelNonOrthCorr = par.nVelNonOrthCorr;
  sp.nNonOrthCorr = par.nNonOrthCorr;
  sp.uRelax = par.uRelax;
  sp.pRelax = par.pRelax;
  sp.corrTol = par.corrTol;
  sp.tolMass = par.tolMass;
  sp.tolVel = par.tolVel;
  sp.velMaxit = par.velMaxit;
  sp.velTol = par.velTol;
  sp.velRelTol = par.velRelTol;
  sp.pMaxit = par.pMaxit;
  sp.pTol = par.pTol;
  sp.pRelTol = par.pRelTol;
  sp.pAmgRebuildEvery = par.pAmgRebuildEvery;
  an_simple_t s = nullptr;
  check(an_simple_create(mesh.h, &sp, tU.data(), tP.data(), ub.data(), vb.data(), wb.data(),
                         pb.data(), inletPatch, &s),
        "SIMPLE set-up");
  struct SimpleGuard {
    an_simple_t s;
    ~SimpleGuard() { an_simple_destroy(s); }
  } simpleGuard{s};

  using clock = std::chrono::steady_clock;
  const auto runStart = clock::now();
  const int maxSteps = par.profileSteps > 0 ? std::min(par.nsteps, par.profileSteps) : par.nsteps;
  an_simple_stats st{};
  int stepConverged = 0;
  for (int step = 1; step <= maxSteps; ++step) {
    const auto t0 = clock::now();
    int conv = 0;
    const an_status rc = an_simple_iterate(s, 1, &st, &conv);
    if (rc != AN_OK && rc != AN_EDIVERGED) check(rc, "SIMPLE iteration");
    const double iterWall = std::chrono::duration<double>(clock::now() - t0).count();
    const double totalWall = std::chrono::duration<double>(clock::now() - runStart).count();
    if (step == 1 || (par.printEvery > 0 && step % par.printEvery == 0))
      std::printf("iter %4d : massRes = %.3e, duRel = %.3e, dvRel = %.3e, dwRel = %.3e, dpRel = "
                  "%.3e, bicgstab=[%d %d %d], pcgLast=%d, pcgTot=%d, iterWall = %.3e s, "
                  "totalWall = %.3e s\n",
                  step, st.massRes, st.duRel, st.dvRel, st.dwRel, st.dpRel, st.bicgstabIts[0],
                  st.bicgstabIts[1], st.bicgstabIts[2], st.pcgLast, st.pcgTotal, iterWall,
                  totalWall);
    if (par.writeEvery > 0 && step % par.writeEvery == 0 && par.writeVtu) {
      const auto u = field(s, "u", (size_t)nC), v = field(s, "v", (size_t)nC),
                 w = field(s, "w", (size_t)nC), p = field(s, "p", (size_t)nC),
                 dc = field(s, "div", (size_t)nC);
      std::vector<double> um((size_t)nC);
      for (int c = 0; c < nC; ++c)
        um[(size_t)c] = std::sqrt(u[(size_t)c] * u[(size_t)c] + v[(size_t)c] * v[(size_t)c] +
                                  w[(size_t)c] * w[(size_t)c]);
      std::ostringstream name;
      name << par.outPrefix << "_iter" << std::setw(6) << std::setfill('0') << step << ".vtu";
      write_vtu(name.str(), mesh, {"p", "umag", "cell_volume", "divCorr"},
                {&p, &um, &mesh.vol, &dc}, u, v, w);
    }
    if (rc == AN_EDIVERGED || !std::isfinite(st.massRes) || st.massRes > 1e50) {
      std::fprintf(stderr, "SIMPLE-like steady solve diverged.\n");
      break;
    }
    stepConverged = step;
    if (conv) {
      std::printf("Converged at iteration %d : massRes = %.3e\n", step, st.massRes);
      break;
    }
  }
  const double loopWall = std::chrono::duration<double>(clock::now() - runStart).count();

  const auto u = field(s, "u", (size_t)nC), v = field(s, "v", (size_t)nC),
             w = field(s, "w", (size_t)nC), p = field(s, "p", (size_t)nC),
             dc = field(s, "div", (size_t)nC);
  std::vector<double> umag((size_t)nC);
  double umax = 0, vmaxf = 0, wmaxf = 0, pmax = 0;
  for (int c = 0; c < nC; ++c) {
    const size_t i = (size_t)c;
    umag[i] = std::sqrt(u[i] * u[i] + v[i] * v[i] + w[i] * w[i]);
    umax = std::max(umax, std::fabs(u[i]));
    vmaxf = std::max(vmaxf, std::fabs(v[i]));
    wmaxf = std::max(wmaxf, std::fabs(w[i]));
    pmax = std::max(pmax, std::fabs(p[i]));
  }
  // cylinder reports: channel height 0.41 and Ubar = 4/9 Umean as in the
  // reference (main.cu:3647-3648)
  const double benchH = 0.41, Ubar = (4.0 / 9.0) * par.Umean;
  const PaperForce cyl =
      paper_force(mesh, cylPatch, u, v, w, p, par.rho, mu, par.pipeDiameter, benchH, Ubar);
  const Traction cylVec = traction(s, mesh, cylPatch, -1);
  const double cylDenom = par.rho * Ubar * Ubar * par.pipeDiameter * benchH;
  V3 cylC{{0.0, 0.0, 0.0}};
  if (cylVec.valid && cylDenom > 1.0e-300)
    for (int d = 0; d < 3; ++d) cylC[(size_t)d] = 2.0 * cylVec.F[(size_t)d] / cylDenom;
  const int forcePatch = mesh.patch(par.forcePatch);
  const Traction pf = par.forceEnable ? traction(s, mesh, forcePatch, par.forceNormalSign)
                                      : Traction{};

  std::printf("\nFinal summary:\n");
  std::printf("------------------------------------------------------------\n");
  std::printf("Iterations    = %d\n", stepConverged);
  std::printf("massRes       = %.8e\n", st.massRes);
  std::printf("max|u|        = %.8e\n", umax);
  std::printf("max|v|        = %.8e\n", vmaxf);
  std::printf("max|w|        = %.8e\n", wmaxf);
  std::printf("max|p|        = %.8e\n", pmax);
  if (cyl.valid) {
    std::printf("------------------------------------------------------------\n");
    std::printf("Cylinder force postprocess, paper formula\n");
    std::printf("force patch   = %s\n", cyl.patch.c_str());
    std::printf("force faces   = %d\n", cyl.nFaces);
    std::printf("force area    = %.12e\n", cyl.area);
    std::printf("rho, mu       = %.12e  %.12e\n", cyl.rho, cyl.mu);
    std::printf("Ubar, D, H    = %.12e  %.12e  %.12e\n", cyl.Ubar, cyl.D, cyl.H);
    std::printf("coeff denom   = rho*Ubar^2*D*H = %.12e\n", cyl.denom);
    std::printf("FD pressure   = %.12e\n", cyl.FDp);
    std::printf("FD viscous    = %.12e\n", cyl.FDv);
    std::printf("FD total      = %.12e\n", cyl.FD);
    std::printf("FL pressure   = %.12e\n", cyl.FLp);
    std::printf("FL viscous    = %.12e\n", cyl.FLv);
    std::printf("FL total      = %.12e\n", cyl.FL);
    std::printf("CD            = %.12e\n", cyl.CD);
    std::printf("CL            = %.12e\n", cyl.CL);
    std::printf("wall dn min/max = %.12e / %.12e\n", cyl.dnMin, cyl.dnMax);
    std::printf("max|dvt/dn|   = %.12e\n", cyl.maxDvtDn);
  } else {
    std::printf("Cylinder force postprocess skipped: patch_3_0 not found.\n");
  }
  if (cylVec.valid) {
    std::printf("------------------------------------------------------------\n");
    std::printf("Cylinder force postprocess, vector wall-shear traction\n");
    std::printf("force patch   = %s\n", cylVec.patch.c_str());
    std::printf("force faces   = %d\n", cylVec.nFaces);
    std::printf("force area    = %.12e\n", cylVec.area);
    std::printf("rho, mu       = %.12e  %.12e\n", par.rho, mu);
    std::printf("Ubar, D, H    = %.12e  %.12e  %.12e\n", Ubar, par.pipeDiameter, benchH);
    std::printf("coeff denom   = rho*Ubar^2*D*H = %.12e\n", cylDenom);
    std::printf("Fp vector     = [%.12e, %.12e, %.12e]\n", cylVec.Fp[0], cylVec.Fp[1],
                cylVec.Fp[2]);
    std::printf("Fv vector     = [%.12e, %.12e, %.12e]\n", cylVec.Fv[0], cylVec.Fv[1],
                cylVec.Fv[2]);
    std::printf("F  vector     = [%.12e, %.12e, %.12e]\n", cylVec.F[0], cylVec.F[1],
                cylVec.F[2]);
    std::printf("C  vector     = [%.12e, %.12e, %.12e]\n", cylC[0], cylC[1], cylC[2]);
    std::printf("CD_vector     = %.12e\n", cylC[0]);
    std::printf("CL_y_vector   = %.12e\n", cylC[1]);
    std::printf("CL_z_vector   = %.12e\n", cylC[2]);
    std::printf("wall dn min/max = %.12e / %.12e\n", cylVec.dnMin, cylVec.dnMax);
    std::printf("max|Ut|       = %.12e\n", cylVec.maxUt);
    std::printf("max shear mag = %.12e\n", cylVec.maxShear);
  }
  // patch force report with drag/lift/span projections (main.cu:2789-2933)
  const V3 dragDir = unit_or_zero(par.dragDir), liftDir = unit_or_zero(par.liftDir),
           spanDir = unit_or_zero(par.spanDir);
  const double pfDenom = (par.rho > 0.0 && par.forceUref > 0.0 && par.forceAreaRef > 0.0)
                             ? par.rho * par.forceUref * par.forceUref * par.forceAreaRef
                             : 0.0;
  if (par.forceEnable) {
    std::printf("------------------------------------------------------------\n");
    if (pf.valid) {
      std::printf("Patch force postprocess, raw Cartesian wall traction\n");
      std::printf("force patch       = %s\n", pf.patch.c_str());
      std::printf("force faces       = %d\n", pf.nFaces);
      std::printf("force area        = %.12e\n", pf.area);
      std::printf("normalSign        = %d\n", par.forceNormalSign);
      std::printf("rho, mu           = %.12e  %.12e\n", par.rho, mu);
      std::printf("pressureForce     = [%.12e, %.12e, %.12e]\n", pf.Fp[0], pf.Fp[1], pf.Fp[2]);
      std::printf("viscousForce      = [%.12e, %.12e, %.12e]\n", pf.Fv[0], pf.Fv[1], pf.Fv[2]);
      std::printf("totalForce        = [%.12e, %.12e, %.12e]\n", pf.F[0], pf.F[1], pf.F[2]);
      const char* ax = "xyz";
      for (int d = 0; d < 3; ++d) std::printf("pressureForce_%c   = %.12e\n", ax[d], pf.Fp[(size_t)d]);
      for (int d = 0; d < 3; ++d) std::printf("viscousForce_%c    = %.12e\n", ax[d], pf.Fv[(size_t)d]);
      for (int d = 0; d < 3; ++d) std::printf("totalForce_%c      = %.12e\n", ax[d], pf.F[(size_t)d]);
      std::printf("wall dn min/max   = %.12e / %.12e\n", pf.dnMin, pf.dnMax);
      std::printf("max|Ut|           = %.12e\n", pf.maxUt);
      std::printf("max shear mag     = %.12e\n", pf.maxShear);
      const V3* dirs[3] = {&dragDir, &liftDir, &spanDir};
      const char* dn[3] = {"drag", "lift", "span"};
      for (int k = 0; k < 3; ++k) {
        const double fpk = dot(pf.Fp, *dirs[k]), fvk = dot(pf.Fv, *dirs[k]),
                     fk = dot(pf.F, *dirs[k]);
        std::printf("%s dir [%.6e, %.6e, %.6e]: pressure %.12e viscous %.12e total %.12e",
                    dn[k], (*dirs[k])[0], (*dirs[k])[1], (*dirs[k])[2], fpk, fvk, fk);
        if (pfDenom > 1.0e-300) std::printf(" C = %.12e", 2.0 * fk / pfDenom);
        std::printf("\n");
      }
    } else {
      std::printf("Patch force postprocess requested but skipped. Check forcePatch.\n");
      std::printf("requested patch  = %s\n", par.forcePatch.c_str());
      std::printf("forceUref        = %.12e\n", par.forceUref);
      std::printf("forceAreaRef     = %.12e\n", par.forceAreaRef);
    }
  }
  double tm[8];
  check(an_simple_timings(s, tm), "timings");
  std::printf("last bicgstab it = [%d %d %d]\n", st.bicgstabIts[0], st.bicgstabIts[1],
              st.bicgstabIts[2]);
  std::printf("last pcg it   = %d\n", st.pcgLast);
  std::printf("assemble time : %.6e s\n", tm[0]);
  std::printf("mom solve time: %.6e s\n", tm[1]);
  std::printf("p setup time  : %.6e s\n", tm[2]);
  std::printf("p solve time  : %.6e s\n", tm[3]);
  std::printf("solve loop wall: %.6e s\n", loopWall);
  std::printf("------------------------------------------------------------\n");
  std::printf("GPU timing breakdown (HIP events):\n");
  std::printf("%-28s %14s %12s %12s\n", "phase", "total[s]", "avg/iter[s]", "%wall");
  const char* phases[6] = {"momentum assembly", "momentum solve", "pressure assembly+setup",
                           "pressure solve", "other kernels", "total"};
  const double denom = loopWall > 1.0e-30 ? loopWall : 1.0;
  for (int i = 0; i < 6; ++i)
    std::printf("%-28s %14.6e %12.6e %11.2f\n", phases[i], tm[i],
                tm[i] / std::max(stepConverged, 1), 100.0 * tm[i] / denom);
  std::printf("pressure PCG iterations     : %.0f over %.0f solves\n", tm[6], tm[7]);
  std::printf("------------------------------------------------------------\n");

  if (par.writeVtu) {
    const std::string vtu = par.outPrefix + "_final.vtu";
    write_vtu(vtu, mesh, {"p", "umag", "cell_volume", "divCorr"}, {&p, &umag, &mesh.vol, &dc},
              u, v, w);
    std::printf("Wrote VTU         : %s\n", vtu.c_str());
  }
  // <outPrefix>_summary.txt: the reference's keys (main.cu:3844-3904)
  std::ofstream so(par.outPrefix + "_summary.txt");
  so << std::setprecision(16);
  so << "points " << mi.nPoints << "\nfaces " << mi.nFaces << "\ninternalFaces "
     << mi.nInternalFaces << "\ncells " << mi.nCells << "\n";
  so << "maxNonOrthDeg " << mi.maxNonOrthDeg << "\nmu " << mu << "\n";
  so << "steps " << stepConverged << "\nmassRes " << st.massRes << "\n";
  so << "maxAbsU " << umax << "\nmaxAbsV " << vmaxf << "\nmaxAbsW " << wmaxf << "\nmaxAbsP "
     << pmax << "\n";
  so << "lastBiCGSTABU " << st.bicgstabIts[0] << "\nlastBiCGSTABV " << st.bicgstabIts[1]
     << "\nlastBiCGSTABW " << st.bicgstabIts[2] << "\nlastPCG " << st.pcgLast << "\n";
  if (cyl.valid) {
    so << "cylinderForcePatch " << cyl.patch << "\ncylinderForceFaces " << cyl.nFaces
       << "\ncylinderForceArea " << cyl.area << "\nforceUbar " << cyl.Ubar << "\nforceD "
       << cyl.D << "\nforceH " << cyl.H << "\nforceCoeffDenom " << cyl.denom << "\nFDPressure "
       << cyl.FDp << "\nFDViscous " << cyl.FDv << "\nFDTotal " << cyl.FD << "\nFLPressure "
       << cyl.FLp << "\nFLViscous " << cyl.FLv << "\nFLTotal " << cyl.FL << "\nCD " << cyl.CD
       << "\nCL " << cyl.CL << "\nwallDnMin " << cyl.dnMin << "\nwallDnMax " << cyl.dnMax
       << "\nmaxAbsDvtDn " << cyl.maxDvtDn << "\n";
  }
  if (cylVec.valid) {
    so << "CDVector " << cylC[0] << "\nCLyVector " << cylC[1] << "\nCLzVector " << cylC[2]
       << "\n";
    const char* ax = "xyz";
    for (int d = 0; d < 3; ++d) so << "F" << ax[d] << "PressureVector " << cylVec.Fp[(size_t)d] << "\n";
    for (int d = 0; d < 3; ++d) so << "F" << ax[d] << "ViscousVector " << cylVec.Fv[(size_t)d] << "\n";
    for (int d = 0; d < 3; ++d) so << "F" << ax[d] << "TotalVector " << cylVec.F[(size_t)d] << "\n";
  }
  if (pf.valid) {
    so << "patchForcePatch " << pf.patch << "\npatchForceFaces " << pf.nFaces
       << "\npatchForceArea " << pf.area << "\npatchForceNormalSign " << par.forceNormalSign
       << "\n";
    const char* ax = "xyz";
    for (int d = 0; d < 3; ++d) so << "pressureForce_" << ax[d] << " " << pf.Fp[(size_t)d] << "\n";
    for (int d = 0; d < 3; ++d) so << "viscousForce_" << ax[d] << " " << pf.Fv[(size_t)d] << "\n";
    for (int d = 0; d < 3; ++d) so << "totalForce_" << ax[d] << " " << pf.F[(size_t)d] << "\n";
  }
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  try {
    return run(argc, argv);
  } catch (const std::exception& e) {
    std::fflush(stdout);
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
}
