// host_case.cpp -- runtime boundary conditions of generic_simple_v1 (see
// host_case.hpp for the reference lines each function follows).
#include "host_case.hpp"

#include <algorithm>
#include <array>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <unordered_map>
#include <unordered_set>

namespace an {

namespace {

// ------------------------------------------------------------ grammar ---
struct LineCtx {
  const std::string& source;
  int line;
  [[noreturn]] void fail(const std::string& what) const {
    throw std::runtime_error("BC config parse error in '" + source + "' at line " +
                             std::to_string(line) + ": " + what);
  }
  double number(const std::string& tok, const char* what) const {
    size_t used = 0;
    double v = 0.0;
    bool ok = true;
    try {
      v = std::stod(tok, &used);
    } catch (const std::exception&) {
      ok = false;
    }
    if (!ok || used != tok.size())
      fail(std::string("invalid numeric token for ") + what + ": '" + tok + "'");
    return v;
  }
  NormalMode mode(const std::string& tok) const {
    if (tok == "average_patch_normal") return NormalMode::AveragePatch;
    if (tok == "local_face_normal") return NormalMode::LocalFace;
    fail("normal mode must be 'average_patch_normal' or 'local_face_normal', got '" + tok + "'");
  }
  int coord(const std::string& tok) const {
    static const char* const names[3][3] = {{"x", "X", "0"}, {"y", "Y", "1"}, {"z", "Z", "2"}};
    for (int c = 0; c < 3; ++c)
      for (const char* n : names[c])
        if (tok == n) return c;
    fail("coordinate must be x/y/z or 0/1/2, got '" + tok + "'");
  }
};

std::string strip(const std::string& raw) {
  std::string s = raw.substr(0, raw.find('#'));
  size_t a = 0, b = s.size();
  while (a < b && std::isspace((unsigned char)s[a])) ++a;
  while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
  return s.substr(a, b - a);
}

std::vector<std::string> words(const std::string& s) {
  std::istringstream is(s);
  std::vector<std::string> w;
  for (std::string t; is >> t;) w.push_back(t);
  return w;
}

const char* kModeHint = " [average_patch_normal|local_face_normal]";

VelSpec velocity_line(const LineCtx& c, const std::vector<std::string>& t) {
  if (t.size() < 3) c.fail("velocity line must be: velocity <patch> <type> ...");
  VelSpec v;
  v.patch = t[1];
  const std::string& type = t[2];
  const size_t n = t.size();
  auto optional_mode = [&](size_t at) {
    return n > at ? c.mode(t[at]) : NormalMode::AveragePatch;
  };
  if (type == "wall_noslip" || type == "zero_gradient") {
    if (n != 3) c.fail(type + " takes no extra arguments");
    v.kind = type == "wall_noslip" ? VelKind::WallNoSlip : VelKind::ZeroGradient;
  } else if (type == "fixed_uniform_vector") {
    if (n != 6)
      c.fail("fixed_uniform_vector syntax: velocity <patch> fixed_uniform_vector <ux> <uy> <uz>");
    v.kind = VelKind::FixedUniformVector;
    static const char* const comp[3] = {"ux", "uy", "uz"};
    for (int d = 0; d < 3; ++d) v.vec[d] = c.number(t[3 + (size_t)d], comp[d]);
  } else if (type == "fixed_normal_speed" || type == "fixed_flow_rate") {
    if (n != 4 && n != 5)
      c.fail(type + " syntax: velocity <patch> " + type + " <value>" + kModeHint);
    if (type == "fixed_normal_speed") {
      v.kind = VelKind::FixedNormalSpeed;
      v.normalSpeed = c.number(t[3], "normal speed");
    } else {
      v.kind = VelKind::FixedFlowRate;
      v.flowRate = c.number(t[3], "flow rate");
    }
    v.mode = optional_mode(4);
  } else if (type == "parabolic_box_inlet") {
    if (n != 7 && n != 8)
      c.fail(std::string("parabolic_box_inlet syntax: velocity <patch> parabolic_box_inlet <H> "
                         "<Umax> <coord1> <coord2>") + kModeHint);
    v.kind = VelKind::ParabolicBoxInlet;
    v.H = c.number(t[3], "parabolic H");
    v.Umax = c.number(t[4], "parabolic Umax");
    v.coordA = c.coord(t[5]);
    v.coordB = c.coord(t[6]);
    if (v.coordA == v.coordB)
      c.fail("parabolic_box_inlet requires two different transverse coordinates");
    v.mode = optional_mode(7);
  } else {
    c.fail("unknown velocity BC type: '" + type + "'");
  }
  return v;
}

PresSpec pressure_line(const LineCtx& c, const std::vector<std::string>& t) {
  if (t.size() < 3) c.fail("pressure line must be: pressure <patch> <type> ...");
  PresSpec p;
  p.patch = t[1];
  const std::string& type = t[2];
  if (type == "zero_gradient" || type == "open") {
    if (t.size() != 3) c.fail(type + " takes no extra arguments");
    p.kind = type == "open" ? PresKind::Open : PresKind::ZeroGradient;
  } else if (type == "fixed_value") {
    if (t.size() != 4) c.fail("fixed_value syntax: pressure <patch> fixed_value <value>");
    p.kind = PresKind::FixedValue;
    p.value = c.number(t[3], "pressure value");
  } else {
    c.fail("unknown pressure BC type: '" + type + "'");
  }
  return p;
}

// ----------------------------------------------------------- geometry ---
using V3 = std::array<double, 3>;
V3 at3(const std::vector<double>& a, int i) { return {a[3 * (size_t)i], a[3 * (size_t)i + 1], a[3 * (size_t)i + 2]}; }
double dot(const V3& a, const V3& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
V3 cross(const V3& a, const V3& b) {
  return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}
// a / |a|, or zero when |a| <= 1e-30 (patch_geometry.cu:36-40)
V3 unit_or_zero(const V3& a) {
  const double n = std::sqrt(dot(a, a));
  if (n <= 1.0e-30) return {0.0, 0.0, 0.0};
  return {a[0] / n, a[1] / n, a[2] / n};
}
void put(double* dst, const V3& v) {
  dst[0] = v[0];
  dst[1] = v[1];
  dst[2] = v[2];
}

// a / |a| with |a| = sqrt(max(a.a, 0)); throws below 1e-30
// (velocity_bc_eval.cu:19-26)
V3 unit_or_throw(const V3& a, const std::string& what) {
  const double n = std::sqrt(std::max(dot(a, a), 0.0));
  if (n <= 1.0e-30) throw std::runtime_error("Cannot normalize vector for " + what);
  return {a[0] / n, a[1] / n, a[2] / n};
}

V3 spec_normal(const VelSpec& s, const PatchGeometry& g, const V3& faceNormal) {
  if (s.mode == NormalMode::AveragePatch)
    return unit_or_throw({g.avgNormal[0], g.avgNormal[1], g.avgNormal[2]},
                         "AveragePatchNormal on patch '" + g.name + "'");
  return unit_or_throw(faceNormal, "LocalFaceNormal on patch '" + g.name + "'");
}

// Face velocity of one spec at face centre x with face normal nf
// (velocity_bc_eval.cu:142-181).
V3 face_velocity(const VelSpec& s, const PatchGeometry& g, const V3& x, const V3& nf) {
  switch (s.kind) {
    case VelKind::WallNoSlip:
    case VelKind::ZeroGradient:
      return {0.0, 0.0, 0.0};
    case VelKind::FixedUniformVector:
      return {s.vec[0], s.vec[1], s.vec[2]};
    case VelKind::ParabolicBoxInlet: {
      // |U| = 16 Umax a b (H-a)(H-b) / H^4 along the inward normal, a and b
      // the two transverse coordinates clamped to [0, H]
      const double H = s.H;
      if (H <= 0.0)
        throw std::runtime_error("ParabolicBoxInlet has non-positive H on patch '" + s.patch +
                                 "'");
      if (s.coordA < 0 || s.coordA > 2 || s.coordB < 0 || s.coordB > 2 || s.coordA == s.coordB)
        throw std::runtime_error("ParabolicBoxInlet has invalid coordinates on patch '" +
                                 s.patch + "'");
      const double a = std::min(std::max(x[(size_t)s.coordA], 0.0), H);
      const double b = std::min(std::max(x[(size_t)s.coordB], 0.0), H);
      const double H2 = H * H, H4 = H2 * H2;
      const double mag = 16.0 * s.Umax * a * b * (H - a) * (H - b) / H4;
      const V3 n = spec_normal(s, g, nf);
      return {-mag * n[0], -mag * n[1], -mag * n[2]};
    }
    case VelKind::FixedNormalSpeed:
    case VelKind::FixedFlowRate: {
      if (s.kind == VelKind::FixedFlowRate && s.mode == NormalMode::AveragePatch &&
          g.planarity < 0.98)
        throw std::runtime_error(
            "FixedFlowRate with AveragePatchNormal requires an approximately planar patch. "
            "Patch '" + g.name + "' has planarity = " + std::to_string(g.planarity) +
            ". Use LocalFaceNormal instead.");
      double un = s.normalSpeed;
      if (s.kind == VelKind::FixedFlowRate) {
        if (g.area <= 1.0e-30)
          throw std::runtime_error("Patch '" + g.name + "' has near-zero area for FixedFlowRate");
        un = s.flowRate / g.area;
      }
      const V3 n = spec_normal(s, g, nf);
      return {un * n[0], un * n[1], un * n[2]};
    }
  }
  throw std::runtime_error("Unhandled VelocityBCType");
}

int patch_of(const HostMesh& m, const std::string& name) {
  for (size_t k = 0; k < m.patchNames.size(); ++k)
    if (m.patchNames[k] == name) return (int)k;
  throw std::runtime_error("Patch '" + name + "' not found in LegacyBCMeshView");
}

template <typename Spec>
std::vector<std::string> names_of(const std::vector<Spec>& specs) {
  std::vector<std::string> out;
  for (const auto& s : specs) out.push_back(s.patch);
  return out;
}

void throw_duplicates(const BcConfig& cfg) {
  const char* what[2] = {"velocity", "pressure"};
  const std::vector<std::string> lists[2] = {names_of(cfg.vel), names_of(cfg.pres)};
  for (int i = 0; i < 2; ++i) {
    const auto dup = duplicate_patches(lists[i]);
    if (dup.empty()) continue;
    std::string msg = std::string("Duplicate ") + what[i] + " BC specs found:\n";
    for (const auto& n : dup) msg += "  " + n + "\n";
    throw std::runtime_error(msg);
  }
}

}  // namespace

BcConfig parse_bc_config(std::istream& in, const std::string& source) {
  BcConfig cfg;
  std::string raw;
  for (int line = 1; std::getline(in, raw); ++line) {
    const std::string s = strip(raw);
    if (s.empty()) continue;
    const auto t = words(s);
    const LineCtx c{source, line};
    if (t[0] == "velocity")
      cfg.vel.push_back(velocity_line(c, t));
    else if (t[0] == "pressure")
      cfg.pres.push_back(pressure_line(c, t));
    else
      c.fail("line must start with 'velocity' or 'pressure'");
  }
  return cfg;
}

BcConfig load_bc_config(const std::string& path) {
  std::ifstream in(path);
  if (!in) throw std::runtime_error("Could not open BC config file: " + path);
  return parse_bc_config(in, path);
}

void check_bc_patches(const BcConfig& cfg, const std::vector<std::string>& patchNames) {
  const std::unordered_set<std::string> known(patchNames.begin(), patchNames.end());
  for (const auto& s : cfg.vel)
    if (!known.count(s.patch))
      throw std::runtime_error("Velocity BC refers to unknown patch '" + s.patch + "'");
  for (const auto& s : cfg.pres)
    if (!known.count(s.patch))
      throw std::runtime_error("Pressure BC refers to unknown patch '" + s.patch + "'");
}

std::vector<std::string> duplicate_patches(const std::vector<std::string>& names) {
  std::unordered_map<std::string, int> seen;
  std::vector<std::string> dup;
  for (const auto& n : names)
    if (++seen[n] == 2) dup.push_back(n);
  return dup;
}

std::vector<PatchGeometry> patch_geometry_table(const HostMesh& m) {
  if (!m.hasGeometry) throw std::runtime_error("mesh has no geometry");
  std::vector<PatchGeometry> out(m.patchNames.size());
  for (size_t k = 0; k < out.size(); ++k) {
    PatchGeometry& g = out[k];
    g.index = (int)k;
    g.name = m.patchNames[k];
    const int f0 = m.patchStart[k], n = m.patchNFaces[k];
    if (n < 0 || f0 < m.nInternalFaces || f0 + n > m.nFaces)
      throw std::runtime_error("Invalid patch face range for patch '" + g.name + "'");
    g.nFaces = n;
    // serial sums in face order: area, area-weighted centres, face vectors
    V3 c{0.0, 0.0, 0.0}, s{0.0, 0.0, 0.0};
    double area = 0.0;
    for (int f = f0; f < f0 + n; ++f) {
      const double A = m.Af[(size_t)f];
      const V3 x = at3(m.xf, f), S = at3(m.Sf, f);
      area += A;
      c = {c[0] + A * x[0], c[1] + A * x[1], c[2] + A * x[2]};
      s = {s[0] + S[0], s[1] + S[1], s[2] + S[2]};
    }
    g.area = area;
    put(g.sumSf, s);
    if (area > 1.0e-30) {
      const double inv = 1.0 / area;
      put(g.centroid, {inv * c[0], inv * c[1], inv * c[2]});
    }
    const V3 an = unit_or_zero(s);
    put(g.avgNormal, an);
    g.planarity = area > 1.0e-30 ? std::sqrt(dot(s, s)) / area : 0.0;
    if (std::sqrt(dot(an, an)) > 0.0) {
      // tangent seed: the axis least aligned with the normal (x before y
      // before z on ties), patch_geometry.cu:49-60
      const double ax = std::fabs(an[0]), ay = std::fabs(an[1]), az = std::fabs(an[2]);
      const V3 seed = (ax <= ay && ax <= az) ? V3{1.0, 0.0, 0.0}
                      : (ay <= ax && ay <= az) ? V3{0.0, 1.0, 0.0}
                                               : V3{0.0, 0.0, 1.0};
      const V3 t1 = unit_or_zero(cross(an, seed));
      put(g.tangent1, t1);
      put(g.tangent2, unit_or_zero(cross(an, t1)));
    } else {
      put(g.tangent1, {1.0, 0.0, 0.0});
      put(g.tangent2, {0.0, 1.0, 0.0});
    }
  }
  return out;
}

BcArrays apply_bc(const HostMesh& m, const BcConfig& cfg) {
  check_bc_patches(cfg, m.patchNames);
  throw_duplicates(cfg);
  const auto geo = patch_geometry_table(m);
  const size_t nP = m.patchNames.size(), nF = (size_t)m.nFaces;
  BcArrays b;
  b.uType.assign(nP, 0);  // "Neumann" until a spec says otherwise
  b.pType.assign(nP, 0);
  b.u.assign(nF, 0.0);
  b.v.assign(nF, 0.0);
  b.w.assign(nF, 0.0);
  b.p.assign(nF, 0.0);
  for (const auto& s : cfg.vel) {
    const int k = patch_of(m, s.patch);
    if (s.kind == VelKind::ZeroGradient) {
      b.uType[(size_t)k] = 0;
      continue;
    }
    b.uType[(size_t)k] = 1;  // every other kind is Dirichlet in all three components
    const PatchGeometry& g = geo[(size_t)k];
    for (int f = m.patchStart[(size_t)k]; f < m.patchStart[(size_t)k] + m.patchNFaces[(size_t)k];
         ++f) {
      const V3 U = face_velocity(s, g, at3(m.xf, f), at3(m.nf, f));
      b.u[(size_t)f] = U[0];
      b.v[(size_t)f] = U[1];
      b.w[(size_t)f] = U[2];
    }
  }
  for (const auto& s : cfg.pres) {
    const int k = patch_of(m, s.patch);
    if (s.kind != PresKind::FixedValue) {
      b.pType[(size_t)k] = 0;
      continue;
    }
    b.pType[(size_t)k] = 1;
    for (int f = m.patchStart[(size_t)k]; f < m.patchStart[(size_t)k] + m.patchNFaces[(size_t)k];
         ++f)
      b.p[(size_t)f] = s.value;
  }
  return b;
}

Coverage bc_coverage(const HostMesh& m, const BcConfig& cfg) {
  throw_duplicates(cfg);
  Coverage c;
  std::vector<std::string> missV, missP;
  for (const auto& name : m.patchNames) {
    bool hv = false, hp = false;
    for (const auto& s : cfg.vel) hv = hv || s.patch == name;
    for (const auto& s : cfg.pres) hp = hp || s.patch == name;
    char line[512];
    std::snprintf(line, sizeof line, "  %-32s velocity=%s pressure=%s\n", name.c_str(),
                  hv ? "yes" : "NO", hp ? "yes" : "NO");
    c.table += line;
    if (!hv) missV.push_back(name);
    if (!hp) missP.push_back(name);
  }
  if (missV.empty() && missP.empty()) return c;
  c.complete = false;
  c.report = "\nERROR: incomplete runtime BC coverage.\n";
  if (!missV.empty()) {
    c.report += "Missing velocity BC for patches:\n";
    for (const auto& n : missV) c.report += "  " + n + "\n";
  }
  if (!missP.empty()) {
    c.report += "Missing pressure BC for patches:\n";
    for (const auto& n : missP) c.report += "  " + n + "\n";
  }
  c.report +=
      "\nEvery polyMesh boundary patch must have exactly one velocity BC and one pressure BC in "
      "generic_simple_v1.\n";
  return c;
}

}  // namespace an
