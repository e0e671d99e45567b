// host_case.hpp -- runtime boundary conditions of generic_simple_v1 on the
// host: the .bc grammar, the patch geometry table, the evaluation of BC
// specs into per-patch types and per-face values, and the patch coverage
// table.  Messages are the reference's texts (they are part of the
// driver's interface).
//
//   parse_bc_config / load_bc_config   GSV1 bc_runtime_config.cu:83-210
//   check_bc_patches                   bc_runtime_config.cu:212-230
//   duplicate_patches                  bc_specs.cu:118-140
//   patch_geometry_table               patch_geometry.cu:83-140
//   apply_bc                           velocity_bc_eval.cu:197-290 (+ bc_to_int, main.cu:1058)
//   bc_coverage                        main.cu:3178-3233
#pragma once

#include <istream>
#include <string>
#include <vector>

#include "host_mesh.hpp"

namespace an {

enum class VelKind { WallNoSlip, ZeroGradient, FixedUniformVector, FixedNormalSpeed, FixedFlowRate,
                     ParabolicBoxInlet };
enum class PresKind { ZeroGradient, Open, FixedValue };
enum class NormalMode { AveragePatch, LocalFace };

struct VelSpec {
  std::string patch;
  VelKind kind = VelKind::ZeroGradient;
  double vec[3] = {0.0, 0.0, 0.0};  // FixedUniformVector
  double normalSpeed = 0.0;         // signed, along the outward normal
  double flowRate = 0.0;            // signed, along the outward normal
  double H = 1.0, Umax = 1.0;       // ParabolicBoxInlet
  int coordA = 1, coordB = 2;
  NormalMode mode = NormalMode::AveragePatch;
};

struct PresSpec {
  std::string patch;
  PresKind kind = PresKind::ZeroGradient;
  double value = 0.0;
};

struct BcConfig {
  std::vector<VelSpec> vel;
  std::vector<PresSpec> pres;
};

// `velocity <patch> <type> ...` / `pressure <patch> <type> ...` lines; '#'
// starts a comment.  Errors: std::runtime_error("BC config parse error in
// '<source>' at line <n>: <what>").
BcConfig parse_bc_config(std::istream& in, const std::string& source);
BcConfig load_bc_config(const std::string& path);

// Every spec names a patch of the mesh ("Velocity BC refers to unknown
// patch '<name>'").
void check_bc_patches(const BcConfig& cfg, const std::vector<std::string>& patchNames);
// Patch names that occur more than once, in the order of their second
// occurrence.
std::vector<std::string> duplicate_patches(const std::vector<std::string>& names);

struct PatchGeometry {
  int index = -1;
  std::string name;
  int nFaces = 0;
  double area = 0.0;
  double centroid[3] = {0.0, 0.0, 0.0};
  double sumSf[3] = {0.0, 0.0, 0.0};
  double avgNormal[3] = {0.0, 0.0, 0.0};
  double planarity = 0.0;
  double tangent1[3] = {0.0, 0.0, 0.0}, tangent2[3] = {0.0, 0.0, 0.0};
};
std::vector<PatchGeometry> patch_geometry_table(const HostMesh& m);

// Per-patch types (1 = Dirichlet, 0 = otherwise) and per-face values
// (nFaces entries; internal faces and non-Dirichlet patches hold 0).
struct BcArrays {
  std::vector<int> uType, pType;
  std::vector<double> u, v, w, p;
};
// Validates (unknown patches, duplicates) and evaluates the specs.
BcArrays apply_bc(const HostMesh& m, const BcConfig& cfg);

struct Coverage {
  std::string table;   // one "  <name> velocity=yes|NO pressure=yes|NO" line per patch
  std::string report;  // the error report when a patch lacks a BC, else empty
  bool complete = true;
};
// Throws on duplicate specs ("Duplicate velocity BC specs found:\n  <name>\n").
Coverage bc_coverage(const HostMesh& m, const BcConfig& cfg);

}  // namespace an
