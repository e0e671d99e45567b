// host_mesh.hpp -- host-side polyMesh container, reader/writer, generators
// and geometry for the anabasis MI355X library.
//
// The geometry follows LP/src/mesh.cu:160-288 operation for operation
// (compiled with -ffp-contract=off) so that cc/vol/xf/nf/Sf/Af are the
// reference's bits.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace an {

struct HostMesh {
  int nPoints = 0, nFaces = 0, nInternalFaces = 0, nCells = 0;
  std::vector<double> P;                    // 3*nPoints
  std::vector<int> faceOff, faceVerts;      // ragged faces
  std::vector<int> owner, neigh;            // nFaces, nInternalFaces
  std::vector<std::string> patchNames, patchTypes;
  std::vector<int> patchStart, patchNFaces;
  std::vector<int> bPatch;                  // nFaces, patch index + 1
  // geometry (interleaved 3-vectors)
  std::vector<double> cc, vol, xf, nf, Sf, Af;
  double maxNonOrthDeg = 0.0;
  bool hasGeometry = false;
  bool hasPoints = false;  // false for owner/neighbour-only meshes
};

// Reads an ASCII OpenFOAM polyMesh directory (boundary, points, faces, owner,
// neighbour), LP/src/mesh.cu:3-186 semantics.  Throws std::runtime_error.
HostMesh read_polymesh(const std::string& dir);
void write_polymesh(const HostMesh& m, const std::string& dir);

// Fills bPatch and nCells from owner/neigh/patches.
void finish_topology(HostMesh& m);
// LP/src/mesh.cu:188-285.  Throws on non-positive volume / degenerate face.
void compute_geometry(HostMesh& m);

// Generators (the reference ships no mesh generator).
HostMesh gen_hexbox(int nx, int ny, int nz, double lx, double ly, double lz,
                    double jitter, uint64_t seed);
HostMesh gen_tetcube(int nblocks, double length);
struct CylinderParams {
  int ns = 16;       // cells along each side of the inner square / quarter arc
  int nr = 8;        // radial cells of the O-grid ring
  int nx1 = 32;      // x cells upstream   [0, xc-s]
  int nx3 = 128;     // x cells downstream [xc+s, L]
  int ny1 = 8;       // y cells below      [0, yc-s]
  int ny3 = 9;       // y cells above      [yc+s, H]
  int nz = 33;       // layers in z
  double L = 2.2, H = 0.41, W = 0.41;  // channel length / height / depth
  double xc = 0.5, yc = 0.2, R = 0.05, s = 0.1;  // cylinder, square half-width
  double radialRatio = 1.08;  // geometric grading of the O-grid ring
};
HostMesh gen_cylinder_channel(const CylinderParams& p);
// Scales ns and derived counts so the mesh has about targetCells cells.
CylinderParams cylinder_params_for_cells(long long targetCells);

}  // namespace an
