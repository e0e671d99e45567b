// capi.hip -- the extern "C" boundary (include/anabasis_hip.h).
//
// Every entry point catches C++ exceptions and turns them into an an_status
// plus a thread-local message; nothing throws across the ABI.
#include <cstring>
#include <sstream>
#include <memory>
#include <string>

#include "../../include/anabasis_hip.h"
#include "kernels.hpp"
#include "simple.hpp"
#include "solver.hpp"
#include "batch.hpp"
#include "host_case.hpp"

namespace an {
static thread_local std::string g_err;
void set_last_error(const std::string& m) { g_err = m; }
}  // namespace an

using namespace an;

struct an_mesh_s {
  DevMesh dm;
};
struct an_solver_s {
  std::unique_ptr<PcgSolver> pcg;
};
struct an_simple_s {
  std::unique_ptr<SimpleSolver> sim;
};
struct an_bc_s {
  BcConfig cfg;
};

static hipMemcpyKind kind_in(an_memspace m) {
  return m == AN_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
}
static hipMemcpyKind kind_out(an_memspace m) {
  return m == AN_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
}

// Device copy of a caller array (or a borrowed pointer when already on the
// device).
template <typename T>
struct InArr {
  DevBuf<T> buf;
  const T* p = nullptr;
  InArr(const T* src, size_t n, an_memspace m, hipStream_t s) {
    if (!src) {
      if (n) throw Error(AN_EINVAL, "required array is NULL");
      return;
    }
    if (m == AN_DEVICE) {
      p = src;
    } else {
      buf.upload(src, n, s);
      p = buf.p;
    }
  }
};
template <typename T>
struct OutArr {
  DevBuf<T> buf;
  T* p = nullptr;
  T* dst;
  size_t n;
  an_memspace m;
  OutArr(T* d, size_t count, an_memspace mm) : dst(d), n(count), m(mm) {
    if (!d) return;
    if (m == AN_DEVICE)
      p = d;
    else {
      buf.alloc(count ? count : 1);
      p = buf.p;
    }
  }
  void finish(hipStream_t s) {
    if (dst && m == AN_HOST && n) AN_HIP(hipMemcpyAsync(dst, p, sizeof(T) * n, hipMemcpyDeviceToHost, s));
  }
};

static DevMesh& devmesh(an_mesh_t m, bool needGeometry) {
  if (!m) throw Error(AN_EINVAL, "mesh handle is NULL");
  if (needGeometry && !m->dm.host.hasGeometry)
    throw Error(AN_EINVAL, "mesh has no points/faces: geometry unavailable");
  m->dm.to_device(0);
  return m->dm;
}

extern "C" {

const char* an_last_error(void) { return g_err.c_str(); }
int an_version(void) { return 100; }

an_status an_set_device(int device) {
  AN_TRY
  AN_HIP(hipSetDevice(device));
  return AN_OK;
  AN_CATCH
}

// ------------------------------------------------------------- mesh ----
static an_status wrap_mesh(HostMesh&& hm, an_mesh_t* out) {
  if (!out) throw Error(AN_EINVAL, "out is NULL");
  auto* m = new an_mesh_s;
  m->dm.host = std::move(hm);
  *out = m;
  return AN_OK;
}

an_status an_mesh_read_polymesh(const char* dir, an_mesh_t* out) {
  AN_TRY
  if (!dir) throw Error(AN_EINVAL, "dir is NULL");
  HostMesh hm;
  try {
    hm = read_polymesh(dir);
  } catch (const std::runtime_error& e) {
    std::string w = e.what();
    if (w.rfind("Could not open", 0) == 0) throw Error(AN_EIO, w);
    throw Error(AN_EINVAL, w);
  }
  return wrap_mesh(std::move(hm), out);
  AN_CATCH
}

an_status an_mesh_from_arrays(int nPoints, const double* points, int nFaces,
                              const int* faceOffsets, const int* faceVerts, const int* owner,
                              int nInternalFaces, const int* neigh, int nPatches,
                              const char* const* patchNames, const int* patchStart,
                              const int* patchNFaces, an_mesh_t* out) {
  AN_TRY
  if (nFaces < 0 || nInternalFaces < 0 || nInternalFaces > nFaces || !owner ||
      (nInternalFaces && !neigh) || nPatches < 0)
    throw Error(AN_EINVAL, "invalid face arrays");
  HostMesh m;
  m.nFaces = nFaces;
  m.owner.assign(owner, owner + nFaces);
  m.neigh.assign(neigh, neigh + nInternalFaces);
  for (int k = 0; k < nPatches; ++k) {
    m.patchNames.push_back(patchNames && patchNames[k] ? patchNames[k] : ("patch" + std::to_string(k)));
    m.patchTypes.push_back("patch");
    m.patchStart.push_back(patchStart[k]);
    m.patchNFaces.push_back(patchNFaces[k]);
  }
  for (int f = 0; f < nFaces; ++f)
    if (owner[f] < 0) throw Error(AN_EINVAL, "negative owner index");
  if (nPoints > 0 && points && faceOffsets && faceVerts) {
    m.nPoints = nPoints;
    m.P.assign(points, points + 3 * (size_t)nPoints);
    m.faceOff.assign(faceOffsets, faceOffsets + nFaces + 1);
    m.faceVerts.assign(faceVerts, faceVerts + faceOffsets[nFaces]);
    for (int v : m.faceVerts)
      if (v < 0 || v >= nPoints) throw Error(AN_EINVAL, "face vertex out of range");
    m.hasPoints = true;
  }
  finish_topology(m);
  if (m.hasPoints) compute_geometry(m);
  return wrap_mesh(std::move(m), out);
  AN_CATCH
}

an_status an_mesh_write_polymesh(an_mesh_t mesh, const char* dir) {
  AN_TRY
  if (!mesh || !dir) throw Error(AN_EINVAL, "NULL argument");
  if (!mesh->dm.host.hasPoints) throw Error(AN_EINVAL, "mesh has no points/faces");
  try {
    write_polymesh(mesh->dm.host, dir);
  } catch (const std::runtime_error& e) {
    throw Error(AN_EIO, e.what());
  }
  return AN_OK;
  AN_CATCH
}

an_status an_mesh_gen_hexbox(int nx, int ny, int nz, double lx, double ly, double lz,
                             double jitter, uint64_t seed, an_mesh_t* out) {
  AN_TRY
  if (!(jitter >= 0.0 && jitter < 0.5)) throw Error(AN_EINVAL, "jitter must be in [0, 0.5)");
  return wrap_mesh(gen_hexbox(nx, ny, nz, lx, ly, lz, jitter, seed), out);
  AN_CATCH
}

an_status an_mesh_gen_tetcube(int nblocks, double length, an_mesh_t* out) {
  AN_TRY
  return wrap_mesh(gen_tetcube(nblocks, length), out);
  AN_CATCH
}

an_status an_mesh_gen_cylinder(long long targetCells, an_mesh_t* out) {
  AN_TRY
  if (targetCells < 1000) throw Error(AN_EINVAL, "targetCells too small");
  return wrap_mesh(gen_cylinder_channel(cylinder_params_for_cells(targetCells)), out);
  AN_CATCH
}

an_status an_mesh_destroy(an_mesh_t mesh) {
  AN_TRY
  delete mesh;
  return AN_OK;
  AN_CATCH
}

an_status an_mesh_get_info(an_mesh_t mesh, an_mesh_info* info) {
  AN_TRY
  if (!mesh || !info) throw Error(AN_EINVAL, "NULL argument");
  const HostMesh& m = mesh->dm.host;
  info->nPoints = m.nPoints;
  info->nFaces = m.nFaces;
  info->nInternalFaces = m.nInternalFaces;
  info->nCells = m.nCells;
  info->nPatches = (int)m.patchNames.size();
  info->nBoundaryFaces = m.nFaces - m.nInternalFaces;
  info->maxNonOrthDeg = m.maxNonOrthDeg;
  info->hasGeometry = m.hasGeometry ? 1 : 0;
  info->nnz = mesh->dm.onDevice ? mesh->dm.addr.nnz : -1;
  return AN_OK;
  AN_CATCH
}

an_status an_mesh_get_array(an_mesh_t mesh, const char* name, void* out) {
  AN_TRY
  if (!mesh || !name || !out) throw Error(AN_EINVAL, "NULL argument");
  const HostMesh& m = mesh->dm.host;
  auto cpi = [&](const std::vector<int>& v) { std::memcpy(out, v.data(), v.size() * sizeof(int)); };
  auto cpd = [&](const std::vector<double>& v) {
    std::memcpy(out, v.data(), v.size() * sizeof(double));
  };
  std::string n = name;
  if (n == "points") cpd(m.P);
  else if (n == "faceOffsets") cpi(m.faceOff);
  else if (n == "faceVerts") cpi(m.faceVerts);
  else if (n == "owner") cpi(m.owner);
  else if (n == "neigh") cpi(m.neigh);
  else if (n == "bPatch") cpi(m.bPatch);
  else if (!m.hasGeometry) throw Error(AN_EINVAL, "mesh has no geometry");
  else if (n == "cc") cpd(m.cc);
  else if (n == "vol") cpd(m.vol);
  else if (n == "xf") cpd(m.xf);
  else if (n == "nf") cpd(m.nf);
  else if (n == "Sf") cpd(m.Sf);
  else if (n == "Af") cpd(m.Af);
  else throw Error(AN_EINVAL, "unknown mesh array '" + n + "'");
  return AN_OK;
  AN_CATCH
}

an_status an_mesh_get_patch(an_mesh_t mesh, int k, char* name, int cap, int* startFace,
                            int* nFaces) {
  AN_TRY
  if (!mesh) throw Error(AN_EINVAL, "NULL mesh");
  const HostMesh& m = mesh->dm.host;
  if (k < 0 || k >= (int)m.patchNames.size()) throw Error(AN_EINVAL, "patch index out of range");
  if (name && cap > 0) {
    std::strncpy(name, m.patchNames[k].c_str(), (size_t)cap - 1);
    name[cap - 1] = 0;
  }
  if (startFace) *startFace = m.patchStart[k];
  if (nFaces) *nFaces = m.patchNFaces[k];
  return AN_OK;
  AN_CATCH
}

// ------------------------------------------------ runtime BCs ----
// std::runtime_error from the BC layer carries the reference's message
#define AN_BC_CALL(stmt)                          \
  try {                                           \
    stmt;                                         \
  } catch (const std::runtime_error& e) {         \
    throw Error(AN_EINVAL, e.what());             \
  }

an_status an_bc_load(const char* path, an_bc_t* out) {
  AN_TRY
  if (!path || !out) throw Error(AN_EINVAL, "NULL argument");
  auto b = std::make_unique<an_bc_s>();
  AN_BC_CALL(b->cfg = load_bc_config(path));
  *out = b.release();
  return AN_OK;
  AN_CATCH
}

an_status an_bc_parse(const char* text, const char* sourceName, an_bc_t* out) {
  AN_TRY
  if (!text || !out) throw Error(AN_EINVAL, "NULL argument");
  std::istringstream in(text);
  auto b = std::make_unique<an_bc_s>();
  AN_BC_CALL(b->cfg = parse_bc_config(in, sourceName ? sourceName : "<string>"));
  *out = b.release();
  return AN_OK;
  AN_CATCH
}

an_status an_bc_destroy(an_bc_t bc) {
  delete bc;
  return AN_OK;
}

an_status an_patch_geometry(an_mesh_t mesh, int k, double* out) {
  AN_TRY
  if (!mesh || !out) throw Error(AN_EINVAL, "NULL argument");
  const HostMesh& m = mesh->dm.host;
  if (k < 0 || k >= (int)m.patchNames.size()) throw Error(AN_EINVAL, "patch index out of range");
  std::vector<PatchGeometry> g;
  AN_BC_CALL(g = patch_geometry_table(m));
  const PatchGeometry& p = g[(size_t)k];
  double* o = out;
  *o++ = p.nFaces;
  *o++ = p.area;
  for (const double* v : {p.centroid, p.sumSf, p.avgNormal}) {
    for (int d = 0; d < 3; ++d) *o++ = v[d];
  }
  *o++ = p.planarity;
  for (const double* v : {p.tangent1, p.tangent2}) {
    for (int d = 0; d < 3; ++d) *o++ = v[d];
  }
  return AN_OK;
  AN_CATCH
}

an_status an_bc_coverage(an_mesh_t mesh, an_bc_t bc, int which, char* text, int cap,
                         int* length, int* complete) {
  AN_TRY
  if (!mesh || !bc) throw Error(AN_EINVAL, "NULL argument");
  if (which != 0 && which != 1) throw Error(AN_EINVAL, "which must be 0 (table) or 1 (report)");
  Coverage c;
  AN_BC_CALL(c = bc_coverage(mesh->dm.host, bc->cfg));
  const std::string& s = which == 0 ? c.table : c.report;
  if (length) *length = (int)s.size();
  if (complete) *complete = c.complete ? 1 : 0;
  if (text && cap > 0) {
    const size_t n = std::min(s.size(), (size_t)cap - 1);
    std::memcpy(text, s.data(), n);
    text[n] = 0;
  }
  return AN_OK;
  AN_CATCH
}

an_status an_bc_apply(an_mesh_t mesh, an_bc_t bc, int* patchUType, int* patchPType,
                      double* uFaceBC, double* vFaceBC, double* wFaceBC, double* pFaceBC) {
  AN_TRY
  if (!mesh || !bc || !patchUType || !patchPType || !uFaceBC || !vFaceBC || !wFaceBC || !pFaceBC)
    throw Error(AN_EINVAL, "NULL argument");
  BcArrays a;
  AN_BC_CALL(a = apply_bc(mesh->dm.host, bc->cfg));
  std::memcpy(patchUType, a.uType.data(), a.uType.size() * sizeof(int));
  std::memcpy(patchPType, a.pType.data(), a.pType.size() * sizeof(int));
  std::memcpy(uFaceBC, a.u.data(), a.u.size() * sizeof(double));
  std::memcpy(vFaceBC, a.v.data(), a.v.size() * sizeof(double));
  std::memcpy(wFaceBC, a.w.data(), a.w.size() * sizeof(double));
  std::memcpy(pFaceBC, a.p.data(), a.p.size() * sizeof(double));
  return AN_OK;
  AN_CATCH
}

// ------------------------------------------------------- addressing ----
static void copy_addr_out(const DevAddressing& A, int* rowOffsets, int* cols, int* diagPos,
                          int* facePP, int* facePN, int* faceNP, int* faceNN, an_memspace mem) {
  hipMemcpyKind k = kind_out(mem);
  auto cp = [&](int* dst, const DevBuf<int>& src, size_t n) {
    if (dst && n) AN_HIP(hipMemcpy(dst, src.p, sizeof(int) * n, k));
  };
  cp(rowOffsets, A.rowOff, (size_t)A.nC + 1);
  cp(cols, A.cols, (size_t)A.nnz);
  cp(diagPos, A.diagPos, (size_t)A.nC);
  cp(facePP, A.facePP, (size_t)A.nIF);
  cp(facePN, A.facePN, (size_t)A.nIF);
  cp(faceNP, A.faceNP, (size_t)A.nIF);
  cp(faceNN, A.faceNN, (size_t)A.nIF);
}

an_status an_mesh_csr(an_mesh_t mesh, int* rowOffsets, int* cols, int* diagPos, int* facePP,
                      int* facePN, int* faceNP, int* faceNN, an_memspace mem) {
  AN_TRY
  DevMesh& d = devmesh(mesh, false);
  copy_addr_out(d.addr, rowOffsets, cols, diagPos, facePP, facePN, faceNP, faceNN, mem);
  return AN_OK;
  AN_CATCH
}

an_status an_csr_from_faces(int nCells, int nFaces, int nInternalFaces, const int* owner,
                            const int* neigh, int64_t* nnzOut, int* rowOffsets, int* cols,
                            int* diagPos, int* facePP, int* facePN, int* faceNP, int* faceNN,
                            an_memspace mem) {
  AN_TRY
  if (nCells <= 0 || nFaces < nInternalFaces || nInternalFaces < 0)
    throw Error(AN_EINVAL, "invalid sizes");
  InArr<int> o(owner, (size_t)nFaces, mem, 0);
  InArr<int> n(neigh, (size_t)nInternalFaces, mem, 0);
  DevAddressing A;
  build_addressing(A, nCells, nFaces, nInternalFaces, o.p, n.p, 0);
  if (nnzOut) *nnzOut = A.nnz;
  copy_addr_out(A, rowOffsets, cols, diagPos, facePP, facePN, faceNP, faceNN, mem);
  return AN_OK;
  AN_CATCH
}

// -------------------------------------------------------- gradients ----
an_status an_lsq_gradient(an_mesh_t mesh, int mode, const int* bcTypeBF, const double* bcValBF,
                          const double* phi, double* grad, an_memspace mem) {
  AN_TRY
  if (mode != AN_GRAD_GSV1 && mode != AN_GRAD_LP) throw Error(AN_EINVAL, "unknown gradient mode");
  DevMesh& d = devmesh(mesh, true);
  hipStream_t s = 0;
  InArr<int> bt(bcTypeBF, (size_t)d.nBF, mem, s);
  InArr<double> bv(bcValBF, (size_t)d.nBF, mem, s);
  InArr<double> ph(phi, (size_t)d.nC, mem, s);
  OutArr<double> g(grad, 3 * (size_t)d.nC, mem);
  if (!grad) throw Error(AN_EINVAL, "grad is NULL");
  DevBuf<double> soa(3 * (size_t)d.nC);
  launch_lsq_gradient(view_of(d), mode, BCView{bt.p, bv.p}, ph.p, soa.p, soa.p + d.nC,
                      soa.p + 2 * (size_t)d.nC, s);
  launch_soa_to_aos(d.nC, soa.p, soa.p + d.nC, soa.p + 2 * (size_t)d.nC, g.p, s);
  g.finish(s);
  AN_HIP(hipStreamSynchronize(s));
  return AN_OK;
  AN_CATCH
}

// --------------------------------------------------------- assembly ----
an_status an_elliptic_assemble(an_mesh_t mesh, const double* gammaFace, const double* cellSource,
                               const int* bcTypeBF, const double* bcValBF, const double* grad,
                               int includeNonOrth, int useReferenceCell, int refCell,
                               double refValue, double* values, double* rhs, an_memspace mem) {
  AN_TRY
  DevMesh& d = devmesh(mesh, true);
  if (useReferenceCell && (refCell < 0 || refCell >= d.nC))
    throw Error(AN_EINVAL, "Invalid reference cell");
  hipStream_t s = 0;
  InArr<double> gam(gammaFace, (size_t)d.nF, mem, s);
  InArr<double> src(cellSource, (size_t)d.nC, mem, s);
  InArr<int> bt(bcTypeBF, (size_t)d.nBF, mem, s);
  InArr<double> bv(bcValBF, (size_t)d.nBF, mem, s);
  InArr<double> gr(grad, 3 * (size_t)d.nC, mem, s);
  OutArr<double> va(values, (size_t)d.addr.nnz, mem);
  OutArr<double> rh(rhs, (size_t)d.nC, mem);
  if (!values || !rhs) throw Error(AN_EINVAL, "values/rhs is NULL");
  DevBuf<double> soa(3 * (size_t)d.nC);
  launch_aos_to_soa(d.nC, gr.p, soa.p, soa.p + d.nC, soa.p + 2 * (size_t)d.nC, s);
  launch_elliptic_assemble(view_of(d), gam.p, src.p, BCView{bt.p, bv.p}, soa.p, soa.p + d.nC,
                           soa.p + 2 * (size_t)d.nC, includeNonOrth, useReferenceCell, refCell,
                           refValue, va.p, rh.p, s);
  va.finish(s);
  rh.finish(s);
  AN_HIP(hipStreamSynchronize(s));
  return AN_OK;
  AN_CATCH
}

an_status an_transport_assemble(an_mesh_t mesh, const double* faceFlux, const double* gammaFace,
                                const double* Su, const double* Sp, const int* bcTypeBF,
                                const double* bcValBF, const double* grad, int includeNonOrth,
                                int upwind, double* values, double* rhs, an_memspace mem) {
  AN_TRY
  DevMesh& d = devmesh(mesh, true);
  hipStream_t s = 0;
  InArr<double> ff(faceFlux, (size_t)d.nF, mem, s);
  InArr<double> gam(gammaFace, (size_t)d.nF, mem, s);
  InArr<double> su(Su, (size_t)d.nC, mem, s);
  InArr<double> sp(Sp, (size_t)d.nC, mem, s);
  InArr<int> bt(bcTypeBF, (size_t)d.nBF, mem, s);
  InArr<double> bv(bcValBF, (size_t)d.nBF, mem, s);
  InArr<double> gr(grad, 3 * (size_t)d.nC, mem, s);
  OutArr<double> va(values, (size_t)d.addr.nnz, mem);
  OutArr<double> rh(rhs, (size_t)d.nC, mem);
  if (!values || !rhs) throw Error(AN_EINVAL, "values/rhs is NULL");
  DevBuf<double> soa(3 * (size_t)d.nC);
  launch_aos_to_soa(d.nC, gr.p, soa.p, soa.p + d.nC, soa.p + 2 * (size_t)d.nC, s);
  launch_transport_assemble(view_of(d), ff.p, gam.p, su.p, sp.p, BCView{bt.p, bv.p}, soa.p,
                            soa.p + d.nC, soa.p + 2 * (size_t)d.nC, includeNonOrth, upwind, va.p,
                            rh.p, s);
  va.finish(s);
  rh.finish(s);
  AN_HIP(hipStreamSynchronize(s));
  return AN_OK;
  AN_CATCH
}

// ----------------------------------------------------------- solvers ---
void an_pcg_default_options(an_pcg_options* o) {
  if (o) default_pcg_options(o);
}

an_status an_pcg_create(int n, const int* rowOffsets, const int* cols, int64_t nnz,
                        const an_pcg_options* opt, an_memspace mem, an_solver_t* out) {
  AN_TRY
  if (n <= 0 || !rowOffsets || !cols || nnz <= 0 || !out) throw Error(AN_EINVAL, "invalid CSR");
  an_pcg_options o;
  default_pcg_options(&o);
  if (opt) o = *opt;
  if (o.precision < 0 || o.precision > 2) throw Error(AN_EINVAL, "precision must be 0, 1 or 2");
  if (o.smoother != 0 && o.smoother != 1) throw Error(AN_EINVAL, "smoother must be 0 or 1");
  if (o.smoother == 1 && !(o.chebFraction > 0.0 && o.chebFraction < 1.0))
    throw Error(AN_EINVAL, "chebFraction must be in (0, 1)");
  auto* h = new an_solver_s;
  try {
    h->pcg.reset(new PcgSolver(n, rowOffsets, cols, nnz, o, mem == AN_DEVICE, 0));
  } catch (...) {
    delete h;
    throw;
  }
  *out = h;
  return AN_OK;
  AN_CATCH
}

an_status an_pcg_set_values(an_solver_t s, const double* values, int rebuild, an_memspace mem) {
  AN_TRY
  if (!s || !values) throw Error(AN_EINVAL, "NULL argument");
  const hipStream_t st = s->pcg->s;
  InArr<double> v(values, (size_t)s->pcg->A.nnz, mem, st);
  s->pcg->set_values_device(v.p, rebuild != 0);
  AN_HIP(hipStreamSynchronize(st));
  return AN_OK;
  AN_CATCH
}

an_status an_pcg_solve(an_solver_t s, const double* b, double* x, an_memspace mem,
                       an_solve_info* info) {
  AN_TRY
  if (!s || !b || !x) throw Error(AN_EINVAL, "NULL argument");
  const int n = s->pcg->A.n;
  const hipStream_t hs = s->pcg->s;
  InArr<double> bb(b, (size_t)n, mem, hs);
  DevBuf<double> xh;
  double* xd = x;
  if (mem == AN_HOST) {
    xh.upload(x, (size_t)n, hs);
    xd = xh.p;
  }
  an_status st = s->pcg->solve(bb.p, xd, info);
  // the solver returns with its last kernels queued; x is complete when
  // this call returns, as after HYPRE's solve
  if (mem == AN_HOST)
    AN_HIP(hipMemcpyAsync(x, xd, sizeof(double) * n, hipMemcpyDeviceToHost, hs));
  AN_HIP(hipStreamSynchronize(hs));
  if (st == AN_ENOCONV) set_last_error("PCG did not converge");
  return st;
  AN_CATCH
}

an_status an_pcg_set_stream(an_solver_t s, void* hipStream) {
  AN_TRY
  if (!s) throw Error(AN_EINVAL, "NULL argument");
  AN_HIP(hipStreamSynchronize(s->pcg->s));
  s->pcg->s = static_cast<hipStream_t>(hipStream);
  return AN_OK;
  AN_CATCH
}

an_status an_pcg_residual_history(an_solver_t s, double* out, int cap, int* count) {
  AN_TRY
  if (!s || !count || (cap > 0 && !out)) throw Error(AN_EINVAL, "NULL argument");
  const std::vector<double>& h = s->pcg->history;
  *count = (int)h.size();
  for (int i = 0; i < cap && i < (int)h.size(); ++i) out[i] = h[(size_t)i];
  return AN_OK;
  AN_CATCH
}

an_status an_spmv(an_solver_t s, const double* x, double* y, an_memspace mem) {
  AN_TRY
  if (!s || !x || !y) throw Error(AN_EINVAL, "NULL argument");
  PcgSolver& P = *s->pcg;
  InArr<double> xx(x, (size_t)P.A.n, mem, P.s);
  OutArr<double> yy(y, (size_t)P.A.n, mem);
  spmv<double, double>(P.A, xx.p, yy.p, nullptr, P.scal.p, 0, P.red, P.s);
  yy.finish(P.s);
  AN_HIP(hipStreamSynchronize(P.s));
  return AN_OK;
  AN_CATCH
}

an_status an_amg_apply(an_solver_t s, const double* r, double* z, an_memspace mem) {
  AN_TRY
  if (!s || !r || !z) throw Error(AN_EINVAL, "NULL argument");
  PcgSolver& P = *s->pcg;
  if (!P.amg.built) throw Error(AN_EINVAL, "AMG hierarchy not built");
  InArr<double> rr(r, (size_t)P.A.n, mem, P.s);
  OutArr<double> zz(z, (size_t)P.A.n, mem);
  P.apply_prec(rr.p, zz.p);
  zz.finish(P.s);
  AN_HIP(hipStreamSynchronize(P.s));
  return AN_OK;
  AN_CATCH
}

an_status an_amg_aggregates(an_solver_t s, int* agg, int* nCoarse) {
  AN_TRY
  if (!s || !agg) throw Error(AN_EINVAL, "NULL argument");
  PcgSolver& P = *s->pcg;
  if (!P.amg.built) throw Error(AN_EINVAL, "AMG hierarchy not built");
  const AmgLevel& L = P.amg.lv[0];
  if (L.isCoarsest) {
    for (int i = 0; i < L.n; ++i) agg[i] = i;
    if (nCoarse) *nCoarse = L.n;
  } else {
    AN_HIP(hipMemcpy(agg, L.agg.p, sizeof(int) * L.n, hipMemcpyDeviceToHost));
    if (nCoarse) *nCoarse = L.nc;
  }
  return AN_OK;
  AN_CATCH
}

an_status an_solver_destroy(an_solver_t s) {
  AN_TRY
  delete s;
  return AN_OK;
  AN_CATCH
}

an_status an_bicgstab_solve(int n, const int* rowOffsets, const int* cols, int64_t nnz,
                            const double* values, const double* b, double* x, int maxIter,
                            double relTol, double absTol, an_memspace mem, an_solve_info* info) {
  AN_TRY
  if (n <= 0 || !rowOffsets || !cols || !values || !b || !x) throw Error(AN_EINVAL, "NULL argument");
  DevCSR<double> A;
  A.n = n;
  A.nnz = nnz;
  A.rowOff.alloc((size_t)n + 1);
  A.cols.alloc((size_t)nnz);
  A.vals.alloc((size_t)nnz);
  hipMemcpyKind k = kind_in(mem);
  AN_HIP(hipMemcpy(A.rowOff.p, rowOffsets, sizeof(int) * ((size_t)n + 1), k));
  AN_HIP(hipMemcpy(A.cols.p, cols, sizeof(int) * (size_t)nnz, k));
  AN_HIP(hipMemcpy(A.vals.p, values, sizeof(double) * (size_t)nnz, k));
  A.finish_pattern(0);
  InArr<double> bb(b, (size_t)n, mem, 0);
  DevBuf<double> xh;
  double* xd = x;
  if (mem == AN_HOST) {
    xh.upload(x, (size_t)n, 0);
    xd = xh.p;
  }
  BicgWork w;
  an_status st = bicgstab_device(A, bb.p, xd, maxIter, relTol, absTol, w, 0, info);
  if (mem == AN_HOST) AN_HIP(hipMemcpy(x, xd, sizeof(double) * n, hipMemcpyDeviceToHost));
  if (st == AN_ENOCONV) set_last_error("BiCGSTAB did not converge");
  return st;
  AN_CATCH
}

an_status an_bicgstab_solve3(int n, const int* rowOffsets, const int* cols, int64_t nnz,
                             const double* values, const double* const* b, double* const* x,
                             int maxIter, double relTol, double absTol, an_memspace mem,
                             an_solve_info* info) {
  AN_TRY
  if (n <= 0 || !rowOffsets || !cols || !values || !b || !x) throw Error(AN_EINVAL, "NULL argument");
  for (int c = 0; c < 3; ++c)
    if (!b[c] || !x[c]) throw Error(AN_EINVAL, "NULL right-hand side or solution");
  DevCSR<double> A;
  A.n = n;
  A.nnz = nnz;
  A.rowOff.alloc((size_t)n + 1);
  A.cols.alloc((size_t)nnz);
  A.vals.alloc((size_t)nnz);
  hipMemcpyKind k = kind_in(mem);
  AN_HIP(hipMemcpy(A.rowOff.p, rowOffsets, sizeof(int) * ((size_t)n + 1), k));
  AN_HIP(hipMemcpy(A.cols.p, cols, sizeof(int) * (size_t)nnz, k));
  AN_HIP(hipMemcpy(A.vals.p, values, sizeof(double) * (size_t)nnz, k));
  A.finish_pattern(0);
  DevBuf<double> bh[3], xh[3];
  const double* bd[3];
  double* xd[3];
  for (int c = 0; c < 3; ++c) {
    bd[c] = b[c];
    xd[c] = x[c];
    if (mem == AN_HOST) {
      bh[c].upload(b[c], (size_t)n, 0);
      xh[c].upload(x[c], (size_t)n, 0);
      bd[c] = bh[c].p;
      xd[c] = xh[c].p;
    }
  }
  hipEvent_t e0, e1;
  AN_HIP(hipEventCreate(&e0));
  AN_HIP(hipEventCreate(&e1));
  Bicg3Work w;
  an_solve_info inf[3] = {};
  an_status st = bicgstab3_device(A, bd, xd, maxIter, relTol, absTol, w, 0, inf, 7, e0, e1);
  float ms = 0.f;
  const bool timed = inf[0].iterations + inf[1].iterations + inf[2].iterations > 0;
  if (timed) AN_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  for (int c = 0; c < 3; ++c) {
    inf[c].spmvCalls = timed ? 1 : 0;
    inf[c].spmvAvgSeconds = ms * 1e-3;
    if (mem == AN_HOST) AN_HIP(hipMemcpy(x[c], xd[c], sizeof(double) * n, hipMemcpyDeviceToHost));
  }
  if (info)
    for (int c = 0; c < 3; ++c) info[c] = inf[c];
  if (st == AN_ENOCONV) set_last_error("BiCGSTAB did not converge for every right-hand side");
  return st;
  AN_CATCH
}

// ------------------------------------------------- libpoisson / LS ----
void an_elliptic_default_options(an_elliptic_options* o) {
  if (!o) return;
  o->nNonOrthCorr = 2;  // EllipticOptions defaults (LP/include/poisson_library.h:5-13)
  o->nonOrth = 1;
  o->useReferenceCell = 0;
  o->referenceCell = 0;
  o->referenceValue = 0.0;
  default_pcg_options(&o->solver);
}

an_status an_solve_scalar_elliptic(an_mesh_t mesh, const double* gammaFace,
                                   const double* cellSource, const int* bcTypeBF,
                                   const double* bcValBF, const an_elliptic_options* opt,
                                   double* phi, an_memspace mem, an_solve_info* info) {
  AN_TRY
  DevMesh& d = devmesh(mesh, true);
  an_elliptic_options o;
  an_elliptic_default_options(&o);
  if (opt) o = *opt;
  hipStream_t s = 0;
  InArr<double> gam(gammaFace, (size_t)d.nF, mem, s);
  InArr<double> src(cellSource, (size_t)d.nC, mem, s);
  InArr<int> bt(bcTypeBF, (size_t)d.nBF, mem, s);
  InArr<double> bv(bcValBF, (size_t)d.nBF, mem, s);
  if (!phi) throw Error(AN_EINVAL, "phi is NULL");
  // poisson_library.cu:26-32
  std::vector<int> hbt((size_t)d.nBF);
  AN_HIP(hipMemcpy(hbt.data(), bt.p, sizeof(int) * d.nBF, hipMemcpyDeviceToHost));
  bool anyD = false;
  for (int t : hbt) anyD |= (t == AN_BC_DIRICHLET);
  if (!anyD && !o.useReferenceCell)
    throw Error(AN_EINVAL, "Pure-Neumann scalar elliptic problem needs useReferenceCell=true.");
  if (o.useReferenceCell && (o.referenceCell < 0 || o.referenceCell >= d.nC))
    throw Error(AN_EINVAL, "Invalid reference cell");
  const int nOuter = o.nonOrth ? (o.nNonOrthCorr > 0 ? o.nNonOrthCorr : 0) + 1 : 1;
  MeshView mv = view_of(d);
  DevBuf<double> x((size_t)d.nC), g(3 * (size_t)d.nC), vals((size_t)d.addr.nnz), rhs((size_t)d.nC);
  x.zero(s);
  PcgSolver P(d.nC, d.addr.rowOff.p, d.addr.cols.p, d.addr.nnz, o.solver, true, s);
  an_solve_info last{};
  an_status st = AN_OK;
  for (int outer = 0; outer < nOuter; ++outer) {
    launch_lsq_gradient(mv, 1, BCView{bt.p, bv.p}, x.p, g.p, g.p + d.nC, g.p + 2 * (size_t)d.nC, s);
    launch_elliptic_assemble(mv, gam.p, src.p, BCView{bt.p, bv.p}, g.p, g.p + d.nC,
                             g.p + 2 * (size_t)d.nC, o.nonOrth, o.useReferenceCell,
                             o.referenceCell, o.referenceValue, vals.p, rhs.p, s);
    // solve_system_hypre_gpu sets up a fresh solver per call and starts
    // from x0 = 0 (LP/src/hypre_backend.cu:35,49,94-95)
    P.set_values_device(vals.p, true);
    x.zero(s);
    st = P.solve(rhs.p, x.p, &last);
  }
  AN_HIP(hipMemcpyAsync(phi, x.p, sizeof(double) * d.nC, kind_out(mem), s));
  AN_HIP(hipStreamSynchronize(s));
  if (info) *info = last;
  if (st == AN_ENOCONV) set_last_error("PCG did not converge");
  return st;
  AN_CATCH
}

void an_transport_default_options(an_transport_options* o) {
  if (!o) return;
  o->upwind = 0;  // ScalarTransportOptions defaults (LS/include/...:84-96)
  o->nonOrth = 1;
  o->nNonOrthCorr = 2;
  o->maxIter = 4000;
  o->relTol = 0.0;
  o->absTol = 1e-10;
}

an_status an_solve_steady_scalar_transport(an_mesh_t mesh, const double* faceFlux,
                                           const double* gammaFace, const double* Su,
                                           const double* Sp, const int* bcTypeBF,
                                           const double* bcValBF, const an_transport_options* opt,
                                           const double* x0, double* phi, an_memspace mem,
                                           an_solve_info* info) {
  AN_TRY
  DevMesh& d = devmesh(mesh, true);
  an_transport_options o;
  an_transport_default_options(&o);
  if (opt) o = *opt;
  hipStream_t s = 0;
  InArr<double> ff(faceFlux, (size_t)d.nF, mem, s);
  InArr<double> gam(gammaFace, (size_t)d.nF, mem, s);
  InArr<double> su(Su, (size_t)d.nC, mem, s);
  InArr<double> sp(Sp, (size_t)d.nC, mem, s);
  InArr<int> bt(bcTypeBF, (size_t)d.nBF, mem, s);
  InArr<double> bv(bcValBF, (size_t)d.nBF, mem, s);
  if (!phi) throw Error(AN_EINVAL, "phi is NULL");
  MeshView mv = view_of(d);
  DevBuf<double> x((size_t)d.nC), g(3 * (size_t)d.nC), rhs((size_t)d.nC);
  DevCSR<double> A;
  A.n = d.nC;
  A.nnz = d.addr.nnz;
  A.rowOff.alloc((size_t)d.nC + 1);
  A.cols.alloc((size_t)A.nnz);
  A.vals.alloc((size_t)A.nnz);
  AN_HIP(hipMemcpyAsync(A.rowOff.p, d.addr.rowOff.p, sizeof(int) * ((size_t)d.nC + 1),
                        hipMemcpyDeviceToDevice, s));
  AN_HIP(hipMemcpyAsync(A.cols.p, d.addr.cols.p, sizeof(int) * (size_t)A.nnz,
                        hipMemcpyDeviceToDevice, s));
  A.finish_pattern(s);
  if (x0)
    AN_HIP(hipMemcpyAsync(x.p, x0, sizeof(double) * d.nC, kind_in(mem), s));
  else
    x.zero(s);
  const int nOuter = o.nonOrth ? (o.nNonOrthCorr > 0 ? o.nNonOrthCorr : 0) + 1 : 1;
  BicgWork w;
  an_solve_info last{};
  an_status st = AN_OK;
  for (int outer = 0; outer < nOuter; ++outer) {
    launch_lsq_gradient(mv, 1, BCView{bt.p, bv.p}, x.p, g.p, g.p + d.nC, g.p + 2 * (size_t)d.nC, s);
    launch_transport_assemble(mv, ff.p, gam.p, su.p, sp.p, BCView{bt.p, bv.p}, g.p, g.p + d.nC,
                              g.p + 2 * (size_t)d.nC, o.nonOrth, o.upwind, A.vals.p, rhs.p, s);
    // warm start from the previous pass (scalar_transport_library.cu:506)
    st = bicgstab_device(A, rhs.p, x.p, o.maxIter, o.relTol, o.absTol, w, s, &last);
  }
  AN_HIP(hipMemcpyAsync(phi, x.p, sizeof(double) * d.nC, kind_out(mem), s));
  AN_HIP(hipStreamSynchronize(s));
  if (info) *info = last;
  if (st == AN_ENOCONV) set_last_error("BiCGSTAB did not converge");
  return st;
  AN_CATCH
}


// ------------------------------------------------------------ SIMPLE ----
void an_simple_default_params(an_simple_params* p) {
  if (!p) return;
  // generic_simple_v1 Params defaults (main.cu:73-90) with the validated
  // cylinder controls of cases/cylinder.case:26-50
  p->rho = 1.0;
  p->mu = 1e-3;
  p->nVelNonOrthCorr = 1;
  p->nNonOrthCorr = 10;
  p->uRelax = 0.5;
  p->pRelax = 0.3;
  p->corrTol = 1e-10;
  p->tolMass = 1e-6;
  p->tolVel = 1e-6;
  p->velMaxit = 1000;
  p->velTol = 1e-10;
  p->velRelTol = 0.0;
  p->pMaxit = 4000;
  p->pTol = 1e-10;
  p->pRelTol = 0.0;
  p->pAmgRebuildEvery = 10;
  p->precision = 0;
  p->fuseMomentumRhs = 1;  // bit-identical to the component-by-component order
  default_pcg_options(&p->pressure);
  // the SIMPLE pressure corrections are short warm-started solves on a
  // hierarchy rebuilt every pAmgRebuildEvery iterations: one K-cycle level,
  // a lower strength threshold and omega 0.9 give the fastest iteration on
  // the 2.08M-cell cylinder case; four pairwise matchings per level (about
  // 16 cells per aggregate) and a dense coarsest level of up to 1024 rows
  // halve the chain of small, latency-bound coarse levels (2.08M cells:
  // 2,079,000 -> ~130k -> ~8k -> ~500 dense; 23.1 -> 25.6 it/s measured)
  p->pressure.kcycle = 1;
  p->pressure.strength = 0.1;
  p->pressure.jacobiOmega = 0.9;
  p->pressure.pairPasses = 4;
  p->pressure.coarseSize = 1024;
  // the non-orthogonal passes solve one matrix with successive right-hand
  // sides: start each from the projection onto the earlier passes' solutions
  p->pressure.rhsProjection = 6;
}

an_status an_simple_create(an_mesh_t mesh, const an_simple_params* par, const int* patchUType,
                           const int* patchPType, const double* uFaceBC, const double* vFaceBC,
                           const double* wFaceBC, const double* pFaceBC, int inletPatch,
                           an_simple_t* out) {
  AN_TRY
  if (!mesh || !par || !patchUType || !patchPType || !uFaceBC || !vFaceBC || !wFaceBC ||
      !pFaceBC || !out)
    throw Error(AN_EINVAL, "NULL argument");
  if (par->precision < 0 || par->precision > 2) throw Error(AN_EINVAL, "precision must be 0, 1 or 2");
  if (!(par->uRelax > 0.0) || !(par->mu >= 0.0)) throw Error(AN_EINVAL, "invalid uRelax / mu");
  auto* h = new an_simple_s;
  try {
    h->sim.reset(new SimpleSolver(mesh->dm, *par, patchUType, patchPType, uFaceBC, vFaceBC,
                                  wFaceBC, pFaceBC, inletPatch));
  } catch (...) {
    delete h;
    throw;
  }
  *out = h;
  return AN_OK;
  AN_CATCH
}

an_status an_simple_iterate(an_simple_t s, int n, an_simple_stats* last, int* converged) {
  AN_TRY
  if (!s || n < 0) throw Error(AN_EINVAL, "invalid argument");
  an_status st = s->sim->iterate(n, last, converged);
  if (st == AN_EDIVERGED) set_last_error("SIMPLE-like steady solve diverged.");
  return st;
  AN_CATCH
}

an_status an_simple_get_field(an_simple_t s, const char* name, double* out, an_memspace mem) {
  AN_TRY
  if (!s || !name || !out) throw Error(AN_EINVAL, "NULL argument");
  s->sim->get_field(name, out, mem == AN_DEVICE);
  return AN_OK;
  AN_CATCH
}

an_status an_simple_set_field(an_simple_t s, const char* name, const double* in,
                              an_memspace mem) {
  AN_TRY
  if (!s || !name || !in) throw Error(AN_EINVAL, "NULL argument");
  s->sim->set_field(name, in, mem == AN_DEVICE);
  return AN_OK;
  AN_CATCH
}

an_status an_simple_patch_forces(an_simple_t s, int patch, int normalSign, double* out) {
  AN_TRY
  if (!s || !out) throw Error(AN_EINVAL, "NULL argument");
  s->sim->patch_forces(patch, normalSign, out);
  return AN_OK;
  AN_CATCH
}

an_status an_simple_timings(an_simple_t s, double* out) {
  AN_TRY
  if (!s || !out) throw Error(AN_EINVAL, "NULL argument");
  s->sim->timings(out);
  return AN_OK;
  AN_CATCH
}

an_status an_simple_destroy(an_simple_t s) {
  AN_TRY
  delete s;
  return AN_OK;
  AN_CATCH
}

// ----------------------------------------------------------- adjoint ----
an_status an_face_gamma_from_cells(an_mesh_t mesh, const double* kCell, double* gammaFace,
                                   an_memspace mem) {
  AN_TRY
  if (!kCell || !gammaFace) throw Error(AN_EINVAL, "NULL argument");
  DevMesh& d = devmesh(mesh, true);
  hipStream_t s = 0;
  InArr<double> k(kCell, (size_t)d.nC, mem, s);
  DevBuf<double> g((size_t)d.nF);
  launch_gamma_from_cells(view_of(d), k.p, g.p, s);
  AN_HIP(hipMemcpyAsync(gammaFace, g.p, sizeof(double) * d.nF, kind_out(mem), s));
  AN_HIP(hipStreamSynchronize(s));
  return AN_OK;
  AN_CATCH
}

an_status an_dgamma_to_cells(an_mesh_t mesh, const double* dJdgammaFace, double* dJdkCell,
                             an_memspace mem) {
  AN_TRY
  if (!dJdgammaFace || !dJdkCell) throw Error(AN_EINVAL, "NULL argument");
  DevMesh& d = devmesh(mesh, true);
  hipStream_t s = 0;
  InArr<double> dg(dJdgammaFace, (size_t)d.nF, mem, s);
  DevBuf<double> dk((size_t)d.nC);
  launch_dk_from_dgamma(view_of(d), dg.p, dk.p, s);
  AN_HIP(hipMemcpyAsync(dJdkCell, dk.p, sizeof(double) * d.nC, kind_out(mem), s));
  AN_HIP(hipStreamSynchronize(s));
  return AN_OK;
  AN_CATCH
}

an_status an_lsq_gradient_adjoint(an_mesh_t mesh, const int* bcTypeBF, const double* bcValBF,
                                  const double* w, double* phibar, double* phibarBF,
                                  an_memspace mem) {
  AN_TRY
  if (!w || !phibar) throw Error(AN_EINVAL, "NULL argument");
  DevMesh& d = devmesh(mesh, true);
  hipStream_t s = 0;
  InArr<int> bt(bcTypeBF, (size_t)d.nBF, mem, s);
  InArr<double> bv(bcValBF, (size_t)d.nBF, mem, s);
  InArr<double> ww(w, 3 * (size_t)d.nC, mem, s);
  DevBuf<double> z(3 * (size_t)d.nC), pb((size_t)d.nC), pbB((size_t)(d.nBF > 0 ? d.nBF : 1));
  launch_lsq_adjoint(view_of(d), BCView{bt.p, bv.p}, ww.p, z.p, pb.p, pbB.p, s);
  AN_HIP(hipMemcpyAsync(phibar, pb.p, sizeof(double) * d.nC, kind_out(mem), s));
  if (phibarBF && d.nBF > 0)
    AN_HIP(hipMemcpyAsync(phibarBF, pbB.p, sizeof(double) * d.nBF, kind_out(mem), s));
  AN_HIP(hipStreamSynchronize(s));
  return AN_OK;
  AN_CATCH
}

an_status an_transport_adjoint(an_mesh_t mesh, const double* faceFlux, const double* gammaFace,
                               const double* Su, const double* Sp, const int* bcTypeBF,
                               const double* bcValBF, const an_transport_options* opt,
                               const double* phi, const double* dJdphi, double* dJdgammaFace,
                               double* dJdbcBF, double* adjoint, an_memspace mem,
                               an_solve_info* info) {
  AN_TRY
  if (!phi || !dJdphi || !dJdgammaFace) throw Error(AN_EINVAL, "NULL argument");
  DevMesh& d = devmesh(mesh, true);
  an_transport_options o;
  an_transport_default_options(&o);
  if (opt) o = *opt;
  hipStream_t s = 0;
  InArr<double> ff(faceFlux, (size_t)d.nF, mem, s);
  InArr<double> gam(gammaFace, (size_t)d.nF, mem, s);
  InArr<double> su(Su, (size_t)d.nC, mem, s);
  InArr<double> sp(Sp, (size_t)d.nC, mem, s);
  InArr<int> bt(bcTypeBF, (size_t)d.nBF, mem, s);
  InArr<double> bv(bcValBF, (size_t)d.nBF, mem, s);
  InArr<double> ph(phi, (size_t)d.nC, mem, s);
  InArr<double> dj(dJdphi, (size_t)d.nC, mem, s);
  DevBuf<double> dg((size_t)d.nF), db((size_t)(d.nBF > 0 ? d.nBF : 1)), psi((size_t)d.nC);
  const int nOuter = o.nonOrth ? (o.nNonOrthCorr > 0 ? o.nNonOrthCorr : 0) + 1 : 1;
  an_solve_info last{};
  an_status st = transport_adjoint(d, view_of(d), ff.p, gam.p, su.p, sp.p, BCView{bt.p, bv.p},
                                   o.upwind, o.nonOrth, nOuter, o.maxIter, o.relTol, o.absTol,
                                   ph.p, dj.p, dg.p, db.p, psi.p, s, &last);
  AN_HIP(hipMemcpyAsync(dJdgammaFace, dg.p, sizeof(double) * d.nF, kind_out(mem), s));
  if (dJdbcBF && d.nBF > 0)
    AN_HIP(hipMemcpyAsync(dJdbcBF, db.p, sizeof(double) * d.nBF, kind_out(mem), s));
  if (adjoint) AN_HIP(hipMemcpyAsync(adjoint, psi.p, sizeof(double) * d.nC, kind_out(mem), s));
  AN_HIP(hipStreamSynchronize(s));
  if (info) *info = last;
  if (st == AN_ENOCONV) set_last_error("adjoint BiCGSTAB did not converge");
  return st;
  AN_CATCH
}

// ------------------------------------------------------------- batch ----
struct an_batch_s {
  std::unique_ptr<BatchSolver> b;
};

an_status an_batch_create(int nProblems, const int* nRows, const int* const* rowOffsets,
                          const int* const* cols, const an_pcg_options* opt, an_memspace mem,
                          an_batch_t* out) {
  AN_TRY
  if (nProblems <= 0 || !nRows || !rowOffsets || !cols || !out)
    throw Error(AN_EINVAL, "NULL argument");
  an_pcg_options o;
  default_pcg_options(&o);
  if (opt) o = *opt;
  if (o.precision != 0 && o.precision != 1)
    throw Error(AN_EINVAL, "batched solves take precision 0 or 1");
  if (o.smoother != 0) throw Error(AN_EINVAL, "batched solves take the Jacobi smoother");
  std::vector<int> rows(nRows, nRows + nProblems);
  std::vector<const int*> ro(rowOffsets, rowOffsets + nProblems), ci(cols, cols + nProblems);
  auto h = std::make_unique<an_batch_s>();
  h->b.reset(new BatchSolver(nProblems, rows, ro, ci, mem == AN_DEVICE, o));
  *out = h.release();
  return AN_OK;
  AN_CATCH
}

an_status an_batch_set_values(an_batch_t b, const double* const* values, an_memspace mem) {
  AN_TRY
  if (!b || !values) throw Error(AN_EINVAL, "NULL argument");
  std::vector<const double*> v(values, values + b->b->nP);
  b->b->set_values(v, mem == AN_DEVICE);
  return AN_OK;
  AN_CATCH
}

an_status an_batch_solve(an_batch_t b, const double* const* rhs, double* const* x,
                         an_memspace mem, an_solve_info* info) {
  AN_TRY
  if (!b || !rhs || !x) throw Error(AN_EINVAL, "NULL argument");
  const int nP = b->b->nP;
  std::vector<const double*> bb(rhs, rhs + nP);
  std::vector<double*> xx(x, x + nP);
  std::vector<an_solve_info> inf;
  an_status st = b->b->solve(bb, xx, mem == AN_DEVICE, inf);
  if (info)
    for (int p = 0; p < nP; ++p) info[p] = inf[p];
  if (st == AN_ENOCONV) set_last_error("batched PCG: a problem did not converge");
  return st;
  AN_CATCH
}

an_status an_batch_tuning(an_batch_t b, int* rowsPerWave, double* seconds) {
  AN_TRY
  if (!b) throw Error(AN_EINVAL, "NULL argument");
  if (rowsPerWave) *rowsPerWave = b->b->lay.rows;
  if (seconds)
    for (size_t i = 0; i < 3; ++i) seconds[i] = i < b->b->tuneSeconds.size() ? b->b->tuneSeconds[i] : 0.0;
  return AN_OK;
  AN_CATCH
}

an_status an_batch_use_graph(an_batch_t b, int enable) {
  AN_TRY
  if (!b) throw Error(AN_EINVAL, "NULL argument");
  b->b->launch = enable ? AN_BATCH_AUTO : AN_BATCH_EAGER;
  return AN_OK;
  AN_CATCH
}

an_status an_batch_set_launch(an_batch_t b, int launch) {
  AN_TRY
  if (!b) throw Error(AN_EINVAL, "NULL argument");
  if (launch < AN_BATCH_AUTO || launch > AN_BATCH_PERSISTENT)
    throw Error(AN_EINVAL, "unknown batch launch mode");
  b->b->launch = launch;
  return AN_OK;
  AN_CATCH
}

an_status an_batch_get_tuning(an_batch_t b, an_batch_tuning_info* out) {
  AN_TRY
  if (!b || !out) throw Error(AN_EINVAL, "NULL argument");
  const BatchSolver& s = *b->b;
  *out = an_batch_tuning_info{};
  out->rowsPerWave = s.lay.rows;
  for (size_t i = 0; i < 5 && i < s.tuneSeconds.size(); ++i) out->rowSeconds[i] = s.tuneSeconds[i];
  out->launch = s.chosen;
  out->persistentBlocksPerCU = s.persistBlocksPerCU;
  for (size_t i = 0; i < 4 && i < s.launchSeconds.size(); ++i)
    out->launchSeconds[i] = s.launchSeconds[i];
  out->launchesLastSolve = s.lastLaunches;
  return AN_OK;
  AN_CATCH
}

an_status an_batch_destroy(an_batch_t b) {
  AN_TRY
  delete b;
  return AN_OK;
  AN_CATCH
}

}  // extern "C"
