// host_mesh.cpp -- polyMesh I/O, mesh generators and reference geometry.
#include "host_mesh.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <random>
#include <sstream>
#include <stdexcept>
#include <sys/stat.h>
#include <unordered_map>

namespace an {

// ------------------------------------------------------------ reading ---
static std::string slurp(const std::string& fn) {
  std::ifstream in(fn.c_str(), std::ios::binary);
  if (!in) throw std::runtime_error("Could not open " + fn);
  std::ostringstream ss;
  ss << in.rdbuf();
  return ss.str();
}

// Removes /* */ then // comments (LP/src/mesh.cu:11-29).
static std::string strip_comments(const std::string& txt) {
  std::string s;
  s.reserve(txt.size());
  size_t i = 0;
  while (i < txt.size()) {
    if (txt[i] == '/' && i + 1 < txt.size() && txt[i + 1] == '*') {
      size_t e = txt.find("*/", i + 2);
      if (e == std::string::npos) {  // unterminated: reference keeps the rest
        s.append(txt, i, std::string::npos);
        break;
      }
      i = e + 2;
      continue;
    }
    s.push_back(txt[i++]);
  }
  std::string out;
  out.reserve(s.size());
  size_t a = 0;
  while (a <= s.size()) {
    size_t e = s.find('\n', a);
    if (e == std::string::npos) e = s.size();
    // the first "//" of this line (searching only the line keeps the pass
    // linear in the file size)
    size_t stop = e;
    for (size_t k = a; k + 1 < e; ++k)
      if (s[k] == '/' && s[k + 1] == '/') {
        stop = k;
        break;
      }
    out.append(s, a, stop - a);
    out.push_back('\n');
    if (e == s.size()) break;
    a = e + 1;
  }
  return out;
}

// First "<digits><ws>(" ... matching ")" (LP/src/mesh.cu:31-58); returns
// [begin,end) of the list body.
static void main_list(const std::string& t, size_t& b, size_t& e) {
  size_t start = std::string::npos;
  for (size_t i = 0; i < t.size(); ++i) {
    if (std::isdigit((unsigned char)t[i])) {
      size_t j = i;
      while (j < t.size() && std::isdigit((unsigned char)t[j])) ++j;
      while (j < t.size() && std::isspace((unsigned char)t[j])) ++j;
      if (j < t.size() && t[j] == '(') {
        start = j;
        break;
      }
      i = j > i ? j - 1 : i;
    }
  }
  if (start == std::string::npos)
    throw std::runtime_error("Could not locate top-level OpenFOAM list");
  int depth = 0;
  for (size_t i = start; i < t.size(); ++i) {
    if (t[i] == '(')
      ++depth;
    else if (t[i] == ')') {
      if (--depth == 0) {
        b = start + 1;
        e = i;
        return;
      }
    }
  }
  throw std::runtime_error("Failed to match parentheses");
}

static std::vector<int> read_labels(const std::string& fn) {
  std::string t = strip_comments(slurp(fn));
  size_t b, e;
  main_list(t, b, e);
  std::vector<int> v;
  const char* p = t.c_str() + b;
  const char* end = t.c_str() + e;
  while (p < end) {
    while (p < end && std::isspace((unsigned char)*p)) ++p;
    if (p >= end) break;
    char* q;
    long x = std::strtol(p, &q, 10);
    if (q == p) break;  // reference stops at the first non-integer
    v.push_back((int)x);
    p = q;
  }
  return v;
}

static std::vector<double> read_points(const std::string& fn) {
  std::string t = strip_comments(slurp(fn));
  size_t b, e;
  main_list(t, b, e);
  std::vector<double> P;
  size_t pos = b;
  while (true) {
    size_t a = t.find('(', pos);
    if (a == std::string::npos || a >= e) break;
    size_t c = t.find(')', a + 1);
    if (c == std::string::npos || c > e) break;
    const char* p = t.c_str() + a + 1;
    double x[3] = {0, 0, 0};
    for (int k = 0; k < 3; ++k) {
      char* q;
      x[k] = std::strtod(p, &q);
      p = q;
    }
    P.push_back(x[0]);
    P.push_back(x[1]);
    P.push_back(x[2]);
    pos = c + 1;
  }
  return P;
}

static void read_faces(const std::string& fn, std::vector<int>& off,
                       std::vector<int>& verts) {
  std::string t = strip_comments(slurp(fn));
  size_t b, e;
  main_list(t, b, e);
  off.assign(1, 0);
  verts.clear();
  size_t pos = b;
  while (pos < e) {
    while (pos < e && std::isspace((unsigned char)t[pos])) ++pos;
    if (pos >= e) break;
    if (!std::isdigit((unsigned char)t[pos])) {
      ++pos;
      continue;
    }
    size_t q = pos;
    while (q < e && std::isdigit((unsigned char)t[q])) ++q;
    int k = std::atoi(t.substr(pos, q - pos).c_str());
    while (q < e && std::isspace((unsigned char)t[q])) ++q;
    if (q >= e || t[q] != '(') throw std::runtime_error("Malformed faces list");
    size_t r = t.find(')', q + 1);
    if (r == std::string::npos) throw std::runtime_error("Malformed faces list");
    const char* p = t.c_str() + q + 1;
    for (int i = 0; i < k; ++i) {
      char* qq;
      long x = std::strtol(p, &qq, 10);
      verts.push_back((int)x);
      p = qq;
    }
    off.push_back((int)verts.size());
    pos = r + 1;
  }
}

static void read_boundary(const std::string& fn, HostMesh& m) {
  std::string t = strip_comments(slurp(fn));
  size_t b, e;
  main_list(t, b, e);
  std::string inside = t.substr(b, e - b);
  size_t pos = 0;
  while (pos < inside.size()) {
    while (pos < inside.size() && std::isspace((unsigned char)inside[pos])) ++pos;
    if (pos >= inside.size()) break;
    if (!(std::isalpha((unsigned char)inside[pos]) || inside[pos] == '_')) {
      ++pos;
      continue;
    }
    size_t a = pos;
    while (pos < inside.size() &&
           (std::isalnum((unsigned char)inside[pos]) || inside[pos] == '_'))
      ++pos;
    std::string name = inside.substr(a, pos - a);
    while (pos < inside.size() && std::isspace((unsigned char)inside[pos])) ++pos;
    if (pos >= inside.size() || inside[pos] != '{') continue;
    int depth = 1;
    size_t bodyStart = ++pos;
    while (pos < inside.size() && depth > 0) {
      if (inside[pos] == '{')
        ++depth;
      else if (inside[pos] == '}')
        --depth;
      ++pos;
    }
    std::string body = inside.substr(bodyStart, pos - bodyStart - 1);
    auto find_int = [&](const std::string& key) -> int {
      size_t k = body.find(key);
      if (k == std::string::npos) return 0;
      k += key.size();
      while (k < body.size() && !std::isdigit((unsigned char)body[k]) && body[k] != '-') ++k;
      size_t q = k;
      while (q < body.size() && (std::isdigit((unsigned char)body[q]) || body[q] == '-')) ++q;
      return std::atoi(body.substr(k, q - k).c_str());
    };
    auto find_word = [&](const std::string& key) -> std::string {
      size_t k = body.find(key);
      if (k == std::string::npos) return "";
      k += key.size();
      while (k < body.size() && std::isspace((unsigned char)body[k])) ++k;
      size_t q = k;
      while (q < body.size() && (std::isalnum((unsigned char)body[q]) || body[q] == '_')) ++q;
      return body.substr(k, q - k);
    };
    m.patchNames.push_back(name);
    m.patchNFaces.push_back(find_int("nFaces"));
    m.patchStart.push_back(find_int("startFace"));
    m.patchTypes.push_back(find_word("type"));
  }
}

void finish_topology(HostMesh& m) {
  m.nInternalFaces = (int)m.neigh.size();
  m.nCells = 0;
  for (int v : m.owner) m.nCells = std::max(m.nCells, v + 1);
  for (int v : m.neigh) m.nCells = std::max(m.nCells, v + 1);
  m.bPatch.assign(m.nFaces, 0);
  for (size_t k = 0; k < m.patchNames.size(); ++k) {
    int s = m.patchStart[k], n = m.patchNFaces[k];
    if (s < 0 || n < 0 || (long long)s + n > m.nFaces)
      throw std::runtime_error("Invalid patch face range for patch '" + m.patchNames[k] + "'");
    for (int f = s; f < s + n; ++f) m.bPatch[f] = (int)k + 1;
  }
}

// ------------------------------------------------------- binary cache ---
// <dir>/.anabasis_mesh.bin: the parsed mesh and its geometry, written after
// an ASCII parse and used instead of the ASCII files while their sizes and
// modification times are the ones recorded in it.  Best effort both ways:
// an unwritable directory or a stale, truncated or foreign cache file just
// means the ASCII files are parsed.  ANABASIS_MESH_CACHE=0 switches it off.
namespace {
constexpr char kCacheMagic[8] = {'A', 'N', 'M', 'E', 'S', 'H', '0', '1'};
const char* const kMeshFiles[5] = {"boundary", "points", "faces", "owner", "neighbour"};

bool cache_enabled() {
  const char* e = std::getenv("ANABASIS_MESH_CACHE");
  return !(e && std::strcmp(e, "0") == 0);
}

// FNV-1a of the first and last 4 KiB of a file (with its size and mtime:
// a cheap fingerprint for files rewritten within one mtime tick)
int64_t edge_hash(const std::string& fn, int64_t size) {
  std::FILE* f = std::fopen(fn.c_str(), "rb");
  if (!f) return 0;
  uint64_t h = 1469598103934665603ull;
  unsigned char buf[4096];
  for (int part = 0; part < 2; ++part) {
    if (part == 1 && std::fseek(f, (long)std::max<int64_t>(0, size - 4096), SEEK_SET) != 0) break;
    const size_t n = std::fread(buf, 1, sizeof buf, f);
    for (size_t i = 0; i < n; ++i) h = (h ^ buf[i]) * 1099511628211ull;
  }
  std::fclose(f);
  return (int64_t)h;
}

// size, mtime (ns) and edge hash of each ASCII file; false if one is missing
constexpr int kStamps = 15;
bool file_stamps(const std::string& dir, int64_t out[kStamps]) {
  for (int i = 0; i < 5; ++i) {
    const std::string fn = dir + "/" + kMeshFiles[i];
    struct stat st;
    if (::stat(fn.c_str(), &st) != 0) return false;
    out[3 * i] = (int64_t)st.st_size;
    out[3 * i + 1] = (int64_t)st.st_mtim.tv_sec * 1000000000LL + (int64_t)st.st_mtim.tv_nsec;
    out[3 * i + 2] = edge_hash(fn, (int64_t)st.st_size);
  }
  return true;
}

struct CacheWriter {
  std::FILE* f;
  bool ok = true;
  void bytes(const void* p, size_t n) { ok = ok && (n == 0 || std::fwrite(p, 1, n, f) == n); }
  void i64(int64_t v) { bytes(&v, sizeof v); }
  template <typename T>
  void vec(const std::vector<T>& v) {
    i64((int64_t)v.size());
    bytes(v.data(), v.size() * sizeof(T));
  }
  void str(const std::string& s) {
    i64((int64_t)s.size());
    bytes(s.data(), s.size());
  }
};

struct CacheReader {
  std::FILE* f;
  bool ok = true;
  void bytes(void* p, size_t n) { ok = ok && (n == 0 || std::fread(p, 1, n, f) == n); }
  int64_t i64() {
    int64_t v = 0;
    bytes(&v, sizeof v);
    return v;
  }
  template <typename T>
  void vec(std::vector<T>& v, int64_t expect) {
    const int64_t n = i64();
    if (!ok || n != expect || n < 0) {
      ok = false;
      return;
    }
    v.resize((size_t)n);
    bytes(v.data(), (size_t)n * sizeof(T));
  }
  std::string str() {
    const int64_t n = i64();
    if (!ok || n < 0 || n > 4096) {
      ok = false;
      return {};
    }
    std::string s((size_t)n, ' ');
    bytes(&s[0], (size_t)n);
    return s;
  }
};

bool load_cache(const std::string& dir, HostMesh& m) {
  int64_t want[kStamps];
  if (!cache_enabled() || !file_stamps(dir, want)) return false;
  std::FILE* f = std::fopen((dir + "/.anabasis_mesh.bin").c_str(), "rb");
  if (!f) return false;
  CacheReader r{f};
  char magic[8];
  r.bytes(magic, sizeof magic);
  bool ok = r.ok && std::memcmp(magic, kCacheMagic, sizeof magic) == 0;
  for (int i = 0; ok && i < kStamps; ++i) ok = r.i64() == want[i];
  if (ok) {
    m.nPoints = (int)r.i64();
    m.nFaces = (int)r.i64();
    m.nInternalFaces = (int)r.i64();
    m.nCells = (int)r.i64();
    const int64_t nPatches = r.i64(), nFaceVerts = r.i64();
    ok = r.ok && m.nPoints >= 0 && m.nFaces >= 0 && m.nInternalFaces >= 0 && m.nCells >= 0 &&
         nPatches >= 0 && nPatches < 100000 && nFaceVerts >= 0;
    for (int64_t k = 0; ok && k < nPatches; ++k) {
      m.patchNames.push_back(r.str());
      m.patchTypes.push_back(r.str());
      m.patchStart.push_back((int)r.i64());
      m.patchNFaces.push_back((int)r.i64());
      ok = r.ok;
    }
    const int64_t nP = m.nPoints, nF = m.nFaces, nC = m.nCells;
    r.vec(m.P, 3 * nP);
    r.vec(m.faceOff, nF + 1);
    r.vec(m.faceVerts, nFaceVerts);
    r.vec(m.owner, nF);
    r.vec(m.neigh, (int64_t)m.nInternalFaces);
    r.vec(m.bPatch, nF);
    r.vec(m.cc, 3 * nC);
    r.vec(m.vol, nC);
    r.vec(m.xf, 3 * nF);
    r.vec(m.nf, 3 * nF);
    r.vec(m.Sf, 3 * nF);
    r.vec(m.Af, nF);
    r.bytes(&m.maxNonOrthDeg, sizeof m.maxNonOrthDeg);
    char tail = 0;
    ok = ok && r.ok && std::fread(&tail, 1, 1, f) == 0;  // nothing after the record
  }
  std::fclose(f);
  if (!ok) {
    m = HostMesh();
    return false;
  }
  m.hasPoints = m.hasGeometry = true;
  return true;
}

void save_cache(const std::string& dir, const HostMesh& m) {
  int64_t stamps[kStamps];
  if (!cache_enabled() || !file_stamps(dir, stamps)) return;
  const std::string tmp = dir + "/.anabasis_mesh.bin.tmp";
  std::FILE* f = std::fopen(tmp.c_str(), "wb");
  if (!f) return;
  CacheWriter w{f};
  w.bytes(kCacheMagic, sizeof kCacheMagic);
  for (int64_t v : stamps) w.i64(v);
  w.i64(m.nPoints);
  w.i64(m.nFaces);
  w.i64(m.nInternalFaces);
  w.i64(m.nCells);
  w.i64((int64_t)m.patchNames.size());
  w.i64((int64_t)m.faceVerts.size());
  for (size_t k = 0; k < m.patchNames.size(); ++k) {
    w.str(m.patchNames[k]);
    w.str(k < m.patchTypes.size() ? m.patchTypes[k] : std::string());
    w.i64(m.patchStart[k]);
    w.i64(m.patchNFaces[k]);
  }
  w.vec(m.P);
  w.vec(m.faceOff);
  w.vec(m.faceVerts);
  w.vec(m.owner);
  w.vec(m.neigh);
  w.vec(m.bPatch);
  w.vec(m.cc);
  w.vec(m.vol);
  w.vec(m.xf);
  w.vec(m.nf);
  w.vec(m.Sf);
  w.vec(m.Af);
  w.bytes(&m.maxNonOrthDeg, sizeof m.maxNonOrthDeg);
  const bool ok = w.ok && std::fclose(f) == 0;
  if (ok)
    std::rename(tmp.c_str(), (dir + "/.anabasis_mesh.bin").c_str());
  else
    std::remove(tmp.c_str());
}
}  // namespace

HostMesh read_polymesh(const std::string& dir) {
  HostMesh m;
  if (load_cache(dir, m)) return m;
  read_boundary(dir + "/boundary", m);
  m.P = read_points(dir + "/points");
  m.nPoints = (int)(m.P.size() / 3);
  read_faces(dir + "/faces", m.faceOff, m.faceVerts);
  m.nFaces = (int)m.faceOff.size() - 1;
  m.owner = read_labels(dir + "/owner");
  m.neigh = read_labels(dir + "/neighbour");
  if ((int)m.owner.size() < m.nFaces)
    throw std::runtime_error("owner list shorter than faces list");
  m.owner.resize(m.nFaces);
  if ((int)m.neigh.size() > m.nFaces)
    throw std::runtime_error("neighbour list longer than faces list");
  for (int v : m.faceVerts)
    if (v < 0 || v >= m.nPoints) throw std::runtime_error("face vertex out of range");
  m.hasPoints = true;
  finish_topology(m);
  compute_geometry(m);
  save_cache(dir, m);
  return m;
}

// ------------------------------------------------------------ writing ---
static void write_header(std::FILE* f, const char* cls, const char* obj) {
  std::fprintf(f,
               "FoamFile\n{\n    format      ascii;\n    class       %s;\n"
               "    location    \"constant/polyMesh\";\n    object      %s;\n}\n\n",
               cls, obj);
}

void write_polymesh(const HostMesh& m, const std::string& dir) {
  ::mkdir(dir.c_str(), 0755);
  std::remove((dir + "/.anabasis_mesh.bin").c_str());  // describes the old files
  auto open = [&](const char* name) {
    std::FILE* f = std::fopen((dir + "/" + name).c_str(), "w");
    if (!f) throw std::runtime_error("Could not write " + dir + "/" + name);
    return f;
  };
  std::FILE* f = open("points");
  write_header(f, "vectorField", "points");
  std::fprintf(f, "%d\n(\n", m.nPoints);
  for (int i = 0; i < m.nPoints; ++i)
    std::fprintf(f, "(%.17g %.17g %.17g)\n", m.P[3 * i], m.P[3 * i + 1], m.P[3 * i + 2]);
  std::fprintf(f, ")\n");
  std::fclose(f);
  f = open("faces");
  write_header(f, "faceList", "faces");
  std::fprintf(f, "%d\n(\n", m.nFaces);
  for (int i = 0; i < m.nFaces; ++i) {
    std::fprintf(f, "%d(", m.faceOff[i + 1] - m.faceOff[i]);
    for (int k = m.faceOff[i]; k < m.faceOff[i + 1]; ++k)
      std::fprintf(f, k + 1 < m.faceOff[i + 1] ? "%d " : "%d", m.faceVerts[k]);
    std::fprintf(f, ")\n");
  }
  std::fprintf(f, ")\n");
  std::fclose(f);
  f = open("owner");
  write_header(f, "labelList", "owner");
  std::fprintf(f, "%d\n(\n", m.nFaces);
  for (int i = 0; i < m.nFaces; ++i) std::fprintf(f, "%d\n", m.owner[i]);
  std::fprintf(f, ")\n");
  std::fclose(f);
  f = open("neighbour");
  write_header(f, "labelList", "neighbour");
  std::fprintf(f, "%d\n(\n", m.nInternalFaces);
  for (int i = 0; i < m.nInternalFaces; ++i) std::fprintf(f, "%d\n", m.neigh[i]);
  std::fprintf(f, ")\n");
  std::fclose(f);
  f = open("boundary");
  write_header(f, "polyBoundaryMesh", "boundary");
  std::fprintf(f, "%d\n(\n", (int)m.patchNames.size());
  for (size_t k = 0; k < m.patchNames.size(); ++k)
    std::fprintf(f, "    %s\n    {\n        type            %s;\n        nFaces          %d;\n"
                    "        startFace       %d;\n    }\n",
                 m.patchNames[k].c_str(),
                 m.patchTypes.size() > k && !m.patchTypes[k].empty() ? m.patchTypes[k].c_str()
                                                                     : "patch",
                 m.patchNFaces[k], m.patchStart[k]);
  std::fprintf(f, ")\n");
  std::fclose(f);
}

// ------------------------------------------------------------ geometry --
void compute_geometry(HostMesh& m) {
  const int nC = m.nCells, nF = m.nFaces, nIF = m.nInternalFaces;
  if (!m.hasPoints) throw std::runtime_error("mesh has no points/faces: geometry unavailable");
  // cell -> faces: owner faces in face order, then neighbour faces
  std::vector<int> off(nC + 1, 0);
  for (int f = 0; f < nF; ++f) off[m.owner[f] + 1]++;
  for (int f = 0; f < nIF; ++f) off[m.neigh[f] + 1]++;
  for (int c = 0; c < nC; ++c) off[c + 1] += off[c];
  std::vector<int> cf(off[nC]);
  std::vector<signed char> co(off[nC]);
  {
    std::vector<int> fill(off.begin(), off.end() - 1);
    for (int f = 0; f < nF; ++f) {
      int p = m.owner[f];
      cf[fill[p]] = f;
      co[fill[p]++] = 1;
    }
    for (int f = 0; f < nIF; ++f) {
      int n = m.neigh[f];
      cf[fill[n]] = f;
      co[fill[n]++] = -1;
    }
  }
  m.cc.assign(3 * (size_t)nC, 0.0);
  m.vol.assign(nC, 0.0);
  const double* P = m.P.data();
  std::vector<int> verts;
  std::vector<int> fv;
  for (int c = 0; c < nC; ++c) {
    verts.clear();
    for (int j = off[c]; j < off[c + 1]; ++j) {
      int f = cf[j];
      verts.insert(verts.end(), m.faceVerts.begin() + m.faceOff[f],
                   m.faceVerts.begin() + m.faceOff[f + 1]);
    }
    std::sort(verts.begin(), verts.end());
    verts.erase(std::unique(verts.begin(), verts.end()), verts.end());
    double c0[3] = {0, 0, 0};
    for (int v : verts) {
      c0[0] = c0[0] + P[3 * (size_t)v];
      c0[1] = c0[1] + P[3 * (size_t)v + 1];
      c0[2] = c0[2] + P[3 * (size_t)v + 2];
    }
    double s = 1.0 / (double)std::max((int)verts.size(), 1);
    c0[0] = s * c0[0];
    c0[1] = s * c0[1];
    c0[2] = s * c0[2];
    double V = 0.0, M[3] = {0, 0, 0};
    for (int j = off[c]; j < off[c + 1]; ++j) {
      int f = cf[j];
      fv.assign(m.faceVerts.begin() + m.faceOff[f], m.faceVerts.begin() + m.faceOff[f + 1]);
      if (co[j] < 0) std::reverse(fv.begin(), fv.end());
      const double* a = P + 3 * (size_t)fv[0];
      for (size_t i = 1; i + 1 < fv.size(); ++i) {
        const double* b = P + 3 * (size_t)fv[i];
        const double* d = P + 3 * (size_t)fv[i + 1];
        double ra[3] = {a[0] - c0[0], a[1] - c0[1], a[2] - c0[2]};
        double rb[3] = {b[0] - c0[0], b[1] - c0[1], b[2] - c0[2]};
        double rd[3] = {d[0] - c0[0], d[1] - c0[1], d[2] - c0[2]};
        double cr[3] = {rb[1] * rd[2] - rb[2] * rd[1], rb[2] * rd[0] - rb[0] * rd[2],
                        rb[0] * rd[1] - rb[1] * rd[0]};
        double vTet = (ra[0] * cr[0] + ra[1] * cr[1] + ra[2] * cr[2]) / 6.0;
        double cT[3];
        for (int q = 0; q < 3; ++q) cT[q] = 0.25 * ((c0[q] + a[q]) + (b[q] + d[q]));
        V += vTet;
        M[0] = M[0] + vTet * cT[0];
        M[1] = M[1] + vTet * cT[1];
        M[2] = M[2] + vTet * cT[2];
      }
    }
    if (V <= 0.0) throw std::runtime_error("Non-positive cell volume at cell " + std::to_string(c));
    m.vol[c] = V;
    double iv = 1.0 / V;
    m.cc[3 * (size_t)c] = iv * M[0];
    m.cc[3 * (size_t)c + 1] = iv * M[1];
    m.cc[3 * (size_t)c + 2] = iv * M[2];
  }
  m.xf.assign(3 * (size_t)nF, 0.0);
  m.nf.assign(3 * (size_t)nF, 0.0);
  m.Sf.assign(3 * (size_t)nF, 0.0);
  m.Af.assign(nF, 0.0);
  for (int f = 0; f < nF; ++f) {
    const int* fvv = m.faceVerts.data() + m.faceOff[f];
    int k = m.faceOff[f + 1] - m.faceOff[f];
    double x[3] = {0, 0, 0};
    for (int i = 0; i < k; ++i) {
      const double* p = P + 3 * (size_t)fvv[i];
      x[0] = x[0] + p[0];
      x[1] = x[1] + p[1];
      x[2] = x[2] + p[2];
    }
    double s = 1.0 / (double)std::max(k, 1);
    double* xo = &m.xf[3 * (size_t)f];
    xo[0] = s * x[0];
    xo[1] = s * x[1];
    xo[2] = s * x[2];
    const double* a = P + 3 * (size_t)fvv[0];
    double av[3] = {0, 0, 0};
    for (int i = 1; i + 1 < k; ++i) {
      const double* b = P + 3 * (size_t)fvv[i];
      const double* d = P + 3 * (size_t)fvv[i + 1];
      double u[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
      double w[3] = {d[0] - a[0], d[1] - a[1], d[2] - a[2]};
      double cr[3] = {u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2],
                      u[0] * w[1] - u[1] * w[0]};
      av[0] = av[0] + 0.5 * cr[0];
      av[1] = av[1] + 0.5 * cr[1];
      av[2] = av[2] + 0.5 * cr[2];
    }
    double am = std::sqrt(av[0] * av[0] + av[1] * av[1] + av[2] * av[2]);
    if (am <= 1e-30) throw std::runtime_error("Degenerate face area at face " + std::to_string(f));
    double ia = 1.0 / am;
    double n[3] = {ia * av[0], ia * av[1], ia * av[2]};
    int Pc = m.owner[f];
    double dt[3];
    if (f < nIF) {
      int Nc = m.neigh[f];
      for (int q = 0; q < 3; ++q) dt[q] = m.cc[3 * (size_t)Nc + q] - m.cc[3 * (size_t)Pc + q];
    } else {
      for (int q = 0; q < 3; ++q) dt[q] = xo[q] - m.cc[3 * (size_t)Pc + q];
    }
    if (n[0] * dt[0] + n[1] * dt[1] + n[2] * dt[2] < 0.0) {
      n[0] = -1.0 * n[0];
      n[1] = -1.0 * n[1];
      n[2] = -1.0 * n[2];
    }
    m.Af[f] = am;
    for (int q = 0; q < 3; ++q) {
      m.nf[3 * (size_t)f + q] = n[q];
      m.Sf[3 * (size_t)f + q] = am * n[q];
    }
  }
  double mx = 0.0;
  const double kPi = 3.141592653589793238462643383279502884;
  for (int f = 0; f < nIF; ++f) {
    int p = m.owner[f], q = m.neigh[f];
    double d[3] = {m.cc[3 * (size_t)q] - m.cc[3 * (size_t)p],
                   m.cc[3 * (size_t)q + 1] - m.cc[3 * (size_t)p + 1],
                   m.cc[3 * (size_t)q + 2] - m.cc[3 * (size_t)p + 2]};
    double dn = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    double ca = std::fabs(d[0] * m.nf[3 * (size_t)f] + d[1] * m.nf[3 * (size_t)f + 1] +
                          d[2] * m.nf[3 * (size_t)f + 2]) /
                std::max(dn, 1e-30);
    ca = std::min(1.0, std::max(0.0, ca));
    mx = std::max(mx, std::acos(ca) * 180.0 / kPi);
  }
  m.maxNonOrthDeg = mx;
  m.hasGeometry = true;
}

// ---------------------------------------------------------- generators --
// Uniform double in [0,1) from the top 53 bits of a 64-bit Mersenne Twister
// draw (std::mt19937_64 is fully specified by the standard).
static inline double u01(std::mt19937_64& g) {
  return (double)(g() >> 11) * (1.0 / 9007199254740992.0);
}

HostMesh gen_hexbox(int nx, int ny, int nz, double lx, double ly, double lz,
                    double jitter, uint64_t seed) {
  if (nx < 1 || ny < 1 || nz < 1) throw std::runtime_error("hexbox: counts must be >= 1");
  HostMesh m;
  const int px = nx + 1, py = ny + 1, pz = nz + 1;
  m.nPoints = px * py * pz;
  m.P.resize(3 * (size_t)m.nPoints);
  const double hx = lx / nx, hy = ly / ny, hz = lz / nz;
  std::mt19937_64 g(seed);
  for (int k = 0; k < pz; ++k)
    for (int j = 0; j < py; ++j)
      for (int i = 0; i < px; ++i) {
        size_t v = (size_t)i + (size_t)px * ((size_t)j + (size_t)py * k);
        double x = hx * i, y = hy * j, z = hz * k;
        if (jitter > 0.0) {
          double rx = (2.0 * u01(g) - 1.0) * jitter;
          double ry = (2.0 * u01(g) - 1.0) * jitter;
          double rz = (2.0 * u01(g) - 1.0) * jitter;
          if (i > 0 && i < nx) x += rx * hx;
          if (j > 0 && j < ny) y += ry * hy;
          if (k > 0 && k < nz) z += rz * hz;
        }
        m.P[3 * v] = x;
        m.P[3 * v + 1] = y;
        m.P[3 * v + 2] = z;
      }
  auto V = [&](int i, int j, int k) { return i + px * (j + py * k); };
  auto C = [&](int i, int j, int k) { return i + nx * (j + ny * k); };
  const long long nIF = (long long)(nx - 1) * ny * nz + (long long)nx * (ny - 1) * nz +
                        (long long)nx * ny * (nz - 1);
  const long long nBF = 2LL * (ny * nz + nx * nz + nx * ny);
  if (nIF + nBF > 2000000000LL) throw std::runtime_error("hexbox too large");
  m.nFaces = (int)(nIF + nBF);
  m.faceOff.resize(m.nFaces + 1);
  m.faceVerts.resize(4 * (size_t)m.nFaces);
  m.owner.resize(m.nFaces);
  m.neigh.resize(nIF);
  int f = 0;
  auto put = [&](int a, int b, int c, int d, int own) {
    m.faceOff[f] = 4 * f;
    m.faceVerts[4 * (size_t)f] = a;
    m.faceVerts[4 * (size_t)f + 1] = b;
    m.faceVerts[4 * (size_t)f + 2] = c;
    m.faceVerts[4 * (size_t)f + 3] = d;
    m.owner[f] = own;
  };
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        int c = C(i, j, k);
        if (i + 1 < nx) {
          put(V(i + 1, j, k), V(i + 1, j + 1, k), V(i + 1, j + 1, k + 1), V(i + 1, j, k + 1), c);
          m.neigh[f++] = c + 1;
        }
        if (j + 1 < ny) {
          put(V(i, j + 1, k), V(i, j + 1, k + 1), V(i + 1, j + 1, k + 1), V(i + 1, j + 1, k), c);
          m.neigh[f++] = c + nx;
        }
        if (k + 1 < nz) {
          put(V(i, j, k + 1), V(i + 1, j, k + 1), V(i + 1, j + 1, k + 1), V(i, j + 1, k + 1), c);
          m.neigh[f++] = c + nx * ny;
        }
      }
  const char* names[6] = {"xmin", "xmax", "ymin", "ymax", "zmin", "zmax"};
  for (int pch = 0; pch < 6; ++pch) {
    m.patchNames.push_back(names[pch]);
    m.patchTypes.push_back("patch");
    m.patchStart.push_back(f);
    if (pch == 0 || pch == 1) {
      int i = pch == 0 ? 0 : nx;
      for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j) {
          if (pch == 0)
            put(V(0, j, k), V(0, j, k + 1), V(0, j + 1, k + 1), V(0, j + 1, k), C(0, j, k));
          else
            put(V(i, j, k), V(i, j + 1, k), V(i, j + 1, k + 1), V(i, j, k + 1), C(nx - 1, j, k));
          ++f;
        }
    } else if (pch == 2 || pch == 3) {
      int j = pch == 2 ? 0 : ny;
      for (int k = 0; k < nz; ++k)
        for (int i = 0; i < nx; ++i) {
          if (pch == 2)
            put(V(i, 0, k), V(i + 1, 0, k), V(i + 1, 0, k + 1), V(i, 0, k + 1), C(i, 0, k));
          else
            put(V(i, j, k), V(i, j, k + 1), V(i + 1, j, k + 1), V(i + 1, j, k), C(i, ny - 1, k));
          ++f;
        }
    } else {
      int k = pch == 4 ? 0 : nz;
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
          if (pch == 4)
            put(V(i, j, 0), V(i, j + 1, 0), V(i + 1, j + 1, 0), V(i + 1, j, 0), C(i, j, 0));
          else
            put(V(i, j, k), V(i + 1, j, k), V(i + 1, j + 1, k), V(i, j + 1, k), C(i, j, nz - 1));
          ++f;
        }
    }
    m.patchNFaces.push_back(f - m.patchStart.back());
  }
  m.faceOff[m.nFaces] = 4 * m.nFaces;
  m.hasPoints = true;
  finish_topology(m);
  compute_geometry(m);
  return m;
}

// Generic cells-to-polyMesh assembly: every cell lists its faces oriented
// outward; identical vertex sets pair up into internal faces owned by the
// lower cell id; internal faces are ordered by (owner, neighbour) and
// boundary faces by (patch, owner, emission order).
struct CellFace {
  int cell;
  int patch;  // patch for boundary use (classified by the caller), -1 unknown
  int nv;
  int v[4];
};

static HostMesh assemble_cells(int nPoints, std::vector<double>&& P, int nCells,
                               std::vector<CellFace>& faces,
                               const std::vector<std::string>& patchNames) {
  const size_t nE = faces.size();
  std::vector<std::pair<uint64_t, uint64_t>> key(nE);
  for (size_t e = 0; e < nE; ++e) {
    int s[4] = {-1, -1, -1, -1};
    for (int i = 0; i < faces[e].nv; ++i) s[i] = faces[e].v[i];
    std::sort(s, s + faces[e].nv);
    uint64_t a = ((uint64_t)(uint32_t)s[0] << 32) | (uint32_t)s[1];
    uint64_t b = ((uint64_t)(uint32_t)s[2] << 32) | (uint32_t)(faces[e].nv > 3 ? s[3] : -1);
    key[e] = {a, b};
  }
  std::vector<uint32_t> idx(nE);
  for (size_t e = 0; e < nE; ++e) idx[e] = (uint32_t)e;
  std::sort(idx.begin(), idx.end(), [&](uint32_t x, uint32_t y) {
    if (key[x] != key[y]) return key[x] < key[y];
    return x < y;
  });
  struct IF {
    int own, nei;
    uint32_t e;
  };
  std::vector<IF> internal;
  std::vector<std::pair<uint64_t, uint32_t>> bnd;  // (patch<<32|owner, e)
  internal.reserve(nE / 2);
  for (size_t t = 0; t < nE;) {
    size_t u = t + 1;
    while (u < nE && key[idx[u]] == key[idx[t]]) ++u;
    if (u - t == 2) {
      uint32_t e1 = idx[t], e2 = idx[t + 1];
      int c1 = faces[e1].cell, c2 = faces[e2].cell;
      if (c1 == c2) throw std::runtime_error("cell face matched to itself");
      IF f;
      if (c1 < c2) {
        f.own = c1;
        f.nei = c2;
        f.e = e1;
      } else {
        f.own = c2;
        f.nei = c1;
        f.e = e2;
      }
      internal.push_back(f);
    } else if (u - t == 1) {
      uint32_t e = idx[t];
      if (faces[e].patch < 0) throw std::runtime_error("unclassified boundary face");
      bnd.push_back({((uint64_t)faces[e].patch << 32) | (uint32_t)faces[e].cell, e});
    } else {
      throw std::runtime_error("face shared by more than two cells");
    }
    t = u;
  }
  std::sort(internal.begin(), internal.end(), [](const IF& a, const IF& b) {
    if (a.own != b.own) return a.own < b.own;
    if (a.nei != b.nei) return a.nei < b.nei;
    return a.e < b.e;
  });
  std::sort(bnd.begin(), bnd.end());
  HostMesh m;
  m.nPoints = nPoints;
  m.P = std::move(P);
  m.nFaces = (int)(internal.size() + bnd.size());
  m.faceOff.assign(1, 0);
  m.owner.reserve(m.nFaces);
  m.neigh.reserve(internal.size());
  auto emit = [&](uint32_t e) {
    for (int i = 0; i < faces[e].nv; ++i) m.faceVerts.push_back(faces[e].v[i]);
    m.faceOff.push_back((int)m.faceVerts.size());
  };
  for (const IF& f : internal) {
    emit(f.e);
    m.owner.push_back(f.own);
    m.neigh.push_back(f.nei);
  }
  size_t b = 0;
  for (size_t p = 0; p < patchNames.size(); ++p) {
    m.patchNames.push_back(patchNames[p]);
    m.patchTypes.push_back("patch");
    m.patchStart.push_back((int)m.owner.size());
    while (b < bnd.size() && (bnd[b].first >> 32) == p) {
      emit(bnd[b].second);
      m.owner.push_back(faces[bnd[b].second].cell);
      ++b;
    }
    m.patchNFaces.push_back((int)m.owner.size() - m.patchStart.back());
  }
  if (b != bnd.size()) throw std::runtime_error("boundary face with unknown patch id");
  m.hasPoints = true;
  finish_topology(m);
  if (m.nCells != nCells) throw std::runtime_error("cell count mismatch after assembly");
  compute_geometry(m);
  return m;
}

HostMesh gen_tetcube(int n, double L) {
  if (n < 1) throw std::runtime_error("tetcube: nblocks must be >= 1");
  const int p = n + 1;
  std::vector<double> P(3 * (size_t)p * p * p);
  const double h = L / n;
  for (int k = 0; k < p; ++k)
    for (int j = 0; j < p; ++j)
      for (int i = 0; i < p; ++i) {
        size_t v = (size_t)i + (size_t)p * ((size_t)j + (size_t)p * k);
        P[3 * v] = h * i;
        P[3 * v + 1] = h * j;
        P[3 * v + 2] = h * k;
      }
  auto V = [&](int i, int j, int k) { return i + p * (j + p * k); };
  // Kuhn subdivision: one tet per permutation of the axes along the main
  // diagonal; conforming across blocks.
  const int perm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  std::vector<CellFace> faces;
  faces.reserve(24 * (size_t)n * n * n);
  int cell = 0;
  for (int k = 0; k < n; ++k)
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i)
        for (int t = 0; t < 6; ++t) {
          int q[3] = {i, j, k};
          int tv[4];
          tv[0] = V(q[0], q[1], q[2]);
          for (int s = 0; s < 3; ++s) {
            q[perm[t][s]] += 1;
            tv[s + 1] = V(q[0], q[1], q[2]);
          }
          double ct[3] = {0, 0, 0};
          for (int a = 0; a < 4; ++a)
            for (int d = 0; d < 3; ++d) ct[d] += 0.25 * P[3 * (size_t)tv[a] + d];
          const int fl[4][3] = {{1, 2, 3}, {0, 3, 2}, {0, 1, 3}, {0, 2, 1}};
          for (int fi = 0; fi < 4; ++fi) {
            CellFace cf;
            cf.cell = cell;
            cf.nv = 3;
            for (int a = 0; a < 3; ++a) cf.v[a] = tv[fl[fi][a]];
            cf.v[3] = -1;
            const double* a0 = &P[3 * (size_t)cf.v[0]];
            const double* a1 = &P[3 * (size_t)cf.v[1]];
            const double* a2 = &P[3 * (size_t)cf.v[2]];
            double u[3] = {a1[0] - a0[0], a1[1] - a0[1], a1[2] - a0[2]};
            double w[3] = {a2[0] - a0[0], a2[1] - a0[1], a2[2] - a0[2]};
            double nn[3] = {u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2],
                            u[0] * w[1] - u[1] * w[0]};
            double fc[3], dd = 0.0;
            for (int d = 0; d < 3; ++d) {
              fc[d] = (a0[d] + a1[d] + a2[d]) / 3.0;
              dd += nn[d] * (fc[d] - ct[d]);
            }
            if (dd < 0.0) std::swap(cf.v[1], cf.v[2]);
            // boundary classification by the face centroid
            const double eps = 1e-9 * L;
            cf.patch = -1;
            if (std::fabs(fc[0]) < eps && std::fabs(u[0]) < eps && std::fabs(w[0]) < eps) cf.patch = 0;
            else if (std::fabs(fc[0] - L) < eps && std::fabs(u[0]) < eps && std::fabs(w[0]) < eps) cf.patch = 1;
            else if (std::fabs(fc[1]) < eps && std::fabs(u[1]) < eps && std::fabs(w[1]) < eps) cf.patch = 2;
            else if (std::fabs(fc[1] - L) < eps && std::fabs(u[1]) < eps && std::fabs(w[1]) < eps) cf.patch = 3;
            else if (std::fabs(fc[2]) < eps && std::fabs(u[2]) < eps && std::fabs(w[2]) < eps) cf.patch = 4;
            else if (std::fabs(fc[2] - L) < eps && std::fabs(u[2]) < eps && std::fabs(w[2]) < eps) cf.patch = 5;
            faces.push_back(cf);
          }
          ++cell;
        }
  return assemble_cells(p * p * p, std::move(P), cell, faces,
                        {"xmin", "xmax", "ymin", "ymax", "zmin", "zmax"});
}

// --------------------------------------------------- cylinder channel ---
CylinderParams cylinder_params_for_cells(long long target) {
  CylinderParams best;
  long long bestErr = -1;
  for (int ns = 4; ns <= 200; ++ns) {
    CylinderParams p;
    double h = 2.0 * p.s / ns;
    p.ns = ns;
    p.nr = std::max(2, ns / 2);
    p.nx1 = std::max(2, (int)std::lround((p.xc - p.s) / h));
    p.nx3 = std::max(2, (int)std::lround((p.L - p.xc - p.s) / h));
    p.ny1 = std::max(2, (int)std::lround((p.yc - p.s) / h));
    p.ny3 = std::max(2, (int)std::lround((p.H - p.yc - p.s) / h));
    p.nz = std::max(2, (int)std::lround(p.W / h));
    long long n2 = (long long)(p.nx1 + ns + p.nx3) * (p.ny1 + ns + p.ny3) - (long long)ns * ns +
                   4LL * ns * p.nr;
    long long nc = n2 * p.nz;
    long long err = nc > target ? nc - target : target - nc;
    if (bestErr < 0 || err < bestErr) {
      bestErr = err;
      best = p;
    }
  }
  return best;
}

HostMesh gen_cylinder_channel(const CylinderParams& p) {
  if (p.ns < 2 || p.nr < 1 || p.nx1 < 1 || p.nx3 < 1 || p.ny1 < 1 || p.ny3 < 1 || p.nz < 1)
    throw std::runtime_error("cylinder: invalid block counts");
  if (!(p.xc - p.s > 0 && p.xc + p.s < p.L && p.yc - p.s > 0 && p.yc + p.s < p.H && p.R < p.s))
    throw std::runtime_error("cylinder: inconsistent geometry");
  const double kPi = 3.141592653589793238462643383279502884;
  // tensor grid lines
  std::vector<double> X, Y;
  auto uni = [](std::vector<double>& v, double a, double b, int n, bool first) {
    for (int i = first ? 0 : 1; i <= n; ++i) v.push_back(a + (b - a) * (double)i / n);
  };
  auto tanline = [&](std::vector<double>& v, double c) {
    for (int i = 1; i <= p.ns; ++i) {
      double th = -0.25 * kPi + 0.5 * kPi * (double)i / p.ns;
      v.push_back(i == p.ns ? c + p.s : c + p.s * std::tan(th));
    }
  };
  uni(X, 0.0, p.xc - p.s, p.nx1, true);
  tanline(X, p.xc);
  uni(X, p.xc + p.s, p.L, p.nx3, false);
  uni(Y, 0.0, p.yc - p.s, p.ny1, true);
  tanline(Y, p.yc);
  uni(Y, p.yc + p.s, p.H, p.ny3, false);
  const int nX = (int)X.size(), nY = (int)Y.size();  // node counts
  const int i0 = p.nx1, i1 = p.nx1 + p.ns, j0 = p.ny1, j1 = p.ny1 + p.ns;
  // 2D nodes: tensor nodes not strictly inside the square, then O-grid rings
  std::vector<int> tid((size_t)nX * nY, -1);
  std::vector<double> px, py;
  for (int j = 0; j < nY; ++j)
    for (int i = 0; i < nX; ++i) {
      bool inside = (i > i0 && i < i1 && j > j0 && j < j1);
      if (inside) continue;
      tid[(size_t)j * nX + i] = (int)px.size();
      px.push_back(X[i]);
      py.push_back(Y[j]);
    }
  const int nRing = 4 * p.ns;
  // square boundary walk, counterclockwise from the bottom-right corner
  std::vector<int> sq(nRing);
  for (int m = 0; m < nRing; ++m) {
    int side = m / p.ns, t = m % p.ns, i, j;
    if (side == 0) { i = i1; j = j0 + t; }
    else if (side == 1) { i = i1 - t; j = j1; }
    else if (side == 2) { i = i0; j = j1 - t; }
    else { i = i0 + t; j = j0; }
    sq[m] = tid[(size_t)j * nX + i];
  }
  std::vector<int> ring((size_t)nRing * (p.nr + 1));
  double q = p.radialRatio, denom = (std::fabs(q - 1.0) < 1e-12) ? (double)p.nr : (std::pow(q, p.nr) - 1.0);
  for (int r = 0; r <= p.nr; ++r) {
    double g = (std::fabs(q - 1.0) < 1e-12) ? (double)r / p.nr : (std::pow(q, r) - 1.0) / denom;
    for (int m = 0; m < nRing; ++m) {
      if (r == p.nr) {
        ring[(size_t)r * nRing + m] = sq[m];
        continue;
      }
      double th = -0.25 * kPi + 0.5 * kPi * (double)m / p.ns;
      double cx = p.xc + p.R * std::cos(th), cy = p.yc + p.R * std::sin(th);
      double sx = px[sq[m]], sy = py[sq[m]];
      ring[(size_t)r * nRing + m] = (int)px.size();
      px.push_back(cx + g * (sx - cx));
      py.push_back(cy + g * (sy - cy));
    }
  }
  const int n2p = (int)px.size();
  // 2D quads (counterclockwise)
  std::vector<int> quad;
  for (int j = 0; j + 1 < nY; ++j)
    for (int i = 0; i + 1 < nX; ++i) {
      if (i >= i0 && i < i1 && j >= j0 && j < j1) continue;
      quad.push_back(tid[(size_t)j * nX + i]);
      quad.push_back(tid[(size_t)j * nX + i + 1]);
      quad.push_back(tid[(size_t)(j + 1) * nX + i + 1]);
      quad.push_back(tid[(size_t)(j + 1) * nX + i]);
    }
  for (int r = 0; r < p.nr; ++r)
    for (int m = 0; m < nRing; ++m) {
      int m1 = (m + 1) % nRing;
      quad.push_back(ring[(size_t)r * nRing + m]);
      quad.push_back(ring[(size_t)(r + 1) * nRing + m]);
      quad.push_back(ring[(size_t)(r + 1) * nRing + m1]);
      quad.push_back(ring[(size_t)r * nRing + m1]);
    }
  const int n2 = (int)(quad.size() / 4);
  // 2D edges: key (min,max) -> (quad, local edge) ; internal if two quads
  struct E2 {
    uint64_t key;
    int q, a, b;  // quad, directed edge a->b in that quad's ccw order
  };
  std::vector<E2> edges;
  edges.reserve(4 * (size_t)n2);
  for (int c = 0; c < n2; ++c)
    for (int e = 0; e < 4; ++e) {
      int a = quad[4 * c + e], b = quad[4 * c + (e + 1) % 4];
      uint64_t k = ((uint64_t)(uint32_t)std::min(a, b) << 32) | (uint32_t)std::max(a, b);
      edges.push_back({k, c, a, b});
    }
  std::sort(edges.begin(), edges.end(), [](const E2& x, const E2& y) {
    if (x.key != y.key) return x.key < y.key;
    return x.q < y.q;
  });
  // 3D points: layer-major
  const int nz = p.nz;
  std::vector<double> P(3 * (size_t)n2p * (nz + 1));
  for (int k = 0; k <= nz; ++k) {
    double z = p.W * (double)k / nz;
    for (int v = 0; v < n2p; ++v) {
      size_t g = (size_t)k * n2p + v;
      P[3 * g] = px[v];
      P[3 * g + 1] = py[v];
      P[3 * g + 2] = z;
    }
  }
  auto PV = [&](int v, int k) { return k * n2p + v; };
  auto CELL = [&](int qd, int k) { return k * n2 + qd; };
  // patches: inlet, walls(y), outlet, z-walls, cylinder (fixture order)
  const std::vector<std::string> names = {"patch_1_0", "patch_0_0", "patch_2_0", "patch_5_0",
                                          "patch_3_0"};
  struct F3 {
    int own, nei, patch;
    int v[4];
  };
  std::vector<F3> intl, bnd;
  const double eps = 1e-9;
  for (size_t t = 0; t < edges.size();) {
    size_t u = t + 1;
    while (u < edges.size() && edges[u].key == edges[t].key) ++u;
    for (int k = 0; k < nz; ++k) {
      if (u - t == 2) {
        const E2& e1 = edges[t];
        const E2& e2 = edges[t + 1];
        const E2& eo = (e1.q < e2.q) ? e1 : e2;
        const E2& en = (e1.q < e2.q) ? e2 : e1;
        F3 f;
        f.own = CELL(eo.q, k);
        f.nei = CELL(en.q, k);
        f.patch = -1;
        f.v[0] = PV(eo.a, k);
        f.v[1] = PV(eo.b, k);
        f.v[2] = PV(eo.b, k + 1);
        f.v[3] = PV(eo.a, k + 1);
        intl.push_back(f);
      } else if (u - t == 1) {
        const E2& e = edges[t];
        double mx = 0.5 * (px[e.a] + px[e.b]), my = 0.5 * (py[e.a] + py[e.b]);
        int pch;
        if (std::fabs(mx) < eps) pch = 0;
        else if (std::fabs(mx - p.L) < eps) pch = 2;
        else if (std::fabs(my) < eps || std::fabs(my - p.H) < eps) pch = 1;
        else pch = 4;  // cylinder
        F3 f;
        f.own = CELL(e.q, k);
        f.nei = -1;
        f.patch = pch;
        f.v[0] = PV(e.a, k);
        f.v[1] = PV(e.b, k);
        f.v[2] = PV(e.b, k + 1);
        f.v[3] = PV(e.a, k + 1);
        bnd.push_back(f);
      } else {
        throw std::runtime_error("cylinder: edge shared by more than two quads");
      }
    }
    t = u;
  }
  for (int c = 0; c < n2; ++c) {
    const int* qv = &quad[4 * c];
    for (int k = 0; k <= nz; ++k) {
      F3 f;
      if (k == 0 || k == nz) {
        f.own = CELL(c, k == 0 ? 0 : nz - 1);
        f.nei = -1;
        f.patch = 3;
        if (k == 0) {
          f.v[0] = PV(qv[0], 0); f.v[1] = PV(qv[3], 0); f.v[2] = PV(qv[2], 0); f.v[3] = PV(qv[1], 0);
        } else {
          f.v[0] = PV(qv[0], nz); f.v[1] = PV(qv[1], nz); f.v[2] = PV(qv[2], nz); f.v[3] = PV(qv[3], nz);
        }
        bnd.push_back(f);
      } else {
        f.own = CELL(c, k - 1);
        f.nei = CELL(c, k);
        f.patch = -1;
        f.v[0] = PV(qv[0], k); f.v[1] = PV(qv[1], k); f.v[2] = PV(qv[2], k); f.v[3] = PV(qv[3], k);
        intl.push_back(f);
      }
    }
  }
  std::sort(intl.begin(), intl.end(), [](const F3& a, const F3& b) {
    if (a.own != b.own) return a.own < b.own;
    return a.nei < b.nei;
  });
  std::stable_sort(bnd.begin(), bnd.end(), [](const F3& a, const F3& b) {
    if (a.patch != b.patch) return a.patch < b.patch;
    return a.own < b.own;
  });
  HostMesh m;
  m.nPoints = n2p * (nz + 1);
  m.P = std::move(P);
  m.nFaces = (int)(intl.size() + bnd.size());
  m.faceOff.resize(m.nFaces + 1);
  m.faceVerts.resize(4 * (size_t)m.nFaces);
  m.owner.resize(m.nFaces);
  m.neigh.resize(intl.size());
  int f = 0;
  for (const F3& x : intl) {
    m.faceOff[f] = 4 * f;
    for (int a = 0; a < 4; ++a) m.faceVerts[4 * (size_t)f + a] = x.v[a];
    m.owner[f] = x.own;
    m.neigh[f] = x.nei;
    ++f;
  }
  size_t b = 0;
  for (int pch = 0; pch < (int)names.size(); ++pch) {
    m.patchNames.push_back(names[pch]);
    m.patchTypes.push_back("patch");
    m.patchStart.push_back(f);
    while (b < bnd.size() && bnd[b].patch == pch) {
      m.faceOff[f] = 4 * f;
      for (int a = 0; a < 4; ++a) m.faceVerts[4 * (size_t)f + a] = bnd[b].v[a];
      m.owner[f] = bnd[b].own;
      ++f;
      ++b;
    }
    m.patchNFaces.push_back(f - m.patchStart.back());
  }
  m.faceOff[m.nFaces] = 4 * m.nFaces;
  m.hasPoints = true;
  finish_topology(m);
  compute_geometry(m);
  return m;
}

}  // namespace an
