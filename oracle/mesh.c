/*
 * mesh.c -- oracle restatement of polyMesh geometry and CSR addressing.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * or_mesh_geometry   follows LP/src/mesh.cu:188-285.
 *   Pinned by: closed forms on generated meshes (hex volume h^3, tet volume
 *   h^3/6, sum of volumes = box volume, sum of outward S over each closed
 *   cell = 0, centroids of undistorted hexes = box centres) in
 *   tests/test_oracle.py (test_geometry_*, test_product_*).
 * or_pattern_*       follows LP/src/scalar_elliptic.cu:16-58
 *   (== GSV1 build_momentum_pattern main.cu:1315-1340).
 *   Pinned by tests/test_oracle.py (test_pattern_*): the CRC of the CSR
 *   built from the reference's 44,100-cell cylinder owner/neighbour fixture
 *   (nnz = 617,000), an independent python set restatement, and structural
 *   invariants (sorted unique columns, diagonal present, PN/NP transposed,
 *   nnz = nC + 2*nIF for meshes without duplicate cell pairs).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static int cmp_int(const void *a, const void *b) {
  int x = *(const int *)a, y = *(const int *)b;
  return (x > y) - (x < y);
}

/* The unique vertex set of a cell, as std::set<int> iteration (ascending). */
static int unique_sorted(int *v, int n) {
  if (n == 0) return 0;
  qsort(v, (size_t)n, sizeof(int), cmp_int);
  int m = 1;
  for (int i = 1; i < n; ++i)
    if (v[i] != v[m - 1]) v[m++] = v[i];
  return m;
}

int or_mesh_geometry(int nPoints, const double *P, int nFaces,
                     const int *faceOff, const int *faceVerts,
                     const int *owner, int nInternalFaces, const int *neigh,
                     int nCells, double *cc, double *vol, double *xf,
                     double *nf, double *Sf, double *Af,
                     double *maxNonOrthDeg) {
  (void)nPoints;
  /* cellFaces / cellOrient: owner faces in face order, then neighbour faces
   * in face order (mesh.cu:188-199). */
  int *cnt = (int *)calloc((size_t)nCells + 1, sizeof(int));
  for (int f = 0; f < nFaces; ++f) cnt[owner[f] + 1]++;
  for (int f = 0; f < nInternalFaces; ++f) cnt[neigh[f] + 1]++;
  for (int c = 0; c < nCells; ++c) cnt[c + 1] += cnt[c];
  int *cf = (int *)malloc(sizeof(int) * (size_t)cnt[nCells]);
  int *co = (int *)malloc(sizeof(int) * (size_t)cnt[nCells]);
  int *fill = (int *)malloc(sizeof(int) * (size_t)nCells);
  for (int c = 0; c < nCells; ++c) fill[c] = cnt[c];
  for (int f = 0; f < nFaces; ++f) {
    int p = owner[f];
    cf[fill[p]] = f;
    co[fill[p]++] = +1;
  }
  for (int f = 0; f < nInternalFaces; ++f) {
    int n = neigh[f];
    cf[fill[n]] = f;
    co[fill[n]++] = -1;
  }

  int maxFaceVerts = 0;
  for (int f = 0; f < nFaces; ++f) {
    int k = faceOff[f + 1] - faceOff[f];
    if (k > maxFaceVerts) maxFaceVerts = k;
  }
  int vcap = 64;
  int *verts = (int *)malloc(sizeof(int) * (size_t)vcap);
  int *fv = (int *)malloc(sizeof(int) * (size_t)(maxFaceVerts + 1));
  int rc = 0;

  /* cell centroid and volume (mesh.cu:201-230) */
  for (int c = 0; c < nCells && rc == 0; ++c) {
    int nv = 0;
    for (int j = cnt[c]; j < cnt[c + 1]; ++j) {
      int f = cf[j];
      int k = faceOff[f + 1] - faceOff[f];
      if (nv + k > vcap) {
        while (nv + k > vcap) vcap *= 2;
        verts = (int *)realloc(verts, sizeof(int) * (size_t)vcap);
      }
      memcpy(verts + nv, faceVerts + faceOff[f], sizeof(int) * (size_t)k);
      nv += k;
    }
    nv = unique_sorted(verts, nv);
    double c0[3] = {0, 0, 0};
    for (int i = 0; i < nv; ++i) {
      const double *p = P + 3 * (size_t)verts[i];
      c0[0] = c0[0] + p[0];
      c0[1] = c0[1] + p[1];
      c0[2] = c0[2] + p[2];
    }
    double s = 1.0 / (double)(nv > 1 ? nv : 1);
    c0[0] = s * c0[0];
    c0[1] = s * c0[1];
    c0[2] = s * c0[2];

    double V = 0.0, M[3] = {0, 0, 0};
    for (int j = cnt[c]; j < cnt[c + 1]; ++j) {
      int f = cf[j];
      int k = faceOff[f + 1] - faceOff[f];
      for (int i = 0; i < k; ++i) fv[i] = faceVerts[faceOff[f] + i];
      if (co[j] < 0) {
        for (int i = 0; i < k / 2; ++i) {
          int t = fv[i];
          fv[i] = fv[k - 1 - i];
          fv[k - 1 - i] = t;
        }
      }
      const double *a = P + 3 * (size_t)fv[0];
      for (int i = 1; i + 1 < k; ++i) {
        const double *b = P + 3 * (size_t)fv[i];
        const double *d = P + 3 * (size_t)fv[i + 1];
        double ra[3] = {a[0] - c0[0], a[1] - c0[1], a[2] - c0[2]};
        double rb[3] = {b[0] - c0[0], b[1] - c0[1], b[2] - c0[2]};
        double rd[3] = {d[0] - c0[0], d[1] - c0[1], d[2] - c0[2]};
        double cr[3] = {rb[1] * rd[2] - rb[2] * rd[1],
                        rb[2] * rd[0] - rb[0] * rd[2],
                        rb[0] * rd[1] - rb[1] * rd[0]};
        double vTet = (ra[0] * cr[0] + ra[1] * cr[1] + ra[2] * cr[2]) / 6.0;
        /* cTet = 0.25*((c0+a)+(b+d)) */
        double cT[3];
        for (int q = 0; q < 3; ++q)
          cT[q] = 0.25 * ((c0[q] + a[q]) + (b[q] + d[q]));
        V += vTet;
        M[0] = M[0] + vTet * cT[0];
        M[1] = M[1] + vTet * cT[1];
        M[2] = M[2] + vTet * cT[2];
      }
    }
    if (V <= 0.0) {
      rc = -(1 + c);
      break;
    }
    vol[c] = V;
    double iv = 1.0 / V;
    cc[3 * c + 0] = iv * M[0];
    cc[3 * c + 1] = iv * M[1];
    cc[3 * c + 2] = iv * M[2];
  }

  /* face centre, area vector, orientation (mesh.cu:232-260) */
  for (int f = 0; f < nFaces && rc == 0; ++f) {
    int k = faceOff[f + 1] - faceOff[f];
    const int *fvv = faceVerts + faceOff[f];
    double x[3] = {0, 0, 0};
    for (int i = 0; i < k; ++i) {
      const double *p = P + 3 * (size_t)fvv[i];
      x[0] = x[0] + p[0];
      x[1] = x[1] + p[1];
      x[2] = x[2] + p[2];
    }
    double s = 1.0 / (double)(k > 1 ? k : 1);
    xf[3 * f + 0] = s * x[0];
    xf[3 * f + 1] = s * x[1];
    xf[3 * f + 2] = s * x[2];

    const double *a = P + 3 * (size_t)fvv[0];
    double av[3] = {0, 0, 0};
    for (int i = 1; i + 1 < k; ++i) {
      const double *b = P + 3 * (size_t)fvv[i];
      const double *d = P + 3 * (size_t)fvv[i + 1];
      double u[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
      double w[3] = {d[0] - a[0], d[1] - a[1], d[2] - a[2]};
      double cr[3] = {u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2],
                      u[0] * w[1] - u[1] * w[0]};
      av[0] = av[0] + 0.5 * cr[0];
      av[1] = av[1] + 0.5 * cr[1];
      av[2] = av[2] + 0.5 * cr[2];
    }
    double am = sqrt(av[0] * av[0] + av[1] * av[1] + av[2] * av[2]);
    if (am <= 1e-30) {
      rc = -(1 + nCells + f);
      break;
    }
    double ia = 1.0 / am;
    double n[3] = {ia * av[0], ia * av[1], ia * av[2]};
    int Pc = owner[f];
    double dt[3];
    if (f < nInternalFaces) {
      int Nc = neigh[f];
      for (int q = 0; q < 3; ++q) dt[q] = cc[3 * Nc + q] - cc[3 * Pc + q];
    } else {
      for (int q = 0; q < 3; ++q) dt[q] = xf[3 * f + q] - cc[3 * Pc + q];
    }
    if (n[0] * dt[0] + n[1] * dt[1] + n[2] * dt[2] < 0.0) {
      n[0] = -1.0 * n[0];
      n[1] = -1.0 * n[1];
      n[2] = -1.0 * n[2];
    }
    Af[f] = am;
    for (int q = 0; q < 3; ++q) {
      nf[3 * f + q] = n[q];
      Sf[3 * f + q] = am * n[q];
    }
  }

  /* maxNonOrthDeg (mesh.cu:279-285) */
  if (rc == 0 && maxNonOrthDeg) {
    double mx = 0.0;
    const double kPi = 3.141592653589793238462643383279502884;
    for (int f = 0; f < nInternalFaces; ++f) {
      int p = owner[f], q = neigh[f];
      double d[3] = {cc[3 * q] - cc[3 * p], cc[3 * q + 1] - cc[3 * p + 1],
                     cc[3 * q + 2] - cc[3 * p + 2]};
      double dn = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      double ca = fabs(d[0] * nf[3 * f] + d[1] * nf[3 * f + 1] +
                       d[2] * nf[3 * f + 2]) /
                  (dn > 1e-30 ? dn : 1e-30);
      ca = ca < 1.0 ? ca : 1.0;
      ca = ca > 0.0 ? ca : 0.0;
      double deg = acos(ca) * 180.0 / kPi;
      if (deg > mx) mx = deg;
    }
    *maxNonOrthDeg = mx;
  }

  free(cnt);
  free(cf);
  free(co);
  free(fill);
  free(verts);
  free(fv);
  return rc;
}

/* ------------------------------------------------------------ pattern */
/* Row c holds sorted unique {c} U cellNbrs[c] (scalar_elliptic.cu:30-39). */
typedef struct {
  int *off;  /* nCells+1 */
  int *nb;   /* neighbour lists with duplicates, sorted+unique per row */
  int *len;  /* unique length per row */
} nbr_lists;

static void build_nbr_lists(int nCells, int nInternalFaces, const int *owner,
                            const int *neigh, nbr_lists *L) {
  L->off = (int *)calloc((size_t)nCells + 1, sizeof(int));
  L->len = (int *)calloc((size_t)nCells, sizeof(int));
  for (int f = 0; f < nInternalFaces; ++f) {
    L->off[owner[f] + 1]++;
    L->off[neigh[f] + 1]++;
  }
  for (int c = 0; c < nCells; ++c) L->off[c + 1] += L->off[c];
  L->nb = (int *)malloc(sizeof(int) * ((size_t)L->off[nCells] + 1));
  int *fill = (int *)malloc(sizeof(int) * (size_t)nCells);
  for (int c = 0; c < nCells; ++c) fill[c] = L->off[c];
  for (int f = 0; f < nInternalFaces; ++f) {
    L->nb[fill[owner[f]]++] = neigh[f];
    L->nb[fill[neigh[f]]++] = owner[f];
  }
  free(fill);
  for (int c = 0; c < nCells; ++c) {
    int *v = L->nb + L->off[c];
    int n = L->off[c + 1] - L->off[c];
    L->len[c] = unique_sorted(v, n);
  }
}

static void free_nbr_lists(nbr_lists *L) {
  free(L->off);
  free(L->nb);
  free(L->len);
}

/* number of entries of the sorted unique row {c} U nbrs */
static int row_len(const nbr_lists *L, int c) {
  const int *v = L->nb + L->off[c];
  int n = L->len[c], has = 0;
  for (int i = 0; i < n; ++i)
    if (v[i] == c) has = 1;
  return n + (has ? 0 : 1);
}

int64_t or_pattern_nnz(int nCells, int nFaces, int nInternalFaces,
                       const int *owner, const int *neigh) {
  (void)nFaces;
  nbr_lists L;
  build_nbr_lists(nCells, nInternalFaces, owner, neigh, &L);
  int64_t nnz = 0;
  for (int c = 0; c < nCells; ++c) nnz += row_len(&L, c);
  free_nbr_lists(&L);
  return nnz;
}

static int find_pos(const int *cols, int beg, int end, int key) {
  int lo = beg, hi = end - 1;
  while (lo <= hi) {
    int mid = (lo + hi) / 2;
    if (cols[mid] == key) return mid;
    if (cols[mid] < key)
      lo = mid + 1;
    else
      hi = mid - 1;
  }
  return -1;
}

int or_pattern_fill(int nCells, int nFaces, int nInternalFaces,
                    const int *owner, const int *neigh, int *rowOffsets,
                    int *cols, int *diagPos, int *facePP, int *facePN,
                    int *faceNP, int *faceNN) {
  (void)nFaces;
  nbr_lists L;
  build_nbr_lists(nCells, nInternalFaces, owner, neigh, &L);
  rowOffsets[0] = 0;
  for (int c = 0; c < nCells; ++c) {
    const int *v = L.nb + L.off[c];
    int n = L.len[c];
    int o = rowOffsets[c];
    int k = 0, placed = 0;
    for (int i = 0; i < n; ++i) {
      if (!placed && c < v[i]) {
        cols[o + k] = c;
        diagPos[c] = o + k;
        ++k;
        placed = 1;
      }
      if (v[i] == c) {
        if (!placed) {
          cols[o + k] = c;
          diagPos[c] = o + k;
          ++k;
          placed = 1;
        }
        continue;
      }
      cols[o + k++] = v[i];
    }
    if (!placed) {
      cols[o + k] = c;
      diagPos[c] = o + k;
      ++k;
    }
    rowOffsets[c + 1] = o + k;
  }
  int bad = 0;
  for (int f = 0; f < nInternalFaces; ++f) {
    int P = owner[f], N = neigh[f];
    facePP[f] = diagPos[P];
    facePN[f] = find_pos(cols, rowOffsets[P], rowOffsets[P + 1], N);
    faceNP[f] = find_pos(cols, rowOffsets[N], rowOffsets[N + 1], P);
    faceNN[f] = diagPos[N];
    if (facePN[f] < 0 || faceNP[f] < 0) bad = 1;
  }
  free_nbr_lists(&L);
  return bad ? -1 : 0;
}
