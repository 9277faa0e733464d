// Hierarchical Nedelec (first kind) H(curl) basis on the unit tetrahedron,
// uniform order p <= 3, usable on host and device.
//
// Follows petgem/hvfem.py:319-1052 (shape3DETet, AncEE, AncETri, HomLegendre,
// HomIJacobi, PolyLegendre, PolyIJacobi, PolyJacobi, OrientE, OrientTri,
// ProjectTetE, ProjectTetF, AffineTetrahedron), i.e. Fuentes, Keith,
// Demkowicz & Nagaraj (2015).  DOF order inside an element: 6 edges x p,
// 4 faces x p(p-1) (two families interleaved), p(p-1)(p-2)/2 bubbles (three
// families interleaved).  Edge functions depend only on their edge's
// orientation flag and face functions only on their face's orientation code,
// which is what lets the element kernels work from per-orientation tables.
//
// Also holds the symmetric tetrahedral quadrature rules of orders 2/4/6
// (petgem/hvfem.py:1055-1127 with the mapping of hvfem.py:1603-1608).
#pragma once

#ifndef __HIP_DEVICE_COMPILE__
#include <cstddef>
#endif
#include <hip/hip_runtime.h>

#define PG_HD __host__ __device__ inline

namespace pg {

constexpr int kMaxP = 3;
constexpr int kMaxDofs = 45;  // p = 3

PG_HD int ndof_element(int p) { return p * (p + 2) * (p + 3) / 2; }
PG_HD int ndof_edge(int p) { return p; }
PG_HD int ndof_face(int p) { return p * (p - 1); }
PG_HD int ndof_vol(int p) { return p * (p - 1) * (p - 2) / 2; }

struct V3 {
  double x, y, z;
};
PG_HD V3 v3(double a, double b, double c) { return V3{a, b, c}; }
PG_HD V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
PG_HD V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
PG_HD V3 operator-(V3 a) { return V3{-a.x, -a.y, -a.z}; }
PG_HD V3 operator*(double s, V3 a) { return V3{s * a.x, s * a.y, s * a.z}; }
PG_HD V3 cross(V3 a, V3 b) {
  return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// local entities (mesh.py:27-32, 70-73; hvfem.py:149-161, 166-175)
PG_HD int edge_vert(int e, int k) {
  // (0,1),(1,2),(0,2),(0,3),(1,3),(2,3)
  const int t[12] = {0, 1, 1, 2, 0, 2, 0, 3, 1, 3, 2, 3};
  return t[2 * e + k];
}
PG_HD int face_vert(int f, int k) {
  // (0,1,2),(0,1,3),(1,2,3),(0,2,3)
  const int t[12] = {0, 1, 2, 0, 1, 3, 1, 2, 3, 0, 2, 3};
  return t[3 * f + k];
}
PG_HD int face_edge(int f, int k) {
  // local edges of each face: (0,1,2),(0,4,3),(1,5,4),(2,5,3)
  const int t[12] = {0, 1, 2, 0, 4, 3, 1, 5, 4, 2, 5, 3};
  return t[3 * f + k];
}
PG_HD int face_perm(int o, int k) {
  // OrientTri permutations (hvfem.py:849-872)
  const int t[18] = {0, 1, 2, 1, 2, 0, 2, 0, 1, 0, 2, 1, 1, 0, 2, 2, 1, 0};
  return t[3 * o + k];
}

// ---------------------------------------------------------------------------
// polynomial families
// ---------------------------------------------------------------------------
// Shifted scaled Legendre P_i(x;t), i < nord (PolyLegendre).
PG_HD void legendre(double x, double t, int nord, double* P) {
  P[0] = 1.0;
  if (nord >= 2) {
    const double y = 2.0 * x - t;
    const double tt = t * t;
    P[1] = y;
    for (int i = 1; i < nord - 1; ++i) {
      const double v = (2.0 * i + 1.0) * y * P[i] - (double)i * tt * P[i - 1];
      P[i + 1] = v / (double)(i + 1);
    }
  }
}

// Shifted scaled Jacobi P^{alpha_i}_j(x;t), alpha_i = minalpha + 2i (PolyJacobi).
// P is (nord+1) x (nord+1), row-major with stride kMaxP+1.
constexpr int kJ = kMaxP + 1;
PG_HD void jacobi(double x, double t, int nord, double minalpha, double P[kJ][kJ]) {
  for (int i = 0; i < kJ; ++i)
    for (int j = 0; j < kJ; ++j) P[i][j] = 0.0;
  for (int i = 0; i <= nord; ++i) P[i][0] = 1.0;
  const double y = 2.0 * x - t;
  if (nord >= 1)
    for (int i = 0; i < nord; ++i) P[i][1] = y + (minalpha + 2.0 * i) * x;
  if (nord >= 2) {
    const double tt = t * t;
    for (int i = 0; i < nord - 1; ++i) {
      const double al = minalpha + 2.0 * i;
      const double aa = al * al;
      for (int j = 2; j <= nord - i; ++j) {
        const double ai = 2.0 * j * (j + al) * (2.0 * j + al - 2.0);
        const double bi = 2.0 * j + al - 1.0;
        const double ci = (2.0 * j + al) * (2.0 * j + al - 2.0);
        const double di = 2.0 * (j + al - 1.0) * (j - 1.0) * (2.0 * j + al);
        const double v = bi * (ci * y + aa * t) * P[i][j - 1] - di * tt * P[i][j - 2];
        P[i][j] = v / ai;
      }
    }
  }
}

// Integrated Jacobi L, plain P and t-part R, nord x nord (PolyIJacobi).
PG_HD void int_jacobi(double x, double t, int nord, double minalpha, double L[kJ][kJ],
                      double P[kJ][kJ], double R[kJ][kJ]) {
  double pt[kJ][kJ];
  jacobi(x, t, nord, minalpha, pt);
  for (int i = 0; i < kJ; ++i)
    for (int j = 0; j < kJ; ++j) {
      P[i][j] = (i < nord && j < nord) ? pt[i][j] : 0.0;
      L[i][j] = 0.0;
      R[i][j] = 0.0;
    }
  for (int i = 0; i < nord; ++i) L[i][0] = x;
  if (nord >= 2) {
    const double tt = t * t;
    for (int i = 0; i < nord - 1; ++i) {
      const double al = minalpha + 2.0 * i;
      for (int j = 2; j <= nord - i; ++j) {
        const double tia = 2.0 * j + al;
        const double tiam1 = tia - 1.0;
        const double tiam2 = tia - 2.0;
        const double ai = (j + al) / (tiam1 * tia);
        const double bi = al / (tiam2 * tia);
        const double ci = (j - 1.0) / (tiam2 * tiam1);
        L[i][j - 1] = ai * pt[i][j] + bi * t * pt[i][j - 1] - ci * tt * pt[i][j - 2];
        R[i][j - 1] = -(j - 1.0) * (pt[i][j - 1] + t * pt[i][j - 2]) / tiam2;
      }
    }
  }
}

// Homogenised integrated Jacobi and gradients (HomIJacobi).  unit_sum: s0+s1 == 1.
PG_HD void hom_int_jacobi(double s0, double s1, V3 ds0, V3 ds1, int nord, double minalpha,
                          bool unit_sum, double L[kJ][kJ], V3 D[kJ][kJ]) {
  double P[kJ][kJ], R[kJ][kJ];
  if (unit_sum) {
    int_jacobi(s1, 1.0, nord, minalpha, L, P, R);
    for (int i = 0; i < nord; ++i)
      for (int j = 0; j < nord - i; ++j) D[i][j] = P[i][j] * ds1;
  } else {
    int_jacobi(s1, s0 + s1, nord, minalpha, L, P, R);
    const V3 ds01 = ds0 + ds1;
    for (int i = 0; i < nord; ++i)
      for (int j = 0; j < nord - i; ++j) D[i][j] = P[i][j] * ds1 + R[i][j] * ds01;
  }
}

// Edge ancillary functions (AncEE, general case): E_i = P_i(s1; s0+s1)(s0 ds1 - s1 ds0),
// curl E_i = (i+2) P_i ds0 x ds1, i < nord.
PG_HD void anc_edge(double s0, double s1, V3 ds0, V3 ds1, int nord, V3* E, V3* C) {
  double hp[kMaxP + 1];
  legendre(s1, s0 + s1, nord, hp);
  const V3 whi = s0 * ds1 - s1 * ds0;
  const V3 cwhi = cross(ds0, ds1);
  for (int i = 0; i < nord; ++i) {
    E[i] = hp[i] * whi;
    C[i] = ((i + 2.0) * hp[i]) * cwhi;
  }
}

// Triangle ancillary functions (AncETri): ET[j][k-1], j>=0, k>=1, j+k <= nord-1.
PG_HD void anc_tri(const double s[3], const V3 ds[3], int nord, V3 ET[kJ][kJ], V3 CT[kJ][kJ]) {
  V3 E[kMaxP + 1], C[kMaxP + 1];
  anc_edge(s[0], s[1], ds[0], ds[1], nord - 1, E, C);
  double L[kJ][kJ];
  V3 D[kJ][kJ];
  hom_int_jacobi(s[0] + s[1], s[2], ds[0] + ds[1], ds[2], nord - 1, 1.0, false, L, D);
  for (int i = 0; i < nord; ++i)
    for (int j = 0; j < i; ++j) {
      const int k = i - j;
      ET[j][k - 1] = L[j][k - 1] * E[j];
      CT[j][k - 1] = L[j][k - 1] * C[j] + cross(D[j][k - 1], E[j]);
    }
}

// Affine coordinates of the unit tetrahedron and their gradients.
PG_HD void affine(const double X[3], double lam[4], V3 dlam[4]) {
  lam[0] = 1.0 - X[0] - X[1] - X[2];
  lam[1] = X[0];
  lam[2] = X[1];
  lam[3] = X[2];
  dlam[0] = v3(-1.0, -1.0, -1.0);
  dlam[1] = v3(1.0, 0.0, 0.0);
  dlam[2] = v3(0.0, 1.0, 0.0);
  dlam[3] = v3(0.0, 0.0, 1.0);
}

// The p functions of local edge `e` with orientation flag `o`.
PG_HD void edge_functions(const double lam[4], const V3 dlam[4], int e, int o, int p, V3* S,
                          V3* C) {
  int a = edge_vert(e, 0), b = edge_vert(e, 1);
  if (o) {
    const int t = a;
    a = b;
    b = t;
  }
  anc_edge(lam[a], lam[b], dlam[a], dlam[b], p, S, C);
}

// The p(p-1) functions of local face `f` with orientation code `o`.
PG_HD void face_functions(const double lam[4], const V3 dlam[4], int f, int o, int p, V3* S,
                          V3* C) {
  if (p < 2) return;
  int g[3];
  for (int k = 0; k < 3; ++k) g[k] = face_vert(f, face_perm(o, k));
  for (int fam = 0; fam < 2; ++fam) {
    double s[3];
    V3 ds[3];
    for (int k = 0; k < 3; ++k) {
      const int v = g[(fam + k) % 3];
      s[k] = lam[v];
      ds[k] = dlam[v];
    }
    V3 ET[kJ][kJ], CT[kJ][kJ];
    anc_tri(s, ds, p, ET, CT);
    int m = fam;
    for (int k = 1; k < p; ++k)
      for (int r = 0; r < k; ++r) {
        const int pp = k - r;
        S[m] = ET[r][pp - 1];
        C[m] = CT[r][pp - 1];
        m += 2;
      }
  }
}

// The p(p-1)(p-2)/2 interior functions.
PG_HD void bubble_functions(const double lam[4], const V3 dlam[4], int p, V3* S, V3* C) {
  if (p < 3) return;
  for (int fam = 0; fam < 3; ++fam) {
    double s[3];
    V3 ds[3];
    for (int k = 0; k < 3; ++k) {
      s[k] = lam[(fam + k) % 4];
      ds[k] = dlam[(fam + k) % 4];
    }
    const int d = (fam + 3) % 4;
    V3 ET[kJ][kJ], CT[kJ][kJ];
    anc_tri(s, ds, p - 1, ET, CT);
    double Lb[kJ][kJ];
    V3 Db[kJ][kJ];
    hom_int_jacobi(1.0 - lam[d], lam[d], -dlam[d], dlam[d], p - 2, 2.0, true, Lb, Db);
    int m = fam;
    for (int j = 2; j < p; ++j)
      for (int k = 1; k < j; ++k)
        for (int r = 0; r < k; ++r) {
          const int pp = k - r, q = j - k;
          const double hb = Lb[k - 1][q - 1];
          S[m] = hb * ET[r][pp - 1];
          C[m] = hb * CT[r][pp - 1] + cross(Db[k - 1][q - 1], ET[r][pp - 1]);
          m += 3;
        }
  }
}

// All n reference functions and curls of one element at reference point X.
PG_HD void element_functions(const double X[3], const int eo[6], const int fo[4], int p, V3* S,
                             V3* C) {
  double lam[4];
  V3 dlam[4];
  affine(X, lam, dlam);
  int m = 0;
  for (int e = 0; e < 6; ++e) {
    edge_functions(lam, dlam, e, eo[e], p, S + m, C + m);
    m += p;
  }
  for (int f = 0; f < 4; ++f) {
    face_functions(lam, dlam, f, fo[f], p, S + m, C + m);
    m += p * (p - 1);
  }
  bubble_functions(lam, dlam, p, S + m, C + m);
}

// ---------------------------------------------------------------------------
// per-orientation "variants": one index per (entity, orientation, function)
//   edges : v = (e*2 + o)*p + k                        (12p variants)
//   faces : v = 12p + (f*6 + o)*p(p-1) + k            (24p(p-1))
//   bubble: v = 12p + 24p(p-1) + k                    (p(p-1)(p-2)/2)
// ---------------------------------------------------------------------------
PG_HD int n_variants(int p) { return 12 * p + 24 * p * (p - 1) + ndof_vol(p); }

// local entity (0..5 edges, 6..9 faces, 10 volume) -> first local DOF
PG_HD int entity_local_offset(int la, int p) {
  if (la < 6) return la * p;
  if (la < 10) return 6 * p + (la - 6) * p * (p - 1);
  return 6 * p + 4 * p * (p - 1);
}
PG_HD int entity_ndof(int la, int p) {
  if (la < 6) return p;
  if (la < 10) return p * (p - 1);
  return ndof_vol(p);
}
// first variant index of local entity `la` with its orientation `o`
PG_HD int entity_variant_base(int la, int o, int p) {
  if (la < 6) return (la * 2 + o) * p;
  if (la < 10) return 12 * p + ((la - 6) * 6 + o) * p * (p - 1);
  return 12 * p + 24 * p * (p - 1);
}

// packed element orientation: bits 0..5 edge flags, bits 6+3f..8+3f face code f
PG_HD int orient_of(int packed, int la) {
  if (la < 6) return (packed >> la) & 1;
  if (la < 10) return (packed >> (6 + 3 * (la - 6))) & 7;
  return 0;
}

// ---------------------------------------------------------------------------
// quadrature (reference data of petgem/hvfem.py:1065-1127)
// ---------------------------------------------------------------------------
PG_HD int quad_npoints(int p) { return p == 1 ? 4 : (p == 2 ? 11 : 24); }

// rows (t0,t1,t2,w) on the [-1,1] simplex for order 2p
PG_HD void quad_row(int p, int q, double r[4]) {
  const double o2[4][4] = {
      {-0.723606797749979, -0.723606797749979, -0.723606797749979, 0.333333333333333},
      {0.170820393249937, -0.723606797749979, -0.723606797749979, 0.333333333333333},
      {-0.723606797749979, 0.170820393249937, -0.723606797749979, 0.333333333333333},
      {-0.723606797749979, -0.723606797749979, 0.170820393249937, 0.333333333333333}};
  const double o4[11][4] = {
      {-0.500000000000000, -0.500000000000000, -0.500000000000000, -0.105244444444444},
      {-0.857142857142857, -0.857142857142857, -0.857142857142857, 0.060977777777778},
      {-0.857142857142857, -0.857142857142857, 0.571428571428571, 0.060977777777778},
      {-0.857142857142857, 0.571428571428571, -0.857142857142857, 0.060977777777778},
      {0.571428571428571, -0.857142857142857, -0.857142857142857, 0.060977777777778},
      {-0.201192847666402, -0.201192847666402, -0.798807152333598, 0.199111111111111},
      {-0.201192847666402, -0.798807152333598, -0.201192847666402, 0.199111111111111},
      {-0.798807152333598, -0.201192847666402, -0.201192847666402, 0.199111111111111},
      {-0.201192847666402, -0.798807152333598, -0.798807152333598, 0.199111111111111},
      {-0.798807152333598, -0.201192847666402, -0.798807152333598, 0.199111111111111},
      {-0.798807152333598, -0.798807152333598, -0.201192847666402, 0.199111111111111}};
  const double A = -0.570794257481696, B = -0.287617227554912, Wa = 0.053230333677557;
  const double Cc = -0.918652082930777, Dd = 0.755956248792332, Wb = 0.013436281407094;
  const double Ee = -0.355324219715449, F = -0.934027340853653, Wc = 0.073809575391540;
  const double G = -0.872677996249965, H = -0.460655337083368, I = 0.206011329583298,
               Wd = 0.064285714285714;
  const double o6[24][4] = {{A, A, A, Wa},   {B, A, A, Wa},  {A, B, A, Wa},  {A, A, B, Wa},
                            {Cc, Cc, Cc, Wb}, {Dd, Cc, Cc, Wb}, {Cc, Dd, Cc, Wb}, {Cc, Cc, Dd, Wb},
                            {Ee, Ee, Ee, Wc}, {F, Ee, Ee, Wc}, {Ee, F, Ee, Wc}, {Ee, Ee, F, Wc},
                            {G, G, H, Wd},   {G, H, G, Wd},  {G, G, I, Wd},  {G, I, G, Wd},
                            {G, H, I, Wd},   {G, I, H, Wd},  {H, G, G, Wd},  {H, G, I, Wd},
                            {H, I, G, Wd},   {I, G, H, Wd},  {I, G, G, Wd},  {I, H, G, Wd}};
  for (int k = 0; k < 4; ++k) r[k] = p == 1 ? o2[q][k] : (p == 2 ? o4[q][k] : o6[q][k]);
}

// point on the unit tetrahedron and weight (hvfem.py:1603-1608)
PG_HD void quad_point(int p, int q, double X[3], double* w) {
  double r[4];
  quad_row(p, q, r);
  *w = r[3] / 8.0;
  X[0] = (1.0 + r[1]) / 2.0;
  X[1] = -(1.0 + (r[0] + r[1] + r[2])) / 2.0;
  X[2] = (1.0 + r[0]) / 2.0;
}

}  // namespace pg
