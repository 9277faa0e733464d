// Element geometry, element matrices, atomic-free CSR assembly, PEC boundary
// condition, point location, point-dipole source and receiver interpolation
// for gfx950.
//
// Element matrices use the reference-tensor form: with J the Jacobian
// (rows x_k - x_0, petgem/hvfem.py:112-114), G = J^-T diag(sh,sh,sv) J^-1 and
// H = J J^T / det J,
//     M_ij = det J * sum_{a<=b} G'_ab T^M_ab[v_i][v_j]
//     K_ij =         sum_{a<=b} H'_ab T^K_ab[v_i][v_j]
// where v_i is the orientation "variant" of local DOF i and T are quadrature
// contractions of the reference basis computed once per order (tables.cpp
// part below).  This equals the per-quadrature-point loop of
// computeElementalMatrices (petgem/hvfem.py:258-314) up to summation order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "pg_basis.h"
#include "pg_common.h"
#include "pg_rowasm.h"

namespace pg {

constexpr double kMu0 = 4.0 * 3.14159265358979323846 * 1e-7;  // petgem/solver.py:175

// -------------------------------------------------------------------------
// reference tables
// -------------------------------------------------------------------------
static void build_table(int p, std::vector<double>& T) {
  const int V = n_variants(p);
  T.assign((size_t)V * V * 12, 0.0);
  std::vector<V3> S(V), C(V);
  const int nq = quad_npoints(p);
  for (int q = 0; q < nq; ++q) {
    double X[3], w;
    quad_point(p, q, X, &w);
    double lam[4];
    V3 dlam[4];
    affine(X, lam, dlam);
    for (int e = 0; e < 6; ++e)
      for (int o = 0; o < 2; ++o) {
        const int v = entity_variant_base(e, o, p);
        edge_functions(lam, dlam, e, o, p, &S[v], &C[v]);
      }
    if (p >= 2)
      for (int f = 0; f < 4; ++f)
        for (int o = 0; o < 6; ++o) {
          const int v = entity_variant_base(6 + f, o, p);
          face_functions(lam, dlam, f, o, p, &S[v], &C[v]);
        }
    if (p >= 3) {
      const int v = entity_variant_base(10, 0, p);
      bubble_functions(lam, dlam, p, &S[v], &C[v]);
    }
    for (int a = 0; a < V; ++a)
      for (int b = 0; b < V; ++b) {
        double* t = &T[((size_t)a * V + b) * 12];
        const V3 s = S[a], u = S[b], c = C[a], d = C[b];
        t[0] += w * (s.x * u.x);
        t[1] += w * (s.y * u.y);
        t[2] += w * (s.z * u.z);
        t[3] += w * (s.x * u.y + s.y * u.x);
        t[4] += w * (s.x * u.z + s.z * u.x);
        t[5] += w * (s.y * u.z + s.z * u.y);
        t[6] += w * (c.x * d.x);
        t[7] += w * (c.y * d.y);
        t[8] += w * (c.z * d.z);
        t[9] += w * (c.x * d.y + c.y * d.x);
        t[10] += w * (c.x * d.z + c.z * d.x);
        t[11] += w * (c.y * d.z + c.z * d.y);
      }
  }
}

struct Tables {
  int p;
  int V;
  int device;
  double* d_T;
};

// -------------------------------------------------------------------------
// geometry + orientation (one thread per element)
// -------------------------------------------------------------------------
// workspace layout: coef[E][12] (det*G' then H'/det), then orient[E] (int32),
// then one int32 flag (invalid element or conductivity)
__host__ __device__ inline double* ws_coef(void* ws) { return reinterpret_cast<double*>(ws); }
__host__ __device__ inline int32_t* ws_orient(void* ws, int64_t E) {
  return reinterpret_cast<int32_t*>(reinterpret_cast<double*>(ws) + 12 * E);
}
__host__ __device__ inline int32_t* ws_flag(void* ws, int64_t E) { return ws_orient(ws, E) + E; }

__device__ inline int element_orientation(const int32_t* en, const int32_t* ee, const int32_t* edn,
                                          const int32_t* fe, int* eo, int* fo) {
  int packed = 0;
  for (int i = 0; i < 6; ++i) {
    const int a = en[edge_vert(i, 0)], b = en[edge_vert(i, 1)];
    const int flag = (a == edn[2 * i + 1] && b == edn[2 * i + 0]) ? 1 : 0;
    eo[i] = flag;
    packed |= flag << i;
  }
  for (int f = 0; f < 4; ++f) {
    const int g0 = ee[face_edge(f, 0)], g1 = ee[face_edge(f, 1)];
    int k1 = 0, k2 = 0;
    for (int k = 0; k < 3; ++k) {
      if (g0 == fe[3 * f + k]) k1 = k + 1;
      if (g1 == fe[3 * f + k]) k2 = k + 1;
    }
    const int sel = k1 * 10 + k2;
    int code = 0;
    if (sel == 12) code = 0;
    else if (sel == 31) code = 1;
    else if (sel == 23) code = 2;
    else if (sel == 32) code = 3;
    else if (sel == 13) code = 4;
    else if (sel == 21) code = 5;
    fo[f] = code;
    packed |= code << (6 + 3 * f);
  }
  return packed;
}

__device__ inline void jacobian(const double* c, double J[3][3], double iJ[3][3], double* det) {
  for (int r = 0; r < 3; ++r)
    for (int k = 0; k < 3; ++k) J[r][k] = c[3 * (r + 1) + k] - c[k];
  const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
  const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
  const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
  const double d = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
  const double id = 1.0 / d;
  iJ[0][0] = c00 * id;
  iJ[1][0] = c01 * id;
  iJ[2][0] = c02 * id;
  iJ[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
  iJ[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
  iJ[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
  iJ[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
  iJ[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
  iJ[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
  *det = d;
}

__global__ void k_geometry(int64_t E, const double* __restrict__ coords,
                           const int32_t* __restrict__ en, const int32_t* __restrict__ ee,
                           const int32_t* __restrict__ edn, const int32_t* __restrict__ fe,
                           const double* __restrict__ sigma, double* __restrict__ coef,
                           int32_t* __restrict__ orient, int32_t* __restrict__ flag) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  double J[3][3], iJ[3][3], det;
  jacobian(coords + 12 * e, J, iJ, &det);
  const double s[3] = {sigma[2 * e], sigma[2 * e], sigma[2 * e + 1]};
  // the determinant is signed, as in the reference (hvfem.py:265); a degenerate
  // element or a negative / non-finite conductivity is an input error
  if (!(det != 0.0) || !isfinite(det) || !(s[0] >= 0.0) || !(s[2] >= 0.0) || !isfinite(s[0]) ||
      !isfinite(s[2]))
    flag[0] = 1;  // every writer stores the same value
  const int ia[6] = {0, 1, 2, 0, 0, 1}, ib[6] = {0, 1, 2, 1, 2, 2};
  double* out = coef + 12 * e;
  const double idet = 1.0 / det;
  for (int c = 0; c < 6; ++c) {
    const int a = ia[c], b = ib[c];
    double g = 0.0, h = 0.0;
    for (int k = 0; k < 3; ++k) {
      g += s[k] * iJ[k][a] * iJ[k][b];
      h += J[a][k] * J[b][k];
    }
    out[c] = det * g;
    out[6 + c] = h * idet;
  }
  int eo[6], fo[4];
  orient[e] = element_orientation(en + 4 * e, ee + 6 * e, edn + 12 * e, fe + 12 * e, eo, fo);
}

__global__ void k_orientation(int64_t E, const int32_t* __restrict__ en,
                              const int32_t* __restrict__ ee, const int32_t* __restrict__ edn,
                              const int32_t* __restrict__ fe, int32_t* __restrict__ eo_out,
                              int32_t* __restrict__ fo_out) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  int eo[6], fo[4];
  element_orientation(en + 4 * e, ee + 6 * e, edn + 12 * e, fe + 12 * e, eo, fo);
  for (int i = 0; i < 6; ++i) eo_out[6 * e + i] = eo[i];
  for (int i = 0; i < 4; ++i) fo_out[4 * e + i] = fo[i];
}

// -------------------------------------------------------------------------
// element matrices (one thread per (element, i, j))
// -------------------------------------------------------------------------
template <int P>
__device__ inline void dof_entity(int i, int& la, int& k) {
  constexpr int ne = 6 * P, nf = P * (P - 1);
  if (i < ne) {
    la = i / P;
    k = i % P;
    return;
  }
  i -= ne;
  if (nf > 0 && i < 4 * nf) {
    la = 6 + i / (nf > 0 ? nf : 1);
    k = i % (nf > 0 ? nf : 1);
    return;
  }
  i -= 4 * nf;
  la = 10;
  k = i;
}

__device__ inline void contract(const double* __restrict__ t, const double* __restrict__ cf,
                                double* m, double* kk) {
  double a = 0.0, b = 0.0;
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    a += cf[c] * t[c];
    b += cf[6 + c] * t[6 + c];
  }
  *m = a;
  *kk = b;
}

template <int P>
__global__ void k_element_matrices(int64_t E, const double* __restrict__ T,
                                   const double* __restrict__ coef,
                                   const int32_t* __restrict__ orient, double* __restrict__ M,
                                   double* __restrict__ K) {
  constexpr int n = P * (P + 2) * (P + 3) / 2;
  constexpr int V = 12 * P + 24 * P * (P - 1) + P * (P - 1) * (P - 2) / 2;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= E * n * n) return;
  const int64_t e = t / (n * n);
  const int ij = (int)(t - e * n * n);
  const int i = ij / n, j = ij % n;
  const int pk = orient[e];
  int la, ka, lb, kb;
  dof_entity<P>(i, la, ka);
  dof_entity<P>(j, lb, kb);
  const int va = entity_variant_base(la, orient_of(pk, la), P) + ka;
  const int vb = entity_variant_base(lb, orient_of(pk, lb), P) + kb;
  double m, kk;
  contract(T + ((size_t)va * V + vb) * 12, coef + 12 * e, &m, &kk);
  M[t] = m;
  K[t] = kk;
}

// -------------------------------------------------------------------------
// assembly
// -------------------------------------------------------------------------
template <int P>
__device__ inline int entity_ndof_global(int64_t g, int64_t nE, int64_t nF) {
  if (g < nE) return P;
  if (g < nE + nF) return P * (P - 1);
  return P * (P - 1) * (P - 2) / 2;
}

// -------------------------------------------------------------------------
// One-pass assembly (p = 3 plans beyond the row-major bounds): one thread per
// (entity pair, block row)
// that loops over the pair's contributors in ascending element order, contracting
// each entry from the tables and the element's coefficients.  At p = 3 the
// two-phase form below would pad every local block to 6 x 6 (a 5.8 GB scratch
// at C3) and loses to this form.
// -------------------------------------------------------------------------
template <int P>
__global__ void __launch_bounds__(256) k_assemble(
    int64_t npairs, int64_t nE, int64_t nF, const double* __restrict__ T,
    const int32_t* __restrict__ pair_slot0, const int32_t* __restrict__ pair_rowlen,
    const int32_t* __restrict__ pair_cbeg, const int32_t* __restrict__ pair_ab,
    const uint32_t* __restrict__ contrib, const double* __restrict__ coef,
    const int32_t* __restrict__ orient, double omega_mu, const uint8_t* __restrict__ ent_bc,
    double2* __restrict__ values, double* __restrict__ Kout, double* __restrict__ Mout) {
  constexpr int MAXD = P == 1 ? 1 : P * (P - 1);
  constexpr int V = 12 * P + 24 * P * (P - 1) + P * (P - 1) * (P - 2) / 2;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t pair = t / MAXD;
  const int k = (int)(t - pair * MAXD);
  if (pair >= npairs) return;
  const int64_t A = pair_ab[2 * pair], B = pair_ab[2 * pair + 1];
  const int dA = entity_ndof_global<P>(A, nE, nF);
  const int dB = entity_ndof_global<P>(B, nE, nF);
  if (k >= dA) return;
  double re[MAXD], im[MAXD];
#pragma unroll
  for (int l = 0; l < MAXD; ++l) re[l] = im[l] = 0.0;
  const int c0 = pair_cbeg[pair], c1 = pair_cbeg[pair + 1];
  for (int c = c0; c < c1; ++c) {
    const uint32_t u = contrib[c];
    const int64_t e = u >> 8;
    const int la = (u >> 4) & 15, lb = u & 15;
    const int pk = orient[e];
    const double* cf = coef + 12 * e;
    const int va = entity_variant_base(la, orient_of(pk, la), P) + k;
    const int vb = entity_variant_base(lb, orient_of(pk, lb), P);
    const double* trow = T + ((size_t)va * V + vb) * 12;
#pragma unroll
    for (int l = 0; l < MAXD; ++l) {
      if (l < dB) {
        double m, kk;
        contract(trow + 12 * l, cf, &m, &kk);
        re[l] += kk;
        // separately rounded product, as the reference forms i*w*mu0*M_e before adding
        im[l] += Kout ? m : -__dmul_rn(omega_mu, m);
      }
    }
  }
  if (ent_bc != nullptr && (ent_bc[A] | ent_bc[B])) {
#pragma unroll
    for (int l = 0; l < MAXD; ++l) {
      re[l] = (A == B && k == l && ent_bc[A]) ? 1.0 : 0.0;
      im[l] = 0.0;
    }
  }
  const int64_t s0 = (int64_t)pair_slot0[pair] + (int64_t)k * pair_rowlen[pair];
  if (Kout) {
#pragma unroll
    for (int l = 0; l < MAXD; ++l)
      if (l < dB) {
        Kout[s0 + l] = re[l];
        Mout[s0 + l] = im[l];
      }
  } else {
#pragma unroll
    for (int l = 0; l < MAXD; ++l)
      if (l < dB) values[s0 + l] = make_double2(re[l], im[l]);
  }
}

// -------------------------------------------------------------------------
// Two-phase assembly (p = 1, 2), atomic-free, every CSR slot the sum of its
// contributors in the plan's ascending element order (the order of the
// reference's single-rank ADD_VALUES), bitwise the one-pass kernel's result:
//  k_asm_blocks  one thread per (local entity la, local entity lb, element):
//                the lanes of a wave take consecutive elements at the same
//                local pair, so their coefficient loads are contiguous and
//                their table rows repeat with the few orientation patterns of
//                neighbouring elements (cache hits instead of a gather per
//                lane), and the nloc^2 workgroups of one block of elements run
//                one after another (its coefficients read once from L2); each
//                writes its d(la) x d(lb)
//                block of (K_e, -omega mu0 M_e) entries (or (K_e, M_e)):
//                blk[t][k][l], t = (la nloc + lb) E + e, MAXD x MAXD per t;
//  k_asm_sums    one thread per (entity pair, row k of its block): sums its
//                contributors' rows k of those blocks in the plan's ascending
//                element order and stores the pair block's row.
// Traffic: the blocks are written and read once (2.46 GB at C2) besides the
// CSR values, but every load is either a broadcast (an element's coefficients)
// or contiguous, where one thread per slot row gathering coefficients and
// table rows per contributor was bound by divergent gathers.
// -------------------------------------------------------------------------
// streaming (non-temporal) stores of the block scratch: written once, read
// once by the sums, so it should not evict the tables and coefficients from L2
typedef double asm_d2 __attribute__((ext_vector_type(2)));
__device__ inline void nt_store(double2* p, double2 v) {
  __builtin_nontemporal_store(asm_d2{v.x, v.y}, reinterpret_cast<asm_d2*>(p));
}

template <int P>
constexpr int asm_nloc() {
  return P == 1 ? 6 : (P == 2 ? 10 : 11);
}
template <int P>
constexpr int asm_maxd() {
  return P == 1 ? 1 : P * (P - 1);
}
template <int P>
__device__ inline int local_entity_ndof(int la) {
  return la < 6 ? P : (la < 10 ? P * (P - 1) : P * (P - 1) * (P - 2) / 2);
}

template <int P>
__global__ void __launch_bounds__(256) k_asm_blocks(int64_t E, const double* __restrict__ T,
                                                    const double* __restrict__ coef,
                                                    const int32_t* __restrict__ orient,
                                                    double omega_mu, int km,
                                                    double2* __restrict__ blk) {
  constexpr int NL = asm_nloc<P>(), MAXD = asm_maxd<P>();
  constexpr int V = 12 * P + 24 * P * (P - 1) + P * (P - 1) * (P - 2) / 2;
  // p = 2: the workgroup's blocks (4 values each) are staged in LDS and written
  // as one contiguous run, instead of four stores 64 bytes apart per lane
  constexpr bool kStage = P == 2;
  __shared__ double2 stage[kStage ? 256 * MAXD * MAXD : 1];
  // workgroup b: local pair ll = b mod nloc^2 of the element block b / nloc^2,
  // so that the nloc^2 workgroups of one block of 256 elements run together
  // and read its coefficients from L2 (the pair as the slow index re-read all
  // coefficients once per pair)
  const int ll = (int)(blockIdx.x % (NL * NL)), la = ll / NL, lb = ll % NL;
  const int64_t e0 = (int64_t)(blockIdx.x / (NL * NL)) * blockDim.x, e = e0 + threadIdx.x;
  const int64_t t = (int64_t)ll * E + e;
  if (!kStage && e >= E) return;
  if (e < E) {
    const int dA = local_entity_ndof<P>(la), dB = local_entity_ndof<P>(lb);
    const int pk = orient[e];
    const double* cf = coef + 12 * e;
    const int va0 = entity_variant_base(la, orient_of(pk, la), P);
    const int vb = entity_variant_base(lb, orient_of(pk, lb), P);
    double2* out = kStage ? stage + threadIdx.x * (MAXD * MAXD) : blk + t * (MAXD * MAXD);
    for (int k = 0; k < dA; ++k) {
      const double* trow = T + ((size_t)(va0 + k) * V + vb) * 12;
#pragma unroll
      for (int l = 0; l < MAXD; ++l)
        if (l < dB) {
          double m, kk;
          contract(trow + 12 * l, cf, &m, &kk);
          // separately rounded product, as the reference forms i*w*mu0*M_e before adding
          const double2 v = make_double2(kk, km ? m : -__dmul_rn(omega_mu, m));
          if constexpr (kStage) out[k * MAXD + l] = v;
          else nt_store(out + k * MAXD + l, v);
        }
    }
  }
  if constexpr (kStage) {
    __syncthreads();
    const int64_t t0 = (int64_t)ll * E + e0;
    const int64_t n = (E - e0 < (int64_t)blockDim.x ? E - e0 : (int64_t)blockDim.x) *
                      (MAXD * MAXD);
    for (int i = threadIdx.x; i < n; i += blockDim.x) nt_store(blk + t0 * (MAXD * MAXD) + i, stage[i]);
  }
}

template <int P>
__global__ void __launch_bounds__(256) k_asm_sums(
    int64_t npairs, int64_t nE, int64_t nF, int64_t E, const int32_t* __restrict__ pair_slot0,
    const int32_t* __restrict__ pair_rowlen, const int32_t* __restrict__ pair_cbeg,
    const int32_t* __restrict__ pair_ab, const uint32_t* __restrict__ contrib,
    const double2* __restrict__ blk, const uint8_t* __restrict__ ent_bc,
    double2* __restrict__ values, double* __restrict__ Kout, double* __restrict__ Mout) {
  constexpr int NL = asm_nloc<P>(), MAXD = asm_maxd<P>();
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t pair = t / MAXD;
  const int k = (int)(t - pair * MAXD);
  if (pair >= npairs) return;
  const int64_t A = pair_ab[2 * pair], B = pair_ab[2 * pair + 1];
  const int dA = entity_ndof_global<P>(A, nE, nF);
  const int dB = entity_ndof_global<P>(B, nE, nF);
  if (k >= dA) return;
  double re[MAXD], im[MAXD];
#pragma unroll
  for (int l = 0; l < MAXD; ++l) re[l] = im[l] = 0.0;
  const int c0 = pair_cbeg[pair], c1 = pair_cbeg[pair + 1];
  for (int c = c0; c < c1; ++c) {
    const uint32_t u = contrib[c];
    const int64_t e = u >> 8;
    const int la = (u >> 4) & 15, lb = u & 15;
    const double2* row = blk + ((int64_t)(la * NL + lb) * E + e) * (MAXD * MAXD) + k * MAXD;
#pragma unroll
    for (int l = 0; l < MAXD; ++l)
      if (l < dB) {
        const double2 v = row[l];
        re[l] += v.x;
        im[l] += v.y;
      }
  }
  if (ent_bc != nullptr && (ent_bc[A] | ent_bc[B])) {
#pragma unroll
    for (int l = 0; l < MAXD; ++l) {
      re[l] = (A == B && k == l && ent_bc[A]) ? 1.0 : 0.0;
      im[l] = 0.0;
    }
  }
  const int64_t s0 = (int64_t)pair_slot0[pair] + (int64_t)k * pair_rowlen[pair];
  if (Kout) {
#pragma unroll
    for (int l = 0; l < MAXD; ++l)
      if (l < dB) {
        Kout[s0 + l] = re[l];
        Mout[s0 + l] = im[l];
      }
  } else {
#pragma unroll
    for (int l = 0; l < MAXD; ++l)
      if (l < dB) values[s0 + l] = make_double2(re[l], im[l]);
  }
}

// -------------------------------------------------------------------------
// Row-major assembly (p = 1-3; plans within the bounds of pg_rowasm.h): one
// single-wave workgroup per entity A with DOFs keeps A's d(A) CSR rows in LDS
// and walks the elements containing A in ascending order (the contributors of
// the diagonal pair (A, A)).  For each element its lanes form the entries of
// the element's block row la(A) -- row k < d(A), every local DOF j -- from the
// tables and the element's coefficients (one element per step: wave-uniform
// loads) and add each into the LDS row at the column of entity(e, lb).  Every
// slot is thus the sum of its contributors in ascending element order, bitwise
// the two-phase result, with no block scratch; A's rows then leave as one
// contiguous run of d(A) * rowlen values.
// -------------------------------------------------------------------------
template <int P, int LEN>
__global__ void __launch_bounds__(64) k_asm_rows(
    int64_t nE, int64_t nF, const int32_t* __restrict__ epair, const double* __restrict__ T,
    const int32_t* __restrict__ rowptr, const int32_t* __restrict__ pair_slot0,
    const int32_t* __restrict__ pair_cbeg, const int32_t* __restrict__ pair_ab,
    const uint32_t* __restrict__ contrib, const double* __restrict__ coef,
    const int32_t* __restrict__ orient, double omega_mu, int km,
    const uint8_t* __restrict__ ent_bc, double2* __restrict__ values, double* __restrict__ Kout,
    double* __restrict__ Mout) {
  constexpr int NL = asm_nloc<P>(), MAXD = asm_maxd<P>();
  constexpr int DE = P, DF = P * (P - 1), DV = P * (P - 1) * (P - 2) / 2;
  constexpr int NLD = 6 * DE + 4 * DF + DV;
  constexpr int V = 12 * P + 24 * P * (P - 1) + P * (P - 1) * (P - 2) / 2;
  __shared__ double2 acc[MAXD * LEN];
  __shared__ int16_t offs[kRowAsmMaxInc * NL];  // column of entity(e_i, lb) in A's rows
  __shared__ int32_t inc_e[kRowAsmMaxInc];
  __shared__ int32_t inc_o[kRowAsmMaxInc];  // la << 24 | orientation code (18 bits)
  __shared__ int32_t pr_cbeg[kRowAsmMaxPairs + 1];
  __shared__ int16_t pr_off[kRowAsmMaxPairs];
  __shared__ uint8_t colbc[LEN];
  __shared__ int32_t s_pd;
  const int lane = threadIdx.x;
  const int64_t A = blockIdx.x;  // the grid covers the entities with DOFs
  const int dA = entity_ndof_global<P>(A, nE, nF);
  const int64_t r0 = A < nE ? A * DE
                            : (A < nE + nF ? nE * DE + (A - nE) * DF
                                           : nE * DE + nF * DF + (A - nE - nF) * DV);
  const int64_t rowbase = rowptr[r0];
  const int rowlen = (int)(rowptr[r0 + 1] - rowbase);
  const int p0 = epair[A], npr = epair[A + 1] - p0;
  // (the plan export admits only plans within the bounds)
  if (dA == 0 || npr <= 0 || npr > kRowAsmMaxPairs || rowlen > LEN) return;
  const int bcA = ent_bc ? ent_bc[A] : 0;
  // A's pairs: contributor ranges, column offsets, PEC columns, the diagonal pair
  for (int q = lane; q < npr; q += 64) {
    const int pi = p0 + q;
    const int64_t B = pair_ab[2 * pi + 1];
    const int off = (int)(pair_slot0[pi] - rowbase);
    pr_cbeg[q] = pair_cbeg[pi];
    pr_off[q] = (int16_t)off;
    const int dB = entity_ndof_global<P>(B, nE, nF);
    const uint8_t bb = ent_bc ? ent_bc[B] : 0;
    for (int l = 0; l < dB; ++l) colbc[off + l] = bb;
    if (B == A) s_pd = q;
  }
  if (lane == 0) pr_cbeg[npr] = pair_cbeg[p0 + npr];
  for (int q = lane; q < dA * rowlen; q += 64) acc[q] = make_double2(0.0, 0.0);
  __syncthreads();
  const int pd = s_pd;
  const int ci0 = pr_cbeg[pd], ninc = pr_cbeg[pd + 1] - ci0;
  if (ninc > kRowAsmMaxInc) return;  // uniform
  // every contributor (e, la, lb) of A's pairs, loaded while the elements
  // containing A (the diagonal pair's contributors, ascending) are listed with
  // A's local entity and the element's orientation code
  constexpr int CPL = (kRowAsmMaxInc * NL + 63) / 64;  // contributors per lane
  const int c0 = pr_cbeg[0], c1 = pr_cbeg[npr];
  uint32_t cu[CPL];
#pragma unroll
  for (int m = 0; m < CPL; ++m) cu[m] = c0 + lane + 64 * m < c1 ? contrib[c0 + lane + 64 * m] : 0u;
  for (int i = lane; i < ninc; i += 64) {
    const uint32_t u = contrib[ci0 + i];
    const int e = (int)(u >> 8);
    inc_e[i] = e;
    inc_o[i] = (int)(((u >> 4) & 15) << 24) | orient[e];
  }
  __syncthreads();
  // column offset of entity(e, lb) for every contributor of A's pairs
#pragma unroll
  for (int m = 0; m < CPL; ++m) {
    const int c = c0 + lane + 64 * m;
    if (c >= c1) continue;
    const uint32_t u = cu[m];
    const int e = (int)(u >> 8), lb = (int)(u & 15);
    int lo = 0, hi = npr - 1;  // the pair of c: last q with pr_cbeg[q] <= c
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (pr_cbeg[mid] <= c) lo = mid;
      else hi = mid - 1;
    }
    int a = 0, b = ninc - 1;  // the position of e in the ascending element list
    while (a < b) {
      const int mid = (a + b) >> 1;
      if (inc_e[mid] < e) a = mid + 1;
      else b = mid;
    }
    offs[a * NL + lb] = pr_off[lo];
  }
  __syncthreads();
  // element block rows, ascending elements
  for (int i = 0; i < ninc; ++i) {
    const int64_t e = __builtin_amdgcn_readfirstlane(inc_e[i]);
    const int io = __builtin_amdgcn_readfirstlane(inc_o[i]);
    const int la = io >> 24, pk = io & 0xffffff;
    const double* cf = coef + 12 * e;
    const int va0 = entity_variant_base(la, orient_of(pk, la), P);
    for (int q = lane; q < dA * NLD; q += 64) {
      const int k = q / NLD, j = q - k * NLD;
      int lb, l;
      dof_entity<P>(j, lb, l);
      const int vb = entity_variant_base(lb, orient_of(pk, lb), P) + l;
      double m, kk;
      contract(T + ((size_t)(va0 + k) * V + vb) * 12, cf, &m, &kk);
      // separately rounded product, as the reference forms i*w*mu0*M_e before adding
      const double im = km ? m : -__dmul_rn(omega_mu, m);
      double2* sl = acc + k * rowlen + offs[i * NL + lb] + l;
      sl->x += kk;
      sl->y += im;
    }
    __syncthreads();
  }
  // PEC rows and columns (zero, unit diagonal), then A's rows in one run
  const int dg = pr_off[pd];
  for (int q = lane; q < dA * rowlen; q += 64) {
    const int k = q / rowlen, col = q - k * rowlen;
    double2 v = acc[q];
    if (bcA | colbc[col]) v = make_double2((bcA && col - dg == k) ? 1.0 : 0.0, 0.0);
    if (Kout) {
      Kout[rowbase + q] = v.x;
      Mout[rowbase + q] = v.y;
    } else {
      values[rowbase + q] = v;
    }
  }
}

// -------------------------------------------------------------------------
// PEC boundary condition on an assembled CSR (16 lanes per row, so that the
// column and value reads of a row are coalesced)
// -------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_dirichlet(int64_t n, const int32_t* __restrict__ rowptr,
                                                   const int32_t* __restrict__ cols,
                                                   const uint8_t* __restrict__ bc,
                                                   double2* __restrict__ vals,
                                                   double2* __restrict__ b) {
  constexpr int G = 16;
  const int64_t r = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G;
  const int lane = threadIdx.x % G;
  if (r >= n) return;
  const bool br = bc[r] != 0;
  if (vals != nullptr)
    for (int s = rowptr[r] + lane; s < rowptr[r + 1]; s += G) {
      const int c = cols[s];
      if (br) vals[s] = make_double2(c == r ? 1.0 : 0.0, 0.0);
      else if (bc[c]) vals[s] = make_double2(0.0, 0.0);
    }
  if (br && b != nullptr && lane == 0) b[r] = make_double2(0.0, 0.0);
}

// -------------------------------------------------------------------------
// point location: lowest element index whose barycentric coordinates are
// within [-tol, 1+tol]
// -------------------------------------------------------------------------
__device__ inline bool barycentric_inside(const double* c, const double* x, double tol) {
  // columns of T are v_i - v_3 (i = 0..2); solve T lam = x - v_3
  double T[3][3];
  for (int i = 0; i < 3; ++i)
    for (int k = 0; k < 3; ++k) T[k][i] = c[3 * i + k] - c[9 + k];
  const double r[3] = {x[0] - c[9], x[1] - c[10], x[2] - c[11]};
  const double c00 = T[1][1] * T[2][2] - T[1][2] * T[2][1];
  const double c01 = T[1][2] * T[2][0] - T[1][0] * T[2][2];
  const double c02 = T[1][0] * T[2][1] - T[1][1] * T[2][0];
  const double d = T[0][0] * c00 + T[0][1] * c01 + T[0][2] * c02;
  if (!(d != 0.0) || !isfinite(d)) return false;
  const double id = 1.0 / d;
  double inv[3][3];
  inv[0][0] = c00 * id;
  inv[1][0] = c01 * id;
  inv[2][0] = c02 * id;
  inv[0][1] = (T[0][2] * T[2][1] - T[0][1] * T[2][2]) * id;
  inv[1][1] = (T[0][0] * T[2][2] - T[0][2] * T[2][0]) * id;
  inv[2][1] = (T[0][1] * T[2][0] - T[0][0] * T[2][1]) * id;
  inv[0][2] = (T[0][1] * T[1][2] - T[0][2] * T[1][1]) * id;
  inv[1][2] = (T[0][2] * T[1][0] - T[0][0] * T[1][2]) * id;
  inv[2][2] = (T[0][0] * T[1][1] - T[0][1] * T[1][0]) * id;
  double last = 1.0;
  for (int i = 0; i < 3; ++i) {
    const double l = inv[i][0] * r[0] + inv[i][1] * r[1] + inv[i][2] * r[2];
    if (!(l >= -tol && l <= 1.0 + tol)) return false;
    last -= l;
  }
  return last >= -tol && last <= 1.0 + tol;
}

__global__ void k_locate(int64_t E, const double* __restrict__ coords, int64_t npts,
                         const double* __restrict__ pts, double tol, int32_t* __restrict__ out) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  const double* c = coords + 12 * e;
  double lo[3], hi[3];
  for (int k = 0; k < 3; ++k) {
    lo[k] = fmin(fmin(c[k], c[3 + k]), fmin(c[6 + k], c[9 + k]));
    hi[k] = fmax(fmax(c[k], c[3 + k]), fmax(c[6 + k], c[9 + k]));
  }
  for (int64_t q = 0; q < npts; ++q) {
    const double* x = pts + 3 * q;
    const double sl = 1e-6 * (hi[0] - lo[0] + hi[1] - lo[1] + hi[2] - lo[2]);
    if (x[0] < lo[0] - sl || x[0] > hi[0] + sl || x[1] < lo[1] - sl || x[1] > hi[1] + sl ||
        x[2] < lo[2] - sl || x[2] > hi[2] + sl)
      continue;
    if (barycentric_inside(c, x, tol)) atomicMin(&out[q], (int32_t)e);
  }
}

__global__ void k_fill_i32(int64_t n, int32_t* p, int32_t v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_notfound(int64_t n, int32_t* p) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && p[i] == 0x7fffffff) p[i] = -1;
}

// -------------------------------------------------------------------------
// basis at a physical point of one element (source and receivers)
// -------------------------------------------------------------------------
template <int P>
__device__ inline void physical_basis(const double* c, const int32_t* en, const int32_t* ee,
                                      const int32_t* edn, const int32_t* fe, const double* x,
                                      V3* N, V3* Cu) {
  constexpr int n = P * (P + 2) * (P + 3) / 2;
  double J[3][3], iJ[3][3], det;
  jacobian(c, J, iJ, &det);
  // reference coordinates: xi = J^-T (x - x0)
  const double r[3] = {x[0] - c[0], x[1] - c[1], x[2] - c[2]};
  double X[3];
  for (int a = 0; a < 3; ++a) X[a] = r[0] * iJ[0][a] + r[1] * iJ[1][a] + r[2] * iJ[2][a];
  int eo[6], fo[4];
  element_orientation(en, ee, edn, fe, eo, fo);
  V3 S[n], C[n];
  element_functions(X, eo, fo, P, S, C);
  for (int i = 0; i < n; ++i) {
    const V3 s = S[i], u = C[i];
    N[i] = v3(iJ[0][0] * s.x + iJ[0][1] * s.y + iJ[0][2] * s.z,
              iJ[1][0] * s.x + iJ[1][1] * s.y + iJ[1][2] * s.z,
              iJ[2][0] * s.x + iJ[2][1] * s.y + iJ[2][2] * s.z);
    Cu[i] = v3((J[0][0] * u.x + J[1][0] * u.y + J[2][0] * u.z) / det,
               (J[0][1] * u.x + J[1][1] * u.y + J[2][1] * u.z) / det,
               (J[0][2] * u.x + J[1][2] * u.y + J[2][2] * u.z) / det);
  }
}

template <int P>
__global__ void k_source(int32_t e, const double* __restrict__ coords,
                         const int32_t* __restrict__ en, const int32_t* __restrict__ ee,
                         const int32_t* __restrict__ edn, const int32_t* __restrict__ fe,
                         const int32_t* __restrict__ dofs, const double* __restrict__ pos,
                         const double* __restrict__ moment, double omega_mu,
                         double2* __restrict__ b) {
  constexpr int n = P * (P + 2) * (P + 3) / 2;
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  V3 N[n], Cu[n];
  const int64_t E = e;
  physical_basis<P>(coords + 12 * E, en + 4 * E, ee + 6 * E, edn + 12 * E, fe + 12 * E, pos, N,
                    Cu);
  for (int i = 0; i < n; ++i) {
    const double v = moment[0] * N[i].x + moment[1] * N[i].y + moment[2] * N[i].z;
    const int g = dofs[E * n + i];
    double2 bi = b[g];
    bi.y += omega_mu * v;
    b[g] = bi;
  }
}

template <int P>
__global__ void k_receivers(int64_t npts, const double* __restrict__ pts,
                            const int32_t* __restrict__ elem, const double* __restrict__ coords,
                            const int32_t* __restrict__ en, const int32_t* __restrict__ ee,
                            const int32_t* __restrict__ edn, const int32_t* __restrict__ fe,
                            const int32_t* __restrict__ dofs, const double2* __restrict__ x,
                            double omega_mu, double2* __restrict__ out) {
  constexpr int n = P * (P + 2) * (P + 3) / 2;
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= npts) return;
  const int64_t E = elem[q];
  double2* o = out + 6 * q;
  if (E < 0) {
    for (int k = 0; k < 6; ++k) o[k] = make_double2(0.0, 0.0);
    return;
  }
  V3 N[n], Cu[n];
  physical_basis<P>(coords + 12 * E, en + 4 * E, ee + 6 * E, edn + 12 * E, fe + 12 * E,
                    pts + 3 * q, N, Cu);
  double2 f[6];
  for (int k = 0; k < 6; ++k) f[k] = make_double2(0.0, 0.0);
  for (int i = 0; i < n; ++i) {
    const double2 xi = x[dofs[E * n + i]];
    const double nv[3] = {N[i].x, N[i].y, N[i].z};
    const double cv[3] = {Cu[i].x, Cu[i].y, Cu[i].z};
    for (int k = 0; k < 3; ++k) {
      f[k].x += xi.x * nv[k];
      f[k].y += xi.y * nv[k];
      f[3 + k].x += xi.x * cv[k];
      f[3 + k].y += xi.y * cv[k];
    }
  }
  // H = curl E / (i omega mu0): (a + ib) / (i w) = (b - i a) / w
  for (int k = 0; k < 3; ++k) {
    o[k] = f[k];
    o[3 + k] = make_double2(f[3 + k].y / omega_mu, -f[3 + k].x / omega_mu);
  }
}


// -------------------------------------------------------------------------
// adjoint: receiver adjoint source and per-element gradient contraction
// -------------------------------------------------------------------------
// f[dofs_e(r)] += sum_c coef[r][c] * N_c(x_r) for every located receiver r,
// accumulated by one thread in receiver order (deterministic).
template <int P>
__global__ void k_adjoint_rhs(int64_t npts, const double* __restrict__ pts,
                              const int32_t* __restrict__ elem, const double* __restrict__ coords,
                              const int32_t* __restrict__ en, const int32_t* __restrict__ ee,
                              const int32_t* __restrict__ edn, const int32_t* __restrict__ fe,
                              const int32_t* __restrict__ dofs, const double2* __restrict__ coef,
                              double2* __restrict__ f) {
  constexpr int n = P * (P + 2) * (P + 3) / 2;
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  for (int64_t q = 0; q < npts; ++q) {
    const int64_t E = elem[q];
    if (E < 0) continue;
    V3 N[n], Cu[n];
    physical_basis<P>(coords + 12 * E, en + 4 * E, ee + 6 * E, edn + 12 * E, fe + 12 * E,
                      pts + 3 * q, N, Cu);
    const double2 c0 = coef[3 * q], c1 = coef[3 * q + 1], c2 = coef[3 * q + 2];
    for (int i = 0; i < n; ++i) {
      const int g = dofs[E * n + i];
      double2 v = f[g];
      v.x += c0.x * N[i].x + c1.x * N[i].y + c2.x * N[i].z;
      v.y += c0.y * N[i].x + c1.y * N[i].y + c2.y * N[i].z;
      f[g] = v;
    }
  }
}

// g[e] += sum_r -omega_r*mu0 * Im(lam_e^T M_e x_e), M_e recomputed from the tables.
// One thread per element.  Column r of x and lam holds DOF d at d*sd + r*sr
// ((nrhs, N) layout: sd = 1, sr = N; (N, nrhs) layout: sd = nrhs, sr = 1).  Each
// M_ij is contracted once per chunk of RC columns.
template <int P>
__global__ void __launch_bounds__(256) k_adjoint_grad(
    int64_t E, const double* __restrict__ T, const double* __restrict__ coef,
    const int32_t* __restrict__ orient, const int32_t* __restrict__ dofs, int nrhs, int64_t sd,
    int64_t sr, const double2* __restrict__ x, const double2* __restrict__ lam,
    const double* __restrict__ omega_mu, double* __restrict__ g) {
  constexpr int n = P * (P + 2) * (P + 3) / 2;
  constexpr int V = 12 * P + 24 * P * (P - 1) + P * (P - 1) * (P - 2) / 2;
  constexpr int RC = 8;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  const int pk = orient[e];
  const double* cf = coef + 12 * e;
  const int32_t* de = dofs + e * n;
  double acc_g = 0.0;
  for (int r0 = 0; r0 < nrhs; r0 += RC) {
    const int nr = nrhs - r0 < RC ? nrhs - r0 : RC;
    double sim[RC];
#pragma unroll
    for (int t = 0; t < RC; ++t) sim[t] = 0.0;
    for (int i = 0; i < n; ++i) {
      int la, ka;
      dof_entity<P>(i, la, ka);
      const int va = entity_variant_base(la, orient_of(pk, la), P) + ka;
      double are[RC], aim[RC];
#pragma unroll
      for (int t = 0; t < RC; ++t) are[t] = aim[t] = 0.0;
      for (int j = 0; j < n; ++j) {
        int lb, kb;
        dof_entity<P>(j, lb, kb);
        const int vb = entity_variant_base(lb, orient_of(pk, lb), P) + kb;
        double m, kk;
        contract(T + ((size_t)va * V + vb) * 12, cf, &m, &kk);
        const double2* xj = x + (int64_t)de[j] * sd + (int64_t)r0 * sr;
#pragma unroll
        for (int t = 0; t < RC; ++t)
          if (t < nr) {
            const double2 v = xj[t * sr];
            are[t] += m * v.x;
            aim[t] += m * v.y;
          }
      }
      const double2* li = lam + (int64_t)de[i] * sd + (int64_t)r0 * sr;
#pragma unroll
      for (int t = 0; t < RC; ++t)
        if (t < nr) {
          const double2 l = li[t * sr];
          sim[t] += l.x * aim[t] + l.y * are[t];
        }
    }
#pragma unroll
    for (int t = 0; t < RC; ++t)
      if (t < nr) acc_g += -omega_mu[r0 + t] * sim[t];
  }
  g[e] += acc_g;
}

static inline unsigned nblocks(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// ---------------------------------------------------------------------------
// Staging of the adjoint gradient's per-right-hand-side omega*mu0 for the
// entry point that allocates nothing (pg_adjoint_gradient_noalloc).  Every
// device holds its own copy of `scratch`, at its own address; in host code the
// name is only a stand-in, so the address is resolved once per device with
// hipGetSymbolAddress (that device current) and only that address is used.
// Each call overwrites the first nrhs entries before its kernel reads them;
// calls that overlap on one device (several host threads) race on it.
// ---------------------------------------------------------------------------
constexpr int kAdjointMaxRhs = 4096;
static __device__ __align__(256) double scratch[kAdjointMaxRhs];

static std::mutex g_scratch_mu;
static std::vector<double*> g_scratch_ptr;  // by device ordinal; nullptr = not resolved yet

static int scratch_address(double** out) {
  int dev = 0;
  PG_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if ((int)g_scratch_ptr.size() <= dev) g_scratch_ptr.resize(dev + 1, nullptr);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    PG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) PG_FAIL(PG_ERR_HIP, "hipGetSymbolAddress returned a null staging address");
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return PG_OK;
}

}  // namespace pg

using namespace pg;

// =========================================================================
// C ABI
// =========================================================================
extern "C" {

int pg_tables_host(int p, double* out, int64_t out_len) {
  if (p < 1 || p > kMaxP) PG_FAIL(PG_ERR_UNSUPPORTED, "order p=%d not supported (1..3)", p);
  std::vector<double> T;
  build_table(p, T);
  if (out_len < (int64_t)T.size()) PG_FAIL(PG_ERR_ARG, "table needs %zu doubles", T.size());
  std::memcpy(out, T.data(), T.size() * sizeof(double));
  return PG_OK;
}

int pg_tables_create(int p, void** tables) {
  if (!tables) PG_FAIL(PG_ERR_ARG, "tables is NULL");
  if (p < 1 || p > kMaxP) PG_FAIL(PG_ERR_UNSUPPORTED, "order p=%d not supported (1..3)", p);
  std::vector<double> T;
  build_table(p, T);
  Tables* t = new Tables();
  t->p = p;
  t->V = n_variants(p);
  hipError_t err = hipGetDevice(&t->device);
  if (err == hipSuccess) err = hipMalloc(&t->d_T, T.size() * sizeof(double));
  if (err != hipSuccess) {
    delete t;
    PG_FAIL(PG_ERR_HIP, "hipMalloc(tables): %s", hipGetErrorString(err));
  }
  err = hipMemcpy(t->d_T, T.data(), T.size() * sizeof(double), hipMemcpyHostToDevice);
  if (err != hipSuccess) {
    (void)hipFree(t->d_T);
    delete t;
    PG_FAIL(PG_ERR_HIP, "hipMemcpy(tables): %s", hipGetErrorString(err));
  }
  *tables = t;
  return PG_OK;
}

int pg_tables_destroy(void* tables) {
  if (!tables) return PG_OK;
  Tables* t = static_cast<Tables*>(tables);
  (void)hipFree(t->d_T);
  delete t;
  return PG_OK;
}

size_t pg_element_workspace_bytes(int64_t n_elems) {
  return (size_t)n_elems * (12 * sizeof(double) + sizeof(int32_t)) + sizeof(int32_t);
}

int pg_element_geometry(int p, int64_t E, const double* coords, const int32_t* en,
                        const int32_t* ee, const int32_t* edn, const int32_t* fe,
                        const double* sigma, void* ws, void* stream) {
  if (p < 1 || p > kMaxP) PG_FAIL(PG_ERR_UNSUPPORTED, "order p=%d not supported", p);
  if (E <= 0) PG_FAIL(PG_ERR_ARG, "n_elems must be positive");
  if (!coords || !en || !ee || !edn || !fe || !sigma || !ws) PG_FAIL(PG_ERR_ARG, "NULL input");
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, 1, ws_flag(ws, E), 0);
  hipLaunchKernelGGL(k_geometry, dim3(nblocks(E, 256)), dim3(256), 0, s, E, coords, en, ee, edn,
                     fe, sigma, ws_coef(ws), ws_orient(ws, E), ws_flag(ws, E));
  PG_CHECK_LAUNCH();
  int32_t flag = 0;
  PG_HIP(hipMemcpyAsync(&flag, ws_flag(ws, E), sizeof(int32_t), hipMemcpyDeviceToHost, s));
  PG_HIP(hipStreamSynchronize(s));
  if (flag)
    PG_FAIL(PG_ERR_ARG,
            "invalid element input: a degenerate element (zero or non-finite Jacobian "
            "determinant) or a negative / non-finite conductivity");
  return PG_OK;
}

int pg_element_orientation(int64_t E, const int32_t* en, const int32_t* ee, const int32_t* edn,
                           const int32_t* fe, int32_t* eo, int32_t* fo, void* stream) {
  if (E <= 0 || !en || !ee || !edn || !fe || !eo || !fo) PG_FAIL(PG_ERR_ARG, "bad arguments");
  hipLaunchKernelGGL(k_orientation, dim3(nblocks(E, 256)), dim3(256), 0, as_stream(stream), E, en,
                     ee, edn, fe, eo, fo);
  PG_CHECK_LAUNCH();
  return PG_OK;
}

int pg_element_matrices(const void* tables, int64_t E, const void* ws, double* M, double* K,
                        void* stream) {
  if (!tables || !ws || !M || !K || E <= 0) PG_FAIL(PG_ERR_ARG, "bad arguments");
  const Tables* t = static_cast<const Tables*>(tables);
  const int n = ndof_element(t->p);
  const int64_t total = E * n * n;
  void* w = const_cast<void*>(ws);
  const dim3 g(nblocks(total, 256)), b(256);
  hipStream_t s = as_stream(stream);
  switch (t->p) {
    case 1: hipLaunchKernelGGL(k_element_matrices<1>, g, b, 0, s, E, t->d_T, ws_coef(w), ws_orient(w, E), M, K); break;
    case 2: hipLaunchKernelGGL(k_element_matrices<2>, g, b, 0, s, E, t->d_T, ws_coef(w), ws_orient(w, E), M, K); break;
    case 3: hipLaunchKernelGGL(k_element_matrices<3>, g, b, 0, s, E, t->d_T, ws_coef(w), ws_orient(w, E), M, K); break;
  }
  PG_CHECK_LAUNCH();
  return PG_OK;
}

static int assemble_impl(const void* tables, const pg_plan_dev* pl, const void* ws, double omega,
                         const uint8_t* ent_bc, double* values, double* Kout, double* Mout,
                         void* stream) {
  const Tables* t = static_cast<const Tables*>(tables);
  if (pl->p != t->p) PG_FAIL(PG_ERR_ARG, "plan order %d != tables order %d", pl->p, t->p);
  if (pl->n_elems >= (1LL << 24)) PG_FAIL(PG_ERR_UNSUPPORTED, "more than 2^24 elements");
  const int maxd = t->p == 1 ? 1 : t->p * (t->p - 1);
  const int nloc = t->p == 1 ? 6 : (t->p == 2 ? 10 : 11);
  const int64_t threads = pl->npairs * maxd;
  const int64_t nblk = pl->n_elems * nloc * nloc;
  void* w = const_cast<void*>(ws);
  const double om = omega * kMu0;
  const dim3 g(nblocks(threads, 256)), gb(nblocks(pl->n_elems, 256) * nloc * nloc), b(256);
  hipStream_t s = as_stream(stream);
  double2* v = reinterpret_cast<double2*>(values);
  if (pl->entity_pair) {
    const int64_t nent = t->p == 1 ? pl->n_edges
                                   : pl->n_edges + pl->n_faces + (t->p == 3 ? pl->n_elems : 0);
    const int km = Kout ? 1 : 0;
    // the short-row variant when the plan's longest row fits it (more workgroups resident)
    const bool short_rows = pl->entity_row_max > 0 && pl->entity_row_max <= kRowAsmShortLen;
#define PG_ROWS(P, LEN)                                                                        \
  hipLaunchKernelGGL((k_asm_rows<P, LEN>), dim3((unsigned)nent), dim3(64), 0, s, pl->n_edges,   \
                     pl->n_faces, pl->entity_pair, t->d_T, pl->rowptr, pl->pair_slot0,          \
                     pl->pair_cbeg, pl->pair_ab, pl->contrib, ws_coef(w),                       \
                     ws_orient(w, pl->n_elems), om, km, ent_bc, v, Kout, Mout)
    if (t->p == 1 && short_rows) {
      PG_ROWS(1, kRowAsmShortLen);
    } else if (t->p == 1) {
      PG_ROWS(1, kRowAsmMaxLen);
    } else if (t->p == 2 && short_rows) {
      PG_ROWS(2, kRowAsmShortLen);
    } else if (t->p == 2) {
      PG_ROWS(2, kRowAsmMaxLen);
    } else if (short_rows) {
      PG_ROWS(3, kRowAsmShortLen);
    } else {
      PG_ROWS(3, kRowAsmMaxLen);
    }
#undef PG_ROWS
    PG_CHECK_LAUNCH();
    return PG_OK;
  }
  if (t->p == 3) {
    hipLaunchKernelGGL(k_assemble<3>, g, b, 0, s, pl->npairs, pl->n_edges, pl->n_faces, t->d_T,
                       pl->pair_slot0, pl->pair_rowlen, pl->pair_cbeg, pl->pair_ab, pl->contrib,
                       ws_coef(w), ws_orient(w, pl->n_elems), om, ent_bc, v, Kout, Mout);
    PG_CHECK_LAUNCH();
    return PG_OK;
  }
  // the element blocks: stream-ordered scratch (the pool keeps it between calls)
  double2* blk = nullptr;
  PG_HIP(hipMallocAsync(reinterpret_cast<void**>(&blk),
                        (size_t)nblk * maxd * maxd * sizeof(double2), s));
  const int km = Kout ? 1 : 0;
#define PG_ASM(P)                                                                             \
  hipLaunchKernelGGL(k_asm_blocks<P>, gb, b, 0, s, pl->n_elems, t->d_T, ws_coef(w),             \
                     ws_orient(w, pl->n_elems), om, km, blk);                                  \
  hipLaunchKernelGGL(k_asm_sums<P>, g, b, 0, s, pl->npairs, pl->n_edges, pl->n_faces,           \
                     pl->n_elems,                                                              \
                     pl->pair_slot0, pl->pair_rowlen, pl->pair_cbeg, pl->pair_ab, pl->contrib, \
                     blk, ent_bc, v, Kout, Mout)
  if (t->p == 1) {
    PG_ASM(1);
  } else {
    PG_ASM(2);
  }
#undef PG_ASM
  PG_CHECK_LAUNCH();
  PG_HIP(hipFreeAsync(blk, s));
  return PG_OK;
}

int pg_assemble(const void* tables, const pg_plan_dev* pl, const void* ws, double omega,
                const uint8_t* ent_bc, double* values, void* stream) {
  if (!tables || !pl || !ws || !values) PG_FAIL(PG_ERR_ARG, "bad arguments");
  return assemble_impl(tables, pl, ws, omega, ent_bc, values, nullptr, nullptr, stream);
}

int pg_assemble_km(const void* tables, const pg_plan_dev* pl, const void* ws,
                   const uint8_t* ent_bc, double* K, double* M, void* stream) {
  if (!tables || !pl || !ws || !K || !M) PG_FAIL(PG_ERR_ARG, "bad arguments");
  return assemble_impl(tables, pl, ws, 0.0, ent_bc, nullptr, K, M, stream);
}

int pg_apply_dirichlet(int64_t n, const int32_t* rowptr, const int32_t* cols, const uint8_t* bc,
                       double* values, double* b, void* stream) {
  if (n <= 0 || !rowptr || !cols || !bc || (!values && !b)) PG_FAIL(PG_ERR_ARG, "bad arguments");
  hipLaunchKernelGGL(k_dirichlet, dim3(nblocks(16 * n, 256)), dim3(256), 0, as_stream(stream), n,
                     rowptr, cols, bc, reinterpret_cast<double2*>(values),
                     reinterpret_cast<double2*>(b));
  PG_CHECK_LAUNCH();
  return PG_OK;
}

int pg_locate_points(int64_t E, const double* coords, int64_t npts, const double* pts, double tol,
                     int32_t* out, void* stream) {
  if (E <= 0 || npts <= 0 || !coords || !pts || !out) PG_FAIL(PG_ERR_ARG, "bad arguments");
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(k_fill_i32, dim3(nblocks(npts, 256)), dim3(256), 0, s, npts, out,
                     (int32_t)0x7fffffff);
  hipLaunchKernelGGL(k_locate, dim3(nblocks(E, 256)), dim3(256), 0, s, E, coords, npts, pts, tol,
                     out);
  hipLaunchKernelGGL(k_notfound, dim3(nblocks(npts, 256)), dim3(256), 0, s, npts, out);
  PG_CHECK_LAUNCH();
  return PG_OK;
}

int pg_source_vector(int p, int32_t e, const double* coords, const int32_t* en, const int32_t* ee,
                     const int32_t* edn, const int32_t* fe, const int32_t* dofs,
                     const double* pos, const double* moment, double omega, double* b,
                     void* stream) {
  if (p < 1 || p > kMaxP) PG_FAIL(PG_ERR_UNSUPPORTED, "order p=%d not supported", p);
  if (e < 0) PG_FAIL(PG_ERR_NOT_FOUND, "source element not found (outside the mesh)");
  if (!coords || !en || !ee || !edn || !fe || !dofs || !pos || !moment || !b)
    PG_FAIL(PG_ERR_ARG, "NULL input");
  const double om = omega * kMu0;
  hipStream_t s = as_stream(stream);
  double2* bb = reinterpret_cast<double2*>(b);
  switch (p) {
    case 1: hipLaunchKernelGGL(k_source<1>, dim3(1), dim3(64), 0, s, e, coords, en, ee, edn, fe, dofs, pos, moment, om, bb); break;
    case 2: hipLaunchKernelGGL(k_source<2>, dim3(1), dim3(64), 0, s, e, coords, en, ee, edn, fe, dofs, pos, moment, om, bb); break;
    case 3: hipLaunchKernelGGL(k_source<3>, dim3(1), dim3(64), 0, s, e, coords, en, ee, edn, fe, dofs, pos, moment, om, bb); break;
  }
  PG_CHECK_LAUNCH();
  return PG_OK;
}

int pg_receiver_fields(int p, int64_t npts, const double* pts, const int32_t* elem,
                       const double* coords, const int32_t* en, const int32_t* ee,
                       const int32_t* edn, const int32_t* fe, const int32_t* dofs,
                       const double* x, double omega, double* fields, void* stream) {
  if (p < 1 || p > kMaxP) PG_FAIL(PG_ERR_UNSUPPORTED, "order p=%d not supported", p);
  if (npts <= 0 || !pts || !elem || !coords || !en || !ee || !edn || !fe || !dofs || !x ||
      !fields)
    PG_FAIL(PG_ERR_ARG, "bad arguments");
  const double om = omega * kMu0;
  hipStream_t s = as_stream(stream);
  const dim3 g(nblocks(npts, 64)), b(64);
  const double2* xx = reinterpret_cast<const double2*>(x);
  double2* f = reinterpret_cast<double2*>(fields);
  switch (p) {
    case 1: hipLaunchKernelGGL(k_receivers<1>, g, b, 0, s, npts, pts, elem, coords, en, ee, edn, fe, dofs, xx, om, f); break;
    case 2: hipLaunchKernelGGL(k_receivers<2>, g, b, 0, s, npts, pts, elem, coords, en, ee, edn, fe, dofs, xx, om, f); break;
    case 3: hipLaunchKernelGGL(k_receivers<3>, g, b, 0, s, npts, pts, elem, coords, en, ee, edn, fe, dofs, xx, om, f); break;
  }
  PG_CHECK_LAUNCH();
  return PG_OK;
}

int pg_adjoint_rhs(int p, int64_t npts, const double* pts, const int32_t* elem,
                   const double* coords, const int32_t* en, const int32_t* ee, const int32_t* edn,
                   const int32_t* fe, const int32_t* dofs, const double* coef, double* f,
                   void* stream) {
  if (p < 1 || p > kMaxP) PG_FAIL(PG_ERR_UNSUPPORTED, "order p=%d not supported", p);
  if (npts <= 0 || !pts || !elem || !coords || !en || !ee || !edn || !fe || !dofs || !coef || !f)
    PG_FAIL(PG_ERR_ARG, "bad arguments");
  hipStream_t s = as_stream(stream);
  const double2* c = reinterpret_cast<const double2*>(coef);
  double2* ff = reinterpret_cast<double2*>(f);
  switch (p) {
    case 1: hipLaunchKernelGGL(k_adjoint_rhs<1>, dim3(1), dim3(64), 0, s, npts, pts, elem, coords, en, ee, edn, fe, dofs, c, ff); break;
    case 2: hipLaunchKernelGGL(k_adjoint_rhs<2>, dim3(1), dim3(64), 0, s, npts, pts, elem, coords, en, ee, edn, fe, dofs, c, ff); break;
    case 3: hipLaunchKernelGGL(k_adjoint_rhs<3>, dim3(1), dim3(64), 0, s, npts, pts, elem, coords, en, ee, edn, fe, dofs, c, ff); break;
  }
  PG_CHECK_LAUNCH();
  return PG_OK;
}

// static_staging: stage omega*mu0 in the file-scope device array instead of a
// stream-ordered allocation (pg_adjoint_gradient_noalloc)
static int adjoint_gradient(const void* tables, int64_t E, const void* ws, const int32_t* dofs,
                            int nrhs, int64_t ndof, int64_t sd, int64_t sr, const double* x,
                            const double* lam, const double* omega, double* g, void* stream,
                            bool static_staging = false) {
  if (!tables || !ws || !dofs || nrhs <= 0 || ndof <= 0 || !x || !lam || !omega || !g || E <= 0)
    PG_FAIL(PG_ERR_ARG, "bad arguments");
  if (nrhs > kAdjointMaxRhs)
    PG_FAIL(PG_ERR_ARG, "at most %d right-hand sides per call (got %d)", kAdjointMaxRhs, nrhs);
  const Tables* t = static_cast<const Tables*>(tables);
  double om[kAdjointMaxRhs];
  for (int r = 0; r < nrhs; ++r) om[r] = omega[r] * kMu0;
  double* d_om = nullptr;
  hipStream_t s = as_stream(stream);
  if (static_staging) {
    const int rc = scratch_address(&d_om);
    if (rc != PG_OK) return rc;
  } else {
    PG_HIP(hipMallocAsync(reinterpret_cast<void**>(&d_om), nrhs * sizeof(double), s));
  }
  PG_HIP(hipMemcpyAsync(d_om, om, nrhs * sizeof(double), hipMemcpyHostToDevice, s));
  void* w = const_cast<void*>(ws);
  const dim3 gr(nblocks(E, 256)), bl(256);
  const double2* xx = reinterpret_cast<const double2*>(x);
  const double2* ll = reinterpret_cast<const double2*>(lam);
  switch (t->p) {
    case 1: hipLaunchKernelGGL(k_adjoint_grad<1>, gr, bl, 0, s, E, t->d_T, ws_coef(w), ws_orient(w, E), dofs, nrhs, sd, sr, xx, ll, d_om, g); break;
    case 2: hipLaunchKernelGGL(k_adjoint_grad<2>, gr, bl, 0, s, E, t->d_T, ws_coef(w), ws_orient(w, E), dofs, nrhs, sd, sr, xx, ll, d_om, g); break;
    case 3: hipLaunchKernelGGL(k_adjoint_grad<3>, gr, bl, 0, s, E, t->d_T, ws_coef(w), ws_orient(w, E), dofs, nrhs, sd, sr, xx, ll, d_om, g); break;
  }
  PG_CHECK_LAUNCH();
  // the host copy above is synchronous with respect to `om` only once the stream drains
  PG_HIP(hipStreamSynchronize(s));
  if (!static_staging) PG_HIP(hipFreeAsync(d_om, s));
  return PG_OK;
}

int pg_adjoint_gradient(const void* tables, int64_t E, const void* ws, const int32_t* dofs,
                        int nrhs, int64_t ndof, const double* x, const double* lam,
                        const double* omega, double* g, void* stream) {
  return adjoint_gradient(tables, E, ws, dofs, nrhs, ndof, 1, ndof, x, lam, omega, g, stream);
}

int pg_adjoint_gradient_rows(const void* tables, int64_t E, const void* ws, const int32_t* dofs,
                             int nrhs, int64_t ndof, const double* x, const double* lam,
                             const double* omega, double* g, void* stream) {
  return adjoint_gradient(tables, E, ws, dofs, nrhs, ndof, nrhs, 1, x, lam, omega, g, stream);
}

int pg_adjoint_gradient_noalloc(const void* tables, int64_t E, const void* ws,
                                const int32_t* dofs, int nrhs, int64_t ndof, const double* x,
                                const double* lam, const double* omega, double* g,
                                void* stream) {
  return adjoint_gradient(tables, E, ws, dofs, nrhs, ndof, 1, ndof, x, lam, omega, g, stream,
                          true);
}

}  // extern "C"
