// Assembly plan: the symbolic half of the global assembly (host, integers only).
//
// The reference preallocates a PETSc AIJ matrix with a constant row count and
// lets MatSetValues discover the pattern element by element
// (petgem/parallel.py:150-177, petgem/solver.py:186-235).  The pattern that
// results is the union of the element DOF cliques with sorted columns
// (explicit zeros kept).  Here it is built once, at the level of mesh
// entities: every DOF of an entity (edge, face, volume) has the same column
// set, so the CSR is a grid of entity-pair blocks of size d(A) x d(B).
//
// For every entity pair (A, B) the plan stores the CSR slot of DOF pair
// (A,0),(B,0), the row length of A's rows, and the list of (element, local
// entity of A, local entity of B) that contribute, in ascending element order
// -- the order in which the reference's single-rank ADD_VALUES sums them.  The
// numeric kernel then sums each slot in that fixed order with no atomics.
//
// Entity ids: edges [0, nE), faces [nE, nE+nF), volumes [nE+nF, nE+nF+E);
// DOF offsets follow petgem/hvfem.py:52-71 (edges, then faces, then volumes).
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "pg_common.h"
#include "pg_rowasm.h"

namespace pg {

struct Plan {
  int p;
  int64_t E, nE, nF, nEnt;
  int64_t ndof, nnz, npairs, ncontrib;
  int maxd;
  std::vector<int32_t> rowptr, cols, pair_slot0, pair_rowlen, pair_cbeg, pair_ab;
  std::vector<uint32_t> contrib;
  // first pair of every entity (nEnt + 1), and the largest row length, pair
  // count and element count of an entity with DOFs
  std::vector<int32_t> entity_pair;
  int64_t max_rowlen = 0, max_pairs = 0, max_inc = 0;
};

namespace {

struct Ctx {
  int p;
  int64_t E, nE, nF;
  const int32_t* ee;
  const int32_t* ef;
  int nloc;
  int loc[11];
  int d_edge, d_face, d_vol;

  int64_t gid(int64_t e, int la) const {
    if (la < 6) return ee[6 * e + la];
    if (la < 10) return nE + ef[4 * e + (la - 6)];
    return nE + nF + e;
  }
  int ndof(int64_t g) const { return g < nE ? d_edge : (g < nE + nF ? d_face : d_vol); }
  int64_t dof0(int64_t g) const {
    if (g < nE) return g * d_edge;
    if (g < nE + nF) return nE * d_edge + (g - nE) * d_face;
    return nE * d_edge + nF * d_face + (g - nE - nF) * d_vol;
  }
};

struct Item {
  int64_t B;
  int64_t e;
  uint8_t la, lb;
};

template <typename F>
void parallel_for(int64_t n, F f) {
  unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (n < 4096) nt = 1;
  std::vector<std::thread> th;
  const int64_t chunk = (n + nt - 1) / nt;
  for (unsigned t = 0; t < nt; ++t) {
    const int64_t a = t * chunk, b = std::min<int64_t>(n, a + chunk);
    if (a >= b) break;
    th.emplace_back([=, &f]() { f(a, b); });
  }
  for (auto& x : th) x.join();
}

}  // namespace
}  // namespace pg

using namespace pg;

extern "C" {

int pg_plan_build(int p, int64_t E, int64_t nE, int64_t nF, const int32_t* ee,
                  const int32_t* ef, void** out) {
  if (!out) PG_FAIL(PG_ERR_ARG, "out is NULL");
  if (p < 1 || p > 3) PG_FAIL(PG_ERR_UNSUPPORTED, "order p=%d not supported (1..3)", p);
  if (E <= 0 || nE <= 0 || nF <= 0 || !ee || !ef) PG_FAIL(PG_ERR_ARG, "bad mesh arguments");
  if (E >= (1LL << 24)) PG_FAIL(PG_ERR_UNSUPPORTED, "more than 2^24 elements");
  Ctx c;
  c.p = p;
  c.E = E;
  c.nE = nE;
  c.nF = nF;
  c.ee = ee;
  c.ef = ef;
  c.d_edge = p;
  c.d_face = p * (p - 1);
  c.d_vol = p * (p - 1) * (p - 2) / 2;
  c.nloc = 0;
  for (int la = 0; la < 6; ++la) c.loc[c.nloc++] = la;
  if (c.d_face)
    for (int la = 6; la < 10; ++la) c.loc[c.nloc++] = la;
  if (c.d_vol) c.loc[c.nloc++] = 10;
  for (int64_t i = 0; i < 6 * E; ++i)
    if (ee[i] < 0 || ee[i] >= nE) PG_FAIL(PG_ERR_ARG, "edge id out of range");
  for (int64_t i = 0; i < 4 * E; ++i)
    if (ef[i] < 0 || ef[i] >= nF) PG_FAIL(PG_ERR_ARG, "face id out of range");

  const int64_t nEnt = nE + nF + E;
  // entity -> (element, local entity) incidence, ascending element order
  std::vector<int64_t> iptr(nEnt + 1, 0);
  for (int64_t e = 0; e < E; ++e)
    for (int k = 0; k < c.nloc; ++k) iptr[c.gid(e, c.loc[k]) + 1]++;
  for (int64_t g = 0; g < nEnt; ++g) iptr[g + 1] += iptr[g];
  std::vector<int64_t> inc_e(iptr[nEnt]);
  std::vector<uint8_t> inc_l(iptr[nEnt]);
  {
    std::vector<int64_t> pos(iptr.begin(), iptr.end() - 1);
    for (int64_t e = 0; e < E; ++e)
      for (int k = 0; k < c.nloc; ++k) {
        const int64_t g = c.gid(e, c.loc[k]);
        inc_e[pos[g]] = e;
        inc_l[pos[g]] = (uint8_t)c.loc[k];
        pos[g]++;
      }
  }

  // pass 1: per entity, number of neighbour entities and row length
  std::vector<int64_t> nadj(nEnt, 0), rowlen(nEnt, 0);
  auto gather = [&](int64_t A, std::vector<Item>& items) {
    items.clear();
    for (int64_t s = iptr[A]; s < iptr[A + 1]; ++s) {
      const int64_t e = inc_e[s];
      for (int k = 0; k < c.nloc; ++k) {
        const int lb = c.loc[k];
        items.push_back(Item{c.gid(e, lb), e, inc_l[s], (uint8_t)lb});
      }
    }
    std::stable_sort(items.begin(), items.end(),
                     [](const Item& x, const Item& y) { return x.B < y.B; });
  };
  parallel_for(nEnt, [&](int64_t a, int64_t b) {
    std::vector<Item> items;
    for (int64_t A = a; A < b; ++A) {
      if (c.ndof(A) == 0) continue;
      gather(A, items);
      int64_t na = 0, len = 0;
      for (size_t i = 0; i < items.size(); ++i)
        if (i == 0 || items[i].B != items[i - 1].B) {
          na++;
          len += c.ndof(items[i].B);
        }
      nadj[A] = na;
      rowlen[A] = len;
    }
  });

  Plan* P = new Plan();
  P->p = p;
  P->E = E;
  P->nE = nE;
  P->nF = nF;
  P->nEnt = nEnt;
  P->maxd = std::max(c.d_edge, std::max(c.d_face, c.d_vol));
  P->ndof = c.dof0(nEnt - 1) + c.ndof(nEnt - 1);
  if (c.d_vol == 0) P->ndof = c.d_face ? nE * c.d_edge + nF * c.d_face : nE * c.d_edge;

  // prefix sums: pairs, contributions, nnz
  std::vector<int64_t> pbase(nEnt + 1, 0), zbase(nEnt + 1, 0), cbase(nEnt + 1, 0);
  for (int64_t A = 0; A < nEnt; ++A) {
    pbase[A + 1] = pbase[A] + nadj[A];
    zbase[A + 1] = zbase[A] + (int64_t)c.ndof(A) * rowlen[A];
    cbase[A + 1] = cbase[A] + (c.ndof(A) ? (iptr[A + 1] - iptr[A]) * c.nloc : 0);
  }
  P->npairs = pbase[nEnt];
  P->nnz = zbase[nEnt];
  for (int64_t A = 0; A < nEnt; ++A)
    if (c.ndof(A)) {
      P->max_rowlen = std::max(P->max_rowlen, rowlen[A]);
      P->max_pairs = std::max(P->max_pairs, nadj[A]);
      P->max_inc = std::max(P->max_inc, iptr[A + 1] - iptr[A]);
    }
  P->ncontrib = cbase[nEnt];
  if (P->nnz >= (1LL << 31) || P->ncontrib >= (1LL << 31) || P->ndof >= (1LL << 31)) {
    delete P;
    PG_FAIL(PG_ERR_UNSUPPORTED, "system too large for 32-bit CSR indices (nnz=%lld)",
            (long long)zbase[nEnt]);
  }
  P->rowptr.assign(P->ndof + 1, 0);
  P->cols.resize(P->nnz);
  P->pair_slot0.resize(P->npairs);
  P->pair_rowlen.resize(P->npairs);
  P->pair_cbeg.resize(P->npairs + 1);
  P->pair_ab.resize(2 * P->npairs);
  P->contrib.resize(P->ncontrib);
  P->pair_cbeg[P->npairs] = (int32_t)P->ncontrib;
  P->entity_pair.resize(nEnt + 1);
  for (int64_t A = 0; A <= nEnt; ++A) P->entity_pair[A] = (int32_t)pbase[A];

  // rowptr: DOF rows in entity order (DOF numbering is entity-major)
  for (int64_t A = 0; A < nEnt; ++A) {
    const int d = c.ndof(A);
    const int64_t r0 = c.dof0(A);
    for (int k = 0; k < d; ++k) P->rowptr[r0 + k] = (int32_t)(zbase[A] + k * rowlen[A]);
  }
  P->rowptr[P->ndof] = (int32_t)P->nnz;

  // pass 2: fill pairs, contributions and columns
  parallel_for(nEnt, [&](int64_t a, int64_t b) {
    std::vector<Item> items;
    for (int64_t A = a; A < b; ++A) {
      const int dA = c.ndof(A);
      if (dA == 0) continue;
      gather(A, items);
      int64_t pi = pbase[A], ci = cbase[A], off = 0;
      std::vector<int32_t> rowcols;
      rowcols.reserve(rowlen[A]);
      for (size_t i = 0; i < items.size(); ++i) {
        const Item& it = items[i];
        if (i == 0 || it.B != items[i - 1].B) {
          if (i != 0) off += c.ndof(items[i - 1].B);
          P->pair_slot0[pi] = (int32_t)(zbase[A] + off);
          P->pair_rowlen[pi] = (int32_t)rowlen[A];
          P->pair_cbeg[pi] = (int32_t)ci;
          P->pair_ab[2 * pi] = (int32_t)A;
          P->pair_ab[2 * pi + 1] = (int32_t)it.B;
          const int64_t b0 = c.dof0(it.B);
          for (int l = 0; l < c.ndof(it.B); ++l) rowcols.push_back((int32_t)(b0 + l));
          pi++;
        }
        P->contrib[ci++] = ((uint32_t)it.e << 8) | ((uint32_t)it.la << 4) | (uint32_t)it.lb;
      }
      for (int k = 0; k < dA; ++k)
        std::memcpy(&P->cols[zbase[A] + k * rowlen[A]], rowcols.data(),
                    rowcols.size() * sizeof(int32_t));
    }
  });
  *out = P;
  return PG_OK;
}

int pg_plan_info(const void* plan, int64_t* info) {
  if (!plan || !info) PG_FAIL(PG_ERR_ARG, "NULL argument");
  const Plan* P = static_cast<const Plan*>(plan);
  info[0] = P->ndof;
  info[1] = P->nnz;
  info[2] = P->npairs;
  info[3] = P->ncontrib;
  info[4] = P->nEnt;
  info[5] = P->maxd;
  return PG_OK;
}

int pg_plan_export(const void* plan, int32_t* rowptr, int32_t* cols, int32_t* slot0,
                   int32_t* rowlen, int32_t* cbeg, int32_t* ab, uint32_t* contrib) {
  if (!plan) PG_FAIL(PG_ERR_ARG, "plan is NULL");
  const Plan* P = static_cast<const Plan*>(plan);
  auto cp = [](void* dst, const void* src, size_t bytes) {
    if (dst) std::memcpy(dst, src, bytes);
  };
  cp(rowptr, P->rowptr.data(), P->rowptr.size() * 4);
  cp(cols, P->cols.data(), P->cols.size() * 4);
  cp(slot0, P->pair_slot0.data(), P->pair_slot0.size() * 4);
  cp(rowlen, P->pair_rowlen.data(), P->pair_rowlen.size() * 4);
  cp(cbeg, P->pair_cbeg.data(), P->pair_cbeg.size() * 4);
  cp(ab, P->pair_ab.data(), P->pair_ab.size() * 4);
  cp(contrib, P->contrib.data(), P->contrib.size() * 4);
  return PG_OK;
}

int pg_plan_entity_pairs(const void* plan, int32_t* entity_pair, int32_t* max_rowlen) {
  if (!plan || !entity_pair) PG_FAIL(PG_ERR_ARG, "NULL argument");
  const Plan* P = static_cast<const Plan*>(plan);
  if (P->max_rowlen > kRowAsmMaxLen || P->max_pairs > kRowAsmMaxPairs ||
      P->max_inc > kRowAsmMaxInc)
    PG_FAIL(PG_ERR_UNSUPPORTED,
            "entity rows beyond the row-major assembly bounds (row length %lld of %d, "
            "%lld coupled entities of %d, %lld elements of %d)",
            (long long)P->max_rowlen, kRowAsmMaxLen, (long long)P->max_pairs, kRowAsmMaxPairs,
            (long long)P->max_inc, kRowAsmMaxInc);
  std::memcpy(entity_pair, P->entity_pair.data(), P->entity_pair.size() * sizeof(int32_t));
  if (max_rowlen) *max_rowlen = (int32_t)P->max_rowlen;
  return PG_OK;
}

int pg_plan_destroy(void* plan) {
  delete static_cast<Plan*>(plan);
  return PG_OK;
}

}  // extern "C"
