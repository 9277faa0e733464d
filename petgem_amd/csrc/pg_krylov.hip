// Complex CSR SpMV and the Krylov solve for gfx950.
//
// Replaces PETSc's MatMult and KSPSolve as called from Solver.run
// (petgem/solver.py:584-590).  The system matrix is complex symmetric
// (A = A^T: K and M are symmetric), so the default method is COCG
// (conjugate orthogonal CG, van der Vorst & Melissen 1990) with point or
// entity-block Jacobi; GMRES(m) is provided as well.  The stopping test is the
// true-norm residual ||b - A x||_2 / ||b||_2 (PETSc's GMRES default is the
// preconditioned residual, which is weaker).
//
// Device organisation:
//  * one HIP stream; a COCG iteration is
//      p = z + beta p | q = A p (+ partials of p^T q) | finish mu, alpha |
//      x += alpha p, r -= alpha q, z = P r (+ partials of r^T z, |r|^2) |
//      finish rho, beta, convergence
//    with every scalar in device memory;
//  * reductions: each workgroup writes its partial sums with plain stores and
//    a one-workgroup "finish" kernel adds them in a fixed order (bitwise
//    reproducible).  The kernel boundary makes the partials visible, so no
//    workgroup issues a device-scope fence (those force L2 write-backs
//    across the XCDs and cost ~0.2 ms per kernel at this size);
//  * every kernel returns immediately once the device-side `done` flag is
//    set, so the host enqueues batches of iterations and reads the flags only
//    between batches;
//  * each XCD owns a contiguous eighth of the rows; in the SpMV its
//    workgroups sweep that eighth together (row-interleaved) so that the
//    gathers of x stay in the XCD's L2, in the vector kernels each workgroup
//    streams one contiguous chunk;
//  * SpMV kernels (chosen per matrix by timing): "row" = G lanes per CSR row,
//    "entity-row" = G lanes per mesh entity whose d rows share one column
//    list (each index and gathered x value read once for d rows); each lane
//    issues U independent value/index/gather loads before any FMA.
//
// Mixed precision (PG_PREC_MIXED): iterative refinement.  The Krylov
// recursion, SpMV and preconditioner run in complex64 on a complex64 copy of
// A; the correction is accumulated into the FP64 solution and the FP64
// residual b - A x is recomputed whenever the complex64 residual has dropped
// by `inner_rtol` since the last refinement, after which the complex64 COCG
// restarts from that residual.  Dot products are always accumulated in FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "pg_blockinv.h"
#include "pg_common.h"
#include "pg_eop.h"

namespace pg {

constexpr int kBS = 256;
constexpr int kMaxGrid = 2048;
constexpr int kMaxBlock = 8;
constexpr int kMaxRestart = 128;

struct Scal {
  double2 rho, mu, alpha, beta;
  double res2, bnorm2, tol2, ref2;
  double res64;
  int32_t iter, maxit, done, status;
  // stagnation watch of the complex64 recursion (mixed precision): best |r|^2
  // of the current round and the iteration of its last 10 % improvement
  double best2;
  int32_t last_imp, stall_limit, stalled, pad;
};

// a complex64 round whose best residual norm has not improved by 10 % for
// this many iterations (default) has stagnated; the solve then continues in FP64
constexpr int kStallIters = 400;

// ---------------------------------------------------------------------------
// complex helpers
// ---------------------------------------------------------------------------
__device__ inline double2 to_d(double2 v) { return v; }
__device__ inline double2 to_d(float2 v) { return make_double2(v.x, v.y); }
template <typename T>
__device__ inline T from_d(double2 v);
template <>
__device__ inline double2 from_d<double2>(double2 v) { return v; }
template <>
__device__ inline float2 from_d<float2>(double2 v) { return make_float2((float)v.x, (float)v.y); }

__device__ inline double2 cm(double2 a, double2 b) { return zmul(a, b); }
__device__ inline float2 cm(float2 a, float2 b) { return cmul(a, b); }
__device__ inline double2 cfma(double2 a, double2 b, double2 c) {
  return make_double2(fma(a.x, b.x, fma(-a.y, b.y, c.x)), fma(a.x, b.y, fma(a.y, b.x, c.y)));
}
__device__ inline float2 cfma(float2 a, float2 b, float2 c) {
  return make_float2(fmaf(a.x, b.x, fmaf(-a.y, b.y, c.x)), fmaf(a.x, b.y, fmaf(a.y, b.x, c.y)));
}
template <typename T>
__device__ inline T czero();
template <>
__device__ inline double2 czero<double2>() { return make_double2(0.0, 0.0); }
template <>
__device__ inline float2 czero<float2>() { return make_float2(0.f, 0.f); }

// ---------------------------------------------------------------------------
// work distribution and reductions
// ---------------------------------------------------------------------------
__device__ inline int64_t logical_block() {
  const int64_t b = blockIdx.x, nb = gridDim.x;
  if (nb % 8 != 0) return b;
  return (b % 8) * (nb / 8) + b / 8;
}
__device__ inline void chunk(int64_t n, int64_t unit, int64_t* lo, int64_t* hi) {
  const int64_t nb = gridDim.x;
  const int64_t units = (n + unit - 1) / unit;
  const int64_t per = (units + nb - 1) / nb;
  const int64_t lb = logical_block();
  const int64_t a = lb * per * unit, b = (lb + 1) * per * unit;
  *lo = a < n ? a : n;
  *hi = b < n ? b : n;
}

// Split for the SpMV, whose gathers of x have spatial reuse: XCD x (workgroup
// b runs on XCD b % 8, the round-robin dispatch order) owns the x-th eighth of
// the rows, and that XCD's workgroups sweep it together, interleaved by
// `per` rows, so that at any moment the XCD's gathers fall in a narrow moving
// window of x that stays in its L2.  (Contiguous per-workgroup chunks spread
// the XCD's concurrent workgroups over its whole eighth and re-fetch x from
// memory many times over.)
struct Sweep {
  int64_t first, end, stride;
};
__device__ inline Sweep xcd_sweep(int64_t n, int per) {
  const int64_t b = blockIdx.x, nb = gridDim.x;
  if (nb % 8 != 0) return {b * per, n, nb * per};
  const int64_t x = b % 8, nbx = nb / 8;
  return {n * x / 8 + (b / 8) * per, n * (x + 1) / 8, nbx * per};
}

// block reduction of NV doubles; result valid in thread 0
template <int NV>
__device__ inline void block_reduce(double (&v)[NV]) {
  __shared__ double sm[NV][kBS / kWave];
  const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = wave_sum(v[k]);
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) sm[k][w] = v[k];
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double s = 0.0;
      for (int i = 0; i < kBS / kWave; ++i) s += sm[k][i];
      v[k] = s;
    }
  __syncthreads();
}

// this workgroup's partial sums -> part[k * gridDim.x + blockIdx.x]
template <int NV>
__device__ inline void write_partials(double (&v)[NV], double* part) {
  block_reduce<NV>(v);
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) part[(int64_t)k * gridDim.x + blockIdx.x] = v[k];
}

// in a one-workgroup finish kernel: fixed-order totals of np partials; all threads
template <int NV>
__device__ inline void sum_partials(const double* part, int np, double (&tot)[NV]) {
  double acc[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) acc[k] = 0.0;
  for (int i = threadIdx.x; i < np; i += kBS)
#pragma unroll
    for (int k = 0; k < NV; ++k) acc[k] += part[(int64_t)k * np + i];
  block_reduce<NV>(acc);
  __shared__ double res[NV];
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) res[k] = acc[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) tot[k] = res[k];
}

// Matrix loads.  (Non-temporal loads were measured slower here: without L2
// allocation, the neighbouring halves of a cache line that other load
// instructions of the same wave use are fetched from memory again.)
template <typename T>
__device__ inline T ld_stream(const T* p) {
  return *p;
}

// ---------------------------------------------------------------------------
// SpMV.  MODE 0: y = A x.  MODE 1: y = A x and partials of x^T y (COCG mu).
// MODE 2: y = b - A x and partials of |y|^2 (true residual).
// ---------------------------------------------------------------------------
template <int MODE, typename XT>
__device__ inline void spmv_row_out(int64_t row, XT acc, const XT* __restrict__ x,
                                    XT* __restrict__ y, const double2* __restrict__ b,
                                    double (&red)[2]) {
  if (MODE == 2) {
    const double2 bb = b[row];
    const double2 rr = make_double2(bb.x - (double)acc.x, bb.y - (double)acc.y);
    y[row] = from_d<XT>(rr);
    red[0] += rr.x * rr.x + rr.y * rr.y;
  } else {
    y[row] = acc;
    if (MODE == 1) {
      const double2 pv = to_d(x[row]), qv = to_d(acc);
      red[0] += pv.x * qv.x - pv.y * qv.y;
      red[1] += pv.x * qv.y + pv.y * qv.x;
    }
  }
}

// G lanes per CSR row; the next row's extent is fetched before this row's work
template <int G, int U, int MODE, typename VT, typename XT>
__global__ void __launch_bounds__(kBS) k_spmv(const Scal* __restrict__ s, int64_t n,
                                              const int32_t* __restrict__ rowptr,
                                              const int32_t* __restrict__ cols,
                                              const VT* __restrict__ vals,
                                              const XT* __restrict__ x, XT* __restrict__ y,
                                              const double2* __restrict__ b, double* part) {
  if (MODE != 0 && s->done) return;
  constexpr int RPB = kBS / G;
  const Sweep sw = xcd_sweep(n, RPB);
  const int lane = threadIdx.x % G;
  double red[2] = {0.0, 0.0};
  int64_t row = sw.first + threadIdx.x / G;
  int s0n = 0, s1n = 0;
  if (row < sw.end) {
    s0n = rowptr[row];
    s1n = rowptr[row + 1];
  }
  for (; row < sw.end; row += sw.stride) {
    const int s0 = s0n, s1 = s1n;
    const int64_t nx = row + sw.stride;
    if (nx < sw.end) {
      s0n = rowptr[nx];
      s1n = rowptr[nx + 1];
    }
    XT acc = czero<XT>();
    for (int kb = s0 + lane; kb < s1; kb += G * U) {
      int c[U];
      VT a[U];
      XT xv[U];
#pragma unroll
      for (int j = 0; j < U; ++j) {
        const int k = kb + j * G;
        c[j] = k < s1 ? ld_stream(cols + k) : -1;
        a[j] = k < s1 ? ld_stream(vals + k) : VT{0, 0};
      }
#pragma unroll
      for (int j = 0; j < U; ++j) xv[j] = c[j] >= 0 ? x[c[j]] : czero<XT>();
#pragma unroll
      for (int j = 0; j < U; ++j) acc = cfma(XT{a[j].x, a[j].y}, xv[j], acc);
    }
    acc.x = group_sum<G>(acc.x);
    acc.y = group_sum<G>(acc.y);
    if (lane == 0) spmv_row_out<MODE, XT>(row, acc, x, y, b, red);
  }
  if (MODE != 0) write_partials<2>(red, part);
}

// Entity-row SpMV over the solver's entity schedule.  The d rows of an entity
// (DOFs r0 .. r0+d) share one column list of length L; the solver keeps the
// column lists and the d x L value blocks in schedule order (cols_s from
// cbase, vals_s from vbase), so both stream sequentially while the gathered x
// stays spatially local.  em[i] = {r0, d | L << 8, cbase, vbase} (one 16-byte
// load, fetched one entity ahead).
template <int G, int U, int D, int MODE, typename VT, typename XT>
__global__ void __launch_bounds__(kBS) k_spmv_er(const Scal* __restrict__ s, int64_t nent,
                                                 const int4* __restrict__ emeta,
                                                 const int32_t* __restrict__ cols,
                                                 const VT* __restrict__ vals,
                                                 const XT* __restrict__ x, XT* __restrict__ y,
                                                 const double2* __restrict__ b, double* part) {
  if (MODE != 0 && s->done) return;
  constexpr int EPB = kBS / G;
  const Sweep sw = xcd_sweep(nent, EPB);
  const int lane = threadIdx.x % G;
  double red[2] = {0.0, 0.0};
  int64_t ent = sw.first + threadIdx.x / G;
  int4 mn = make_int4(0, 0, 0, 0);
  if (ent < sw.end) mn = emeta[ent];
  for (; ent < sw.end; ent += sw.stride) {
    const int r0 = mn.x, d = mn.y & 0xff, L = mn.y >> 8, cbase = mn.z, vbase = mn.w;
    const int64_t nx = ent + sw.stride;
    if (nx < sw.end) mn = emeta[nx];
    XT acc[D];
#pragma unroll
    for (int k = 0; k < D; ++k) acc[k] = czero<XT>();
    for (int tb = lane; tb < L; tb += G * U) {
      int c[U];
      XT xv[U];
#pragma unroll
      for (int j = 0; j < U; ++j) {
        const int t = tb + j * G;
        c[j] = t < L ? ld_stream(cols + cbase + t) : -1;
      }
#pragma unroll
      for (int j = 0; j < U; ++j) xv[j] = c[j] >= 0 ? x[c[j]] : czero<XT>();
#pragma unroll
      for (int k = 0; k < D; ++k)
        if (k < d)
#pragma unroll
          for (int j = 0; j < U; ++j) {
            const int t = tb + j * G;
            const VT a = t < L ? ld_stream(vals + vbase + k * L + t) : VT{0, 0};
            acc[k] = cfma(XT{a.x, a.y}, xv[j], acc[k]);
          }
    }
#pragma unroll
    for (int k = 0; k < D; ++k) {
      acc[k].x = group_sum<G>(acc[k].x);
      acc[k].y = group_sum<G>(acc[k].y);
    }
    if (lane == 0) {
#pragma unroll
      for (int k = 0; k < D; ++k)
        if (k < d) spmv_row_out<MODE, XT>(r0 + k, acc[k], x, y, b, red);
    }
  }
  if (MODE != 0) write_partials<2>(red, part);
}

// Entity-row SpMV for meshes whose entities all carry two DOFs (p = 2): every
// entity has d = 2 rows and its columns come in aligned pairs (c, c+1) of one
// column entity, so each lane loads one column index, one 2-value slice of x
// and one 2-value slice of each row per column pair (16-byte loads for
// complex64).  cols2 holds the first column of each pair (cbase / 2 onwards).
template <typename T>
struct PairOf;
template <>
struct PairOf<float2> {
  using type = float4;
};
template <>
struct PairOf<double2> {
  using type = double4;
};

template <int G, int U, int MODE, typename VT, typename XT>
__global__ void __launch_bounds__(kBS) k_spmv_er2(const Scal* __restrict__ s, int64_t nent,
                                                  const int4* __restrict__ emeta,
                                                  const int32_t* __restrict__ cols2,
                                                  const VT* __restrict__ vals,
                                                  const XT* __restrict__ x, XT* __restrict__ y,
                                                  const double2* __restrict__ b, double* part) {
  using VP = typename PairOf<VT>::type;
  using XP = typename PairOf<XT>::type;
  if (MODE != 0 && s->done) return;
  constexpr int EPB = kBS / G;
  const Sweep sw = xcd_sweep(nent, EPB);
  const int lane = threadIdx.x % G;
  double red[2] = {0.0, 0.0};
  int64_t ent = sw.first + threadIdx.x / G;
  int4 mn = make_int4(0, 0, 0, 0);
  if (ent < sw.end) mn = emeta[ent];
  for (; ent < sw.end; ent += sw.stride) {
    const int r0 = mn.x, L = mn.y >> 8, c2 = mn.z >> 1, vbase = mn.w;
    const int L2 = L >> 1;
    const int64_t nx = ent + sw.stride;
    if (nx < sw.end) mn = emeta[nx];
    const VP* v0 = reinterpret_cast<const VP*>(vals + vbase);
    const VP* v1 = reinterpret_cast<const VP*>(vals + vbase + L);
    XT acc0 = czero<XT>(), acc1 = czero<XT>();
    for (int tb = lane; tb < L2; tb += G * U) {
      int c[U];
      XP xv[U];
#pragma unroll
      for (int j = 0; j < U; ++j) {
        const int t = tb + j * G;
        c[j] = t < L2 ? cols2[c2 + t] : -1;
      }
#pragma unroll
      for (int j = 0; j < U; ++j)
        xv[j] = c[j] >= 0 ? *reinterpret_cast<const XP*>(x + c[j]) : XP{0, 0, 0, 0};
#pragma unroll
      for (int j = 0; j < U; ++j) {
        const int t = tb + j * G;
        const VP a = t < L2 ? v0[t] : VP{0, 0, 0, 0};
        const VP e = t < L2 ? v1[t] : VP{0, 0, 0, 0};
        acc0 = cfma(XT{a.x, a.y}, XT{xv[j].x, xv[j].y}, acc0);
        acc0 = cfma(XT{a.z, a.w}, XT{xv[j].z, xv[j].w}, acc0);
        acc1 = cfma(XT{e.x, e.y}, XT{xv[j].x, xv[j].y}, acc1);
        acc1 = cfma(XT{e.z, e.w}, XT{xv[j].z, xv[j].w}, acc1);
      }
    }
    acc0.x = group_sum<G>(acc0.x);
    acc0.y = group_sum<G>(acc0.y);
    acc1.x = group_sum<G>(acc1.x);
    acc1.y = group_sum<G>(acc1.y);
    if (lane == 0) {
      spmv_row_out<MODE, XT>(r0, acc0, x, y, b, red);
      spmv_row_out<MODE, XT>(r0 + 1, acc1, x, y, b, red);
    }
  }
  if (MODE != 0) write_partials<2>(red, part);
}

// column pairs: all entities two rows, every column list made of aligned pairs
__global__ void k_check_pairs(int64_t nent, const int4* __restrict__ em,
                              const int32_t* __restrict__ cols_s, int32_t* bad) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nent) return;
  const int4 m = em[i];
  const int d = m.y & 0xff, L = m.y >> 8;
  if (d != 2 || (L & 1) || (m.z & 1) || (m.w & 1)) {
    bad[0] = 1;
    return;
  }
  for (int j = 0; j < L; j += 2) {
    const int c = cols_s[m.z + j];
    if ((c & 1) || cols_s[m.z + j + 1] != c + 1) {
      bad[0] = 1;
      return;
    }
  }
}
__global__ void k_make_pairs(int64_t n2, const int32_t* __restrict__ cols_s,
                             int32_t* __restrict__ cols2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n2) cols2[i] = cols_s[2 * i];
}

// scheduled copies: entity i's column list (L entries from CSR slot src[i]) and
// its d x L value block (contiguous in the CSR) into the schedule-order arrays
__global__ void __launch_bounds__(kBS) k_sched_cols(int64_t nent, const int4* __restrict__ em,
                                                    const int32_t* __restrict__ src,
                                                    const int32_t* __restrict__ cols,
                                                    int32_t* __restrict__ cols_s) {
  constexpr int G = 32;
  const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G;
  if (i >= nent) return;
  const int4 m = em[i];
  const int L = m.y >> 8, base = src[i];
  for (int t = threadIdx.x % G; t < L; t += G) cols_s[m.z + t] = cols[base + t];
}
template <typename VT>
__global__ void __launch_bounds__(kBS) k_sched_vals(int64_t nent, const int4* __restrict__ em,
                                                    const int32_t* __restrict__ src,
                                                    const double2* __restrict__ vals,
                                                    VT* __restrict__ vals_s) {
  constexpr int G = 32;
  const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G;
  if (i >= nent) return;
  const int4 m = em[i];
  const int len = (m.y & 0xff) * (m.y >> 8), base = src[i];
  for (int t = threadIdx.x % G; t < len; t += G) vals_s[m.w + t] = from_d<VT>(vals[base + t]);
}

// finish mu = p^T q, alpha = rho / mu
__global__ void __launch_bounds__(kBS) k_fin_mu(Scal* __restrict__ s, const double* part,
                                                int np) {
  if (s->done) return;
  double tot[2];
  sum_partials<2>(part, np, tot);
  if (threadIdx.x != 0) return;
  const double2 mu = make_double2(tot[0], tot[1]);
  s->mu = mu;
  const double2 al = zdiv(s->rho, mu);
  if (!(mu.x != 0.0 || mu.y != 0.0) || !isfinite(al.x) || !isfinite(al.y)) {
    s->status = PG_ERR_BREAKDOWN;
    s->done = 1;
    s->alpha = make_double2(0.0, 0.0);
  } else {
    s->alpha = al;
  }
}

// finish |b - A x|^2
__global__ void __launch_bounds__(kBS) k_fin_res64(Scal* __restrict__ s, const double* part,
                                                   int np) {
  if (s->done) return;
  double tot[1];
  sum_partials<1>(part, np, tot);
  if (threadIdx.x == 0) s->res64 = tot[0];
}

// ---------------------------------------------------------------------------
// COCG vector kernels
// ---------------------------------------------------------------------------
// r -= alpha q; z = Dinv r; partials of r^T z and |r|^2.  The step x += alpha p
// of the same iteration is taken by the next k_pupdate (which reads p anyway)
// or, once the recursion pauses, by k_xstep.
template <typename XT>
__global__ void __launch_bounds__(kBS) k_update(const Scal* __restrict__ s, int64_t n,
                                                XT* __restrict__ r, const XT* __restrict__ q,
                                                const XT* __restrict__ dinv,
                                                XT* __restrict__ z, double* part) {
  if (s->done) return;
  const double2 al = s->alpha;
  const XT na = from_d<XT>(make_double2(-al.x, -al.y));
  int64_t lo, hi;
  chunk(n, kBS, &lo, &hi);
  double red[3] = {0.0, 0.0, 0.0};
  for (int64_t i = lo + threadIdx.x; i < hi; i += kBS) {
    const XT ri = cfma(na, q[i], r[i]);
    r[i] = ri;
    const XT zi = cm(dinv[i], ri);
    z[i] = zi;
    const double2 rd = to_d(ri), zd = to_d(zi);
    red[0] += rd.x * zd.x - rd.y * zd.y;
    red[1] += rd.x * zd.y + rd.y * zd.x;
    red[2] += rd.x * rd.x + rd.y * rd.y;
  }
  write_partials<3>(red, part);
}

// same with the entity-block preconditioner: one thread per block
template <typename XT>
__global__ void __launch_bounds__(kBS) k_update_blk(
    const Scal* __restrict__ s, int64_t nb, const int32_t* __restrict__ bptr,
    const int32_t* __restrict__ boff, XT* __restrict__ r, const XT* __restrict__ q,
    const XT* __restrict__ binv, XT* __restrict__ z, double* part) {
  if (s->done) return;
  const double2 al = s->alpha;
  const XT na = from_d<XT>(make_double2(-al.x, -al.y));
  int64_t lo, hi;
  chunk(nb, kBS, &lo, &hi);
  double red[3] = {0.0, 0.0, 0.0};
  for (int64_t k = lo + threadIdx.x; k < hi; k += kBS) {
    const int i0 = bptr[k], d = bptr[k + 1] - i0;
    const XT* B = binv + boff[k];
    XT rl[kMaxBlock];
#pragma unroll
    for (int t = 0; t < kMaxBlock; ++t)
      if (t < d) {
        const int i = i0 + t;
        rl[t] = cfma(na, q[i], r[i]);
        r[i] = rl[t];
      }
#pragma unroll
    for (int t = 0; t < kMaxBlock; ++t)
      if (t < d) {
        XT zt = czero<XT>();
#pragma unroll
        for (int u = 0; u < kMaxBlock; ++u)
          if (u < d) zt = cfma(B[t * d + u], rl[u], zt);
        z[i0 + t] = zt;
        const double2 rd = to_d(rl[t]), zd = to_d(zt);
        red[0] += rd.x * zd.x - rd.y * zd.y;
        red[1] += rd.x * zd.y + rd.y * zd.x;
        red[2] += rd.x * rd.x + rd.y * rd.y;
      }
  }
  write_partials<3>(red, part);
}

// The same update when every block has two DOFs (p = 2: block k is DOFs 2k,
// 2k+1 and its inverse the four entries at 4k): one 16-byte (complex64) or
// 32-byte (complex128) access per vector and block, and no block index reads.
template <typename XT>
__global__ void __launch_bounds__(kBS) k_update_blk2(const Scal* __restrict__ s, int64_t nb,
                                                     XT* __restrict__ r,
                                                     const XT* __restrict__ q,
                                                     const XT* __restrict__ binv,
                                                     XT* __restrict__ z, double* part) {
  using V = typename PairOf<XT>::type;
  if (s->done) return;
  const double2 al = s->alpha;
  const XT na = from_d<XT>(make_double2(-al.x, -al.y));
  int64_t lo, hi;
  chunk(nb, kBS, &lo, &hi);
  double red[3] = {0.0, 0.0, 0.0};
  V* r2 = reinterpret_cast<V*>(r);
  V* z2 = reinterpret_cast<V*>(z);
  const V* q2 = reinterpret_cast<const V*>(q);
  const V* b2 = reinterpret_cast<const V*>(binv);
  for (int64_t k = lo + threadIdx.x; k < hi; k += kBS) {
    const V rv = r2[k], qv = q2[k], B0 = b2[2 * k], B1 = b2[2 * k + 1];
    const XT ra = cfma(na, XT{qv.x, qv.y}, XT{rv.x, rv.y});
    const XT rb = cfma(na, XT{qv.z, qv.w}, XT{rv.z, rv.w});
    r2[k] = V{ra.x, ra.y, rb.x, rb.y};
    // z = B r, B row-major 2 x 2: B0 = (B00, B01), B1 = (B10, B11)
    const XT za = cfma(XT{B0.z, B0.w}, rb, cfma(XT{B0.x, B0.y}, ra, czero<XT>()));
    const XT zb = cfma(XT{B1.z, B1.w}, rb, cfma(XT{B1.x, B1.y}, ra, czero<XT>()));
    z2[k] = V{za.x, za.y, zb.x, zb.y};
    const double2 r0 = to_d(ra), r1 = to_d(rb), z0 = to_d(za), z1 = to_d(zb);
    red[0] += r0.x * z0.x - r0.y * z0.y;
    red[1] += r0.x * z0.y + r0.y * z0.x;
    red[2] += r0.x * r0.x + r0.y * r0.y;
    red[0] += r1.x * z1.x - r1.y * z1.y;
    red[1] += r1.x * z1.y + r1.y * z1.x;
    red[2] += r1.x * r1.x + r1.y * r1.y;
  }
  write_partials<3>(red, part);
}

// finish rho, beta, |r|^2 and the convergence / pause decision
__global__ void __launch_bounds__(kBS) k_fin_update(Scal* __restrict__ s, const double* part,
                                                    int np) {
  if (s->done) return;
  double tot[3];
  sum_partials<3>(part, np, tot);
  if (threadIdx.x != 0) return;
  const double2 rho_new = make_double2(tot[0], tot[1]);
  const double2 beta = zdiv(rho_new, s->rho);
  s->beta = beta;
  s->rho = rho_new;
  s->res2 = tot[2];
  s->iter += 1;
  if (!isfinite(tot[2]) || !isfinite(beta.x) || !isfinite(beta.y)) {
    s->status = PG_ERR_BREAKDOWN;
    s->done = 1;
  } else if (tot[2] <= s->tol2 * s->bnorm2 || tot[2] <= s->ref2 || s->iter >= s->maxit) {
    s->done = 1;  // converged recursion, refinement point, or iteration limit
  } else if (s->stall_limit > 0) {
    if (tot[2] < 0.81 * s->best2) {
      s->best2 = tot[2];
      s->last_imp = s->iter;
    } else if (s->iter - s->last_imp >= s->stall_limit) {
      s->stalled = 1;  // stagnated round
      s->done = 1;
    }
  }
}

// x += alpha p (the previous iteration's step: alpha is still that iteration's,
// and after a (re)start p = 0), then p = z + beta p
template <typename XT>
__global__ void __launch_bounds__(kBS) k_pupdate(const Scal* __restrict__ s, int64_t n,
                                                 const XT* __restrict__ z, XT* __restrict__ p,
                                                 XT* __restrict__ x) {
  if (s->done) return;
  const XT a = from_d<XT>(s->alpha);
  const XT be = from_d<XT>(s->beta);
  int64_t lo, hi;
  chunk(n, kBS, &lo, &hi);
  for (int64_t i = lo + threadIdx.x; i < hi; i += kBS) {
    const XT pv = p[i];
    x[i] = cfma(a, pv, x[i]);
    p[i] = cfma(be, pv, z[i]);
  }
}

// x += alpha p: the step of the last iteration before the recursion paused
// (alpha is 0 after a breakdown)
template <typename XT>
__global__ void k_xstep(const Scal* __restrict__ s, int64_t n, const XT* __restrict__ p,
                        XT* __restrict__ x) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const XT a = from_d<XT>(s->alpha);
  x[i] = cfma(a, p[i], x[i]);
}

// (re)start: r = r64 (demoted), z = P r, p = 0; partials of r^T z, |r|^2, |b|^2
template <typename XT>
__global__ void __launch_bounds__(kBS) k_reset(int64_t n, const double2* __restrict__ r64,
                                               XT* __restrict__ r, const XT* __restrict__ dinv,
                                               XT* __restrict__ z, XT* __restrict__ p,
                                               const double2* __restrict__ b, double* part) {
  int64_t lo, hi;
  chunk(n, kBS, &lo, &hi);
  double red[4] = {0.0, 0.0, 0.0, 0.0};
  for (int64_t i = lo + threadIdx.x; i < hi; i += kBS) {
    const XT ri = from_d<XT>(r64[i]);
    r[i] = ri;
    const XT zi = cm(dinv[i], ri);
    z[i] = zi;
    p[i] = czero<XT>();
    const double2 rd = to_d(ri), zd = to_d(zi);
    red[0] += rd.x * zd.x - rd.y * zd.y;
    red[1] += rd.x * zd.y + rd.y * zd.x;
    red[2] += rd.x * rd.x + rd.y * rd.y;
    if (b) {
      const double2 bb = b[i];
      red[3] += bb.x * bb.x + bb.y * bb.y;
    }
  }
  write_partials<4>(red, part);
}

template <typename XT>
__global__ void __launch_bounds__(kBS) k_reset_blk(int64_t nb, const int32_t* __restrict__ bptr,
                                                   const int32_t* __restrict__ boff,
                                                   const double2* __restrict__ r64,
                                                   XT* __restrict__ r,
                                                   const XT* __restrict__ binv,
                                                   XT* __restrict__ z, XT* __restrict__ p,
                                                   const double2* __restrict__ b, double* part) {
  int64_t lo, hi;
  chunk(nb, kBS, &lo, &hi);
  double red[4] = {0.0, 0.0, 0.0, 0.0};
  for (int64_t k = lo + threadIdx.x; k < hi; k += kBS) {
    const int i0 = bptr[k], d = bptr[k + 1] - i0;
    const XT* B = binv + boff[k];
    XT rl[kMaxBlock];
#pragma unroll
    for (int t = 0; t < kMaxBlock; ++t)
      if (t < d) {
        const int i = i0 + t;
        rl[t] = from_d<XT>(r64[i]);
        r[i] = rl[t];
        p[i] = czero<XT>();
        if (b) {
          const double2 bb = b[i];
          red[3] += bb.x * bb.x + bb.y * bb.y;
        }
      }
#pragma unroll
    for (int t = 0; t < kMaxBlock; ++t)
      if (t < d) {
        XT zt = czero<XT>();
#pragma unroll
        for (int u = 0; u < kMaxBlock; ++u)
          if (u < d) zt = cfma(B[t * d + u], rl[u], zt);
        z[i0 + t] = zt;
        const double2 rd = to_d(rl[t]), zd = to_d(zt);
        red[0] += rd.x * zd.x - rd.y * zd.y;
        red[1] += rd.x * zd.y + rd.y * zd.x;
        red[2] += rd.x * rd.x + rd.y * rd.y;
      }
  }
  write_partials<4>(red, part);
}

// finish a (re)start: rho, |r|^2 (and |b|^2); beta = 0, resume iterating
__global__ void __launch_bounds__(kBS) k_fin_reset(Scal* __restrict__ s, const double* part,
                                                   int np, int has_b) {
  double tot[4];
  sum_partials<4>(part, np, tot);
  if (threadIdx.x != 0) return;
  s->rho = make_double2(tot[0], tot[1]);
  s->beta = make_double2(0.0, 0.0);
  s->res2 = tot[2];
  if (has_b) s->bnorm2 = tot[3];
  s->best2 = tot[2];
  s->last_imp = s->iter;
  s->stalled = 0;
  s->done = 0;
}

// x64 += promote(x32); x32 = 0
__global__ void k_accumulate(int64_t n, double2* __restrict__ x, float2* __restrict__ c) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float2 v = c[i];
  x[i] = make_double2(x[i].x + (double)v.x, x[i].y + (double)v.y);
  c[i] = make_float2(0.f, 0.f);
}

// ---------------------------------------------------------------------------
// preconditioners
// ---------------------------------------------------------------------------
template <typename VT>
__global__ void k_jacobi_inv(int64_t n, const int32_t* __restrict__ rowptr,
                             const int32_t* __restrict__ cols, const double2* __restrict__ vals,
                             VT* __restrict__ dinv, int32_t* __restrict__ bad) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  int lo = rowptr[r], hi = rowptr[r + 1] - 1;
  double2 d = make_double2(0.0, 0.0);
  while (lo <= hi) {
    const int mid = (lo + hi) >> 1;
    const int c = cols[mid];
    if (c == r) {
      d = vals[mid];
      break;
    }
    if (c < r) lo = mid + 1;
    else hi = mid - 1;
  }
  if (d.x == 0.0 && d.y == 0.0) {
    bad[0] = 1;
    dinv[r] = from_d<VT>(make_double2(1.0, 0.0));
    return;
  }
  dinv[r] = from_d<VT>(zdiv(make_double2(1.0, 0.0), d));
}

__global__ void k_ones(int64_t n, double2* __restrict__ d) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = make_double2(1.0, 0.0);
}
__global__ void k_ones32(int64_t n, float2* __restrict__ d) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = make_float2(1.f, 0.f);
}
__global__ void k_demote(int64_t n, const double* __restrict__ a, float* __restrict__ b) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

// entity-block Jacobi set-up: one thread per block gathers A_BB from the CSR and
// inverts it; row-major at boff[k]
__global__ void k_bjacobi_setup(int64_t nb, const int32_t* __restrict__ bptr,
                                const int32_t* __restrict__ boff,
                                const int32_t* __restrict__ rowptr,
                                const int32_t* __restrict__ cols,
                                const double2* __restrict__ vals, double2* __restrict__ binv,
                                int32_t* __restrict__ bad) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nb) return;
  const int i0 = bptr[k], d = bptr[k + 1] - i0;
  if (d > kMaxBlock || !invert_block(rowptr, cols, vals, i0, d, binv + boff[k])) bad[0] = 1;
}

// z = P v (block), one thread per block
__global__ void k_bmul(int64_t nb, const int32_t* __restrict__ bptr,
                       const int32_t* __restrict__ boff, const double2* __restrict__ binv,
                       const double2* __restrict__ v, double2* __restrict__ z) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nb) return;
  const int i0 = bptr[k], d = bptr[k + 1] - i0;
  const double2* B = binv + boff[k];
  for (int t = 0; t < d; ++t) {
    double2 acc = make_double2(0.0, 0.0);
    for (int u = 0; u < d; ++u) acc = zadd(acc, zmul(B[t * d + u], v[i0 + u]));
    z[i0 + t] = acc;
  }
}

// entity rows must be consecutive rows of equal length sharing one column list
__global__ void k_check_entity_rows(int64_t nent, const int32_t* __restrict__ bptr,
                                    const int32_t* __restrict__ rowptr,
                                    const int32_t* __restrict__ cols, int32_t* bad) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nent) return;
  const int r0 = bptr[e], d = bptr[e + 1] - r0;
  const int base = rowptr[r0], L = rowptr[r0 + 1] - base;
  for (int k = 1; k <= d; ++k)
    if (rowptr[r0 + k] != base + k * L) {
      bad[0] = 1;
      return;
    }
  for (int k = 1; k < d; ++k)
    for (int t = 0; t < L; ++t)
      if (cols[base + k * L + t] != cols[base + t]) {
        bad[0] = 1;
        return;
      }
}

__global__ void k_init_scalars(Scal* s, int maxit, double tol2) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  Scal z{};
  z.maxit = maxit;
  z.tol2 = tol2;
  *s = z;
}
__global__ void k_set_done(Scal* s, int v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) s->done = v;
}
__global__ void k_set_ref2(Scal* s, double v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) s->ref2 = v;
}
__global__ void k_set_stall(Scal* s, int limit) {
  if (threadIdx.x == 0 && blockIdx.x == 0) s->stall_limit = limit;
}

// ---------------------------------------------------------------------------
// GMRES(m) kernels (complex128, Hermitian inner products)
// ---------------------------------------------------------------------------
// partials of V_i^H w for i < k: part[(2i+c) * grid + block]
// dots of w with V_0 .. V_{k-1} (bilinear: conj(V_i)^T w), eight vectors per
// sweep over w; with `norm`, also the partials of |w|^2 (slot 2k).  A set
// `skip` flag (the second Gram-Schmidt pass found unnecessary) returns at once.
constexpr int kDotGroup = 8;
__global__ void __launch_bounds__(kBS) k_mdot(int64_t n, const double2* __restrict__ V,
                                              int64_t ldv, int k, const double2* __restrict__ w,
                                              double* part, int norm,
                                              const int32_t* __restrict__ skip) {
  if (skip && *skip) return;
  int64_t lo, hi;
  chunk(n, kBS, &lo, &hi);
  const int64_t G = gridDim.x;
  for (int g0 = 0; g0 < k; g0 += kDotGroup) {
    const int gk = k - g0 < kDotGroup ? k - g0 : kDotGroup;
    double red[2 * kDotGroup + 1];
#pragma unroll
    for (int q = 0; q < 2 * kDotGroup + 1; ++q) red[q] = 0.0;
    for (int64_t j = lo + threadIdx.x; j < hi; j += kBS) {
      const double2 ww = w[j];
#pragma unroll
      for (int q = 0; q < kDotGroup; ++q)
        if (q < gk) {
          const double2 v = V[(int64_t)(g0 + q) * ldv + j];
          red[2 * q] += v.x * ww.x + v.y * ww.y;
          red[2 * q + 1] += v.x * ww.y - v.y * ww.x;
        }
      if (norm && g0 == 0) red[2 * kDotGroup] += ww.x * ww.x + ww.y * ww.y;
    }
    block_reduce<2 * kDotGroup + 1>(red);
    if (threadIdx.x == 0) {
#pragma unroll
      for (int q = 0; q < kDotGroup; ++q)  // unrolled: red stays in registers
        if (q < gk) {
          part[(int64_t)(2 * (g0 + q)) * G + blockIdx.x] = red[2 * q];
          part[(int64_t)(2 * (g0 + q) + 1) * G + blockIdx.x] = red[2 * q + 1];
        }
      if (norm && g0 == 0) part[(int64_t)(2 * k) * G + blockIdx.x] = red[2 * kDotGroup];
    }
  }
}

// one workgroup per vector i: h[i] = sum of its partials (fixed order); zero
// when the pass is skipped
__global__ void __launch_bounds__(kBS) k_fin_mdot(const double* part, int np,
                                                  double2* __restrict__ h,
                                                  const int32_t* __restrict__ skip) {
  const int i = blockIdx.x;
  if (skip && *skip) {
    if (threadIdx.x == 0) h[i] = make_double2(0.0, 0.0);
    return;
  }
  double tot[2];
  sum_partials<2>(part + (int64_t)(2 * i) * np, np, tot);
  if (threadIdx.x == 0) h[i] = make_double2(tot[0], tot[1]);
}

// w -= sum_i coef[i] V_i ; partials of |w|^2 (returns at once when skipped)
__global__ void __launch_bounds__(kBS) k_maxpy_norm(int64_t n, const double2* __restrict__ V,
                                                    int64_t ldv, int k,
                                                    const double2* __restrict__ coef,
                                                    double2* __restrict__ w, double* part,
                                                    const int32_t* __restrict__ skip) {
  if (skip && *skip) return;
  int64_t lo, hi;
  chunk(n, kBS, &lo, &hi);
  double red[1] = {0.0};
  for (int64_t j = lo + threadIdx.x; j < hi; j += kBS) {
    double2 acc = w[j];
    for (int i = 0; i < k; ++i) {
      const double2 c = coef[i], v = V[(int64_t)i * ldv + j];
      acc.x -= c.x * v.x - c.y * v.y;
      acc.y -= c.x * v.y + c.y * v.x;
    }
    w[j] = acc;
    red[0] += acc.x * acc.x + acc.y * acc.y;
  }
  write_partials<1>(red, part);
}

// the second Gram-Schmidt pass is needed when the first left less than
// 1/sqrt(2) of |w| (the Daniel-Gragg-Kaufman-Stewart test): skip otherwise
__global__ void __launch_bounds__(kBS) k_reorth_test(const double* before, const double* after,
                                                     int np, int32_t* skip) {
  double b[1], a[1];
  sum_partials<1>(before, np, b);
  sum_partials<1>(after, np, a);
  if (threadIdx.x == 0) *skip = (a[0] > 0.5 * b[0]) ? 1 : 0;
}

__global__ void __launch_bounds__(kBS) k_fin_sum1(const double* part, int np, double* out) {
  double tot[1];
  sum_partials<1>(part, np, tot);
  if (threadIdx.x == 0) *out = tot[0];
}

__global__ void k_dmul(int64_t n, const double2* __restrict__ d, const double2* __restrict__ v,
                       double2* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = zmul(d[i], v[i]);
}
__global__ void k_rscale(int64_t n, double2* __restrict__ v, double s) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = make_double2(v[i].x * s, v[i].y * s);
}
__global__ void k_vcomb(int64_t n, const double2* __restrict__ V, int64_t ldv, int k,
                        const double2* __restrict__ y, double2* __restrict__ z) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  double2 acc = make_double2(0.0, 0.0);
  for (int i = 0; i < k; ++i) acc = zadd(acc, zmul(y[i], V[(int64_t)i * ldv + j]));
  z[j] = acc;
}
__global__ void k_axpy1(int64_t n, const double2* __restrict__ v, double2* __restrict__ x) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) x[j] = zadd(x[j], v[j]);
}

static inline unsigned nblk(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// ---------------------------------------------------------------------------
// SpMV configurations and dispatch
// ---------------------------------------------------------------------------
struct SpmvCfg {
  int kind = 0;  // 0 = row, 1 = entity-row, 2 = entity-row with column pairs,
                 // 3 = element-by-element operator
  int G = 16, U = 2, D = 1;
  int grid = 1;
};

static int grid_for(int64_t n, int G) {
  const int64_t per = kBS / G;
  int64_t g = (n + per - 1) / per;
  g = std::min<int64_t>(g, kMaxGrid);
  if (g >= 64) g = (g / 8) * 8;
  return (int)std::max<int64_t>(g, 1);
}
static int vgrid_for(int64_t n) { return grid_for(n, 1); }

// the (G, U) pairs instantiated for both kernels
#define PG_GU_LIST(X) X(8, 4) X(16, 2) X(16, 4) X(32, 1) X(32, 2) X(64, 1)

template <int MODE, typename VT, typename XT>
static void spmv_row(const SpmvCfg& c, hipStream_t st, const Scal* s, int64_t n,
                     const int32_t* rp, const int32_t* ci, const VT* v, const XT* x, XT* y,
                     const double2* b, double* part) {
  const dim3 g(c.grid), bl(kBS);
#define PG_ROW(GG, UU)                                                                         \
  if (c.G == GG && c.U == UU) {                                                                \
    hipLaunchKernelGGL((k_spmv<GG, UU, MODE, VT, XT>), g, bl, 0, st, s, n, rp, ci, v, x, y, b, \
                       part);                                                                  \
    return;                                                                                    \
  }
  PG_GU_LIST(PG_ROW)
#undef PG_ROW
  hipLaunchKernelGGL((k_spmv<16, 2, MODE, VT, XT>), g, bl, 0, st, s, n, rp, ci, v, x, y, b, part);
}

template <int MODE, typename VT, typename XT, int D>
static void spmv_er_d(const SpmvCfg& c, hipStream_t st, const Scal* s, int64_t nent,
                      const int4* em, const int32_t* ci, const VT* v, const XT* x, XT* y,
                      const double2* b, double* part) {
  const dim3 g(c.grid), bl(kBS);
#define PG_ER(GG, UU)                                                                       \
  if (c.G == GG && c.U == UU) {                                                             \
    hipLaunchKernelGGL((k_spmv_er<GG, UU, D, MODE, VT, XT>), g, bl, 0, st, s, nent, em, ci, \
                       v, x, y, b, part);                                                   \
    return;                                                                                 \
  }
  PG_GU_LIST(PG_ER)
#undef PG_ER
  hipLaunchKernelGGL((k_spmv_er<16, 2, D, MODE, VT, XT>), g, bl, 0, st, s, nent, em, ci, v, x, y,
                     b, part);
}

template <int MODE, typename VT, typename XT>
static void spmv_er2(const SpmvCfg& c, hipStream_t st, const Scal* s, int64_t nent, const int4* em,
                     const int32_t* ci, const VT* v, const XT* x, XT* y, const double2* b,
                     double* part) {
  const dim3 g(c.grid), bl(kBS);
#define PG_ER2(GG, UU)                                                                        \
  if (c.G == GG && c.U == UU) {                                                               \
    hipLaunchKernelGGL((k_spmv_er2<GG, UU, MODE, VT, XT>), g, bl, 0, st, s, nent, em, ci, v, x, \
                       y, b, part);                                                           \
    return;                                                                                   \
  }
  PG_GU_LIST(PG_ER2)
#undef PG_ER2
  hipLaunchKernelGGL((k_spmv_er2<16, 2, MODE, VT, XT>), g, bl, 0, st, s, nent, em, ci, v, x, y, b,
                     part);
}

template <int MODE, typename VT, typename XT>
static void spmv_launch(const SpmvCfg& c, hipStream_t st, const Scal* s, int64_t n, int64_t nent,
                        const int4* em, const int32_t* rp, const int32_t* ci, const VT* v,
                        const XT* x, XT* y, const double2* b, double* part) {
  if (c.kind == 0) {
    spmv_row<MODE, VT, XT>(c, st, s, n, rp, ci, v, x, y, b, part);
  } else if (c.kind == 2) {
    spmv_er2<MODE, VT, XT>(c, st, s, nent, em, ci, v, x, y, b, part);
  } else if (c.D == 1) {
    spmv_er_d<MODE, VT, XT, 1>(c, st, s, nent, em, ci, v, x, y, b, part);
  } else if (c.D == 2) {
    spmv_er_d<MODE, VT, XT, 2>(c, st, s, nent, em, ci, v, x, y, b, part);
  } else {
    spmv_er_d<MODE, VT, XT, 6>(c, st, s, nent, em, ci, v, x, y, b, part);
  }
}

// ---------------------------------------------------------------------------
// solver object
// ---------------------------------------------------------------------------
struct Solver {
  int64_t n = 0, nnz = 0;
  pg_solve_opts o{};
  const int32_t* rowptr = nullptr;
  const int32_t* cols = nullptr;
  const double2* vals = nullptr;  // caller's complex128 values
  float2* vals32 = nullptr;       // owned complex64 copy (mixed)
  SpmvCfg c64, c32;               // chosen SpMV kernels
  SpmvCfg cres;                   // CSR kernel of the FP64 true residual
  const ElemOp* eop = nullptr;    // caller's element-by-element operator (optional)
  int gv = 1;
  Scal* s = nullptr;
  Scal* h = nullptr;  // pinned host mirror
  double* part = nullptr;
  double2 *r = nullptr, *p = nullptr, *q = nullptr, *z = nullptr, *dinv = nullptr, *r64 = nullptr;
  float2 *r32 = nullptr, *p32 = nullptr, *q32 = nullptr, *z32 = nullptr, *x32 = nullptr,
         *dinv32 = nullptr;
  int32_t* bad = nullptr;
  // mesh entities (DOF ranges): entity-row SpMV and entity-block Jacobi
  int64_t nblocks = 0;
  int gb = 1, maxblk = 0;
  bool blk2 = false;  // every block two DOFs: the vectorised block update
  bool er_ok = false;
  int32_t *bptr = nullptr, *boff = nullptr;
  // entity-row SpMV in schedule order: descriptors {r0, d | L << 8, cbase, vbase},
  // each scheduled entity's CSR slot, and the scheduled column lists and values
  int4* emeta = nullptr;
  int32_t* esrc = nullptr;
  int32_t* cols_s = nullptr;
  int32_t* cols2 = nullptr;  // first column of each pair (every entity two DOFs)
  bool pairs_ok = false;
  double2* vals_s64 = nullptr;
  double2* binv = nullptr;
  float2* binv32 = nullptr;
  // GMRES(m)
  double2* V = nullptr;
  double2* hdev = nullptr;
  double2* hhost = nullptr;  // pinned
  double* partg = nullptr;
  double* dnorm = nullptr;
  int32_t* skip = nullptr;  // GMRES: second Gram-Schmidt pass not needed
  int restart = 30;
  // SpMV sampling inside the solve: before, between the element operator's two
  // kernels, after
  hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
  std::vector<void*> owned;

  template <typename T>
  int alloc(T** ptr, size_t count) {
    void* v = nullptr;
    hipError_t e = hipMalloc(&v, std::max<size_t>(count, 1) * sizeof(T));
    if (e != hipSuccess)
      PG_FAIL(PG_ERR_HIP, "hipMalloc(%zu): %s", count * sizeof(T), hipGetErrorString(e));
    owned.push_back(v);
    *ptr = static_cast<T*>(v);
    return PG_OK;
  }
  ~Solver() {  // release errors cannot be reported from here
    for (hipEvent_t e : ev)
      if (e) (void)hipEventDestroy(e);
    for (void* v : owned) (void)hipFree(v);
    if (h) (void)hipHostFree(h);
    if (hhost) (void)hipHostFree(hhost);
  }
};

#define PG_TRY(x)                 \
  do {                            \
    int _rc = (x);                \
    if (_rc != PG_OK) return _rc; \
  } while (0)

static int read_scalars(Solver* S, hipStream_t st) {
  PG_HIP(hipMemcpyAsync(S->h, S->s, sizeof(Scal), hipMemcpyDeviceToHost, st));
  PG_HIP(hipStreamSynchronize(st));
  return PG_OK;
}

// the entity-row kind reads the scheduled copies (complex128: vals_s64;
// complex64: vals32, which is kept in schedule order whenever entity rows exist)
static inline const double2* sched_values(Solver* S, const double2*) { return S->vals_s64; }
static inline const float2* sched_values(Solver* S, const float2*) { return S->vals32; }

template <int MODE, typename VT, typename XT>
static void spmv(Solver* S, const SpmvCfg& c, hipStream_t st, const VT* v, const XT* x, XT* y,
                 const double2* b, hipEvent_t mid = nullptr) {
  if (c.kind == 3)
    eop_launch(S->eop, MODE, MODE != 0 ? &S->s->done : nullptr, x, y, b, S->part, st, mid);
  else if (c.kind >= 1)
    spmv_launch<MODE, VT, XT>(c, st, S->s, S->n, S->nblocks, S->emeta, S->rowptr,
                              c.kind == 2 ? S->cols2 : S->cols_s, sched_values(S, v), x, y, b,
                              S->part);
  else
    spmv_launch<MODE, VT, XT>(c, st, S->s, S->n, S->nblocks, nullptr, S->rowptr, S->cols, v, x, y,
                              b, S->part);
}

// r64 = b - A x (FP64) and |r64|^2 into s->res64, always with the assembled
// CSR matrix (the system whose residual the tolerance refers to)
static int true_residual(Solver* S, hipStream_t st, const double2* b, const double2* x) {
  spmv<2, double2, double2>(S, S->cres, st, S->vals, x, S->r64, b);
  hipLaunchKernelGGL(k_fin_res64, dim3(1), dim3(kBS), 0, st, S->s, S->part, S->cres.grid);
  PG_CHECK_LAUNCH();
  return PG_OK;
}

// average duration of the solve's own SpMV (MODE 1) over `reps` launches
template <typename VT, typename XT>
static float time_spmv(Solver* S, const SpmvCfg& c, hipStream_t st, const VT* v, const XT* x,
                       XT* y, int reps) {
  hipEvent_t a = nullptr, b = nullptr;
  float ms = 1e30f;  // an untimed candidate is never picked
  if (hipEventCreate(&a) == hipSuccess && hipEventCreate(&b) == hipSuccess) {
    spmv<1, VT, XT>(S, c, st, v, x, y, nullptr);
    if (hipEventRecord(a, st) == hipSuccess) {
      for (int i = 0; i < reps; ++i) spmv<1, VT, XT>(S, c, st, v, x, y, nullptr);
      if (hipEventRecord(b, st) != hipSuccess || hipEventSynchronize(b) != hipSuccess ||
          hipEventElapsedTime(&ms, a, b) != hipSuccess)
        ms = 1e30f * reps;
    }
  }
  if (a) (void)hipEventDestroy(a);
  if (b) (void)hipEventDestroy(b);
  return ms / reps;
}

// fastest candidate: every candidate timed over 6 launches, the best three
// again over 20 launches each
template <typename VT, typename XT>
static SpmvCfg pick_spmv(Solver* S, const std::vector<SpmvCfg>& cands, hipStream_t st,
                         const VT* v, const XT* x, XT* y) {
  std::vector<std::pair<float, int>> t;
  for (size_t i = 0; i < cands.size(); ++i)
    t.push_back({time_spmv<VT, XT>(S, cands[i], st, v, x, y, 6), (int)i});
  std::sort(t.begin(), t.end());
  const size_t top = std::min<size_t>(3, t.size());
  float best = 1e30f;
  int bi = t[0].second;
  for (size_t k = 0; k < top; ++k) {
    const float ms = time_spmv<VT, XT>(S, cands[t[k].second], st, v, x, y, 20);
    if (ms < best) {
      best = ms;
      bi = t[k].second;
    }
  }
  return cands[bi];
}

// preconditioned (re)start and iteration, point or entity-block Jacobi
template <typename XT>
static void launch_reset(Solver* S, hipStream_t st, XT* r, const XT* dinv, const XT* binv, XT* z,
                         XT* p, const double2* b) {
  int np;
  if (S->o.pc == PG_PC_BJACOBI) {
    hipLaunchKernelGGL(k_reset_blk<XT>, dim3(S->gb), dim3(kBS), 0, st, S->nblocks, S->bptr,
                       S->boff, S->r64, r, binv, z, p, b, S->part);
    np = S->gb;
  } else {
    hipLaunchKernelGGL(k_reset<XT>, dim3(S->gv), dim3(kBS), 0, st, S->n, S->r64, r, dinv, z, p, b,
                       S->part);
    np = S->gv;
  }
  hipLaunchKernelGGL(k_fin_reset, dim3(1), dim3(kBS), 0, st, S->s, S->part, np, b ? 1 : 0);
}

template <typename VT, typename XT>
static void launch_iteration(Solver* S, hipStream_t st, const SpmvCfg& c, const VT* vals, XT* x,
                             XT* r, XT* p, XT* q, const XT* dinv, const XT* binv, XT* z,
                             bool sample) {
  hipLaunchKernelGGL(k_pupdate<XT>, dim3(S->gv), dim3(kBS), 0, st, S->s, S->n, z, p, x);
  // a failed record only loses this timing sample (the elapsed-time query fails)
  if (sample) (void)hipEventRecord(S->ev[0], st);
  spmv<1, VT, XT>(S, c, st, vals, p, q, nullptr, sample && c.kind == 3 ? S->ev[2] : nullptr);
  if (sample) (void)hipEventRecord(S->ev[1], st);
  hipLaunchKernelGGL(k_fin_mu, dim3(1), dim3(kBS), 0, st, S->s, S->part, c.grid);
  int np;
  if (S->o.pc == PG_PC_BJACOBI && S->blk2) {  // r, q, z: the solver's own 256-byte aligned buffers
    hipLaunchKernelGGL(k_update_blk2<XT>, dim3(S->gb), dim3(kBS), 0, st, S->s, S->nblocks, r, q,
                       binv, z, S->part);
    np = S->gb;
  } else if (S->o.pc == PG_PC_BJACOBI) {
    hipLaunchKernelGGL(k_update_blk<XT>, dim3(S->gb), dim3(kBS), 0, st, S->s, S->nblocks, S->bptr,
                       S->boff, r, q, binv, z, S->part);
    np = S->gb;
  } else {
    hipLaunchKernelGGL(k_update<XT>, dim3(S->gv), dim3(kBS), 0, st, S->s, S->n, r, q, dinv, z,
                       S->part);
    np = S->gv;
  }
  hipLaunchKernelGGL(k_fin_update, dim3(1), dim3(kBS), 0, st, S->s, S->part, np);
}

static int cocg_solve(Solver* S, const double2* b, double2* x, pg_solve_info* info,
                      hipStream_t st) {
  const bool mixed = S->o.precision == PG_PREC_MIXED;
  bool low = mixed;  // current recursion in complex64
  const int64_t n = S->n;
  const double tol2 = S->o.rtol * S->o.rtol;
  hipLaunchKernelGGL(k_init_scalars, dim3(1), dim3(64), 0, st, S->s, S->o.maxit, tol2);
  if (mixed) PG_HIP(hipMemsetAsync(S->x32, 0, n * sizeof(float2), st));
  PG_TRY(true_residual(S, st, b, x));
  if (low)
    launch_reset<float2>(S, st, S->r32, S->dinv32, S->binv32, S->z32, S->p32, b);
  else
    launch_reset<double2>(S, st, S->r, S->dinv, S->binv, S->z, S->p, b);
  PG_CHECK_LAUNCH();
  PG_TRY(read_scalars(S, st));
  const double bnorm2 = S->h->bnorm2;
  int refinements = 0, restarts = 0;
  double spmv_ms = 0.0;
  int spmv_n = 0;
  info->status = -1;
  info->refinements = 0;
  info->fp64_fallback = 0;
  info->spmv_ms = 0.0;
  info->spmv_samples = 0;
  if (bnorm2 == 0.0) {
    PG_HIP(hipMemsetAsync(x, 0, n * sizeof(double2), st));
    info->iterations = 0;
    info->relres = 0.0;
    info->status = PG_OK;
    return PG_OK;
  }
  double res64 = S->h->res64;
  if (res64 <= tol2 * bnorm2) {
    info->iterations = 0;
    info->relres = std::sqrt(res64 / bnorm2);
    info->status = PG_OK;
    return PG_OK;
  }
  const double delta2 = S->o.inner_rtol * S->o.inner_rtol;
  double round_res64 = res64;  // |b - A x|^2 when the current complex64 round began
  if (low) {
    hipLaunchKernelGGL(k_set_ref2, dim3(1), dim3(64), 0, st, S->s, res64 * delta2);
    hipLaunchKernelGGL(k_set_stall, dim3(1), dim3(64), 0, st, S->s,
                       S->o.stall_iters > 0 ? S->o.stall_iters : kStallIters);
  }
  const int check = std::max(1, S->o.check_every);
  double first_ms = 0.0;
  int first_n = 0;
  while (true) {
    // the first iteration of every batch runs in full (the flags were clear
    // when the batch was enqueued): time its SpMV with events on this stream
    const bool two_kernels = (low ? S->c32 : S->c64).kind == 3;
    for (int k = 0; k < check; ++k) {
      if (low)
        launch_iteration<float2, float2>(S, st, S->c32, S->vals32, S->x32, S->r32, S->p32, S->q32,
                                         S->dinv32, S->binv32, S->z32, k == 0);
      else
        launch_iteration<double2, double2>(S, st, S->c64, S->vals, x, S->r, S->p, S->q, S->dinv,
                                           S->binv, S->z, k == 0);
    }
    PG_CHECK_LAUNCH();
    PG_TRY(read_scalars(S, st));
    {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, S->ev[0], S->ev[1]) == hipSuccess) {
        spmv_ms += ms;
        ++spmv_n;
      }
      if (two_kernels && hipEventElapsedTime(&ms, S->ev[0], S->ev[2]) == hipSuccess) {
        first_ms += ms;
        ++first_n;
      }
    }
    const Scal hs = *S->h;
    if (hs.status == PG_ERR_BREAKDOWN) {
      info->status = PG_ERR_BREAKDOWN;
      break;
    }
    if (!hs.done) continue;
    // paused: converged recursion, refinement point, stagnation or iteration limit;
    // the last iteration's step x += alpha p is still to be taken
    if (low)
      hipLaunchKernelGGL(k_xstep<float2>, dim3(nblk(n, kBS)), dim3(kBS), 0, st, S->s, n, S->p32,
                         S->x32);
    else
      hipLaunchKernelGGL(k_xstep<double2>, dim3(nblk(n, kBS)), dim3(kBS), 0, st, S->s, n, S->p, x);
    if (low) {
      hipLaunchKernelGGL(k_accumulate, dim3(nblk(n, kBS)), dim3(kBS), 0, st, n, x, S->x32);
      refinements++;
    }
    hipLaunchKernelGGL(k_set_done, dim3(1), dim3(64), 0, st, S->s, 0);
    PG_TRY(true_residual(S, st, b, x));
    PG_TRY(read_scalars(S, st));
    res64 = S->h->res64;
    if (res64 <= tol2 * bnorm2) {
      info->status = PG_OK;
      break;
    }
    if (hs.iter >= S->o.maxit) break;
    if (low && hs.stalled && res64 <= 1e-2 * round_res64) {
      // the complex64 round stagnated after reducing the true residual at least
      // tenfold: one more round from the FP64 residual (iterative refinement
      // whose rounds are limited by complex64 accuracy)
      launch_reset<float2>(S, st, S->r32, S->dinv32, S->binv32, S->z32, S->p32, nullptr);
      hipLaunchKernelGGL(k_set_ref2, dim3(1), dim3(64), 0, st, S->s, res64 * delta2);
      round_res64 = res64;
    } else if (low && hs.stalled) {
      // the complex64 recursion stagnated without progress: continue in FP64 from x
      low = false;
      info->fp64_fallback = 1;
      hipLaunchKernelGGL(k_set_ref2, dim3(1), dim3(64), 0, st, S->s, 0.0);
      hipLaunchKernelGGL(k_set_stall, dim3(1), dim3(64), 0, st, S->s, 0);
      launch_reset<double2>(S, st, S->r, S->dinv, S->binv, S->z, S->p, nullptr);
    } else if (!low) {
      // recursive residual converged but the true one did not: restart from x
      if (++restarts > 20) break;
      launch_reset<double2>(S, st, S->r, S->dinv, S->binv, S->z, S->p, nullptr);
    } else {
      // restart the complex64 recursion from the FP64 residual (keeping the old
      // search direction across a residual replacement stalls COCG)
      launch_reset<float2>(S, st, S->r32, S->dinv32, S->binv32, S->z32, S->p32, nullptr);
      hipLaunchKernelGGL(k_set_ref2, dim3(1), dim3(64), 0, st, S->s, res64 * delta2);
      round_res64 = res64;
    }
    PG_CHECK_LAUNCH();
  }
  PG_TRY(read_scalars(S, st));
  info->iterations = S->h->iter;
  info->refinements = refinements;
  info->spmv_samples = spmv_n;
  info->spmv_ms = spmv_n ? spmv_ms / spmv_n : 0.0;
  info->spmv_first_ms = first_n ? first_ms / first_n : 0.0;
  info->relres = std::sqrt(res64 / bnorm2);
  return PG_OK;
}

// ---------------------------------------------------------------------------
// GMRES(m), right-preconditioned, CGS with one re-orthogonalisation pass
// ---------------------------------------------------------------------------
static void apply_pc(Solver* S, hipStream_t st, const double2* v, double2* out) {
  if (S->o.pc == PG_PC_BJACOBI)
    hipLaunchKernelGGL(k_bmul, dim3(nblk(S->nblocks, kBS)), dim3(kBS), 0, st, S->nblocks,
                       S->bptr, S->boff, S->binv, v, out);
  else
    hipLaunchKernelGGL(k_dmul, dim3(nblk(S->n, kBS)), dim3(kBS), 0, st, S->n, S->dinv, v, out);
}

static int gmres_solve(Solver* S, const double2* b, double2* x, pg_solve_info* info,
                       hipStream_t st) {
  const int64_t n = S->n;
  const int m = S->restart;
  const double tol = S->o.rtol;
  const unsigned nb = nblk(n, kBS);
  hipLaunchKernelGGL(k_init_scalars, dim3(1), dim3(64), 0, st, S->s, S->o.maxit, tol * tol);
  // |b|^2: w0 = b, no projection
  PG_HIP(hipMemcpyAsync(S->V, b, n * sizeof(double2), hipMemcpyDeviceToDevice, st));
  hipLaunchKernelGGL(k_maxpy_norm, dim3(S->gv), dim3(kBS), 0, st, n, S->V, n, 0, S->hdev, S->V,
                     S->part, nullptr);
  hipLaunchKernelGGL(k_fin_sum1, dim3(1), dim3(kBS), 0, st, S->part, S->gv, &S->s->bnorm2);
  PG_TRY(true_residual(S, st, b, x));
  PG_TRY(read_scalars(S, st));
  const double bnorm = std::sqrt(S->h->bnorm2);
  double beta = std::sqrt(S->h->res64);
  info->status = -1;
  info->refinements = 0;
  int iters = 0;
  if (bnorm == 0.0) {
    PG_HIP(hipMemsetAsync(x, 0, n * sizeof(double2), st));
    info->iterations = 0;
    info->relres = 0.0;
    info->status = PG_OK;
    return PG_OK;
  }
  std::vector<double2> H((size_t)(m + 1) * m), g(m + 1), sn(m);
  std::vector<double> cs(m);
  auto Hij = [&](int i, int j) -> double2& { return H[(size_t)i * m + j]; };
  while (beta > tol * bnorm && iters < S->o.maxit) {
    PG_HIP(hipMemcpyAsync(S->V, S->r64, n * sizeof(double2), hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(k_rscale, dim3(nb), dim3(kBS), 0, st, n, S->V, 1.0 / beta);
    std::fill(H.begin(), H.end(), make_double2(0.0, 0.0));
    std::fill(g.begin(), g.end(), make_double2(0.0, 0.0));
    g[0] = make_double2(beta, 0.0);
    int k = 0;
    for (int j = 0; j < m && iters < S->o.maxit; ++j) {
      double2* Vj = S->V + (int64_t)j * n;
      double2* w = S->V + (int64_t)(j + 1) * n;
      apply_pc(S, st, Vj, S->z);
      spmv<0, double2, double2>(S, S->c64, st, S->vals, S->z, w, nullptr);
      // classical Gram-Schmidt; the second pass only when the first lost more
      // than half of |w|^2 (decided on the device, no host round trip)
      for (int pass = 0; pass < 2; ++pass) {
        double2* hcol = S->hdev + pass * (m + 1);
        const int32_t* skip = pass ? S->skip : nullptr;
        hipLaunchKernelGGL(k_mdot, dim3(S->gv), dim3(kBS), 0, st, n, S->V, n, j + 1, w, S->partg,
                           pass == 0 ? 1 : 0, skip);
        hipLaunchKernelGGL(k_fin_mdot, dim3(j + 1), dim3(kBS), 0, st, S->partg, S->gv, hcol, skip);
        hipLaunchKernelGGL(k_maxpy_norm, dim3(S->gv), dim3(kBS), 0, st, n, S->V, n, j + 1, hcol,
                           w, S->part, skip);
        if (pass == 0)
          hipLaunchKernelGGL(k_reorth_test, dim3(1), dim3(kBS), 0, st,
                             S->partg + (int64_t)(2 * (j + 1)) * S->gv, S->part, S->gv, S->skip);
      }
      hipLaunchKernelGGL(k_fin_sum1, dim3(1), dim3(kBS), 0, st, S->part, S->gv, S->dnorm);
      PG_CHECK_LAUNCH();
      PG_HIP(hipMemcpyAsync(S->hhost, S->hdev, 2 * (m + 1) * sizeof(double2),
                            hipMemcpyDeviceToHost, st));
      PG_HIP(hipMemcpyAsync(&S->h->res2, S->dnorm, sizeof(double), hipMemcpyDeviceToHost, st));
      PG_HIP(hipStreamSynchronize(st));
      const double hn = std::sqrt(S->h->res2);
      for (int i = 0; i <= j; ++i) Hij(i, j) = zadd(S->hhost[i], S->hhost[(m + 1) + i]);
      Hij(j + 1, j) = make_double2(hn, 0.0);
      for (int i = 0; i < j; ++i) {  // previous rotations
        const double2 a = Hij(i, j), bb = Hij(i + 1, j);
        const double2 sc = make_double2(sn[i].x, -sn[i].y);
        const double2 t = zadd(make_double2(cs[i] * a.x, cs[i] * a.y), zmul(sn[i], bb));
        Hij(i + 1, j) = zsub(make_double2(cs[i] * bb.x, cs[i] * bb.y), zmul(sc, a));
        Hij(i, j) = t;
      }
      {  // new rotation [c s; -conj(s) c] zeroing H(j+1,j)
        const double2 a = Hij(j, j), bb = Hij(j + 1, j);
        const double na = std::hypot(a.x, a.y), nbb = std::hypot(bb.x, bb.y);
        const double rho = std::hypot(na, nbb);
        if (na == 0.0) {
          cs[j] = 0.0;
          sn[j] = make_double2(1.0, 0.0);
        } else {
          cs[j] = na / rho;
          const double2 ph = make_double2(a.x / na, a.y / na);
          sn[j] = zmul(ph, make_double2(bb.x / rho, -bb.y / rho));
        }
        Hij(j, j) = zadd(make_double2(cs[j] * a.x, cs[j] * a.y), zmul(sn[j], bb));
        Hij(j + 1, j) = make_double2(0.0, 0.0);
        const double2 sc = make_double2(sn[j].x, -sn[j].y);
        g[j + 1] = zmul(make_double2(-sc.x, -sc.y), g[j]);
        g[j] = make_double2(cs[j] * g[j].x, cs[j] * g[j].y);
      }
      ++iters;
      k = j + 1;
      if (hn > 0.0) hipLaunchKernelGGL(k_rscale, dim3(nb), dim3(kBS), 0, st, n, w, 1.0 / hn);
      if (std::hypot(g[j + 1].x, g[j + 1].y) <= tol * bnorm || hn == 0.0) break;
    }
    // y = H^-1 g (upper triangular); x += P (V y)
    std::vector<double2> y(k);
    for (int i = k - 1; i >= 0; --i) {
      double2 acc = g[i];
      for (int l = i + 1; l < k; ++l) acc = zsub(acc, zmul(Hij(i, l), y[l]));
      y[i] = zdiv(acc, Hij(i, i));
    }
    for (int i = 0; i < k; ++i) S->hhost[i] = y[i];
    PG_HIP(hipMemcpyAsync(S->hdev, S->hhost, k * sizeof(double2), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_vcomb, dim3(nb), dim3(kBS), 0, st, n, S->V, n, k, S->hdev, S->z);
    apply_pc(S, st, S->z, S->q);
    hipLaunchKernelGGL(k_axpy1, dim3(nb), dim3(kBS), 0, st, n, S->q, x);
    PG_TRY(true_residual(S, st, b, x));
    PG_TRY(read_scalars(S, st));
    beta = std::sqrt(S->h->res64);
    if (!std::isfinite(beta)) {
      info->status = PG_ERR_BREAKDOWN;
      break;
    }
  }
  if (beta <= tol * bnorm) info->status = PG_OK;
  info->iterations = iters;
  info->relres = beta / bnorm;
  return PG_OK;
}

// (Re)build what depends on the caller's matrix values: Jacobi or block
// inverses, the schedule-order copies of A (entity rows) and, for mixed
// precision, the complex64 copy of A.
static int refresh_values(Solver* S, hipStream_t st) {
  const int64_t n = S->n, nnz = S->nnz;
  const bool mixed = S->o.precision == PG_PREC_MIXED;
  PG_HIP(hipMemsetAsync(S->bad, 0, sizeof(int32_t), st));
  if (S->er_ok) {
    const dim3 g(nblk(32 * S->nblocks, kBS));
    hipLaunchKernelGGL(k_sched_vals<double2>, g, dim3(kBS), 0, st, S->nblocks, S->emeta, S->esrc,
                       S->vals, S->vals_s64);
    if (mixed)
      hipLaunchKernelGGL(k_sched_vals<float2>, g, dim3(kBS), 0, st, S->nblocks, S->emeta, S->esrc,
                         S->vals, S->vals32);
  } else if (mixed) {
    hipLaunchKernelGGL(k_demote, dim3(nblk(2 * nnz, kBS)), dim3(kBS), 0, st, 2 * nnz,
                       reinterpret_cast<const double*>(S->vals),
                       reinterpret_cast<float*>(S->vals32));
  }
  if (S->o.pc == PG_PC_BJACOBI) {
    hipLaunchKernelGGL(k_bjacobi_setup, dim3(nblk(S->nblocks, kBS)), dim3(kBS), 0, st,
                       S->nblocks, S->bptr, S->boff, S->rowptr, S->cols, S->vals, S->binv,
                       S->bad);
    if (mixed) {
      int32_t tot = 0;
      PG_HIP(hipMemcpyAsync(&tot, S->boff + S->nblocks, sizeof(int32_t), hipMemcpyDeviceToHost,
                            st));
      PG_HIP(hipStreamSynchronize(st));
      hipLaunchKernelGGL(k_demote, dim3(nblk(2 * (int64_t)tot, kBS)), dim3(kBS), 0, st,
                         2 * (int64_t)tot, reinterpret_cast<const double*>(S->binv),
                         reinterpret_cast<float*>(S->binv32));
    }
  } else if (S->o.pc == PG_PC_JACOBI) {
    hipLaunchKernelGGL(k_jacobi_inv<double2>, dim3(nblk(n, kBS)), dim3(kBS), 0, st, n, S->rowptr,
                       S->cols, S->vals, S->dinv, S->bad);
    if (mixed)
      hipLaunchKernelGGL(k_jacobi_inv<float2>, dim3(nblk(n, kBS)), dim3(kBS), 0, st, n, S->rowptr,
                         S->cols, S->vals, S->dinv32, S->bad);
  } else {
    hipLaunchKernelGGL(k_ones, dim3(nblk(n, kBS)), dim3(kBS), 0, st, n, S->dinv);
    if (mixed) hipLaunchKernelGGL(k_ones32, dim3(nblk(n, kBS)), dim3(kBS), 0, st, n, S->dinv32);
  }
  PG_CHECK_LAUNCH();
  PG_HIP(hipMemcpyAsync(&S->h->iter, S->bad, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  PG_HIP(hipStreamSynchronize(st));
  if (S->h->iter != 0)
    PG_FAIL(PG_ERR_ARG, "singular diagonal entry or block: preconditioner undefined");
  return PG_OK;
}

// Entity rows: check that the rows of every entity (block_ptr range) share one
// column list, then lay out the schedule (entity_order, or natural order):
// descriptors, column lists and value blocks in schedule order.
static int setup_entities(Solver* S, const int32_t* bp, const int32_t* order, hipStream_t st) {
  const int64_t nb = S->nblocks;
  if (S->o.spmv == PG_SPMV_ROW || S->o.spmv == PG_SPMV_ELEMENT) {  // no schedule copies
    S->er_ok = false;
    return PG_OK;
  }
  PG_HIP(hipMemsetAsync(S->bad, 0, sizeof(int32_t), st));
  hipLaunchKernelGGL(k_check_entity_rows, dim3(nblk(nb, kBS)), dim3(kBS), 0, st, nb, S->bptr,
                     S->rowptr, S->cols, S->bad);
  PG_CHECK_LAUNCH();
  PG_HIP(hipMemcpyAsync(&S->h->iter, S->bad, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  PG_HIP(hipStreamSynchronize(st));
  S->er_ok = S->h->iter == 0;
  if (!S->er_ok) return PG_OK;
  std::vector<int32_t> rp(S->n + 1);
  PG_HIP(hipMemcpy(rp.data(), S->rowptr, (S->n + 1) * sizeof(int32_t), hipMemcpyDeviceToHost));
  if (order) {
    std::vector<char> seen(nb, 0);
    for (int64_t i = 0; i < nb; ++i) {
      if (order[i] < 0 || order[i] >= nb || seen[order[i]])
        PG_FAIL(PG_ERR_ARG, "entity_order must be a permutation of 0..nblocks-1");
      seen[order[i]] = 1;
    }
  }
  std::vector<int4> em(nb);
  std::vector<int32_t> src(nb);
  int64_t cb = 0, vb = 0;
  for (int64_t i = 0; i < nb; ++i) {
    const int e = order ? order[i] : (int)i;
    const int r0 = bp[e], d = bp[e + 1] - r0;
    const int base = rp[r0], L = rp[r0 + 1] - base;
    if (L >= (1 << 23)) PG_FAIL(PG_ERR_ARG, "entity row too long");
    em[i] = make_int4(r0, d | (L << 8), (int)cb, (int)vb);
    src[i] = base;
    cb += L;
    vb += (int64_t)d * L;
    if (cb >= (1LL << 31) || vb >= (1LL << 31)) PG_FAIL(PG_ERR_ARG, "matrix too large");
  }
  if (vb != S->nnz) PG_FAIL(PG_ERR_ARG, "the entity rows do not cover the matrix");
  int rc;
  if ((rc = S->alloc(&S->esrc, nb)) != PG_OK || (rc = S->alloc(&S->cols_s, cb)) != PG_OK ||
      (rc = S->alloc(&S->vals_s64, S->nnz)) != PG_OK)
    return rc;
  PG_HIP(hipMemcpy(S->emeta, em.data(), nb * sizeof(int4), hipMemcpyHostToDevice));
  PG_HIP(hipMemcpy(S->esrc, src.data(), nb * sizeof(int32_t), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_sched_cols, dim3(nblk(32 * nb, kBS)), dim3(kBS), 0, st, nb, S->emeta,
                     S->esrc, S->cols, S->cols_s);
  PG_CHECK_LAUNCH();
  // column pairs (every entity two DOFs, p = 2)
  PG_HIP(hipMemsetAsync(S->bad, 0, sizeof(int32_t), st));
  hipLaunchKernelGGL(k_check_pairs, dim3(nblk(nb, kBS)), dim3(kBS), 0, st, nb, S->emeta, S->cols_s,
                     S->bad);
  PG_CHECK_LAUNCH();
  PG_HIP(hipMemcpyAsync(&S->h->iter, S->bad, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  PG_HIP(hipStreamSynchronize(st));
  if (S->h->iter == 0) {
    if ((rc = S->alloc(&S->cols2, cb / 2)) != PG_OK) return rc;
    hipLaunchKernelGGL(k_make_pairs, dim3(nblk(cb / 2, kBS)), dim3(kBS), 0, st, cb / 2, S->cols_s,
                       S->cols2);
    PG_CHECK_LAUNCH();
    S->pairs_ok = true;
  }
  return PG_OK;
}

// pick the fastest SpMV configuration for this matrix (row / entity-row, G, U,
// element operator), or copy another solver's choices
static int autotune(Solver* S, hipStream_t st) {
  const bool mixed = S->o.precision == PG_PREC_MIXED;
  if (S->o.like) {
    const Solver* L = static_cast<const Solver*>(S->o.like);
    if (L->n != S->n || L->nnz != S->nnz || L->er_ok != S->er_ok || L->pairs_ok != S->pairs_ok)
      PG_FAIL(PG_ERR_ARG, "`like` is a solver of another system or layout");
    const auto usable = [&](const SpmvCfg& c, bool c128) {
      return c.kind != 3 || (S->eop && eop_supports(S->eop, c128));
    };
    if (!usable(L->c64, true) || (mixed && !usable(L->c32, false)))
      PG_FAIL(PG_ERR_ARG, "`like` uses an element operator this solver does not have");
    if (mixed && L->o.precision != PG_PREC_MIXED)
      PG_FAIL(PG_ERR_ARG, "`like` has no complex64 configuration");
    S->cres = L->cres;
    S->c64 = L->c64;
    if (S->c64.kind == 3) S->c64.grid = eop_grid(S->eop);
    if (mixed) {
      S->c32 = L->c32;
      if (S->c32.kind == 3) S->c32.grid = eop_grid(S->eop);
    }
    return PG_OK;
  }
  if ((S->o.spmv == PG_SPMV_ENTITY || S->o.spmv == PG_SPMV_PAIRS) && !S->er_ok)
    PG_FAIL(PG_ERR_ARG, "entity-row SpMV needs block_ptr ranges whose rows share one column list");
  if (S->o.spmv == PG_SPMV_PAIRS && !S->pairs_ok)
    PG_FAIL(PG_ERR_ARG, "the column-pair SpMV needs every entity to carry two DOFs (p = 2)");
  const bool only_eop = S->o.spmv == PG_SPMV_ELEMENT;
  std::vector<SpmvCfg> cands;
  const int gu[6][2] = {{8, 4}, {16, 2}, {16, 4}, {32, 1}, {32, 2}, {64, 1}};
  for (const auto& g : gu) {
    SpmvCfg c;
    c.kind = 0;
    c.G = g[0];
    c.U = g[1];
    c.grid = grid_for(S->n, c.G);
    if (S->o.spmv == PG_SPMV_AUTO || S->o.spmv == PG_SPMV_ROW || only_eop) cands.push_back(c);
    if (S->er_ok && S->o.spmv != PG_SPMV_ROW && !only_eop) {
      SpmvCfg e = c;
      e.kind = 1;
      e.D = S->maxblk <= 1 ? 1 : (S->maxblk <= 2 ? 2 : 6);
      e.grid = grid_for(S->nblocks, e.G);
      if (S->o.spmv != PG_SPMV_PAIRS) cands.push_back(e);
      if (S->pairs_ok && S->o.spmv != PG_SPMV_ENTITY) {
        e.kind = 2;
        cands.push_back(e);
      }
    }
  }
  // timing inputs: x = 1 (finite), y = FP64 scratch q (allocated in every mode);
  // scalars cleared so that the MODE 1 kernels run to completion
  hipLaunchKernelGGL(k_init_scalars, dim3(1), dim3(64), 0, st, S->s, 1, 0.0);
  hipLaunchKernelGGL(k_ones, dim3(nblk(S->n, kBS)), dim3(kBS), 0, st, S->n, S->r64);
  // the CSR kernels: the true residual's, and the candidates against the
  // element-by-element operator (kind 3), which is tried when the caller gave one
  S->cres = pick_spmv<double2, double2>(S, cands, st, S->vals, S->r64, S->q);
  SpmvCfg ce;
  ce.kind = 3;
  ce.G = ce.U = 0;
  if (S->eop) ce.grid = eop_grid(S->eop);
  S->c64 = S->cres;
  // (an operator without a complex128 product, p = 3, serves complex64 only;
  // the FP64 recursion then keeps the CSR kernel)
  const bool eop64 = S->eop && eop_supports(S->eop, true);
  if (only_eop && !eop64 && !mixed)
    PG_FAIL(PG_ERR_UNSUPPORTED, "the element operator at p = 3 has no complex128 product");
  if (only_eop && eop64) {
    S->c64 = ce;
  } else if (eop64 && S->o.spmv == PG_SPMV_AUTO) {
    const float t_csr = time_spmv<double2, double2>(S, S->cres, st, S->vals, S->r64, S->q, 20);
    const float t_eop = time_spmv<double2, double2>(S, ce, st, S->vals, S->r64, S->q, 20);
    if (t_eop < t_csr) S->c64 = ce;
  }
  if (mixed) {
    hipLaunchKernelGGL(k_ones32, dim3(nblk(S->n, kBS)), dim3(kBS), 0, st, S->n, S->p32);
    if (only_eop) {
      S->c32 = ce;
    } else {
      // the complex64 copy has one layout: schedule order when entity rows exist
      std::vector<SpmvCfg> c32;
      for (const SpmvCfg& c : cands)
        if ((c.kind >= 1) == S->er_ok) c32.push_back(c);
      if (c32.empty()) PG_FAIL(PG_ERR_ARG, "no SpMV kernel for the complex64 layout");
      S->c32 = pick_spmv<float2, float2>(S, c32, st, S->vals32, S->p32, S->q32);
      if (S->eop && S->o.spmv == PG_SPMV_AUTO) {
        const float t_csr =
            time_spmv<float2, float2>(S, S->c32, st, S->vals32, S->p32, S->q32, 20);
        const float t_eop = time_spmv<float2, float2>(S, ce, st, S->vals32, S->p32, S->q32, 20);
        if (t_eop < t_csr) S->c32 = ce;
      }
    }
  }
  PG_CHECK_LAUNCH();
  return PG_OK;
}

}  // namespace pg

using namespace pg;

extern "C" {

int pg_spmv_z(int64_t n, const int32_t* rp, const int32_t* ci, const double* v, const double* x,
              double* y, void* stream) {
  if (n <= 0 || !rp || !ci || !v || !x || !y) PG_FAIL(PG_ERR_ARG, "bad arguments");
  SpmvCfg c;
  c.G = 16;
  c.U = 2;
  c.grid = grid_for(n, 16);
  spmv_launch<0, double2, double2>(c, as_stream(stream), nullptr, n, 0, nullptr, rp, ci,
                                   reinterpret_cast<const double2*>(v),
                                   reinterpret_cast<const double2*>(x),
                                   reinterpret_cast<double2*>(y), nullptr, nullptr);
  PG_CHECK_LAUNCH();
  return PG_OK;
}

int pg_spmv_c(int64_t n, const int32_t* rp, const int32_t* ci, const float* v, const float* x,
              float* y, void* stream) {
  if (n <= 0 || !rp || !ci || !v || !x || !y) PG_FAIL(PG_ERR_ARG, "bad arguments");
  SpmvCfg c;
  c.G = 16;
  c.U = 2;
  c.grid = grid_for(n, 16);
  spmv_launch<0, float2, float2>(c, as_stream(stream), nullptr, n, 0, nullptr, rp, ci,
                                 reinterpret_cast<const float2*>(v),
                                 reinterpret_cast<const float2*>(x), reinterpret_cast<float2*>(y),
                                 nullptr, nullptr);
  PG_CHECK_LAUNCH();
  return PG_OK;
}

int pg_demote_z2c(int64_t n, const double* src, float* dst, void* stream) {
  if (n <= 0 || !src || !dst) PG_FAIL(PG_ERR_ARG, "bad arguments");
  hipLaunchKernelGGL(k_demote, dim3(nblk(2 * n, kBS)), dim3(kBS), 0, as_stream(stream), 2 * n,
                     src, dst);
  PG_CHECK_LAUNCH();
  return PG_OK;
}

int pg_solver_create(int64_t n, int64_t nnz, const int32_t* rowptr, const int32_t* cols,
                     const double* values, const pg_solve_opts* opts, void** out, void* stream) {
  if (!out || !opts || n <= 0 || nnz <= 0 || !rowptr || !cols || !values)
    PG_FAIL(PG_ERR_ARG, "bad arguments");
  if (opts->ksp != PG_KSP_COCG && opts->ksp != PG_KSP_GMRES)
    PG_FAIL(PG_ERR_UNSUPPORTED, "ksp type %d not supported", opts->ksp);
  if (opts->ksp == PG_KSP_GMRES && opts->precision != PG_PREC_FP64)
    PG_FAIL(PG_ERR_UNSUPPORTED, "GMRES runs in FP64 only (use COCG for mixed precision)");
  if (opts->ksp == PG_KSP_GMRES && (opts->restart < 1 || opts->restart > kMaxRestart))
    PG_FAIL(PG_ERR_ARG, "GMRES restart must be in 1..%d", kMaxRestart);
  if (opts->pc != PG_PC_NONE && opts->pc != PG_PC_JACOBI && opts->pc != PG_PC_BJACOBI)
    PG_FAIL(PG_ERR_UNSUPPORTED, "pc type %d not supported", opts->pc);
  if (!(opts->rtol > 0.0)) PG_FAIL(PG_ERR_ARG, "rtol must be positive");
  if (opts->spmv < PG_SPMV_AUTO || opts->spmv > PG_SPMV_ELEMENT)
    PG_FAIL(PG_ERR_ARG, "spmv kind %d not supported", opts->spmv);
  if (opts->spmv == PG_SPMV_ELEMENT && !opts->eop)
    PG_FAIL(PG_ERR_ARG, "PG_SPMV_ELEMENT needs an element operator (pg_eop_create)");
  if (opts->eop && eop_ndof(static_cast<const ElemOp*>(opts->eop)) != n)
    PG_FAIL(PG_ERR_ARG, "the element operator has a different number of DOFs");
  std::vector<int32_t> hboff;
  int maxblk = 0, minblk = 1 << 30;
  if (opts->pc == PG_PC_BJACOBI && (!opts->block_ptr || opts->nblocks <= 0))
    PG_FAIL(PG_ERR_ARG, "block Jacobi needs block_ptr and nblocks");
  if (opts->block_ptr && opts->nblocks > 0) {
    const int32_t* bp = opts->block_ptr;
    if (bp[0] != 0 || bp[opts->nblocks] != n)
      PG_FAIL(PG_ERR_ARG, "block_ptr must start at 0 and end at n");
    hboff.resize(opts->nblocks + 1);
    hboff[0] = 0;
    for (int64_t k = 0; k < opts->nblocks; ++k) {
      const int d = bp[k + 1] - bp[k];
      if (d < 1 || d > kMaxBlock) PG_FAIL(PG_ERR_ARG, "block sizes must be in 1..%d", kMaxBlock);
      maxblk = std::max(maxblk, d);
      minblk = std::min(minblk, d);
      if ((int64_t)hboff[k] + d * d >= (1LL << 31)) PG_FAIL(PG_ERR_ARG, "block storage too large");
      hboff[k + 1] = hboff[k] + d * d;
    }
  }
  hipStream_t st = as_stream(stream);
  Solver* S = new Solver();
  S->n = n;
  S->nnz = nnz;
  S->o = *opts;
  S->o.block_ptr = nullptr;  // the host pointers are not kept
  S->o.entity_order = nullptr;
  S->eop = static_cast<const ElemOp*>(opts->eop);
  if (S->o.maxit <= 0) S->o.maxit = 100000;
  if (S->o.check_every <= 0) S->o.check_every = 25;
  if (!(S->o.inner_rtol > 0.0) || S->o.inner_rtol >= 1.0) S->o.inner_rtol = 1e-4;
  S->rowptr = rowptr;
  S->cols = cols;
  S->vals = reinterpret_cast<const double2*>(values);
  S->gv = vgrid_for(n);
  const bool mixed = opts->precision == PG_PREC_MIXED;
  int rc = PG_OK;
#define PG_A(ptr, cnt)                           \
  if ((rc = S->alloc(&(ptr), (cnt))) != PG_OK) { \
    delete S;                                    \
    return rc;                                   \
  }
  PG_A(S->s, 1);
  PG_A(S->part, 4 * kMaxGrid);
  PG_A(S->r64, n);
  PG_A(S->q, n);
  PG_A(S->bad, 1);
  PG_A(S->dnorm, 1);
  PG_A(S->skip, 1);
  // FP64 recursion vectors (also in mixed precision: the fallback when a
  // complex64 round stagnates)
  PG_A(S->r, n);
  PG_A(S->p, n);
  PG_A(S->z, n);
  PG_A(S->dinv, n);
  if (mixed) {
    PG_A(S->vals32, nnz);
    PG_A(S->r32, n);
    PG_A(S->p32, n);
    PG_A(S->q32, n);
    PG_A(S->z32, n);
    PG_A(S->x32, n);
    PG_A(S->dinv32, n);
  }
  if (opts->block_ptr && opts->nblocks > 0) {
    S->nblocks = opts->nblocks;
    S->maxblk = maxblk;
    S->blk2 = maxblk == 2 && minblk == 2;
    S->gb = vgrid_for(S->nblocks);
    PG_A(S->bptr, S->nblocks + 1);
    PG_A(S->boff, S->nblocks + 1);
    PG_A(S->emeta, S->nblocks);
    if (opts->pc == PG_PC_BJACOBI) {
      PG_A(S->binv, hboff[S->nblocks]);
      if (mixed) PG_A(S->binv32, hboff[S->nblocks]);
    }
    if (hipMemcpy(S->bptr, opts->block_ptr, (S->nblocks + 1) * sizeof(int32_t),
                  hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(S->boff, hboff.data(), (S->nblocks + 1) * sizeof(int32_t),
                  hipMemcpyHostToDevice) != hipSuccess) {
      delete S;
      PG_FAIL(PG_ERR_HIP, "block upload failed");
    }
  }
  if (opts->ksp == PG_KSP_GMRES) {
    S->restart = opts->restart;
    PG_A(S->V, (size_t)(S->restart + 1) * n);
    PG_A(S->hdev, 2 * (S->restart + 1));
    PG_A(S->partg, (size_t)2 * (S->restart + 1) * kMaxGrid);
    if (hipHostMalloc(reinterpret_cast<void**>(&S->hhost),
                      2 * (S->restart + 1) * sizeof(double2), 0) != hipSuccess) {
      delete S;
      PG_FAIL(PG_ERR_HIP, "hipHostMalloc failed");
    }
  }
#undef PG_A
  if (hipHostMalloc(reinterpret_cast<void**>(&S->h), sizeof(Scal), 0) != hipSuccess ||
      hipEventCreate(&S->ev[0]) != hipSuccess || hipEventCreate(&S->ev[1]) != hipSuccess ||
      hipEventCreate(&S->ev[2]) != hipSuccess) {
    delete S;
    PG_FAIL(PG_ERR_HIP, "hipHostMalloc / hipEventCreate failed");
  }
  hipLaunchKernelGGL(k_init_scalars, dim3(1), dim3(64), 0, st, S->s, 0, 0.0);
  if (S->nblocks > 0 &&
      (rc = setup_entities(S, opts->block_ptr, opts->entity_order, st)) != PG_OK) {
    delete S;
    return rc;
  }
  if ((rc = refresh_values(S, st)) != PG_OK || (rc = autotune(S, st)) != PG_OK) {
    delete S;
    return rc;
  }
  if (hipStreamSynchronize(st) != hipSuccess) {
    delete S;
    PG_FAIL(PG_ERR_HIP, "solver setup failed");
  }
  *out = S;
  return PG_OK;
}

int pg_solver_refresh(void* solver, void* stream) {
  if (!solver) PG_FAIL(PG_ERR_ARG, "solver is NULL");
  return refresh_values(static_cast<Solver*>(solver), as_stream(stream));
}

int pg_solver_solve(void* solver, const double* b, double* x, pg_solve_info* info, void* stream) {
  if (!solver || !b || !x || !info) PG_FAIL(PG_ERR_ARG, "bad arguments");
  Solver* S = static_cast<Solver*>(solver);
  *info = pg_solve_info{};
  if (S->o.ksp == PG_KSP_GMRES)
    return gmres_solve(S, reinterpret_cast<const double2*>(b), reinterpret_cast<double2*>(x),
                       info, as_stream(stream));
  return cocg_solve(S, reinterpret_cast<const double2*>(b), reinterpret_cast<double2*>(x), info,
                    as_stream(stream));
}

int pg_solver_profile(void* solver, int reps, double* ms64, double* ms32, int32_t* g64,
                      int32_t* g32, void* stream) {
  if (!solver || reps <= 0 || !ms64 || !ms32 || !g64 || !g32) PG_FAIL(PG_ERR_ARG, "bad arguments");
  Solver* S = static_cast<Solver*>(solver);
  hipStream_t st = as_stream(stream);
  const bool mixed = S->o.precision == PG_PREC_MIXED;
  hipEvent_t a, b;
  PG_HIP(hipEventCreate(&a));
  PG_HIP(hipEventCreate(&b));
  // the solve's own SpMV kernel (MODE 1: q = A p and the mu partials) on a finite
  // p, with the scalars reset so that the early-exit flag is clear
  hipLaunchKernelGGL(k_init_scalars, dim3(1), dim3(64), 0, st, S->s, 1, 0.0);
  float t = 0.f;
  *ms32 = 0.0;
  if (mixed) {
    hipLaunchKernelGGL(k_ones32, dim3(nblk(S->n, kBS)), dim3(kBS), 0, st, S->n, S->p32);
    PG_HIP(hipEventRecord(a, st));
    for (int i = 0; i < reps; ++i)
      spmv<1, float2, float2>(S, S->c32, st, S->vals32, S->p32, S->q32, nullptr);
    PG_HIP(hipEventRecord(b, st));
    PG_HIP(hipEventSynchronize(b));
    PG_HIP(hipEventElapsedTime(&t, a, b));
    *ms32 = t / reps;
  }
  hipLaunchKernelGGL(k_ones, dim3(nblk(S->n, kBS)), dim3(kBS), 0, st, S->n, S->r64);
  PG_HIP(hipEventRecord(a, st));
  for (int i = 0; i < reps; ++i)
    spmv<1, double2, double2>(S, S->c64, st, S->vals, S->r64, S->q, nullptr);
  PG_HIP(hipEventRecord(b, st));
  PG_HIP(hipEventSynchronize(b));
  PG_HIP(hipEventElapsedTime(&t, a, b));
  *ms64 = t / reps;
  // configuration code: kind*1000 + G*10 + U
  *g64 = S->c64.kind * 1000 + S->c64.G * 10 + S->c64.U;
  *g32 = S->c32.kind * 1000 + S->c32.G * 10 + S->c32.U;
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  PG_CHECK_LAUNCH();
  return PG_OK;
}

int pg_solver_destroy(void* solver) {
  delete static_cast<Solver*>(solver);
  return PG_OK;
}

}  // extern "C"
