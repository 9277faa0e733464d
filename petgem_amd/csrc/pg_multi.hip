// Multi-mesh batched solve for gfx950: many independent meshes of uneven size
// (each with its own assembled complex system and right-hand side) advance
// together, one COCG iteration of every mesh per set of kernel launches.
//
// No counterpart in the reference (one KSP per problem, petgem/solver.py:584-590).
// The systems are concatenated into one block-diagonal CSR.  Workgroups are
// dealt to the meshes in proportion to their rows (at least one each), so
// every workgroup belongs to exactly one mesh: its partial sums go to that
// mesh's slots, and a one-workgroup-per-mesh finish kernel reduces them in a
// fixed order (bitwise reproducible, no device-scope fences).  Each mesh runs
// its own COCG + Jacobi recurrence (its own alpha, beta, rho, convergence) and
// drops out of the work as soon as it has converged.
//
// Persistence: the `check` iterations between host convergence checks (six
// kernels each) are captured once into a HIP graph at creation and replayed
// on the batch's own stream until every mesh has converged, so a batch of
// small meshes costs one graph launch per `check` iterations instead of
// 6 * check kernel launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <type_traits>
#include <vector>

#include "pg_blockinv.h"
#include "pg_common.h"
#include "pg_eop.h"

namespace pg {
namespace multi {

constexpr int kBS = 256;
constexpr int kRW = 256;  // rows per workgroup
constexpr int kG = 16;    // lanes per row in the SpMV

struct Seg {  // per-mesh scalars
  double2 rho, mu, alpha, beta;
  double res2, bnorm2;
  int32_t iter, done, status, pad;
  // mixed precision: FP64 |b - A x|^2 at the last refresh, the complex64
  // round's target |r|^2 and its stagnation watch; conv = converged (or
  // stopped), fp64 = its complex64 rounds stopped gaining (the mesh finishes in
  // FP64), refreshed = the FP64 residual has been formed once
  double res64, target2, best2;
  int32_t last_imp, conv, fp64, refreshed;
};

struct Ctl {
  double tol2;
  int32_t maxit, nseg, all_done, pad;
};

// the rows of workgroup w: segment wseg[w], rows [lo, hi)
struct Span {
  int s;
  int64_t lo, hi;
};
__device__ inline Span span_of(const int32_t* __restrict__ wseg, const int32_t* __restrict__ wg0,
                               const int64_t* __restrict__ row0) {
  const int w = blockIdx.x;
  const int s = wseg[w];
  const int64_t lo = row0[s] + (int64_t)(w - wg0[s]) * kRW;
  const int64_t hi = lo + kRW < row0[s + 1] ? lo + kRW : row0[s + 1];
  return {s, lo, hi};
}

// this workgroup's NV partial sums -> part[k * W + w]
template <int NV>
__device__ inline void write_partials(double (&v)[NV], double* part) {
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
      part[(int64_t)k * gridDim.x + blockIdx.x] = s;
    }
}

// in a finish kernel (one workgroup per segment s = blockIdx.x): fixed-order
// totals of the partials of workgroups [wg0[s], wg0[s+1]) out of W
template <int NV>
__device__ inline void seg_totals(const double* part, int W, int w0, int w1, double (&tot)[NV]) {
  double acc[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) acc[k] = 0.0;
  for (int i = w0 + threadIdx.x; i < w1; i += kBS)
#pragma unroll
    for (int k = 0; k < NV; ++k) acc[k] += part[(int64_t)k * W + i];
  __shared__ double sm[NV][kBS / kWave];
  const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
#pragma unroll
  for (int k = 0; k < NV; ++k) acc[k] = wave_sum(acc[k]);
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) sm[k][w] = acc[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double s = 0.0;
    for (int i = 0; i < kBS / kWave; ++i) s += sm[k][i];
    tot[k] = s;
  }
}

__device__ inline double2 cfma(double2 a, double2 b, double2 c) {
  return make_double2(fma(a.x, b.x, fma(-a.y, b.y, c.x)), fma(a.x, b.y, fma(a.y, b.x, c.y)));
}
__device__ inline float2 cfma(float2 a, float2 b, float2 c) {
  return make_float2(fmaf(a.x, b.x, fmaf(-a.y, b.y, c.x)), fmaf(a.x, b.y, fmaf(a.y, b.x, c.y)));
}
__device__ inline double2 to_d(double2 v) { return v; }
__device__ inline double2 to_d(float2 v) { return make_double2(v.x, v.y); }
template <typename T>
__device__ inline T from_d(double2 v);
template <>
__device__ inline double2 from_d<double2>(double2 v) { return v; }
template <>
__device__ inline float2 from_d<float2>(double2 v) { return make_float2((float)v.x, (float)v.y); }
template <typename T>
struct PairOf;  // two consecutive complex values
template <>
struct PairOf<double2> {
  using type = double4;
};
template <>
struct PairOf<float2> {
  using type = float4;
};

// y = A x over this workgroup's rows (kG lanes per row); MODE 0: y = A x and
// partials of x^T y; MODE 1: y = b - A x and partials of |y|^2
template <int MODE>
__device__ inline void spmv_rows(const Span& sp, const int32_t* __restrict__ rowptr,
                                 const int32_t* __restrict__ cols,
                                 const double2* __restrict__ vals, const double2* __restrict__ x,
                                 double2* __restrict__ y, const double2* __restrict__ b,
                                 double (&red)[2]) {
  const int lane = threadIdx.x % kG;
  for (int64_t row = sp.lo + threadIdx.x / kG; row < sp.hi; row += kBS / kG) {
    double2 acc = make_double2(0.0, 0.0);
    const int s0 = rowptr[row], s1 = rowptr[row + 1];
    for (int kb = s0 + lane; kb < s1; kb += 2 * kG) {
      const int k1 = kb + kG;
      const int c0 = cols[kb], c1 = k1 < s1 ? cols[k1] : -1;
      const double2 a0 = vals[kb], a1 = k1 < s1 ? vals[k1] : make_double2(0.0, 0.0);
      const double2 x0 = x[c0], x1 = c1 >= 0 ? x[c1] : make_double2(0.0, 0.0);
      acc = cfma(a0, x0, acc);
      acc = cfma(a1, x1, acc);
    }
    acc.x = group_sum<kG>(acc.x);
    acc.y = group_sum<kG>(acc.y);
    if (lane == 0) {
      if (MODE == 1) {
        const double2 r = make_double2(b[row].x - acc.x, b[row].y - acc.y);
        y[row] = r;
        red[0] += r.x * r.x + r.y * r.y;
      } else {
        y[row] = acc;
        const double2 p = x[row];
        red[0] += p.x * acc.x - p.y * acc.y;
        red[1] += p.x * acc.y + p.y * acc.x;
      }
    }
  }
}

// p = z + beta_s p
template <typename VT>
__global__ void __launch_bounds__(kBS) k_mpupdate(const Ctl* ctl, const Seg* seg,
                                                  const int32_t* __restrict__ wseg,
                                                  const int32_t* __restrict__ wg0,
                                                  const int64_t* __restrict__ row0,
                                                  const VT* __restrict__ z, VT* __restrict__ p) {
  if (ctl->all_done) return;
  const Span sp = span_of(wseg, wg0, row0);
  if (seg[sp.s].done) return;
  const VT be = from_d<VT>(seg[sp.s].beta);
  for (int64_t i = sp.lo + threadIdx.x; i < sp.hi; i += kBS) p[i] = cfma(be, p[i], z[i]);
}

// q = A p, partials of p^T q
__global__ void __launch_bounds__(kBS) k_mspmv(const Ctl* ctl, const Seg* seg,
                                               const int32_t* __restrict__ wseg,
                                               const int32_t* __restrict__ wg0,
                                               const int64_t* __restrict__ row0,
                                               const int32_t* __restrict__ rowptr,
                                               const int32_t* __restrict__ cols,
                                               const double2* __restrict__ vals,
                                               const double2* __restrict__ p,
                                               double2* __restrict__ q, double* part) {
  if (ctl->all_done) return;
  const Span sp = span_of(wseg, wg0, row0);
  if (seg[sp.s].done) return;
  double red[2] = {0.0, 0.0};
  spmv_rows<0>(sp, rowptr, cols, vals, p, q, nullptr, red);
  write_partials<2>(red, part);
}

__global__ void __launch_bounds__(kBS) k_mfin_mu(const Ctl* ctl, Seg* seg,
                                                 const int32_t* __restrict__ wg0,
                                                 const double* part, int W) {
  if (ctl->all_done) return;
  const int s = blockIdx.x;
  if (seg[s].done) return;
  double tot[2];
  seg_totals<2>(part, W, wg0[s], wg0[s + 1], tot);
  if (threadIdx.x != 0) return;
  Seg& c = seg[s];
  const double2 mu = make_double2(tot[0], tot[1]);
  c.mu = mu;
  const double2 al = zdiv(c.rho, mu);
  if (!(mu.x != 0.0 || mu.y != 0.0) || !isfinite(al.x) || !isfinite(al.y)) {
    c.status = PG_ERR_BREAKDOWN;
    c.done = 1;
    c.alpha = make_double2(0.0, 0.0);
  } else {
    c.alpha = al;
  }
}

// x += alpha p; r -= alpha q; z = dinv r; partials of r^T z, |r|^2
__global__ void __launch_bounds__(kBS) k_mupdate(const Ctl* ctl, const Seg* seg,
                                                 const int32_t* __restrict__ wseg,
                                                 const int32_t* __restrict__ wg0,
                                                 const int64_t* __restrict__ row0,
                                                 double2* __restrict__ x, double2* __restrict__ r,
                                                 const double2* __restrict__ p,
                                                 const double2* __restrict__ q,
                                                 const double2* __restrict__ dinv,
                                                 double2* __restrict__ z, double* part) {
  if (ctl->all_done) return;
  const Span sp = span_of(wseg, wg0, row0);
  if (seg[sp.s].done) return;
  const double2 a = seg[sp.s].alpha;
  const double2 na = make_double2(-a.x, -a.y);
  double red[3] = {0.0, 0.0, 0.0};
  for (int64_t i = sp.lo + threadIdx.x; i < sp.hi; i += kBS) {
    x[i] = cfma(a, p[i], x[i]);
    const double2 ri = cfma(na, q[i], r[i]);
    r[i] = ri;
    const double2 zi = zmul(dinv[i], ri);
    z[i] = zi;
    red[0] += ri.x * zi.x - ri.y * zi.y;
    red[1] += ri.x * zi.y + ri.y * zi.x;
    red[2] += ri.x * ri.x + ri.y * ri.y;
  }
  write_partials<3>(red, part);
}

__global__ void __launch_bounds__(kBS) k_mfin_update(const Ctl* ctl, Seg* seg,
                                                     const int32_t* __restrict__ wg0,
                                                     const double* part, int W, int stall_limit) {
  if (ctl->all_done) return;
  const int s = blockIdx.x;
  if (seg[s].done) return;
  double tot[3];
  seg_totals<3>(part, W, wg0[s], wg0[s + 1], tot);
  if (threadIdx.x != 0) return;
  Seg& c = seg[s];
  const double2 rho = make_double2(tot[0], tot[1]);
  const double2 beta = zdiv(rho, c.rho);
  c.beta = beta;
  c.rho = rho;
  c.res2 = tot[2];
  c.iter += 1;
  if (!isfinite(tot[2]) || !isfinite(beta.x) || !isfinite(beta.y)) {
    c.status = PG_ERR_BREAKDOWN;
    c.done = 1;
    c.conv = 1;
  } else if (stall_limit > 0) {  // a complex64 round: ends at its target or stagnated
    if (tot[2] <= c.target2 || c.iter >= ctl->maxit) {
      c.done = 1;
    } else if (tot[2] < 0.81 * c.best2) {
      c.best2 = tot[2];
      c.last_imp = c.iter;
    } else if (c.iter - c.last_imp >= stall_limit) {
      c.done = 1;
    }
  } else if (tot[2] <= ctl->tol2 * c.bnorm2 || c.iter >= ctl->maxit) {
    c.done = 1;
  }
}

// all_done = every mesh done (one workgroup)
__global__ void __launch_bounds__(kBS) k_mfin_all(Ctl* ctl, const Seg* seg, int nseg) {
  if (ctl->all_done) return;
  __shared__ int notdone;
  if (threadIdx.x == 0) notdone = 0;
  __syncthreads();
  int mine = 0;
  for (int s = threadIdx.x; s < nseg; s += kBS) mine |= seg[s].done ? 0 : 1;
  if (mine) notdone = 1;  // every writer stores the same value
  __syncthreads();
  if (threadIdx.x == 0) ctl->all_done = notdone ? 0 : 1;
}

// (re)start from the current x: r = b - A x, z = dinv r, p = 0; partials of r^T z, |r|^2
__global__ void __launch_bounds__(kBS) k_mrestart(const int32_t* __restrict__ wseg,
                                                  const int32_t* __restrict__ wg0,
                                                  const int64_t* __restrict__ row0,
                                                  const int32_t* __restrict__ rowptr,
                                                  const int32_t* __restrict__ cols,
                                                  const double2* __restrict__ vals,
                                                  const double2* __restrict__ x,
                                                  const double2* __restrict__ b,
                                                  double2* __restrict__ r,
                                                  const double2* __restrict__ dinv,
                                                  double2* __restrict__ z, double2* __restrict__ p,
                                                  double* part) {
  const Span sp = span_of(wseg, wg0, row0);
  double red2[2] = {0.0, 0.0};
  spmv_rows<1>(sp, rowptr, cols, vals, x, r, b, red2);
  __syncthreads();  // r of this workgroup's rows is complete
  double red[3] = {0.0, 0.0, 0.0};
  for (int64_t i = sp.lo + threadIdx.x; i < sp.hi; i += kBS) {
    const double2 ri = r[i];
    const double2 zi = zmul(dinv[i], ri);
    z[i] = zi;
    p[i] = make_double2(0.0, 0.0);
    red[0] += ri.x * zi.x - ri.y * zi.y;
    red[1] += ri.x * zi.y + ri.y * zi.x;
    red[2] += ri.x * ri.x + ri.y * ri.y;
  }
  write_partials<3>(red, part);
}

// finish a (re)start: rho, |r|^2 (and |b|^2 the first time); meshes that meet
// the tolerance, stopped on breakdown or used up their iterations stay done
__global__ void __launch_bounds__(kBS) k_mfin_restart(Ctl* ctl, Seg* seg,
                                                      const int32_t* __restrict__ wg0,
                                                      const double* part, int W, int first) {
  const int s = blockIdx.x;
  double tot[3];
  seg_totals<3>(part, W, wg0[s], wg0[s + 1], tot);
  if (threadIdx.x != 0) return;
  Seg& c = seg[s];
  if (first) {
    c = Seg{};
    c.bnorm2 = tot[2];  // x = 0 on the first start
  }
  c.rho = make_double2(tot[0], tot[1]);
  c.res2 = tot[2];
  c.beta = make_double2(0.0, 0.0);
  c.alpha = make_double2(0.0, 0.0);
  c.done = (c.status != 0 || c.iter >= ctl->maxit || tot[2] <= ctl->tol2 * c.bnorm2) ? 1 : 0;
  if (s == 0) ctl->all_done = 0;
}

// |b - A x|^2 per mesh (after the solve)
__global__ void __launch_bounds__(kBS) k_mtrue(const int32_t* __restrict__ wseg,
                                               const int32_t* __restrict__ wg0,
                                               const int64_t* __restrict__ row0,
                                               const int32_t* __restrict__ rowptr,
                                               const int32_t* __restrict__ cols,
                                               const double2* __restrict__ vals,
                                               const double2* __restrict__ x,
                                               const double2* __restrict__ b,
                                               double2* __restrict__ r, double* part) {
  const Span sp = span_of(wseg, wg0, row0);
  double red[2] = {0.0, 0.0};
  spmv_rows<1>(sp, rowptr, cols, vals, x, r, b, red);
  double v[1] = {red[0]};
  write_partials<1>(v, part);
}
__global__ void __launch_bounds__(kBS) k_mfin_true(Seg* seg, const int32_t* __restrict__ wg0,
                                                   const double* part, int W) {
  const int s = blockIdx.x;
  double tot[1];
  seg_totals<1>(part, W, wg0[s], wg0[s + 1], tot);
  if (threadIdx.x == 0) seg[s].res2 = tot[0];
}

__global__ void k_mctl_init(Ctl* ctl, int maxit, double tol2, int nseg) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  Ctl c{};
  c.tol2 = tol2;
  c.maxit = maxit;
  c.nseg = nseg;
  *ctl = c;
}

// concatenation: rowptr of segment s shifted by its first non-zero, columns by its first row
__global__ void k_shift_i32(int64_t n, const int32_t* __restrict__ src, int32_t add,
                            int32_t* __restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i] + add;
}

// Jacobi: dinv = 1 / A_ii (the diagonal of the concatenated, column-shifted CSR)
__global__ void k_mjacobi(int64_t n, const int32_t* __restrict__ rowptr,
                          const int32_t* __restrict__ cols, const double2* __restrict__ vals,
                          double2* __restrict__ dinv, int32_t* bad) {
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
    dinv[r] = make_double2(1.0, 0.0);
    return;
  }
  dinv[r] = zdiv(make_double2(1.0, 0.0), d);
}

// ---------------------------------------------------------------------------
// Element products and entity-block Jacobi (pg_multi_set_element_operator).
// The meshes' entities are numbered mesh after mesh, so segment s owns the
// entities [epos0[s], epos0[s+1]) both in index order (the DOF ranges bptr,
// the block inverses at boff) and in the operator's schedule (gm, gslot).
// Workgroups of kBS entities are dealt to the segments, ewseg / ewg0 as
// wseg / wg0 for the rows.
// ---------------------------------------------------------------------------
__device__ inline Span espan_of(const int32_t* __restrict__ ewseg,
                                const int32_t* __restrict__ ewg0,
                                const int64_t* __restrict__ epos0) {
  const int w = blockIdx.x;
  const int s = ewseg[w];
  const int64_t lo = epos0[s] + (int64_t)(w - ewg0[s]) * kBS;
  const int64_t hi = lo + kBS < epos0[s + 1] ? lo + kBS : epos0[s + 1];
  return {s, lo, hi};
}

// q = A p from the local results: per scheduled entity the sum of its
// incidences in ascending slot order (PEC entity: identity row); partials of p^T q
template <int D, typename VT>
__global__ void __launch_bounds__(kBS) k_esum(const Ctl* ctl, const Seg* seg,
                                              const int32_t* __restrict__ ewseg,
                                              const int32_t* __restrict__ ewg0,
                                              const int64_t* __restrict__ epos0,
                                              const int4* __restrict__ gm,
                                              const int32_t* __restrict__ gslot,
                                              const VT* __restrict__ ye, const VT* __restrict__ p,
                                              VT* __restrict__ q, double* part) {
  if (ctl->all_done) return;
  const Span sp = espan_of(ewseg, ewg0, epos0);
  if (seg[sp.s].done) return;
  double red[2] = {0.0, 0.0};
  const int64_t i = sp.lo + threadIdx.x;
  if (i < sp.hi) {
    const int4 m = gm[i];
    const int r0 = m.x, d = m.y, s0 = m.z, s1 = m.w;
    const VT zero = from_d<VT>(make_double2(0.0, 0.0));
    VT pr[D], acc[D];
#pragma unroll
    for (int t = 0; t < D; ++t) {
      pr[t] = t < d ? p[r0 + t] : zero;
      acc[t] = s0 == s1 ? pr[t] : zero;
    }
    int k = s0;
    for (; k + 2 <= s1; k += 2) {  // two incidences' loads in flight together
      const int b0 = gslot[k], b1 = gslot[k + 1];
      VT v0[D], v1[D];
#pragma unroll
      for (int t = 0; t < D; ++t)
        if (t < d) {
          v0[t] = ye[b0 + t * kWave];
          v1[t] = ye[b1 + t * kWave];
        }
#pragma unroll
      for (int t = 0; t < D; ++t)
        if (t < d) {
          acc[t].x += v0[t].x;
          acc[t].y += v0[t].y;
          acc[t].x += v1[t].x;
          acc[t].y += v1[t].y;
        }
    }
    if (k < s1) {
      const int b0 = gslot[k];
#pragma unroll
      for (int t = 0; t < D; ++t)
        if (t < d) {
          const VT v = ye[b0 + t * kWave];
          acc[t].x += v.x;
          acc[t].y += v.y;
        }
    }
#pragma unroll
    for (int t = 0; t < D; ++t)
      if (t < d) {
        q[r0 + t] = acc[t];
        const double2 pd = to_d(pr[t]), qd = to_d(acc[t]);
        red[0] += pd.x * qd.x - pd.y * qd.y;
        red[1] += pd.x * qd.y + pd.y * qd.x;
      }
  }
  write_partials<2>(red, part);
}

// x += alpha p; r -= alpha q; z = B r (entity blocks); partials of r^T z, |r|^2
template <int D, typename VT>
__global__ void __launch_bounds__(kBS) k_eupdate(const Ctl* ctl, const Seg* seg,
                                                 const int32_t* __restrict__ ewseg,
                                                 const int32_t* __restrict__ ewg0,
                                                 const int64_t* __restrict__ epos0,
                                                 const int32_t* __restrict__ bptr,
                                                 const int32_t* __restrict__ boff,
                                                 const VT* __restrict__ binv, VT* __restrict__ x,
                                                 VT* __restrict__ r, const VT* __restrict__ p,
                                                 const VT* __restrict__ q, VT* __restrict__ z,
                                                 double* part) {
  if (ctl->all_done) return;
  const Span sp = espan_of(ewseg, ewg0, epos0);
  if (seg[sp.s].done) return;
  const double2 ad = seg[sp.s].alpha;
  const VT a = from_d<VT>(ad), na = from_d<VT>(make_double2(-ad.x, -ad.y));
  double red[3] = {0.0, 0.0, 0.0};
  const int64_t k = sp.lo + threadIdx.x;
  if (k < sp.hi) {
    const int i0 = bptr[k], d = bptr[k + 1] - i0;
    const VT* B = binv + boff[k];
    VT rl[D];
#pragma unroll
    for (int t = 0; t < D; ++t)
      if (t < d) {
        x[i0 + t] = cfma(a, p[i0 + t], x[i0 + t]);
        rl[t] = cfma(na, q[i0 + t], r[i0 + t]);
        r[i0 + t] = rl[t];
      }
#pragma unroll
    for (int t = 0; t < D; ++t)
      if (t < d) {
        VT zt = from_d<VT>(make_double2(0.0, 0.0));
#pragma unroll
        for (int u = 0; u < D; ++u)
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

// the same when every entity has two DOFs 2k, 2k + 1 (p = 2; inverse at 4k):
// one 32-byte (complex128) or 16-byte (complex64) access per vector and
// entity, no index reads
template <typename VT>
__global__ void __launch_bounds__(kBS) k_eupdate2(const Ctl* ctl, const Seg* seg,
                                                  const int32_t* __restrict__ ewseg,
                                                  const int32_t* __restrict__ ewg0,
                                                  const int64_t* __restrict__ epos0,
                                                  const VT* __restrict__ binv,
                                                  VT* __restrict__ x, VT* __restrict__ r,
                                                  const VT* __restrict__ p,
                                                  const VT* __restrict__ q, VT* __restrict__ z,
                                                  double* part) {
  using V = typename PairOf<VT>::type;
  if (ctl->all_done) return;
  const Span sp = espan_of(ewseg, ewg0, epos0);
  if (seg[sp.s].done) return;
  const double2 ad = seg[sp.s].alpha;
  const VT a = from_d<VT>(ad), na = from_d<VT>(make_double2(-ad.x, -ad.y));
  const VT zero = from_d<VT>(make_double2(0.0, 0.0));
  double red[3] = {0.0, 0.0, 0.0};
  const int64_t k = sp.lo + threadIdx.x;
  if (k < sp.hi) {
    V* x2 = reinterpret_cast<V*>(x);
    V* r2 = reinterpret_cast<V*>(r);
    V* z2 = reinterpret_cast<V*>(z);
    const V xv = x2[k], rv = r2[k];
    const V pv = reinterpret_cast<const V*>(p)[k];
    const V qv = reinterpret_cast<const V*>(q)[k];
    const V B0 = reinterpret_cast<const V*>(binv)[2 * k];
    const V B1 = reinterpret_cast<const V*>(binv)[2 * k + 1];
    const VT xa = cfma(a, VT{pv.x, pv.y}, VT{xv.x, xv.y});
    const VT xb = cfma(a, VT{pv.z, pv.w}, VT{xv.z, xv.w});
    const VT ra = cfma(na, VT{qv.x, qv.y}, VT{rv.x, rv.y});
    const VT rb = cfma(na, VT{qv.z, qv.w}, VT{rv.z, rv.w});
    x2[k] = V{xa.x, xa.y, xb.x, xb.y};
    r2[k] = V{ra.x, ra.y, rb.x, rb.y};
    // z = B r, B row-major 2 x 2: B0 = (B00, B01), B1 = (B10, B11)
    const VT za = cfma(VT{B0.z, B0.w}, rb, cfma(VT{B0.x, B0.y}, ra, zero));
    const VT zb = cfma(VT{B1.z, B1.w}, rb, cfma(VT{B1.x, B1.y}, ra, zero));
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

// ---------------------------------------------------------------------------
// Mixed precision (pg_multi_set_mixed): each mesh iterates complex64 COCG
// rounds on a correction x32; a round stops at its target |r|^2 (inner
// tolerance times the round's first) or when it stagnates.  The host then
// refreshes the meshes whose round stopped: x += x32, r64 = b - A x in FP64
// (the CSR), and either the mesh has converged, or its round gained less than
// a factor 2 in |r| (it finishes in FP64 afterwards), or a new round starts
// from r64 rounded to complex64.
// ---------------------------------------------------------------------------
// x += x32, x32 = 0 over the rows of the meshes being refreshed
__global__ void __launch_bounds__(kBS) k_saccum(const int32_t* __restrict__ wseg,
                                                const int32_t* __restrict__ wg0,
                                                const int64_t* __restrict__ row0,
                                                double2* __restrict__ x, float2* __restrict__ x32) {
  const Span sp = span_of(wseg, wg0, row0);
  for (int64_t i = sp.lo + threadIdx.x; i < sp.hi; i += kBS) {
    const float2 c = x32[i];
    x[i] = make_double2(x[i].x + (double)c.x, x[i].y + (double)c.y);
    x32[i] = make_float2(0.f, 0.f);
  }
}

// the refreshed |b - A x|^2 of each mesh being refreshed (row partials of k_mtrue)
__global__ void __launch_bounds__(kBS) k_fin_srefresh(Ctl* ctl, Seg* seg,
                                                      const int32_t* __restrict__ wg0,
                                                      const double* part, int W) {
  const int s = blockIdx.x;
  if (wg0[s] == wg0[s + 1]) return;  // not being refreshed
  double tot[1];
  seg_totals<1>(part, W, wg0[s], wg0[s + 1], tot);
  if (threadIdx.x != 0) return;
  Seg& c = seg[s];
  const double r64 = tot[0];
  if (!c.refreshed) c.bnorm2 = r64;  // x = 0 at the first refresh
  const double prev = c.res64;
  c.res64 = r64;
  c.res2 = r64;
  c.done = 1;
  if (!isfinite(r64)) {
    c.status = PG_ERR_BREAKDOWN;
    c.conv = 1;
  } else if (r64 <= ctl->tol2 * c.bnorm2 || c.iter >= ctl->maxit) {
    c.conv = 1;
  } else if (c.refreshed && r64 > 0.25 * prev) {
    c.fp64 = 1;
  } else {
    c.done = 0;  // a new complex64 round
  }
  c.refreshed = 1;
}

// new round of the meshes being refreshed that go on: r32 = r64 rounded,
// z32 = B32 r32, p32 = 0; partials of r^T z and |r|^2
template <int D>
__global__ void __launch_bounds__(kBS) k_sround(const Seg* seg, const int32_t* __restrict__ ewseg,
                                                const int32_t* __restrict__ ewg0,
                                                const int64_t* __restrict__ epos0,
                                                const int32_t* __restrict__ bptr,
                                                const int32_t* __restrict__ boff,
                                                const float2* __restrict__ binv,
                                                const double2* __restrict__ r64,
                                                float2* __restrict__ r, float2* __restrict__ z,
                                                float2* __restrict__ p, double* part) {
  const Span sp = espan_of(ewseg, ewg0, epos0);
  if (seg[sp.s].done) return;
  double red[3] = {0.0, 0.0, 0.0};
  const int64_t k = sp.lo + threadIdx.x;
  if (k < sp.hi) {
    const int i0 = bptr[k], d = bptr[k + 1] - i0;
    const float2* B = binv + boff[k];
    float2 rl[D];
#pragma unroll
    for (int t = 0; t < D; ++t)
      if (t < d) {
        rl[t] = from_d<float2>(r64[i0 + t]);
        r[i0 + t] = rl[t];
        p[i0 + t] = make_float2(0.f, 0.f);
      }
#pragma unroll
    for (int t = 0; t < D; ++t)
      if (t < d) {
        float2 zt = make_float2(0.f, 0.f);
#pragma unroll
        for (int u = 0; u < D; ++u)
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

__global__ void __launch_bounds__(kBS) k_fin_sround(Ctl* ctl, Seg* seg,
                                                    const int32_t* __restrict__ ewg0,
                                                    const double* part, int We, double inner2) {
  const int s = blockIdx.x;
  if (ewg0[s] == ewg0[s + 1] || seg[s].done) return;
  double tot[3];
  seg_totals<3>(part, We, ewg0[s], ewg0[s + 1], tot);
  if (threadIdx.x != 0) return;
  Seg& c = seg[s];
  c.rho = make_double2(tot[0], tot[1]);
  c.beta = make_double2(0.0, 0.0);
  c.alpha = make_double2(0.0, 0.0);
  c.target2 = inner2 * tot[2];
  c.best2 = tot[2];
  c.last_imp = c.iter;
  ctl->all_done = 0;  // every resuming mesh stores the same value
}

// (re)start after r = b - A x: z = B r, p = 0; partials of r^T z, |r|^2
template <int D>
__global__ void __launch_bounds__(kBS) k_erestart(const int32_t* __restrict__ ewseg,
                                                  const int32_t* __restrict__ ewg0,
                                                  const int64_t* __restrict__ epos0,
                                                  const int32_t* __restrict__ bptr,
                                                  const int32_t* __restrict__ boff,
                                                  const double2* __restrict__ binv,
                                                  const double2* __restrict__ r,
                                                  double2* __restrict__ z, double2* __restrict__ p,
                                                  double* part) {
  const Span sp = espan_of(ewseg, ewg0, epos0);
  double red[3] = {0.0, 0.0, 0.0};
  const int64_t k = sp.lo + threadIdx.x;
  if (k < sp.hi) {
    const int i0 = bptr[k], d = bptr[k + 1] - i0;
    const double2* B = binv + boff[k];
    double2 rl[D];
#pragma unroll
    for (int t = 0; t < D; ++t) rl[t] = t < d ? r[i0 + t] : make_double2(0.0, 0.0);
#pragma unroll
    for (int t = 0; t < D; ++t)
      if (t < d) {
        double2 zt = make_double2(0.0, 0.0);
#pragma unroll
        for (int u = 0; u < D; ++u)
          if (u < d) zt = cfma(B[t * d + u], rl[u], zt);
        z[i0 + t] = zt;
        p[i0 + t] = make_double2(0.0, 0.0);
        red[0] += rl[t].x * zt.x - rl[t].y * zt.y;
        red[1] += rl[t].x * zt.y + rl[t].y * zt.x;
        red[2] += rl[t].x * rl[t].x + rl[t].y * rl[t].y;
      }
  }
  write_partials<3>(red, part);
}

// block inverses of the concatenated CSR's entity blocks (index order)
__global__ void k_ebinv(int64_t nent, const int32_t* __restrict__ bptr,
                        const int32_t* __restrict__ boff, const int32_t* __restrict__ rowptr,
                        const int32_t* __restrict__ cols, const double2* __restrict__ vals,
                        double2* __restrict__ binv, int32_t* __restrict__ bad) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nent) return;
  const int i0 = bptr[k];
  if (!invert_block(rowptr, cols, vals, i0, bptr[k + 1] - i0, binv + boff[k])) bad[0] = 1;
}

// Work tables of one set of launches (element mode): the workgroups of the row
// kernels (wseg, wg0, W) and of the entity kernels (ewseg, ewg0, We) dealt to
// a set of meshes (the others get none, an empty range in wg0 / ewg0), and the
// wave ranges [w0, w1) of their element products.  The full set covers every
// mesh; as meshes converge the solve launches a compacted set.
struct Tabs {
  int W = 0, We = 0;
  int32_t *wseg = nullptr, *wg0 = nullptr, *ewseg = nullptr, *ewg0 = nullptr;
  std::vector<std::pair<int64_t, int64_t>> runs;
  int64_t work = 0;  // entities covered
};

struct Multi {
  int nseg = 0, W = 0, check = 25, maxit = 100000;
  int64_t n = 0, nnz = 0;
  double rtol = 1e-8;
  std::vector<int64_t> hrow0;
  int32_t *rowptr = nullptr, *cols = nullptr, *wseg = nullptr, *wg0 = nullptr, *bad = nullptr;
  int64_t* row0 = nullptr;
  double2 *vals = nullptr, *x = nullptr, *b = nullptr, *r = nullptr, *p = nullptr, *q = nullptr,
          *z = nullptr, *dinv = nullptr;
  double* part = nullptr;
  Ctl* ctl = nullptr;
  Seg* seg = nullptr;
  Ctl* hctl = nullptr;
  Seg* hseg = nullptr;
  hipStream_t st = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool use_graph = false;
  // element products + entity-block Jacobi (pg_multi_set_element_operator)
  const ElemOp* eop = nullptr;
  EopSchedule es{};
  int D = 1;
  bool pairs = false;  // every entity two DOFs 2k, 2k + 1 (p = 2)
  int32_t *ebptr = nullptr, *eboff = nullptr;
  int64_t* epos0 = nullptr;
  double2* binv = nullptr;
  int64_t nbinv = 0;
  std::vector<int64_t> spos, swave;  // per mesh: first entity, first wave (nseg + 1)
  Tabs full, cur, ref;  // every mesh; a compacted set; the meshes being refreshed
  const Tabs* act = nullptr;  // the set the iterations launch
  hipGraph_t cgraph = nullptr;  // iterations on the compacted set
  hipGraphExec_t cexec = nullptr;
  // mixed precision (pg_multi_set_mixed): complex64 iteration vectors and
  // block inverses; phase32 = the iterations being launched are complex64
  bool mixed = false, phase32 = false;
  double inner2 = 1e-8;
  int stall = 400;
  float2 *x32 = nullptr, *r32 = nullptr, *p32 = nullptr, *q32 = nullptr, *z32 = nullptr,
         *binv32 = nullptr;
  hipGraph_t mgraph = nullptr;  // complex64 iterations on every mesh
  hipGraphExec_t mexec = nullptr;
  std::vector<void*> owned;
  ~Multi() {  // release errors cannot be reported from here
    if (mexec) (void)hipGraphExecDestroy(mexec);
    if (mgraph) (void)hipGraphDestroy(mgraph);
    if (cexec) (void)hipGraphExecDestroy(cexec);
    if (cgraph) (void)hipGraphDestroy(cgraph);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (st) (void)hipStreamDestroy(st);
    for (void* v : owned) (void)hipFree(v);
    if (hctl) (void)hipHostFree(hctl);
    if (hseg) (void)hipHostFree(hseg);
  }
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
};

static const int32_t* seg_done(const Multi* B) {
  return reinterpret_cast<const int32_t*>(reinterpret_cast<const char*>(B->seg) +
                                          offsetof(Seg, done));
}
static_assert(sizeof(Seg) % sizeof(int32_t) == 0, "Seg stride in int32 units");
constexpr int kSegStride = (int)(sizeof(Seg) / sizeof(int32_t));

template <int D, typename VT>
static void enqueue_iteration_e(Multi* B, const Tabs& T, VT* x, VT* r, VT* p, VT* q, VT* z,
                                const VT* binv) {
  const dim3 g(T.W), ge(T.We), b(kBS), gs(B->nseg);
  hipLaunchKernelGGL(k_mpupdate<VT>, g, b, 0, B->st, B->ctl, B->seg, T.wseg, T.wg0, B->row0, z,
                     p);
  const VT* ye = nullptr;
  for (const auto& run : T.runs)
    ye = eop_local_segments(B->eop, seg_done(B), kSegStride, p, B->st, run.first, run.second);
  hipLaunchKernelGGL((k_esum<D, VT>), ge, b, 0, B->st, B->ctl, B->seg, T.ewseg, T.ewg0, B->epos0,
                     B->es.gm, B->es.gslot, ye, p, q, B->part);
  hipLaunchKernelGGL(k_mfin_mu, gs, b, 0, B->st, B->ctl, B->seg, T.ewg0, B->part, T.We);
  if (D == 2 && B->pairs)
    hipLaunchKernelGGL(k_eupdate2<VT>, ge, b, 0, B->st, B->ctl, B->seg, T.ewseg, T.ewg0,
                       B->epos0, binv, x, r, p, q, z, B->part);
  else
    hipLaunchKernelGGL((k_eupdate<D, VT>), ge, b, 0, B->st, B->ctl, B->seg, T.ewseg, T.ewg0,
                       B->epos0, B->ebptr, B->eboff, binv, x, r, p, q, z, B->part);
  hipLaunchKernelGGL(k_mfin_update, gs, b, 0, B->st, B->ctl, B->seg, T.ewg0, B->part, T.We,
                     std::is_same<VT, float2>::value ? B->stall : 0);
  hipLaunchKernelGGL(k_mfin_all, dim3(1), b, 0, B->st, B->ctl, B->seg, B->nseg);
}

template <int D>
static void enqueue_iteration_e(Multi* B, const Tabs& T) {
  if (B->phase32)
    enqueue_iteration_e<D, float2>(B, T, B->x32, B->r32, B->p32, B->q32, B->z32, B->binv32);
  else
    enqueue_iteration_e<D, double2>(B, T, B->x, B->r, B->p, B->q, B->z, B->binv);
}

// refresh of the meshes in `ref` (complex64 rounds): x += x32, r64 = b - A x,
// then a new round for those that go on (rows of r64 in B->r)
template <int D>
static void enqueue_refresh(Multi* B) {
  const Tabs& T = B->ref;
  hipLaunchKernelGGL(k_saccum, dim3(T.W), dim3(kBS), 0, B->st, T.wseg, T.wg0, B->row0, B->x,
                     B->x32);
  hipLaunchKernelGGL(k_mtrue, dim3(T.W), dim3(kBS), 0, B->st, T.wseg, T.wg0, B->row0, B->rowptr,
                     B->cols, B->vals, B->x, B->b, B->r, B->part);
  hipLaunchKernelGGL(k_fin_srefresh, dim3(B->nseg), dim3(kBS), 0, B->st, B->ctl, B->seg, T.wg0,
                     B->part, T.W);
  hipLaunchKernelGGL(k_sround<D>, dim3(T.We), dim3(kBS), 0, B->st, B->seg, T.ewseg, T.ewg0,
                     B->epos0, B->ebptr, B->eboff, B->binv32, B->r, B->r32, B->z32, B->p32,
                     B->part);
  hipLaunchKernelGGL(k_fin_sround, dim3(B->nseg), dim3(kBS), 0, B->st, B->ctl, B->seg, T.ewg0,
                     B->part, T.We, B->inner2);
}

template <int D>
static void enqueue_restart_e(Multi* B, int first) {
  const Tabs& T = B->full;
  hipLaunchKernelGGL(k_mtrue, dim3(T.W), dim3(kBS), 0, B->st, T.wseg, T.wg0, B->row0, B->rowptr,
                     B->cols, B->vals, B->x, B->b, B->r, B->part);
  hipLaunchKernelGGL(k_erestart<D>, dim3(T.We), dim3(kBS), 0, B->st, T.ewseg, T.ewg0, B->epos0,
                     B->ebptr, B->eboff, B->binv, B->r, B->z, B->p, B->part);
  hipLaunchKernelGGL(k_mfin_restart, dim3(B->nseg), dim3(kBS), 0, B->st, B->ctl, B->seg, T.ewg0,
                     B->part, T.We, first);
  hipLaunchKernelGGL(k_mfin_all, dim3(1), dim3(kBS), 0, B->st, B->ctl, B->seg, B->nseg);
}

// one COCG iteration of every mesh (six launches on the batch's stream; with
// the element operator seven, plus one per further run of waves)
static void enqueue_iteration(Multi* B) {
  if (B->eop) {
    if (B->D == 1) enqueue_iteration_e<1>(B, *B->act);
    else if (B->D == 2) enqueue_iteration_e<2>(B, *B->act);
    else enqueue_iteration_e<6>(B, *B->act);
    return;
  }
  const dim3 g(B->W), b(kBS), gs(B->nseg);
  hipLaunchKernelGGL(k_mpupdate<double2>, g, b, 0, B->st, B->ctl, B->seg, B->wseg, B->wg0,
                     B->row0, B->z, B->p);
  hipLaunchKernelGGL(k_mspmv, g, b, 0, B->st, B->ctl, B->seg, B->wseg, B->wg0, B->row0, B->rowptr,
                     B->cols, B->vals, B->p, B->q, B->part);
  hipLaunchKernelGGL(k_mfin_mu, gs, b, 0, B->st, B->ctl, B->seg, B->wg0, B->part, B->W);
  hipLaunchKernelGGL(k_mupdate, g, b, 0, B->st, B->ctl, B->seg, B->wseg, B->wg0, B->row0, B->x,
                     B->r, B->p, B->q, B->dinv, B->z, B->part);
  hipLaunchKernelGGL(k_mfin_update, gs, b, 0, B->st, B->ctl, B->seg, B->wg0, B->part, B->W,
                     0);
  hipLaunchKernelGGL(k_mfin_all, dim3(1), b, 0, B->st, B->ctl, B->seg, B->nseg);
}

static void enqueue_restart(Multi* B, int first) {
  if (B->eop) {
    if (B->D == 1) enqueue_restart_e<1>(B, first);
    else if (B->D == 2) enqueue_restart_e<2>(B, first);
    else enqueue_restart_e<6>(B, first);
    return;
  }
  hipLaunchKernelGGL(k_mrestart, dim3(B->W), dim3(kBS), 0, B->st, B->wseg, B->wg0, B->row0,
                     B->rowptr, B->cols, B->vals, B->x, B->b, B->r, B->dinv, B->z, B->p, B->part);
  hipLaunchKernelGGL(k_mfin_restart, dim3(B->nseg), dim3(kBS), 0, B->st, B->ctl, B->seg, B->wg0,
                     B->part, B->W, first);
  hipLaunchKernelGGL(k_mfin_all, dim3(1), dim3(kBS), 0, B->st, B->ctl, B->seg, B->nseg);
}

// capture `check` iterations of the current work set (when the batch uses a
// graph) into *g / *ge, replacing what they held
static int capture(Multi* B, hipGraph_t* g, hipGraphExec_t* ge) {
  if (!B->use_graph) return PG_OK;
  if (*ge) (void)hipGraphExecDestroy(*ge);
  if (*g) (void)hipGraphDestroy(*g);
  *ge = nullptr;
  *g = nullptr;
  if (hipStreamBeginCapture(B->st, hipStreamCaptureModeThreadLocal) != hipSuccess)
    PG_FAIL(PG_ERR_HIP, "graph capture failed to start");
  for (int k = 0; k < B->check; ++k) enqueue_iteration(B);
  if (hipStreamEndCapture(B->st, g) != hipSuccess ||
      hipGraphInstantiate(ge, *g, nullptr, nullptr, 0) != hipSuccess)
    PG_FAIL(PG_ERR_HIP, "graph capture failed");
  return PG_OK;
}

// host tables of the meshes with active[s] set (element mode)
static void build_tabs(const Multi* B, const std::vector<char>& active, std::vector<int32_t>& wseg,
                       std::vector<int32_t>& wg0, std::vector<int32_t>& ewseg,
                       std::vector<int32_t>& ewg0, Tabs* T) {
  const int ns = B->nseg;
  wseg.clear();
  ewseg.clear();
  wg0.assign(ns + 1, 0);
  ewg0.assign(ns + 1, 0);
  T->runs.clear();
  T->work = 0;
  for (int s = 0; s < ns; ++s) {
    wg0[s] = (int32_t)wseg.size();
    ewg0[s] = (int32_t)ewseg.size();
    if (!active[s]) continue;
    const int64_t nr = B->hrow0[s + 1] - B->hrow0[s], ne = B->spos[s + 1] - B->spos[s];
    for (int64_t w = 0; w < std::max<int64_t>(1, (nr + kRW - 1) / kRW); ++w) wseg.push_back(s);
    for (int64_t w = 0; w < std::max<int64_t>(1, (ne + kBS - 1) / kBS); ++w) ewseg.push_back(s);
    T->work += ne;
    const int64_t w0 = B->swave[s], w1 = B->swave[s + 1];
    if (w1 > w0) {
      if (!T->runs.empty() && T->runs.back().second == w0) T->runs.back().second = w1;
      else T->runs.emplace_back(w0, w1);
    }
  }
  wg0[ns] = (int32_t)wseg.size();
  ewg0[ns] = (int32_t)ewseg.size();
  T->W = (int)wseg.size();
  T->We = (int)ewseg.size();
}

// tables of the meshes with active[s] set, uploaded to T's device buffers (no
// launch in flight that reads them)
static int upload_tabs(Multi* B, const std::vector<char>& active, Tabs* T) {
  std::vector<int32_t> wseg, wg0, ewseg, ewg0;
  build_tabs(B, active, wseg, wg0, ewseg, ewg0, T);
  PG_HIP(hipMemcpyAsync(T->wseg, wseg.data(), std::max<size_t>(1, wseg.size()) * sizeof(int32_t),
                        hipMemcpyHostToDevice, B->st));
  PG_HIP(hipMemcpyAsync(T->wg0, wg0.data(), wg0.size() * sizeof(int32_t), hipMemcpyHostToDevice,
                        B->st));
  PG_HIP(hipMemcpyAsync(T->ewseg, ewseg.data(),
                        std::max<size_t>(1, ewseg.size()) * sizeof(int32_t),
                        hipMemcpyHostToDevice, B->st));
  PG_HIP(hipMemcpyAsync(T->ewg0, ewg0.data(), ewg0.size() * sizeof(int32_t),
                        hipMemcpyHostToDevice, B->st));
  PG_HIP(hipStreamSynchronize(B->st));
  return PG_OK;
}

// launch only the meshes with active[s] from now on (compacted tables,
// iterations captured again)
static int compact(Multi* B, const std::vector<char>& active) {
  int rc = upload_tabs(B, active, &B->cur);
  if (rc != PG_OK) return rc;
  B->act = &B->cur;
  return capture(B, &B->cgraph, &B->cexec);
}

// refresh the meshes with mask[s] (complex64 rounds); hseg read back after
static int refresh(Multi* B, const std::vector<char>& mask) {
  int rc = upload_tabs(B, mask, &B->ref);
  if (rc != PG_OK) return rc;
  if (B->D == 1) enqueue_refresh<1>(B);
  else if (B->D == 2) enqueue_refresh<2>(B);
  else enqueue_refresh<6>(B);
  PG_CHECK_LAUNCH();
  PG_HIP(hipMemcpyAsync(B->hseg, B->seg, B->nseg * sizeof(Seg), hipMemcpyDeviceToHost, B->st));
  PG_HIP(hipStreamSynchronize(B->st));
  return PG_OK;
}

__global__ void k_seg_reset(Seg* seg, int nseg) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < nseg) seg[s] = Seg{};
}
__global__ void k_demote(int64_t n, const double2* __restrict__ a, float2* __restrict__ b) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = make_float2((float)a[i].x, (float)a[i].y);
}

// the complex64 rounds of a mixed solve; returns with every mesh converged
// (conv) or handed to FP64 (fp64), x accumulated
static int mixed_rounds(Multi* B, int* nlaunch) {
  hipStream_t st = B->st;
  hipLaunchKernelGGL(k_seg_reset, dim3((B->nseg + kBS - 1) / kBS), dim3(kBS), 0, st, B->seg,
                     B->nseg);
  PG_HIP(hipMemsetAsync(B->x32, 0, (size_t)B->n * sizeof(float2), st));
  PG_CHECK_LAUNCH();
  std::vector<char> mask(B->nseg, 1), live(B->nseg, 1);
  int rc = refresh(B, mask);  // the first refresh: r64 = b, |b|^2, the first rounds
  if (rc != PG_OK) return rc;
  B->phase32 = true;
  B->act = &B->full;
  hipGraphExec_t ex = B->mexec;
  int64_t covered = B->full.work;
  while (true) {
    bool any = false, paused = false;
    int64_t work = 0;
    for (int s = 0; s < B->nseg; ++s) {
      const Seg& c = B->hseg[s];
      live[s] = !c.conv && !c.fp64;
      mask[s] = live[s] && c.done;
      any = any || live[s];
      paused = paused || mask[s];
      if (live[s]) work += B->spos[s + 1] - B->spos[s];
    }
    if (!any) break;
    if (paused) {  // rounds that stopped since the last check
      if ((rc = refresh(B, mask)) != PG_OK) return rc;
      continue;
    }
    if (10 * work < 7 * covered) {  // launch only the meshes still iterating
      if ((rc = compact(B, live)) != PG_OK) return rc;
      ex = B->cexec;
      covered = work;
    }
    if (ex) {
      PG_HIP(hipGraphLaunch(ex, st));
    } else {
      for (int k = 0; k < B->check; ++k) enqueue_iteration(B);
      PG_CHECK_LAUNCH();
    }
    ++*nlaunch;
    PG_HIP(hipMemcpyAsync(B->hseg, B->seg, B->nseg * sizeof(Seg), hipMemcpyDeviceToHost, st));
    PG_HIP(hipStreamSynchronize(st));
  }
  B->phase32 = false;
  return PG_OK;
}

}  // namespace multi
}  // namespace pg

using namespace pg;
using namespace pg::multi;

extern "C" {

int pg_multi_create(int nseg, const int64_t* n, const int64_t* nnz, const int32_t* const* rowptr,
                    const int32_t* const* cols, const double* const* values, double rtol,
                    int maxit, int check_every, int use_graph, void** out, void* stream) {
  if (!out || nseg <= 0 || !n || !nnz || !rowptr || !cols || !values)
    PG_FAIL(PG_ERR_ARG, "bad arguments");
  int64_t N = 0, Z = 0;
  std::vector<int64_t> row0(nseg + 1), nz0(nseg + 1);
  std::vector<int32_t> wg0(nseg + 1);
  int64_t W = 0;
  for (int s = 0; s < nseg; ++s) {
    if (n[s] <= 0 || nnz[s] <= 0 || !rowptr[s] || !cols[s] || !values[s])
      PG_FAIL(PG_ERR_ARG, "mesh %d: bad arguments", s);
    row0[s] = N;
    nz0[s] = Z;
    wg0[s] = (int32_t)W;
    N += n[s];
    Z += nnz[s];
    W += (n[s] + kRW - 1) / kRW;
  }
  row0[nseg] = N;
  nz0[nseg] = Z;
  wg0[nseg] = (int32_t)W;
  if (N >= (1LL << 31) || Z >= (1LL << 31) || W >= (1LL << 31))
    PG_FAIL(PG_ERR_ARG, "the concatenated system is too large for 32-bit indices");
  std::vector<int32_t> wseg(W);
  for (int s = 0; s < nseg; ++s)
    for (int32_t w = wg0[s]; w < wg0[s + 1]; ++w) wseg[w] = s;
  hipStream_t cst = as_stream(stream);
  Multi* B = new Multi();
  B->nseg = nseg;
  B->n = N;
  B->nnz = Z;
  B->W = (int)W;
  B->rtol = rtol > 0 ? rtol : 1e-8;
  B->maxit = maxit > 0 ? maxit : 100000;
  B->check = check_every > 0 ? check_every : 25;
  B->hrow0 = row0;
  int rc;
#define PG_MA(ptr, cnt)                          \
  if ((rc = B->alloc(&(ptr), (cnt))) != PG_OK) { \
    delete B;                                    \
    return rc;                                   \
  }
  PG_MA(B->rowptr, N + 1);
  PG_MA(B->cols, Z);
  PG_MA(B->vals, Z);
  PG_MA(B->row0, nseg + 1);
  PG_MA(B->wg0, nseg + 1);
  PG_MA(B->wseg, W);
  PG_MA(B->x, N);
  PG_MA(B->b, N);
  PG_MA(B->r, N);
  PG_MA(B->p, N);
  PG_MA(B->q, N);
  PG_MA(B->z, N);
  PG_MA(B->dinv, N);
  PG_MA(B->part, 3 * W);
  PG_MA(B->ctl, 1);
  PG_MA(B->seg, nseg);
  PG_MA(B->bad, 1);
#undef PG_MA
  if (hipHostMalloc(reinterpret_cast<void**>(&B->hctl), sizeof(Ctl), 0) != hipSuccess ||
      hipHostMalloc(reinterpret_cast<void**>(&B->hseg), nseg * sizeof(Seg), 0) != hipSuccess ||
      hipStreamCreateWithFlags(&B->st, hipStreamNonBlocking) != hipSuccess) {
    delete B;
    PG_FAIL(PG_ERR_HIP, "host buffers / stream creation failed");
  }
  // the caller's arrays are complete once its stream has drained
  if (hipStreamSynchronize(cst) != hipSuccess) {
    delete B;
    PG_FAIL(PG_ERR_HIP, "stream synchronisation failed");
  }
  hipStream_t st = B->st;
  for (int s = 0; s < nseg; ++s) {
    const unsigned gr = (unsigned)((n[s] + 1 + kBS - 1) / kBS), gc = (unsigned)((nnz[s] + kBS - 1) / kBS);
    hipLaunchKernelGGL(k_shift_i32, dim3(gr), dim3(kBS), 0, st, n[s] + 1, rowptr[s],
                       (int32_t)nz0[s], B->rowptr + row0[s]);
    hipLaunchKernelGGL(k_shift_i32, dim3(gc), dim3(kBS), 0, st, nnz[s], cols[s], (int32_t)row0[s],
                       B->cols + nz0[s]);
    if (hipMemcpyAsync(B->vals + nz0[s], values[s], nnz[s] * sizeof(double2),
                       hipMemcpyDeviceToDevice, st) != hipSuccess) {
      delete B;
      PG_FAIL(PG_ERR_HIP, "value copy failed");
    }
  }
  // row pointers of consecutive meshes overlap in one entry (equal values): the
  // later mesh's first entry is written last, in stream order
  if (hipMemcpyAsync(B->row0, row0.data(), (nseg + 1) * sizeof(int64_t), hipMemcpyHostToDevice,
                     st) != hipSuccess ||
      hipMemcpyAsync(B->wg0, wg0.data(), (nseg + 1) * sizeof(int32_t), hipMemcpyHostToDevice,
                     st) != hipSuccess ||
      hipMemcpyAsync(B->wseg, wseg.data(), W * sizeof(int32_t), hipMemcpyHostToDevice, st) !=
          hipSuccess ||
      hipMemsetAsync(B->bad, 0, sizeof(int32_t), st) != hipSuccess) {
    delete B;
    PG_FAIL(PG_ERR_HIP, "upload failed");
  }
  hipLaunchKernelGGL(k_mjacobi, dim3((unsigned)((N + kBS - 1) / kBS)), dim3(kBS), 0, st, N,
                     B->rowptr, B->cols, B->vals, B->dinv, B->bad);
  hipLaunchKernelGGL(k_mctl_init, dim3(1), dim3(64), 0, st, B->ctl, B->maxit,
                     B->rtol * B->rtol, nseg);
  if (hipGetLastError() != hipSuccess ||
      hipMemcpyAsync(&B->hctl->maxit, B->bad, sizeof(int32_t), hipMemcpyDeviceToHost, st) !=
          hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess) {
    delete B;
    PG_FAIL(PG_ERR_HIP, "multi-mesh setup failed");
  }
  if (B->hctl->maxit != 0) {
    delete B;
    PG_FAIL(PG_ERR_ARG, "zero diagonal entry: Jacobi preconditioner undefined");
  }
  B->use_graph = use_graph != 0;
  if ((rc = capture(B, &B->graph, &B->exec)) != PG_OK) {
    delete B;
    return rc;
  }
  *out = B;
  return PG_OK;
}

int pg_multi_set_element_operator(void* batch, void* op, void* stream) {
  if (!batch || !op) PG_FAIL(PG_ERR_ARG, "bad arguments");
  Multi* B = static_cast<Multi*>(batch);
  const ElemOp* E = static_cast<const ElemOp*>(op);
  if (B->eop) PG_FAIL(PG_ERR_ARG, "the batch already has an element operator");
  const int64_t* spos = nullptr;
  const int nseg = eop_segments(E, &spos, nullptr);
  if (nseg != B->nseg)
    PG_FAIL(PG_ERR_ARG, "element operator of %d meshes for a batch of %d", nseg, B->nseg);
  if (eop_ndof(E) != B->n) PG_FAIL(PG_ERR_ARG, "element operator and batch differ in size");
  if (!eop_supports(E, true))
    PG_FAIL(PG_ERR_UNSUPPORTED, "element operator without a complex128 product");
  const EopSchedule es = eop_schedule(E);
  const int64_t nent = es.nent;
  hipStream_t cst = as_stream(stream);
  PG_HIP(hipStreamSynchronize(cst));
  // entity DOF ranges in index order (the schedule's ranges sorted by first
  // DOF; they must tile 0 .. n) and each mesh's entities within its rows
  std::vector<int4> gm(nent);
  PG_HIP(hipMemcpy(gm.data(), es.gm, nent * sizeof(int4), hipMemcpyDeviceToHost));
  for (int s = 0; s < nseg; ++s)
    for (int64_t i = spos[s]; i < spos[s + 1]; ++i)
      if (gm[i].x < B->hrow0[s] || gm[i].x + gm[i].y > B->hrow0[s + 1])
        PG_FAIL(PG_ERR_ARG, "mesh %d: scheduled entity outside its rows", s);
  std::vector<int2> rg(nent);
  for (int64_t i = 0; i < nent; ++i) rg[i] = make_int2(gm[i].x, gm[i].y);
  std::sort(rg.begin(), rg.end(), [](int2 a, int2 b) { return a.x < b.x; });
  std::vector<int32_t> bp(nent + 1), boff(nent + 1);
  boff[0] = 0;
  for (int64_t k = 0; k < nent; ++k) {
    if (rg[k].x != (k ? bp[k - 1] + rg[k - 1].y : 0) || rg[k].y < 1 || rg[k].y > kMaxInvBlock)
      PG_FAIL(PG_ERR_ARG, "element operator entities do not tile the rows");
    bp[k] = rg[k].x;
    boff[k + 1] = boff[k] + rg[k].y * rg[k].y;
  }
  bp[nent] = (int32_t)B->n;
  if (rg[nent - 1].x + rg[nent - 1].y != B->n)
    PG_FAIL(PG_ERR_ARG, "element operator entities do not tile the rows");
  const int64_t* swave = nullptr;
  eop_segments(E, nullptr, &swave);
  B->spos.assign(spos, spos + nseg + 1);
  B->swave.assign(swave, swave + nseg + 1);
  bool pairs = es.D == 2;
  for (int64_t k = 0; k < nent && pairs; ++k) pairs = bp[k] == 2 * k && rg[k].y == 2;
  // the full work set (every mesh) and room for compacted ones
  std::vector<int32_t> wseg, wg0, ewseg, ewg0;
  Tabs full;
  build_tabs(B, std::vector<char>(nseg, 1), wseg, wg0, ewseg, ewg0, &full);
  int rc;
  if ((rc = B->alloc(&full.ewseg, full.We)) != PG_OK ||
      (rc = B->alloc(&full.ewg0, nseg + 1)) != PG_OK ||
      (rc = B->alloc(&B->cur.wseg, full.W)) != PG_OK ||
      (rc = B->alloc(&B->cur.wg0, nseg + 1)) != PG_OK ||
      (rc = B->alloc(&B->cur.ewseg, full.We)) != PG_OK ||
      (rc = B->alloc(&B->cur.ewg0, nseg + 1)) != PG_OK ||
      (rc = B->alloc(&B->ref.wseg, full.W)) != PG_OK ||
      (rc = B->alloc(&B->ref.wg0, nseg + 1)) != PG_OK ||
      (rc = B->alloc(&B->ref.ewseg, full.We)) != PG_OK ||
      (rc = B->alloc(&B->ref.ewg0, nseg + 1)) != PG_OK ||
      (rc = B->alloc(&B->epos0, nseg + 1)) != PG_OK ||
      (rc = B->alloc(&B->ebptr, nent + 1)) != PG_OK ||
      (rc = B->alloc(&B->eboff, nent + 1)) != PG_OK ||
      (rc = B->alloc(&B->binv, boff[nent])) != PG_OK)
    return rc;
  if (full.W != B->W)
    PG_FAIL(PG_ERR_ARG, "row workgroups of the element mode differ from the batch's");
  full.wseg = B->wseg;  // the batch's own row tables cover every mesh
  full.wg0 = B->wg0;
  if (full.We > B->W && (rc = B->alloc(&B->part, 3 * (size_t)full.We)) != PG_OK) return rc;
  hipStream_t st = B->st;
  PG_HIP(hipMemcpy(full.ewseg, ewseg.data(), full.We * sizeof(int32_t), hipMemcpyHostToDevice));
  PG_HIP(hipMemcpy(full.ewg0, ewg0.data(), (nseg + 1) * sizeof(int32_t), hipMemcpyHostToDevice));
  PG_HIP(hipMemcpy(B->epos0, spos, (nseg + 1) * sizeof(int64_t), hipMemcpyHostToDevice));
  PG_HIP(hipMemcpy(B->ebptr, bp.data(), (nent + 1) * sizeof(int32_t), hipMemcpyHostToDevice));
  PG_HIP(hipMemcpy(B->eboff, boff.data(), (nent + 1) * sizeof(int32_t), hipMemcpyHostToDevice));
  PG_HIP(hipMemsetAsync(B->bad, 0, sizeof(int32_t), st));
  hipLaunchKernelGGL(k_ebinv, dim3((unsigned)((nent + kBS - 1) / kBS)), dim3(kBS), 0, st, nent,
                     B->ebptr, B->eboff, B->rowptr, B->cols, B->vals, B->binv, B->bad);
  PG_CHECK_LAUNCH();
  PG_HIP(hipMemcpyAsync(&B->hctl->maxit, B->bad, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  PG_HIP(hipStreamSynchronize(st));
  if (B->hctl->maxit)
    PG_FAIL(PG_ERR_ARG, "singular or non-contiguous entity block: block Jacobi undefined");
  B->es = es;
  B->D = es.D;
  B->pairs = pairs;
  B->nbinv = boff[nent];
  B->full = full;
  B->act = &B->full;
  B->eop = E;
  // the captured iterations now run the element products
  return capture(B, &B->graph, &B->exec);
}

int pg_multi_solve(void* batch, const double* const* b, double* const* x, int32_t* iterations,
                   double* relres, int32_t* status, int32_t* launches, void* stream) {
  if (!batch || !b || !x) PG_FAIL(PG_ERR_ARG, "bad arguments");
  Multi* B = static_cast<Multi*>(batch);
  hipStream_t cst = as_stream(stream);
  PG_HIP(hipStreamSynchronize(cst));  // the caller's right-hand sides are ready
  hipStream_t st = B->st;
  for (int s = 0; s < B->nseg; ++s) {
    if (!b[s] || !x[s]) PG_FAIL(PG_ERR_ARG, "mesh %d: NULL b or x", s);
    const int64_t ns = B->hrow0[s + 1] - B->hrow0[s];
    PG_HIP(hipMemcpyAsync(B->b + B->hrow0[s], b[s], ns * sizeof(double2), hipMemcpyDeviceToDevice,
                          st));
  }
  PG_HIP(hipMemsetAsync(B->x, 0, B->n * sizeof(double2), st));
  hipLaunchKernelGGL(k_mctl_init, dim3(1), dim3(64), 0, st, B->ctl, B->maxit, B->rtol * B->rtol,
                     B->nseg);
  int nlaunch = 0, rc;
  bool iterate64 = true;
  if (B->mixed) {
    // complex64 rounds; the meshes handed to FP64 then go on from their iterate
    if ((rc = mixed_rounds(B, &nlaunch)) != PG_OK) return rc;
    iterate64 = false;
    for (int s = 0; s < B->nseg; ++s) iterate64 = iterate64 || B->hseg[s].fp64;
    if (iterate64) enqueue_restart(B, 0);
  } else {
    enqueue_restart(B, 1);
  }
  PG_CHECK_LAUNCH();
  for (int restart = 0;; ++restart) {
    // element mode: every mesh after a (re)start; once the meshes still
    // iterating hold under 70 % of the entities the launches cover, the
    // launches are cut to them (tables compacted, iterations captured again)
    B->act = &B->full;
    hipGraphExec_t ex = B->exec;
    int64_t covered = B->full.work;
    if (iterate64 && B->eop) {
      PG_HIP(hipMemcpyAsync(B->hseg, B->seg, B->nseg * sizeof(Seg), hipMemcpyDeviceToHost, st));
      PG_HIP(hipStreamSynchronize(st));
    }
    while (iterate64) {
      if (B->eop) {
        std::vector<char> live(B->nseg);
        int64_t work = 0;
        for (int s = 0; s < B->nseg; ++s) {
          live[s] = !B->hseg[s].done;
          if (live[s]) work += B->spos[s + 1] - B->spos[s];
        }
        if (10 * work < 7 * covered && work > 0) {
          if ((rc = compact(B, live)) != PG_OK) return rc;
          ex = B->cexec;
          covered = work;
        }
      }
      if (ex) {
        PG_HIP(hipGraphLaunch(ex, st));
      } else {
        for (int k = 0; k < B->check; ++k) enqueue_iteration(B);
        PG_CHECK_LAUNCH();
      }
      ++nlaunch;
      PG_HIP(hipMemcpyAsync(B->hctl, B->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, st));
      if (B->eop)
        PG_HIP(hipMemcpyAsync(B->hseg, B->seg, B->nseg * sizeof(Seg), hipMemcpyDeviceToHost, st));
      PG_HIP(hipStreamSynchronize(st));
      if (B->hctl->all_done) break;
    }
    iterate64 = true;
    // true residuals; meshes above the tolerance restart from their iterate
    hipLaunchKernelGGL(k_mtrue, dim3(B->W), dim3(kBS), 0, st, B->wseg, B->wg0, B->row0, B->rowptr,
                       B->cols, B->vals, B->x, B->b, B->r, B->part);
    hipLaunchKernelGGL(k_mfin_true, dim3(B->nseg), dim3(kBS), 0, st, B->seg, B->wg0, B->part,
                       B->W);
    PG_CHECK_LAUNCH();
    PG_HIP(hipMemcpyAsync(B->hseg, B->seg, B->nseg * sizeof(Seg), hipMemcpyDeviceToHost, st));
    PG_HIP(hipStreamSynchronize(st));
    bool again = false;
    for (int s = 0; s < B->nseg; ++s) {
      const Seg& c = B->hseg[s];
      if (c.status == 0 && c.iter < B->maxit && c.res2 > B->rtol * B->rtol * c.bnorm2) again = true;
    }
    if (!again || restart >= 20) break;
    enqueue_restart(B, 0);
    PG_CHECK_LAUNCH();
  }
  for (int s = 0; s < B->nseg; ++s) {
    const Seg& c = B->hseg[s];
    const int64_t ns = B->hrow0[s + 1] - B->hrow0[s];
    PG_HIP(hipMemcpyAsync(x[s], B->x + B->hrow0[s], ns * sizeof(double2), hipMemcpyDeviceToDevice,
                          st));
    const double rel = c.bnorm2 > 0 ? std::sqrt(c.res2 / c.bnorm2) : 0.0;
    if (iterations) iterations[s] = c.iter;
    if (relres) relres[s] = rel;
    if (status) status[s] = c.status != 0 ? c.status : (rel <= B->rtol * 1.0000001 ? 0 : -1);
  }
  PG_HIP(hipStreamSynchronize(st));
  if (launches) *launches = nlaunch;
  return PG_OK;
}

int pg_multi_set_mixed(void* batch, double inner_rtol, int stall_iters, void* stream) {
  if (!batch) PG_FAIL(PG_ERR_ARG, "bad arguments");
  Multi* B = static_cast<Multi*>(batch);
  if (!B->eop)
    PG_FAIL(PG_ERR_UNSUPPORTED, "mixed precision needs the element operator "
                                "(pg_multi_set_element_operator)");
  if (!(inner_rtol > 0.0 && inner_rtol < 1.0)) PG_FAIL(PG_ERR_ARG, "inner_rtol must be in (0, 1)");
  PG_HIP(hipStreamSynchronize(as_stream(stream)));
  int rc;
  if (!B->x32) {
    if ((rc = B->alloc(&B->x32, B->n)) != PG_OK || (rc = B->alloc(&B->r32, B->n)) != PG_OK ||
        (rc = B->alloc(&B->p32, B->n)) != PG_OK || (rc = B->alloc(&B->q32, B->n)) != PG_OK ||
        (rc = B->alloc(&B->z32, B->n)) != PG_OK || (rc = B->alloc(&B->binv32, B->nbinv)) != PG_OK)
      return rc;
    hipLaunchKernelGGL(k_demote, dim3((unsigned)((B->nbinv + kBS - 1) / kBS)), dim3(kBS), 0, B->st,
                       B->nbinv, B->binv, B->binv32);
    PG_CHECK_LAUNCH();
    // the complex64 products of every mesh start from zero local results
    PG_HIP(hipMemsetAsync(B->p32, 0, (size_t)B->n * sizeof(float2), B->st));
    PG_HIP(hipStreamSynchronize(B->st));
  }
  B->inner2 = inner_rtol * inner_rtol;
  B->stall = stall_iters > 0 ? stall_iters : 400;
  B->phase32 = true;
  B->act = &B->full;
  rc = capture(B, &B->mgraph, &B->mexec);
  B->phase32 = false;
  if (rc != PG_OK) return rc;
  B->mixed = true;
  return PG_OK;
}

int pg_multi_destroy(void* batch) {
  delete static_cast<Multi*>(batch);
  return PG_OK;
}

}  // extern "C"
