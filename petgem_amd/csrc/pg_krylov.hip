// Complex CSR SpMV and the Krylov solve for gfx950.
//
// Replaces PETSc's MatMult and KSPSolve as called from Solver.run
// (petgem/solver.py:584-590).  The system matrix is complex symmetric
// (A = A^T: K and M are symmetric), so the default method is COCG
// (conjugate orthogonal CG, van der Vorst & Melissen 1990) with a Jacobi
// preconditioner; the stopping test is the true-norm residual
// ||b - A x||_2 / ||b||_2 (PETSc's GMRES default is the preconditioned
// residual, which is weaker).
//
// Device organisation:
//  * one HIP stream; every iteration is three kernels
//      K3 p = z + beta p,   K1 q = A p with mu = p^T q,   K2 x,r,z updates with
//      rho = r^T z and |r|^2;
//    the scalars (alpha, beta, rho, ...) live in device memory and are
//    finished by the last workgroup of each kernel (fixed-order reduction of
//    per-workgroup partials, so results are bitwise reproducible);
//  * every kernel returns immediately once the device-side `done` flag is
//    set, so the host enqueues batches of iterations and looks at the flags
//    only between batches;
//  * rows are split into one contiguous chunk per workgroup and workgroups
//    are renumbered so that each XCD owns a contiguous eighth of the rows
//    (its L2 then sees a compact window of the gathered vector);
//  * the SpMV assigns G lanes (G in 8..64, chosen per matrix by timing) to a
//    row; lanes stride over the row's non-zeros, then combine with
//    cross-lane shuffles.
//
// Mixed precision (PG_PREC_MIXED): iterative refinement.  The Krylov
// recursion, SpMV and Jacobi run in complex64 on a complex64 copy of A; the
// correction is accumulated into the FP64 solution and the FP64 residual
// b - A x is recomputed whenever the complex64 residual has dropped by
// `inner_rtol` since the last refinement, after which the complex64 COCG
// restarts from that residual.  Dot products are always accumulated in FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

#include "pg_common.h"

namespace pg {

constexpr int kBS = 256;
constexpr int kMaxGrid = 2048;

struct Scal {
  double2 rho, mu, alpha, beta;
  double res2, bnorm2, tol2, ref2;
  double res64;
  int32_t iter, maxit, done, status;
  uint32_t counter[4];
};

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
__device__ inline double2 ca(double2 a, double2 b) { return zadd(a, b); }
__device__ inline float2 ca(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
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

// logical workgroup id: each XCD (blockIdx % 8) gets a contiguous range
__device__ inline int64_t logical_block() {
  const int64_t b = blockIdx.x, nb = gridDim.x;
  if (nb % 8 != 0) return b;
  return (b % 8) * (nb / 8) + b / 8;
}
// contiguous chunk [lo, hi) of n items for this workgroup, in units of `unit`
__device__ inline void chunk(int64_t n, int64_t unit, int64_t* lo, int64_t* hi) {
  const int64_t nb = gridDim.x;
  const int64_t units = (n + unit - 1) / unit;
  const int64_t per = (units + nb - 1) / nb;
  const int64_t lb = logical_block();
  const int64_t a = lb * per * unit, b = (lb + 1) * per * unit;
  *lo = a < n ? a : n;
  *hi = b < n ? b : n;
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
}

// Write this workgroup's NV partials; returns true in the last workgroup to
// arrive, after which `tot` (all threads) holds the fixed-order grand totals.
template <int NV>
__device__ inline bool last_block_total(double (&v)[NV], double* part, uint32_t* counter,
                                        double (&tot)[NV]) {
  __shared__ int am_last;
  block_reduce<NV>(v);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) part[(int64_t)k * gridDim.x + blockIdx.x] = v[k];
    __threadfence();
    const uint32_t prev = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_ACQ_REL,
                                                 __HIP_MEMORY_SCOPE_AGENT);
    am_last = (prev == gridDim.x - 1) ? 1 : 0;
  }
  __syncthreads();
  if (!am_last) return false;
  __threadfence();
  double acc[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) acc[k] = 0.0;
  for (int64_t i = threadIdx.x; i < gridDim.x; i += kBS)
#pragma unroll
    for (int k = 0; k < NV; ++k)
      acc[k] += __hip_atomic_load(part + (int64_t)k * gridDim.x + i, __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_AGENT);
  block_reduce<NV>(acc);
  __shared__ double res[NV];
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) res[k] = acc[k];
    __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) tot[k] = res[k];
  return true;
}

// ---------------------------------------------------------------------------
// SpMV.  MODE 0: y = A x.  MODE 1: y = A x and mu = x^T y (COCG).
// MODE 2: y = b - A x and |y|^2 (true residual).
// ---------------------------------------------------------------------------
template <int G, int MODE, typename VT, typename XT>
__global__ void __launch_bounds__(kBS) k_spmv(Scal* __restrict__ s, int64_t n,
                                              const int32_t* __restrict__ rowptr,
                                              const int32_t* __restrict__ cols,
                                              const VT* __restrict__ vals,
                                              const XT* __restrict__ x, XT* __restrict__ y,
                                              const double2* __restrict__ b, double* part) {
  if (MODE != 0 && s->done) return;
  constexpr int RPB = kBS / G;
  int64_t lo, hi;
  chunk(n, RPB, &lo, &hi);
  const int lane = threadIdx.x % G;
  double red[2] = {0.0, 0.0};
  for (int64_t row = lo + threadIdx.x / G; row < hi; row += RPB) {
    XT acc = czero<XT>();
    const int s0 = rowptr[row], s1 = rowptr[row + 1];
    for (int k = s0 + lane; k < s1; k += G) {
      const VT a = vals[k];
      const XT xv = x[cols[k]];
      acc = cfma(XT{(decltype(acc.x))a.x, (decltype(acc.x))a.y}, xv, acc);
    }
    acc.x = group_sum<G>(acc.x);
    acc.y = group_sum<G>(acc.y);
    if (lane == 0) {
      if (MODE == 2) {
        const double2 bb = b[row];
        const double2 rr = make_double2(bb.x - (double)acc.x, bb.y - (double)acc.y);
        y[row] = from_d<XT>(rr);
        red[0] += rr.x * rr.x + rr.y * rr.y;
      } else {
        y[row] = acc;
        if (MODE == 1) {
          const double2 pv = to_d(x[row]);
          const double2 qv = to_d(acc);
          red[0] += pv.x * qv.x - pv.y * qv.y;
          red[1] += pv.x * qv.y + pv.y * qv.x;
        }
      }
    }
  }
  if (MODE == 0) return;
  double tot[2];
  if (!last_block_total<2>(red, part, &s->counter[0], tot)) return;
  if (threadIdx.x != 0) return;
  if (MODE == 1) {
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
  } else {
    s->res64 = tot[0];
  }
}

// K2: x += alpha p; r -= alpha q; z = Dinv r; rho = r^T z; |r|^2.
template <typename XT>
__global__ void __launch_bounds__(kBS) k_cocg_update(Scal* __restrict__ s, int64_t n,
                                                     XT* __restrict__ x, XT* __restrict__ r,
                                                     const XT* __restrict__ p,
                                                     const XT* __restrict__ q,
                                                     const XT* __restrict__ dinv,
                                                     XT* __restrict__ z, double* part,
                                                     int stop_on_converged) {
  if (s->done) return;
  const double2 al = s->alpha;
  const XT a = from_d<XT>(al);
  const XT na = from_d<XT>(make_double2(-al.x, -al.y));
  int64_t lo, hi;
  chunk(n, kBS, &lo, &hi);
  double red[3] = {0.0, 0.0, 0.0};
  for (int64_t i = lo + threadIdx.x; i < hi; i += kBS) {
    const XT pi = p[i];
    x[i] = cfma(a, pi, x[i]);
    const XT ri = cfma(na, q[i], r[i]);
    r[i] = ri;
    const XT zi = cm(dinv[i], ri);
    z[i] = zi;
    const double2 rd = to_d(ri), zd = to_d(zi);
    red[0] += rd.x * zd.x - rd.y * zd.y;
    red[1] += rd.x * zd.y + rd.y * zd.x;
    red[2] += rd.x * rd.x + rd.y * rd.y;
  }
  double tot[3];
  if (!last_block_total<3>(red, part, &s->counter[1], tot)) return;
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
  } else if (tot[2] <= s->tol2 * s->bnorm2) {
    if (stop_on_converged) s->done = 1;
  } else if (s->iter >= s->maxit) {
    s->done = 1;
  }
  if (stop_on_converged && tot[2] <= s->ref2) s->done = 1;
}

// K3: p = z + beta p
template <typename XT>
__global__ void __launch_bounds__(kBS) k_cocg_pupdate(Scal* __restrict__ s, int64_t n,
                                                      const XT* __restrict__ z,
                                                      XT* __restrict__ p) {
  if (s->done) return;
  const XT be = from_d<XT>(s->beta);
  int64_t lo, hi;
  chunk(n, kBS, &lo, &hi);
  for (int64_t i = lo + threadIdx.x; i < hi; i += kBS) p[i] = cfma(be, p[i], z[i]);
}

// init / residual replacement: r = demote(r64) (or r64 itself), z = Dinv r,
// rho_new = r^T z, |r|^2; optionally |b|^2.  With `restart` the recursion
// restarts (beta = 0, p = 0); otherwise beta = rho_new / rho_old.
template <typename XT>
__global__ void __launch_bounds__(kBS) k_cocg_reset(Scal* __restrict__ s, int64_t n,
                                                    const double2* __restrict__ r64,
                                                    XT* __restrict__ r,
                                                    const XT* __restrict__ dinv,
                                                    XT* __restrict__ z, XT* __restrict__ p,
                                                    const double2* __restrict__ b,
                                                    double* part, int restart) {
  int64_t lo, hi;
  chunk(n, kBS, &lo, &hi);
  double red[4] = {0.0, 0.0, 0.0, 0.0};
  for (int64_t i = lo + threadIdx.x; i < hi; i += kBS) {
    const XT ri = from_d<XT>(r64[i]);
    r[i] = ri;
    const XT zi = cm(dinv[i], ri);
    z[i] = zi;
    if (restart) p[i] = czero<XT>();
    const double2 rd = to_d(ri), zd = to_d(zi);
    red[0] += rd.x * zd.x - rd.y * zd.y;
    red[1] += rd.x * zd.y + rd.y * zd.x;
    red[2] += rd.x * rd.x + rd.y * rd.y;
    if (b) {
      const double2 bb = b[i];
      red[3] += bb.x * bb.x + bb.y * bb.y;
    }
  }
  double tot[4];
  if (!last_block_total<4>(red, part, &s->counter[2], tot)) return;
  if (threadIdx.x != 0) return;
  const double2 rho_new = make_double2(tot[0], tot[1]);
  if (restart) {
    s->beta = make_double2(0.0, 0.0);
  } else {
    s->beta = zdiv(rho_new, s->rho);
  }
  s->rho = rho_new;
  s->res2 = tot[2];
  if (b) s->bnorm2 = tot[3];
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

__global__ void k_demote(int64_t n, const double* __restrict__ a, float* __restrict__ b) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

__global__ void k_ones(int64_t n, double2* __restrict__ d) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = make_double2(1.0, 0.0);
}
__global__ void k_ones32(int64_t n, float2* __restrict__ d) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = make_float2(1.f, 0.f);
}

static inline unsigned nblk(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// ---------------------------------------------------------------------------
// host-side dispatch
// ---------------------------------------------------------------------------
template <int MODE, typename VT, typename XT>
static void launch_spmv(int G, int grid, hipStream_t st, Scal* s, int64_t n, const int32_t* rp,
                        const int32_t* ci, const VT* v, const XT* x, XT* y, const double2* b,
                        double* part) {
  const dim3 g(grid), bl(kBS);
  switch (G) {
    case 8: hipLaunchKernelGGL((k_spmv<8, MODE, VT, XT>), g, bl, 0, st, s, n, rp, ci, v, x, y, b, part); break;
    case 16: hipLaunchKernelGGL((k_spmv<16, MODE, VT, XT>), g, bl, 0, st, s, n, rp, ci, v, x, y, b, part); break;
    case 32: hipLaunchKernelGGL((k_spmv<32, MODE, VT, XT>), g, bl, 0, st, s, n, rp, ci, v, x, y, b, part); break;
    default: hipLaunchKernelGGL((k_spmv<64, MODE, VT, XT>), g, bl, 0, st, s, n, rp, ci, v, x, y, b, part); break;
  }
}

static int grid_for(int64_t n, int G) {
  const int64_t rpb = kBS / G;
  int64_t g = (n + rpb - 1) / rpb;
  g = std::min<int64_t>(g, kMaxGrid);
  if (g >= 64) g = (g / 8) * 8;
  return (int)std::max<int64_t>(g, 1);
}
static int vgrid_for(int64_t n) {
  int64_t g = (n + kBS - 1) / kBS;
  g = std::min<int64_t>(g, kMaxGrid);
  if (g >= 64) g = (g / 8) * 8;
  return (int)std::max<int64_t>(g, 1);
}

struct Solver {
  int64_t n = 0, nnz = 0;
  pg_solve_opts o{};
  const int32_t* rowptr = nullptr;
  const int32_t* cols = nullptr;
  const double2* vals = nullptr;  // caller's complex128 values
  float2* vals32 = nullptr;       // owned complex64 copy (mixed)
  int G64 = 16, G32 = 16;
  int gs64 = 1, gs32 = 1, gv = 1;
  Scal* s = nullptr;
  Scal* h = nullptr;  // pinned host mirror
  double* part = nullptr;
  // vectors
  double2 *r = nullptr, *p = nullptr, *q = nullptr, *z = nullptr, *dinv = nullptr, *r64 = nullptr;
  float2 *r32 = nullptr, *p32 = nullptr, *q32 = nullptr, *z32 = nullptr, *x32 = nullptr,
         *dinv32 = nullptr;
  int32_t* bad = nullptr;
  // GMRES(m): Krylov basis V[(m+1) x n], Hessenberg column buffer, partials
  double2* V = nullptr;
  double2* hdev = nullptr;
  double2* hhost = nullptr;  // pinned
  double* partg = nullptr;
  int restart = 30;
  std::vector<void*> owned;

  template <typename T>
  int alloc(T** ptr, size_t count) {
    void* v = nullptr;
    hipError_t e = hipMalloc(&v, std::max<size_t>(count, 1) * sizeof(T));
    if (e != hipSuccess) PG_FAIL(PG_ERR_HIP, "hipMalloc(%zu): %s", count * sizeof(T), hipGetErrorString(e));
    owned.push_back(v);
    *ptr = static_cast<T*>(v);
    return PG_OK;
  }
  ~Solver() {
    for (void* v : owned) hipFree(v);
    if (h) hipHostFree(h);
    if (hhost) hipHostFree(hhost);
  }
};

#define PG_TRY(x)                \
  do {                           \
    int _rc = (x);               \
    if (_rc != PG_OK) return _rc; \
  } while (0)

// time one SpMV configuration (milliseconds, averaged)
template <typename VT, typename XT>
static float time_spmv(Solver* S, int G, hipStream_t st, const VT* v, const XT* x, XT* y) {
  hipEvent_t a, b;
  hipEventCreate(&a);
  hipEventCreate(&b);
  const int grid = grid_for(S->n, G);
  launch_spmv<0, VT, XT>(G, grid, st, S->s, S->n, S->rowptr, S->cols, v, x, y, nullptr, S->part);
  hipEventRecord(a, st);
  const int reps = 5;
  for (int i = 0; i < reps; ++i)
    launch_spmv<0, VT, XT>(G, grid, st, S->s, S->n, S->rowptr, S->cols, v, x, y, nullptr,
                           S->part);
  hipEventRecord(b, st);
  hipEventSynchronize(b);
  float ms = 0.f;
  hipEventElapsedTime(&ms, a, b);
  hipEventDestroy(a);
  hipEventDestroy(b);
  return ms / reps;
}

static int read_scalars(Solver* S, hipStream_t st) {
  PG_HIP(hipMemcpyAsync(S->h, S->s, sizeof(Scal), hipMemcpyDeviceToHost, st));
  PG_HIP(hipStreamSynchronize(st));
  return PG_OK;
}

// true FP64 residual r64 = b - A x, returns |r|^2 in S->h->res64
static int true_residual(Solver* S, hipStream_t st, const double2* b, const double2* x) {
  // MODE 2 checks s->done; clear it for this launch through a tiny kernel-free path:
  // the residual kernel is only launched when done == 0 on the device.
  launch_spmv<2, double2, double2>(S->G64, S->gs64, st, S->s, S->n, S->rowptr, S->cols, S->vals,
                                   x, S->r64, b, S->part);
  PG_CHECK_LAUNCH();
  return PG_OK;
}

__global__ void k_set_done(Scal* s, int v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) s->done = v;
}
__global__ void k_set_ref2(Scal* s, double v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) s->ref2 = v;
}
__global__ void k_init_scalars(Scal* s, int maxit, double tol2) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  Scal z{};
  z.maxit = maxit;
  z.tol2 = tol2;
  *s = z;
}

static int solve(Solver* S, const double2* b, double2* x, pg_solve_info* info, hipStream_t st) {
  const bool mixed = S->o.precision == PG_PREC_MIXED;
  const int64_t n = S->n;
  hipLaunchKernelGGL(k_init_scalars, dim3(1), dim3(64), 0, st, S->s, S->o.maxit,
                     S->o.rtol * S->o.rtol);
  if (mixed) PG_HIP(hipMemsetAsync(S->x32, 0, n * sizeof(float2), st));
  // r64 = b - A x; bnorm
  PG_TRY(true_residual(S, st, b, x));
  if (mixed)
    hipLaunchKernelGGL(k_cocg_reset<float2>, dim3(S->gv), dim3(kBS), 0, st, S->s, n, S->r64,
                       S->r32, S->dinv32, S->z32, S->p32, b, S->part, 1);
  else
    hipLaunchKernelGGL(k_cocg_reset<double2>, dim3(S->gv), dim3(kBS), 0, st, S->s, n, S->r64,
                       S->r, S->dinv, S->z, S->p, b, S->part, 1);
  PG_CHECK_LAUNCH();
  PG_TRY(read_scalars(S, st));
  const double bnorm2 = S->h->bnorm2;
  const double tol2 = S->o.rtol * S->o.rtol;
  int refinements = 0, restarts = 0;
  info->status = -1;
  if (bnorm2 == 0.0) {
    PG_HIP(hipMemsetAsync(x, 0, n * sizeof(double2), st));
    info->iterations = 0;
    info->refinements = 0;
    info->relres = 0.0;
    info->status = PG_OK;
    return PG_OK;
  }
  double res64 = S->h->res2;
  if (res64 <= tol2 * bnorm2) {
    info->iterations = 0;
    info->refinements = 0;
    info->relres = std::sqrt(res64 / bnorm2);
    info->status = PG_OK;
    return PG_OK;
  }
  // residual-replacement threshold for mixed precision (squared, relative to last ref)
  const double delta2 = S->o.inner_rtol * S->o.inner_rtol;
  if (mixed) hipLaunchKernelGGL(k_set_ref2, dim3(1), dim3(64), 0, st, S->s, res64 * delta2);
  const int check = std::max(1, S->o.check_every);
  while (true) {
    for (int k = 0; k < check; ++k) {
      if (mixed) {
        hipLaunchKernelGGL(k_cocg_pupdate<float2>, dim3(S->gv), dim3(kBS), 0, st, S->s, n, S->z32, S->p32);
        launch_spmv<1, float2, float2>(S->G32, S->gs32, st, S->s, n, S->rowptr, S->cols, S->vals32, S->p32, S->q32, nullptr, S->part);
        hipLaunchKernelGGL(k_cocg_update<float2>, dim3(S->gv), dim3(kBS), 0, st, S->s, n, S->x32, S->r32, S->p32, S->q32, S->dinv32, S->z32, S->part, 1);
      } else {
        hipLaunchKernelGGL(k_cocg_pupdate<double2>, dim3(S->gv), dim3(kBS), 0, st, S->s, n, S->z, S->p);
        launch_spmv<1, double2, double2>(S->G64, S->gs64, st, S->s, n, S->rowptr, S->cols, S->vals, S->p, S->q, nullptr, S->part);
        hipLaunchKernelGGL(k_cocg_update<double2>, dim3(S->gv), dim3(kBS), 0, st, S->s, n, x, S->r, S->p, S->q, S->dinv, S->z, S->part, 1);
      }
    }
    PG_CHECK_LAUNCH();
    PG_TRY(read_scalars(S, st));
    Scal hs = *S->h;
    if (hs.status == PG_ERR_BREAKDOWN) {
      info->status = PG_ERR_BREAKDOWN;
      break;
    }
    if (!hs.done) continue;
    // paused: converged recursion, replacement threshold, or maxit
    if (mixed) {
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
    if (!mixed) {
      // recursive residual converged but the true one did not: restart from x
      if (++restarts > 20) break;
      hipLaunchKernelGGL(k_cocg_reset<double2>, dim3(S->gv), dim3(kBS), 0, st, S->s, n, S->r64,
                         S->r, S->dinv, S->z, S->p, (const double2*)nullptr, S->part, 1);
    } else {
      // restart the complex64 recursion from the FP64 residual (keeping the old
      // search direction across a residual replacement stalls COCG)
      hipLaunchKernelGGL(k_cocg_reset<float2>, dim3(S->gv), dim3(kBS), 0, st, S->s, n, S->r64,
                         S->r32, S->dinv32, S->z32, S->p32, (const double2*)nullptr, S->part, 1);
      hipLaunchKernelGGL(k_set_ref2, dim3(1), dim3(64), 0, st, S->s, res64 * delta2);
    }
    PG_CHECK_LAUNCH();
  }
  PG_TRY(read_scalars(S, st));
  info->iterations = S->h->iter;
  info->refinements = refinements;
  info->relres = std::sqrt(res64 / bnorm2);
  return PG_OK;
}


// ---------------------------------------------------------------------------
// GMRES(m) kernels (complex128, Hermitian inner products)
// ---------------------------------------------------------------------------
constexpr int kMaxRestart = 128;

// h[i] = V_i^H w, i < k; fixed-order last-block reduction into h_out
__global__ void __launch_bounds__(kBS) k_mdot(int64_t n, const double2* __restrict__ V,
                                              int64_t ldv, int k, const double2* __restrict__ w,
                                              double* part, uint32_t* counter,
                                              double2* __restrict__ h_out) {
  int64_t lo, hi;
  chunk(n, kBS, &lo, &hi);
  const int64_t G = gridDim.x;
  for (int i = 0; i < k; ++i) {
    double red[2] = {0.0, 0.0};
    const double2* Vi = V + (int64_t)i * ldv;
    for (int64_t j = lo + threadIdx.x; j < hi; j += kBS) {
      const double2 v = Vi[j], ww = w[j];
      red[0] += v.x * ww.x + v.y * ww.y;
      red[1] += v.x * ww.y - v.y * ww.x;
    }
    block_reduce<2>(red);
    if (threadIdx.x == 0) {
      part[(2 * i) * G + blockIdx.x] = red[0];
      part[(2 * i + 1) * G + blockIdx.x] = red[1];
    }
    __syncthreads();
  }
  __shared__ int am_last;
  if (threadIdx.x == 0) {
    __threadfence();
    const uint32_t prev =
        __hip_atomic_fetch_add(counter, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    am_last = (prev == gridDim.x - 1) ? 1 : 0;
  }
  __syncthreads();
  if (!am_last) return;
  __threadfence();
  for (int i = 0; i < k; ++i) {
    double red[2] = {0.0, 0.0};
    for (int64_t b = threadIdx.x; b < G; b += kBS) {
      red[0] += __hip_atomic_load(part + (2 * i) * G + b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      red[1] += __hip_atomic_load(part + (2 * i + 1) * G + b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    block_reduce<2>(red);
    if (threadIdx.x == 0) h_out[i] = make_double2(red[0], red[1]);
    __syncthreads();
  }
  if (threadIdx.x == 0) __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// w -= sum_i coef[i] V_i ; optional |w|^2 into coef_out (last block)
__global__ void __launch_bounds__(kBS) k_maxpy_norm(int64_t n, const double2* __restrict__ V,
                                                    int64_t ldv, int k,
                                                    const double2* __restrict__ coef,
                                                    double2* __restrict__ w, double* part,
                                                    uint32_t* counter, double* norm_out) {
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
  double tot[1];
  if (!last_block_total<1>(red, part, counter, tot)) return;
  if (threadIdx.x == 0) *norm_out = tot[0];
}

// out = dinv * v
__global__ void k_dmul(int64_t n, const double2* __restrict__ d, const double2* __restrict__ v,
                       double2* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = zmul(d[i], v[i]);
}

// v *= s (real)
__global__ void k_rscale(int64_t n, double2* __restrict__ v, double s) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = make_double2(v[i].x * s, v[i].y * s);
}

// x += dinv * sum_i y_i V_i
__global__ void k_gmres_update(int64_t n, double2* __restrict__ x, const double2* __restrict__ d,
                               const double2* __restrict__ V, int64_t ldv, int k,
                               const double2* __restrict__ y) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  double2 acc = make_double2(0.0, 0.0);
  for (int i = 0; i < k; ++i) acc = zadd(acc, zmul(y[i], V[(int64_t)i * ldv + j]));
  x[j] = zadd(x[j], zmul(d[j], acc));
}

static int gmres_solve(Solver* S, const double2* b, double2* x, pg_solve_info* info,
                       hipStream_t st) {
  const int64_t n = S->n;
  const int m = S->restart;
  const double tol = S->o.rtol;
  const unsigned nb = nblk(n, kBS);
  hipLaunchKernelGGL(k_init_scalars, dim3(1), dim3(64), 0, st, S->s, S->o.maxit, tol * tol);
  // |b|^2 through the residual kernel with x = 0 would need a zero vector: use w = b copy
  double2* w0 = S->V;  // scratch: V_0
  PG_HIP(hipMemcpyAsync(w0, b, n * sizeof(double2), hipMemcpyDeviceToDevice, st));
  hipLaunchKernelGGL(k_maxpy_norm, dim3(S->gv), dim3(kBS), 0, st, n, S->V, n, 0, S->hdev, w0,
                     S->part, &S->s->counter[3], &S->s->bnorm2);
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
    // V_0 = r / beta
    PG_HIP(hipMemcpyAsync(S->V, S->r64, n * sizeof(double2), hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(k_rscale, dim3(nb), dim3(kBS), 0, st, n, S->V, 1.0 / beta);
    std::fill(H.begin(), H.end(), make_double2(0.0, 0.0));
    std::fill(g.begin(), g.end(), make_double2(0.0, 0.0));
    g[0] = make_double2(beta, 0.0);
    int k = 0;
    for (int j = 0; j < m && iters < S->o.maxit; ++j) {
      double2* Vj = S->V + (int64_t)j * n;
      double2* w = S->V + (int64_t)(j + 1) * n;
      hipLaunchKernelGGL(k_dmul, dim3(nb), dim3(kBS), 0, st, n, S->dinv, Vj, S->z);
      launch_spmv<0, double2, double2>(S->G64, S->gs64, st, S->s, n, S->rowptr, S->cols, S->vals,
                                       S->z, w, nullptr, S->part);
      // classical Gram-Schmidt, twice
      for (int pass = 0; pass < 2; ++pass) {
        double2* hcol = S->hdev + pass * (m + 1);
        hipLaunchKernelGGL(k_mdot, dim3(S->gv), dim3(kBS), 0, st, n, S->V, n, j + 1, w, S->partg,
                           &S->s->counter[2], hcol);
        hipLaunchKernelGGL(k_maxpy_norm, dim3(S->gv), dim3(kBS), 0, st, n, S->V, n, j + 1, hcol,
                           w, S->part, &S->s->counter[3], &S->s->res2);
      }
      PG_CHECK_LAUNCH();
      PG_HIP(hipMemcpyAsync(S->hhost, S->hdev, 2 * (m + 1) * sizeof(double2),
                            hipMemcpyDeviceToHost, st));
      PG_TRY(read_scalars(S, st));
      const double hn = std::sqrt(S->h->res2);
      for (int i = 0; i <= j; ++i)
        Hij(i, j) = zadd(S->hhost[i], S->hhost[(m + 1) + i]);
      Hij(j + 1, j) = make_double2(hn, 0.0);
      for (int i = 0; i < j; ++i) {  // apply previous rotations
        const double2 a = Hij(i, j), bb = Hij(i + 1, j);
        const double2 sc = make_double2(sn[i].x, -sn[i].y);
        const double2 t = zadd(make_double2(cs[i] * a.x, cs[i] * a.y), zmul(sn[i], bb));
        Hij(i + 1, j) = zsub(make_double2(cs[i] * bb.x, cs[i] * bb.y), zmul(sc, a));
        Hij(i, j) = t;
      }
      {
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
      const double est = std::hypot(g[j + 1].x, g[j + 1].y);
      if (est <= tol * bnorm || hn == 0.0) break;
    }
    // y = H^-1 g (upper triangular), x += Dinv V y
    std::vector<double2> y(k);
    for (int i = k - 1; i >= 0; --i) {
      double2 acc = g[i];
      for (int l = i + 1; l < k; ++l) acc = zsub(acc, zmul(Hij(i, l), y[l]));
      y[i] = zdiv(acc, Hij(i, i));
    }
    for (int i = 0; i < k; ++i) S->hhost[i] = y[i];
    PG_HIP(hipMemcpyAsync(S->hdev, S->hhost, k * sizeof(double2), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_gmres_update, dim3(nb), dim3(kBS), 0, st, n, x, S->dinv, S->V, n, k,
                       S->hdev);
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

// (Re)build everything derived from the caller's matrix values: the Jacobi
// inverse diagonal and, for mixed precision, the complex64 copy of A.
static int refresh_values(Solver* S, hipStream_t st) {
  const int64_t n = S->n, nnz = S->nnz;
  const bool mixed = S->o.precision == PG_PREC_MIXED;
  PG_HIP(hipMemsetAsync(S->bad, 0, sizeof(int32_t), st));
  if (mixed) {
    hipLaunchKernelGGL(k_demote, dim3(nblk(2 * nnz, kBS)), dim3(kBS), 0, st, 2 * nnz,
                       reinterpret_cast<const double*>(S->vals),
                       reinterpret_cast<float*>(S->vals32));
    if (S->o.pc == PG_PC_JACOBI)
      hipLaunchKernelGGL(k_jacobi_inv<float2>, dim3(nblk(n, kBS)), dim3(kBS), 0, st, n,
                         S->rowptr, S->cols, S->vals, S->dinv32, S->bad);
    else
      hipLaunchKernelGGL(k_ones32, dim3(nblk(n, kBS)), dim3(kBS), 0, st, n, S->dinv32);
  } else {
    if (S->o.pc == PG_PC_JACOBI)
      hipLaunchKernelGGL(k_jacobi_inv<double2>, dim3(nblk(n, kBS)), dim3(kBS), 0, st, n,
                         S->rowptr, S->cols, S->vals, S->dinv, S->bad);
    else
      hipLaunchKernelGGL(k_ones, dim3(nblk(n, kBS)), dim3(kBS), 0, st, n, S->dinv);
  }
  PG_CHECK_LAUNCH();
  PG_HIP(hipMemcpyAsync(&S->h->iter, S->bad, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  PG_HIP(hipStreamSynchronize(st));
  if (S->h->iter != 0) PG_FAIL(PG_ERR_ARG, "zero diagonal entry: Jacobi preconditioner undefined");
  return PG_OK;
}
}  // namespace pg

using namespace pg;

extern "C" {

int pg_spmv_z(int64_t n, const int32_t* rp, const int32_t* ci, const double* v, const double* x,
              double* y, void* stream) {
  if (n <= 0 || !rp || !ci || !v || !x || !y) PG_FAIL(PG_ERR_ARG, "bad arguments");
  launch_spmv<0, double2, double2>(16, grid_for(n, 16), as_stream(stream), nullptr, n, rp, ci,
                                   reinterpret_cast<const double2*>(v),
                                   reinterpret_cast<const double2*>(x),
                                   reinterpret_cast<double2*>(y), nullptr, nullptr);
  PG_CHECK_LAUNCH();
  return PG_OK;
}

int pg_spmv_c(int64_t n, const int32_t* rp, const int32_t* ci, const float* v, const float* x,
              float* y, void* stream) {
  if (n <= 0 || !rp || !ci || !v || !x || !y) PG_FAIL(PG_ERR_ARG, "bad arguments");
  launch_spmv<0, float2, float2>(16, grid_for(n, 16), as_stream(stream), nullptr, n, rp, ci,
                                 reinterpret_cast<const float2*>(v),
                                 reinterpret_cast<const float2*>(x),
                                 reinterpret_cast<float2*>(y), nullptr, nullptr);
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
  if (opts->pc != PG_PC_NONE && opts->pc != PG_PC_JACOBI)
    PG_FAIL(PG_ERR_UNSUPPORTED, "pc type %d not supported", opts->pc);
  if (!(opts->rtol > 0.0)) PG_FAIL(PG_ERR_ARG, "rtol must be positive");
  hipStream_t st = as_stream(stream);
  Solver* S = new Solver();
  S->n = n;
  S->nnz = nnz;
  S->o = *opts;
  if (S->o.maxit <= 0) S->o.maxit = 100000;
  if (S->o.check_every <= 0) S->o.check_every = 25;
  if (!(S->o.inner_rtol > 0.0) || S->o.inner_rtol >= 1.0) S->o.inner_rtol = 1e-4;
  S->rowptr = rowptr;
  S->cols = cols;
  S->vals = reinterpret_cast<const double2*>(values);
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
  PG_A(S->bad, 1);
  if (mixed) {
    PG_A(S->vals32, nnz);
    PG_A(S->r32, n);
    PG_A(S->p32, n);
    PG_A(S->q32, n);
    PG_A(S->z32, n);
    PG_A(S->x32, n);
    PG_A(S->dinv32, n);
  } else {
    PG_A(S->r, n);
    PG_A(S->p, n);
    PG_A(S->q, n);
    PG_A(S->z, n);
    PG_A(S->dinv, n);
  }
  if (opts->ksp == PG_KSP_GMRES) {
    S->restart = opts->restart;
    PG_A(S->V, (size_t)(S->restart + 1) * n);
    PG_A(S->hdev, 2 * (S->restart + 1));
    PG_A(S->partg, (size_t)2 * (S->restart + 1) * kMaxGrid);
    if (hipHostMalloc(reinterpret_cast<void**>(&S->hhost), 2 * (S->restart + 1) * sizeof(double2), 0) != hipSuccess) {
      delete S;
      PG_FAIL(PG_ERR_HIP, "hipHostMalloc failed");
    }
  }
#undef PG_A
  if (hipHostMalloc(reinterpret_cast<void**>(&S->h), sizeof(Scal), 0) != hipSuccess) {
    delete S;
    PG_FAIL(PG_ERR_HIP, "hipHostMalloc failed");
  }
  hipLaunchKernelGGL(k_init_scalars, dim3(1), dim3(64), 0, st, S->s, 0, 0.0);
  if ((rc = refresh_values(S, st)) != PG_OK) {
    delete S;
    return rc;
  }
  // autotune the SpMV lane-group size on this matrix (vectors hold scratch data)
  const int cands[4] = {8, 16, 32, 64};
  float best = 1e30f;
  for (int G : cands) {
    const float t = time_spmv<double2, double2>(S, G, st, S->vals, S->r64, mixed ? S->r64 : S->q);
    if (t < best) {
      best = t;
      S->G64 = G;
    }
  }
  if (mixed) {
    best = 1e30f;
    for (int G : cands) {
      const float t = time_spmv<float2, float2>(S, G, st, S->vals32, S->r32, S->q32);
      if (t < best) {
        best = t;
        S->G32 = G;
      }
    }
  }
  S->gs64 = grid_for(n, S->G64);
  S->gs32 = grid_for(n, S->G32);
  S->gv = vgrid_for(n);
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
  if (S->o.ksp == PG_KSP_GMRES)
    return gmres_solve(S, reinterpret_cast<const double2*>(b), reinterpret_cast<double2*>(x),
                       info, as_stream(stream));
  return solve(S, reinterpret_cast<const double2*>(b), reinterpret_cast<double2*>(x), info,
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
  // the solver's own COCG SpMV kernel (q = A p, mu = p^T q), on its own vectors
  hipLaunchKernelGGL(k_init_scalars, dim3(1), dim3(64), 0, st, S->s, 1, 0.0);
  float t = 0.f;
  if (!mixed) {
    hipMemsetAsync(S->p, 0, S->n * sizeof(double2), st);
    hipEventRecord(a, st);
    for (int i = 0; i < reps; ++i)
      launch_spmv<1, double2, double2>(S->G64, S->gs64, st, S->s, S->n, S->rowptr, S->cols,
                                       S->vals, S->p, S->q, nullptr, S->part);
    hipEventRecord(b, st);
    hipEventSynchronize(b);
    hipEventElapsedTime(&t, a, b);
    *ms64 = t / reps;
    *ms32 = 0.0;
  } else {
    hipMemsetAsync(S->p32, 0, S->n * sizeof(float2), st);
    hipEventRecord(a, st);
    for (int i = 0; i < reps; ++i)
      launch_spmv<1, float2, float2>(S->G32, S->gs32, st, S->s, S->n, S->rowptr, S->cols,
                                     S->vals32, S->p32, S->q32, nullptr, S->part);
    hipEventRecord(b, st);
    hipEventSynchronize(b);
    hipEventElapsedTime(&t, a, b);
    *ms32 = t / reps;
    hipEventRecord(a, st);
    for (int i = 0; i < reps; ++i)
      launch_spmv<2, double2, double2>(S->G64, S->gs64, st, S->s, S->n, S->rowptr, S->cols,
                                       S->vals, S->r64, S->r64 /*unused y*/, reinterpret_cast<const double2*>(S->r64), S->part);
    hipEventRecord(b, st);
    hipEventSynchronize(b);
    hipEventElapsedTime(&t, a, b);
    *ms64 = t / reps;
  }
  *g64 = S->G64;
  *g32 = S->G32;
  hipEventDestroy(a);
  hipEventDestroy(b);
  PG_CHECK_LAUNCH();
  return PG_OK;
}

int pg_solver_destroy(void* solver) {
  delete static_cast<Solver*>(solver);
  return PG_OK;
}

}  // extern "C"
