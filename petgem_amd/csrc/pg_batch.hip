// Batched multi-RHS solve for gfx950: many sources and frequencies on one mesh.
//
// The matrix is stored once as its stiffness and mass parts (K, M: one real
// value each per non-zero), and every right-hand side r carries its own
// frequency: A_r = K - i*omega_r*mu0*M is formed on the fly inside the SpMM,
// so one pass over the matrix serves every column.  Vectors are laid out
// row-major [N][R] (the R values of a row are contiguous), which makes the
// gathers x[col][0..R) coalesced.  Each column runs its own COCG recurrence
// (its own alpha, beta, rho, convergence flag) with Jacobi preconditioning on
// its own diagonal K_ii - i*omega_r*mu0*M_ii; all columns share each launch.
// Reductions are per-workgroup partials per column written with plain
// stores, finished in a fixed order by a one-workgroup-per-column kernel
// (bitwise reproducible, no device-scope fences).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "pg_common.h"
#include "pg_eop.h"

namespace pg {
namespace batch {

constexpr int kBS = 256;
constexpr int kMaxGrid = 1024;

struct Col {  // per-column scalars
  double2 rho, mu, alpha, beta;
  double res2, bnorm2;
  double ref2;    // mixed precision: inner target |r|^2 (0 in FP64)
  double res64;   // mixed precision: FP64 |b - A x|^2 at the last refinement
  double best2;   // mixed precision: best |r|^2 of the current complex64 round
  int32_t iter, done, status, outer;  // outer: finished with complex64 (mixed precision)
  int32_t last_imp, stalled, pad0, pad1;  // stagnation watch of the complex64 round
};

// complex helpers for the two vector precisions
__device__ inline double2 to_d(double2 v) { return v; }
__device__ inline double2 to_d(float2 v) { return make_double2(v.x, v.y); }
template <typename T>
__device__ inline T from_d(double2 v);
template <>
__device__ inline double2 from_d<double2>(double2 v) { return v; }
template <>
__device__ inline float2 from_d<float2>(double2 v) { return make_float2((float)v.x, (float)v.y); }

struct Ctl {
  double tol2;
  int32_t maxit, ncols, all_done, stall_limit;  // stall_limit: 0 = no watch (FP64)
};

// chunk of rows for this workgroup (XCD-contiguous renumbering)
__device__ inline void rows_of_block(int64_t n, int64_t* lo, int64_t* hi) {
  const int64_t b = blockIdx.x, nb = gridDim.x;
  const int64_t lb = (nb % 8 == 0) ? (b % 8) * (nb / 8) + b / 8 : b;
  const int64_t per = (n + nb - 1) / nb;
  const int64_t a = lb * per, c = (lb + 1) * per;
  *lo = a < n ? a : n;
  *hi = c < n ? c : n;
}

// rows of the SpMM for this workgroup: XCD x (= blockIdx.x % 8, the dispatch
// order) owns the x-th eighth of the rows and its workgroups sweep it
// together, interleaved by `per` rows, so the gathers of P stay in its L2
struct Sweep {
  int64_t first, end, stride;
};
__device__ inline Sweep xcd_sweep(int64_t n, int per) {
  const int64_t b = blockIdx.x, nb = gridDim.x;
  if (nb % 8 != 0) return {b * per, n, nb * per};
  const int64_t x = b % 8, nbx = nb / 8;
  return {n * x / 8 + (b / 8) * per, n * (x + 1) / 8, nbx * per};
}

// Threads are arranged (rsub, c): c = column lane (0..CT), rsub = row lane.
// Column r handled by lane c is r = c + t*CT, t < R/CT.
// Per-column partial sums (NV doubles each) are reduced over rsub in shared
// memory and written with plain stores to part[(k*R + r)*grid + block]; a
// finish kernel with one workgroup per column adds them in a fixed order
// after the kernel boundary has made them visible.
template <int CT, int NV, int TMAX>
__device__ inline void column_partials(int R, double (&v)[TMAX][NV], double* part) {
  constexpr int RPB = kBS / CT;
  __shared__ double sm[RPB][CT][NV];
  const int c = threadIdx.x % CT, rs = threadIdx.x / CT;
  const int T = R / CT;
  for (int t = 0; t < T; ++t) {
#pragma unroll
    for (int k = 0; k < NV; ++k) sm[rs][c][k] = v[t][k];
    __syncthreads();
    if (rs == 0) {
      for (int k = 0; k < NV; ++k) {
        double acc = 0.0;
        for (int i = 0; i < RPB; ++i) acc += sm[i][c][k];
        part[((int64_t)k * R + (c + t * CT)) * gridDim.x + blockIdx.x] = acc;
      }
    }
    __syncthreads();
  }
}

// in a finish kernel (one workgroup per column r = blockIdx.x): fixed-order
// totals of that column's np partials; valid in every thread
template <int NV>
__device__ inline void column_totals(const double* part, int R, int np, double (&tot)[NV]) {
  const int r = blockIdx.x;
  double acc[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) acc[k] = 0.0;
  for (int i = threadIdx.x; i < np; i += kBS)
#pragma unroll
    for (int k = 0; k < NV; ++k) acc[k] += part[((int64_t)k * R + r) * np + i];
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

// q = A_r p for every column, partials of mu_r = p_r^T q_r
template <int CT, int TMAX>
__global__ void __launch_bounds__(kBS) k_spmm(const Ctl* ctl, int64_t n, int R,
                                              const int32_t* __restrict__ rowptr,
                                              const int32_t* __restrict__ cols,
                                              const double* __restrict__ K,
                                              const double* __restrict__ M,
                                              const double* __restrict__ omu,
                                              const double2* __restrict__ P,
                                              double2* __restrict__ Q, double* part) {
  if (ctl->all_done) return;
  constexpr int RPB = kBS / CT;
  const int c = threadIdx.x % CT, rs = threadIdx.x / CT;
  const int T = R / CT;
  const Sweep sw = xcd_sweep(n, RPB);
  double v[TMAX][2];
  double om[TMAX];
#pragma unroll
  for (int t = 0; t < TMAX; ++t) {
    v[t][0] = v[t][1] = 0.0;
    om[t] = t < T ? omu[c + t * CT] : 0.0;
  }
  for (int64_t row = sw.first + rs; row < sw.end; row += sw.stride) {
    double2 acc[TMAX];
#pragma unroll
    for (int t = 0; t < TMAX; ++t) acc[t] = make_double2(0.0, 0.0);
    const int s0 = rowptr[row], s1 = rowptr[row + 1];
    for (int k = s0; k < s1; ++k) {
      const double kv = K[k], mv = M[k];
      const int64_t cj = (int64_t)cols[k] * R;
#pragma unroll
      for (int t = 0; t < TMAX; ++t)
        if (t < T) {
          const double2 x = P[cj + c + t * CT];
          const double ai = -om[t] * mv;
          acc[t].x += kv * x.x - ai * x.y;
          acc[t].y += kv * x.y + ai * x.x;
        }
    }
#pragma unroll
    for (int t = 0; t < TMAX; ++t)
      if (t < T) {
        const int64_t idx = row * R + c + t * CT;
        Q[idx] = acc[t];
        const double2 pv = P[idx];
        v[t][0] += pv.x * acc[t].x - pv.y * acc[t].y;
        v[t][1] += pv.x * acc[t].y + pv.y * acc[t].x;
      }
  }
  column_partials<CT, 2, TMAX>(R, v, part);
}

// Entity-row SpMM over the batch's entity schedule: the d rows of an entity
// (DOFs r0 .. r0+d) share one column list of length L; column lists and the
// d x L blocks of K and M are kept in schedule order (from cbase / vbase), so
// they stream sequentially, and each gathered row of P serves all d rows.
// em[i] = {r0, d | L << 8, cbase, vbase}.  Lanes c of an entity slot cover the
// columns; U column indices are loaded before their gathers, and the next
// entity's descriptor is fetched one step ahead.
template <int CT, int TMAX, int D, typename VT, typename MT>
__global__ void __launch_bounds__(kBS) k_spmm_er(const Ctl* ctl, int64_t nent, int R,
                                                 const int4* __restrict__ em,
                                                 const int32_t* __restrict__ cols,
                                                 const MT* __restrict__ K,
                                                 const MT* __restrict__ M,
                                                 const double* __restrict__ omu,
                                                 const VT* __restrict__ P,
                                                 VT* __restrict__ Q, double* part) {
  using RT = decltype(VT::x);
  if (ctl->all_done) return;
  constexpr int EPB = kBS / CT;
  constexpr int U = 4;
  const int c = threadIdx.x % CT, es = threadIdx.x / CT;
  const int T = R / CT;
  const Sweep sw = xcd_sweep(nent, EPB);
  double v[TMAX][2];
  RT om[TMAX];
#pragma unroll
  for (int t = 0; t < TMAX; ++t) {
    v[t][0] = v[t][1] = 0.0;
    om[t] = t < T ? (RT)omu[c + t * CT] : (RT)0;
  }
  int64_t ent = sw.first + es;
  int4 mn = make_int4(0, 0, 0, 0);
  if (ent < sw.end) mn = em[ent];
  for (; ent < sw.end; ent += sw.stride) {
    const int r0 = mn.x, d = mn.y & 0xff, L = mn.y >> 8, cbase = mn.z, vbase = mn.w;
    const int64_t nx = ent + sw.stride;
    if (nx < sw.end) mn = em[nx];
    VT acc[D][TMAX];
#pragma unroll
    for (int k = 0; k < D; ++k)
#pragma unroll
      for (int t = 0; t < TMAX; ++t) acc[k][t] = VT{0, 0};
    for (int j0 = 0; j0 < L; j0 += U) {
      int64_t cj[U];
#pragma unroll
      for (int u = 0; u < U; ++u) cj[u] = j0 + u < L ? (int64_t)cols[cbase + j0 + u] * R : -1;
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (cj[u] < 0) continue;
        VT x[TMAX];
#pragma unroll
        for (int t = 0; t < TMAX; ++t) x[t] = t < T ? P[cj[u] + c + t * CT] : VT{0, 0};
#pragma unroll
        for (int k = 0; k < D; ++k)
          if (k < d) {
            const RT kv = K[vbase + k * L + j0 + u], mv = M[vbase + k * L + j0 + u];
#pragma unroll
            for (int t = 0; t < TMAX; ++t)
              if (t < T) {
                const RT ai = -om[t] * mv;
                acc[k][t].x += kv * x[t].x - ai * x[t].y;
                acc[k][t].y += kv * x[t].y + ai * x[t].x;
              }
          }
      }
    }
#pragma unroll
    for (int k = 0; k < D; ++k)
      if (k < d)
#pragma unroll
        for (int t = 0; t < TMAX; ++t)
          if (t < T) {
            const int64_t idx = (int64_t)(r0 + k) * R + c + t * CT;
            Q[idx] = acc[k][t];
            const double2 pv = to_d(P[idx]), qv = to_d(acc[k][t]);
            v[t][0] += pv.x * qv.x - pv.y * qv.y;
            v[t][1] += pv.x * qv.y + pv.y * qv.x;
          }
  }
  column_partials<CT, 2, TMAX>(R, v, part);
}

// q_r = A_r p_r from the element operator's local results: for every entity of
// its schedule the sum of the entity's local results over its elements, in
// ascending slot order (atomic-free, fixed order), PEC entities identity rows;
// with the PEC rows' partials of mu_r = p_r^T q_r (p_i^2 there: the elements'
// share is summed by the local-product kernel, so p is read only at PEC
// entities).  gm[i] = {r0, d, s0, s1}; the local result of slot g, entity DOF
// t, column r is ye[(g + 64 t) R + r].
template <int CT, int TMAX, int D, typename VT>
__global__ void __launch_bounds__(kBS) k_meop_sum(const Ctl* ctl, int64_t nent, int R,
                                                  const int4* __restrict__ gm,
                                                  const int32_t* __restrict__ gslot,
                                                  const VT* __restrict__ ye,
                                                  const VT* __restrict__ P,
                                                  VT* __restrict__ Q, double* part) {
  if (ctl->all_done) return;
  constexpr int EPB = kBS / CT;
  const int c = threadIdx.x % CT, es = threadIdx.x / CT;
  const int T = R / CT;
  const Sweep sw = xcd_sweep(nent, EPB);
  double v[TMAX][2];
#pragma unroll
  for (int t = 0; t < TMAX; ++t) v[t][0] = v[t][1] = 0.0;
  for (int64_t ent = sw.first + es; ent < sw.end; ent += sw.stride) {
    const int4 m = gm[ent];
    const int r0 = m.x, d = m.y, s0 = m.z, s1 = m.w;
    // all column groups of an incidence together (up to TMAX * D independent
    // loads per step), two incidences per step, in ascending slot order
    VT acc[TMAX][D];
#pragma unroll
    for (int t = 0; t < TMAX; ++t)
#pragma unroll
      for (int k = 0; k < D; ++k) acc[t][k] = VT{0, 0};
    if (s0 == s1) {
#pragma unroll
      for (int t = 0; t < TMAX; ++t)
#pragma unroll
        for (int k = 0; k < D; ++k)
          if (t < T && k < d) acc[t][k] = P[(int64_t)(r0 + k) * R + c + t * CT];
    } else {
      int q = s0;
      for (; q + 2 <= s1; q += 2) {
        const int64_t g0 = gslot[q], g1 = gslot[q + 1];
        VT a[TMAX][D], b[TMAX][D];
#pragma unroll
        for (int t = 0; t < TMAX; ++t)
#pragma unroll
          for (int k = 0; k < D; ++k)
            if (t < T && k < d) {
              a[t][k] = ye[(g0 + (int64_t)kWave * k) * R + c + t * CT];
              b[t][k] = ye[(g1 + (int64_t)kWave * k) * R + c + t * CT];
            }
#pragma unroll
        for (int t = 0; t < TMAX; ++t)
#pragma unroll
          for (int k = 0; k < D; ++k)
            if (t < T && k < d) {
              acc[t][k].x += a[t][k].x;
              acc[t][k].y += a[t][k].y;
              acc[t][k].x += b[t][k].x;
              acc[t][k].y += b[t][k].y;
            }
      }
      if (q < s1) {
        const int64_t g0 = gslot[q];
#pragma unroll
        for (int t = 0; t < TMAX; ++t)
#pragma unroll
          for (int k = 0; k < D; ++k)
            if (t < T && k < d) {
              const VT a = ye[(g0 + (int64_t)kWave * k) * R + c + t * CT];
              acc[t][k].x += a.x;
              acc[t][k].y += a.y;
            }
      }
    }
#pragma unroll
    for (int t = 0; t < TMAX; ++t)
#pragma unroll
      for (int k = 0; k < D; ++k)
        if (t < T && k < d) {
          Q[(int64_t)(r0 + k) * R + c + t * CT] = acc[t][k];
          if (s0 == s1) {  // PEC row: q = p
            const double2 pv = to_d(acc[t][k]);
            v[t][0] += pv.x * pv.x - pv.y * pv.y;
            v[t][1] += 2.0 * pv.x * pv.y;
          }
        }
  }
  column_partials<CT, 2, TMAX>(R, v, part);
}

// schedule-order copies: entity i's column list and its d x L blocks of K, M
// (contiguous in the CSR from slot src[i])
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
template <typename MT>
__global__ void __launch_bounds__(kBS) k_sched_km(int64_t nent, const int4* __restrict__ em,
                                                  const int32_t* __restrict__ src,
                                                  const double* __restrict__ K,
                                                  const double* __restrict__ M,
                                                  MT* __restrict__ Ks, MT* __restrict__ Ms) {
  constexpr int G = 32;
  const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G;
  if (i >= nent) return;
  const int4 m = em[i];
  const int len = (m.y & 0xff) * (m.y >> 8), base = src[i];
  for (int t = threadIdx.x % G; t < len; t += G) {
    Ks[m.w + t] = (MT)K[base + t];
    Ms[m.w + t] = (MT)M[base + t];
  }
}

// entity rows must be consecutive rows of equal length sharing one column list
__global__ void k_check_entities(int64_t nent, const int32_t* __restrict__ bptr,
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

// finish mu_r and alpha_r = rho_r / mu_r (one workgroup per column); part2
// (nullable): a second set of np2 partials per column added after the first
__global__ void __launch_bounds__(kBS) k_fin_mu(const Ctl* ctl, Col* col, int R,
                                                const double* part, int np,
                                                const double* part2, int np2) {
  if (ctl->all_done) return;
  double tot[2];
  column_totals<2>(part, R, np, tot);
  if (part2) {
    double t2[2];
    __syncthreads();  // column_totals' shared scratch is reused
    column_totals<2>(part2, R, np2, t2);
    tot[0] += t2[0];
    tot[1] += t2[1];
  }
  if (threadIdx.x != 0) return;
  Col& cc = col[blockIdx.x];
  const double2 mu = make_double2(tot[0], tot[1]);
  cc.mu = mu;
  if (cc.done) {
    cc.alpha = make_double2(0.0, 0.0);
    return;
  }
  const double2 al = zdiv(cc.rho, mu);
  if (!(mu.x != 0.0 || mu.y != 0.0) || !isfinite(al.x) || !isfinite(al.y)) {
    cc.status = PG_ERR_BREAKDOWN;
    cc.done = 1;
    cc.alpha = make_double2(0.0, 0.0);
  } else {
    cc.alpha = al;
  }
}

// z = r / diag_r, as every kernel that forms the preconditioned residual does
template <typename VT>
__device__ inline VT jacobi_z(VT rv, double kd, double md, double om) {
  return from_d<VT>(zdiv(to_d(rv), make_double2(kd, -om * md)));
}

// r -= alpha q; partials of rho = r^T z and |r|^2 with z = r / diag_r.  z is
// not stored: k_pupdate forms it again from r (the same operations, so the
// same values), and takes this iteration's step x += alpha p there, where p is
// read anyway (or k_bxstep takes it once the batch pauses).
template <int CT, int TMAX, typename VT>
__global__ void __launch_bounds__(kBS) k_update(const Ctl* ctl, const Col* col, int64_t n, int R,
                                                const double* __restrict__ dK,
                                                const double* __restrict__ dM,
                                                const double* __restrict__ omu,
                                                VT* __restrict__ Rr, const VT* __restrict__ Q,
                                                double* part) {
  if (ctl->all_done) return;
  constexpr int RPB = kBS / CT;
  const int c = threadIdx.x % CT, rs = threadIdx.x / CT;
  const int T = R / CT;
  int64_t lo, hi;
  rows_of_block(n, &lo, &hi);
  double v[TMAX][3];
  VT al[TMAX];
  double om[TMAX];
#pragma unroll
  for (int t = 0; t < TMAX; ++t) {
    v[t][0] = v[t][1] = v[t][2] = 0.0;
    al[t] = from_d<VT>(t < T ? col[c + t * CT].alpha : make_double2(0.0, 0.0));
    om[t] = t < T ? omu[c + t * CT] : 0.0;
  }
  for (int64_t row = lo + rs; row < hi; row += RPB) {
    const double kd = dK[row], md = dM[row];
#pragma unroll
    for (int t = 0; t < TMAX; ++t)
      if (t < T) {
        const int64_t idx = row * R + c + t * CT;
        const VT qv = Q[idx];
        VT rv = Rr[idx];
        rv.x -= al[t].x * qv.x - al[t].y * qv.y;
        rv.y -= al[t].x * qv.y + al[t].y * qv.x;
        Rr[idx] = rv;
        const double2 rd = to_d(rv);
        const double2 zd = to_d(jacobi_z(rv, kd, md, om[t]));
        v[t][0] += rd.x * zd.x - rd.y * zd.y;
        v[t][1] += rd.x * zd.y + rd.y * zd.x;
        v[t][2] += rd.x * rd.x + rd.y * rd.y;
      }
  }
  column_partials<CT, 3, TMAX>(R, v, part);
}

// finish rho_r, |r_r|^2, beta_r and column r's convergence (one workgroup per column)
__global__ void __launch_bounds__(kBS) k_fin_update(const Ctl* ctl, Col* col, int R,
                                                    const double* part, int np) {
  if (ctl->all_done) return;
  double tot[3];
  column_totals<3>(part, R, np, tot);
  if (threadIdx.x != 0) return;
  Col& cc = col[blockIdx.x];
  if (cc.done) {
    cc.beta = make_double2(0.0, 0.0);
    return;
  }
  const double2 rho_new = make_double2(tot[0], tot[1]);
  const double2 beta = zdiv(rho_new, cc.rho);
  cc.beta = beta;
  cc.rho = rho_new;
  cc.res2 = tot[2];
  cc.iter += 1;
  if (!isfinite(tot[2]) || !isfinite(beta.x) || !isfinite(beta.y)) {
    cc.status = PG_ERR_BREAKDOWN;
    cc.done = 1;
  } else if (tot[2] <= ctl->tol2 * cc.bnorm2 || tot[2] <= cc.ref2 || cc.iter >= ctl->maxit) {
    cc.done = 1;  // converged, refinement point (mixed), or iteration limit
  } else if (ctl->stall_limit > 0) {
    if (tot[2] < 0.81 * cc.best2) {
      cc.best2 = tot[2];
      cc.last_imp = cc.iter;
    } else if (cc.iter - cc.last_imp >= ctl->stall_limit) {
      cc.stalled = 1;  // complex64 round stagnated (see k_fin_refine)
      cc.done = 1;
    }
  }
}

// all_done = every column done (one workgroup)
__global__ void __launch_bounds__(kBS) k_fin_all(Ctl* ctl, const Col* col, int R) {
  if (ctl->all_done) return;
  __shared__ int notdone;
  if (threadIdx.x == 0) notdone = 0;
  __syncthreads();
  int mine = 0;
  for (int r = threadIdx.x; r < R; r += kBS) mine |= col[r].done ? 0 : 1;
  if (mine) notdone = 1;  // benign race: every writer stores 1
  __syncthreads();
  if (threadIdx.x == 0) ctl->all_done = notdone ? 0 : 1;
}

// x += alpha p, the previous iteration's step (alpha is that iteration's; it is
// 0 for columns finished before it and after a (re)start, where p = 0 too);
// then z = r / diag_r and p = z + beta p
template <typename VT>
__device__ inline void step_and_direction(VT* X, VT* P, int64_t idx, VT al, VT be, VT zv) {
  const VT pv = P[idx];
  VT xv = X[idx];
  xv.x += al.x * pv.x - al.y * pv.y;
  xv.y += al.x * pv.y + al.y * pv.x;
  X[idx] = xv;
  P[idx] = VT{zv.x + be.x * pv.x - be.y * pv.y, zv.y + be.x * pv.y + be.y * pv.x};
}

template <int CT, int TMAX, typename VT>
__global__ void __launch_bounds__(kBS) k_pupdate(const Ctl* ctl, const Col* col, int64_t n, int R,
                                                 const double* __restrict__ dK,
                                                 const double* __restrict__ dM,
                                                 const double* __restrict__ omu,
                                                 VT* __restrict__ X, const VT* __restrict__ Rr,
                                                 VT* __restrict__ P) {
  if (ctl->all_done) return;
  constexpr int RPB = kBS / CT;
  const int c = threadIdx.x % CT, rs = threadIdx.x / CT;
  const int T = R / CT;
  int64_t lo, hi;
  rows_of_block(n, &lo, &hi);
  VT al[TMAX], be[TMAX];
  double om[TMAX];
#pragma unroll
  for (int t = 0; t < TMAX; ++t) {
    al[t] = from_d<VT>(t < T ? col[c + t * CT].alpha : make_double2(0.0, 0.0));
    be[t] = from_d<VT>(t < T ? col[c + t * CT].beta : make_double2(0.0, 0.0));
    om[t] = t < T ? omu[c + t * CT] : 0.0;
  }
  for (int64_t row = lo + rs; row < hi; row += RPB) {
    const double kd = dK[row], md = dM[row];
#pragma unroll
    for (int t = 0; t < TMAX; ++t)
      if (t < T) {
        const int64_t idx = row * R + c + t * CT;
        step_and_direction(X, P, idx, al[t], be[t], jacobi_z(Rr[idx], kd, md, om[t]));
      }
  }
}

// x += alpha p for every column: the step of the last iteration before the
// batch paused (alpha is 0 for columns that finished earlier)
template <typename VT>
__global__ void k_bxstep(const Col* col, int64_t nr, int R, const VT* __restrict__ P,
                         VT* __restrict__ X) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nr) return;
  const VT al = from_d<VT>(col[i % R].alpha);
  const VT pv = P[i];
  VT xv = X[i];
  xv.x += al.x * pv.x - al.y * pv.y;
  xv.y += al.x * pv.y + al.y * pv.x;
  X[i] = xv;
}

// init: x = 0, r = b, p = 0; partials of rho = r^T z (z = r/diag), |r|^2 (= |b|^2)
template <int CT, int TMAX>
__global__ void __launch_bounds__(kBS) k_init(int64_t n, int R, const double* __restrict__ dK,
                                              const double* __restrict__ dM,
                                              const double* __restrict__ omu,
                                              const double2* __restrict__ B,
                                              double2* __restrict__ X, double2* __restrict__ Rr,
                                              double2* __restrict__ P, double* part) {
  constexpr int RPB = kBS / CT;
  const int c = threadIdx.x % CT, rs = threadIdx.x / CT;
  const int T = R / CT;
  int64_t lo, hi;
  rows_of_block(n, &lo, &hi);
  double v[TMAX][3];
  double om[TMAX];
#pragma unroll
  for (int t = 0; t < TMAX; ++t) {
    v[t][0] = v[t][1] = v[t][2] = 0.0;
    om[t] = t < T ? omu[c + t * CT] : 0.0;
  }
  for (int64_t row = lo + rs; row < hi; row += RPB) {
    const double kd = dK[row], md = dM[row];
#pragma unroll
    for (int t = 0; t < TMAX; ++t)
      if (t < T) {
        const int64_t idx = row * R + c + t * CT;
        const double2 bv = B[idx];
        X[idx] = make_double2(0.0, 0.0);
        Rr[idx] = bv;
        const double2 zv = jacobi_z(bv, kd, md, om[t]);
        P[idx] = make_double2(0.0, 0.0);
        v[t][0] += bv.x * zv.x - bv.y * zv.y;
        v[t][1] += bv.x * zv.y + bv.y * zv.x;
        v[t][2] += bv.x * bv.x + bv.y * bv.y;
      }
  }
  column_partials<CT, 3, TMAX>(R, v, part);
}

__global__ void __launch_bounds__(kBS) k_fin_init(Col* col, int R, const double* part, int np) {
  double tot[3];
  column_totals<3>(part, R, np, tot);
  if (threadIdx.x != 0) return;
  Col& cc = col[blockIdx.x];
  cc = Col{};  // every field, including the refinement target (0: none in FP64)
  cc.rho = make_double2(tot[0], tot[1]);
  cc.res2 = tot[2];
  cc.bnorm2 = tot[2];
  cc.done = (cc.bnorm2 == 0.0) ? 1 : 0;
}

// true residual per column: partials of |b - A x|^2
template <int CT, int TMAX>
__global__ void __launch_bounds__(kBS) k_true_res(int64_t n, int R,
                                                  const int32_t* __restrict__ rowptr,
                                                  const int32_t* __restrict__ cols,
                                                  const double* __restrict__ K,
                                                  const double* __restrict__ M,
                                                  const double* __restrict__ omu,
                                                  const double2* __restrict__ X,
                                                  const double2* __restrict__ B, double* part) {
  constexpr int RPB = kBS / CT;
  const int c = threadIdx.x % CT, rs = threadIdx.x / CT;
  const int T = R / CT;
  const Sweep sw = xcd_sweep(n, RPB);
  double v[TMAX][1];
  double om[TMAX];
#pragma unroll
  for (int t = 0; t < TMAX; ++t) {
    v[t][0] = 0.0;
    om[t] = t < T ? omu[c + t * CT] : 0.0;
  }
  for (int64_t row = sw.first + rs; row < sw.end; row += sw.stride) {
    double2 acc[TMAX];
#pragma unroll
    for (int t = 0; t < TMAX; ++t) acc[t] = make_double2(0.0, 0.0);
    for (int k = rowptr[row]; k < rowptr[row + 1]; ++k) {
      const double kv = K[k], mv = M[k];
      const int64_t cj = (int64_t)cols[k] * R;
#pragma unroll
      for (int t = 0; t < TMAX; ++t)
        if (t < T) {
          const double2 x = X[cj + c + t * CT];
          const double ai = -om[t] * mv;
          acc[t].x += kv * x.x - ai * x.y;
          acc[t].y += kv * x.y + ai * x.x;
        }
    }
#pragma unroll
    for (int t = 0; t < TMAX; ++t)
      if (t < T) {
        const double2 bv = B[row * R + c + t * CT];
        const double rx = bv.x - acc[t].x, ry = bv.y - acc[t].y;
        v[t][0] += rx * rx + ry * ry;
      }
  }
  column_partials<CT, 1, TMAX>(R, v, part);
}

__global__ void __launch_bounds__(kBS) k_fin_true(Col* col, int R, const double* part, int np) {
  double tot[1];
  column_totals<1>(part, R, np, tot);
  if (threadIdx.x == 0) col[blockIdx.x].res2 = tot[0];
}

// restart from the current iterate: r = b - A x, p = 0; partials of
// rho = r^T z (z = r / diag) and |r|^2 per column
template <int CT, int TMAX>
__global__ void __launch_bounds__(kBS) k_reinit(int64_t n, int R,
                                                const int32_t* __restrict__ rowptr,
                                                const int32_t* __restrict__ cols,
                                                const double* __restrict__ K,
                                                const double* __restrict__ M,
                                                const double* __restrict__ dK,
                                                const double* __restrict__ dM,
                                                const double* __restrict__ omu,
                                                const double2* __restrict__ X,
                                                const double2* __restrict__ B,
                                                double2* __restrict__ Rr,
                                                double2* __restrict__ P, double* part) {
  constexpr int RPB = kBS / CT;
  const int c = threadIdx.x % CT, rs = threadIdx.x / CT;
  const int T = R / CT;
  const Sweep sw = xcd_sweep(n, RPB);
  double v[TMAX][3];
  double om[TMAX];
#pragma unroll
  for (int t = 0; t < TMAX; ++t) {
    v[t][0] = v[t][1] = v[t][2] = 0.0;
    om[t] = t < T ? omu[c + t * CT] : 0.0;
  }
  for (int64_t row = sw.first + rs; row < sw.end; row += sw.stride) {
    double2 acc[TMAX];
#pragma unroll
    for (int t = 0; t < TMAX; ++t) acc[t] = make_double2(0.0, 0.0);
    for (int k = rowptr[row]; k < rowptr[row + 1]; ++k) {
      const double kv = K[k], mv = M[k];
      const int64_t cj = (int64_t)cols[k] * R;
#pragma unroll
      for (int t = 0; t < TMAX; ++t)
        if (t < T) {
          const double2 x = X[cj + c + t * CT];
          const double ai = -om[t] * mv;
          acc[t].x += kv * x.x - ai * x.y;
          acc[t].y += kv * x.y + ai * x.x;
        }
    }
    const double kd = dK[row], md = dM[row];
#pragma unroll
    for (int t = 0; t < TMAX; ++t)
      if (t < T) {
        const int64_t idx = row * R + c + t * CT;
        const double2 bv = B[idx];
        const double2 rv = make_double2(bv.x - acc[t].x, bv.y - acc[t].y);
        Rr[idx] = rv;
        const double2 zv = jacobi_z(rv, kd, md, om[t]);
        P[idx] = make_double2(0.0, 0.0);
        v[t][0] += rv.x * zv.x - rv.y * zv.y;
        v[t][1] += rv.x * zv.y + rv.y * zv.x;
        v[t][2] += rv.x * rv.x + rv.y * rv.y;
      }
  }
  column_partials<CT, 3, TMAX>(R, v, part);
}

// finish a restart: columns whose true residual already meets the tolerance
// (or that stopped on breakdown or the iteration limit) stay done
__global__ void __launch_bounds__(kBS) k_fin_reinit(const Ctl* ctl, Col* col, int R,
                                                    const double* part, int np) {
  double tot[3];
  column_totals<3>(part, R, np, tot);
  if (threadIdx.x != 0) return;
  Col& cc = col[blockIdx.x];
  cc.rho = make_double2(tot[0], tot[1]);
  cc.res2 = tot[2];
  cc.beta = make_double2(0.0, 0.0);
  cc.alpha = make_double2(0.0, 0.0);
  cc.ref2 = 0.0;
  cc.done = (cc.status != 0 || cc.iter >= ctl->maxit || tot[2] <= ctl->tol2 * cc.bnorm2) ? 1 : 0;
}

__global__ void k_ctl_resume(Ctl* ctl) {
  if (threadIdx.x == 0 && blockIdx.x == 0) ctl->all_done = 0;
}
__global__ void k_ctl_stall(Ctl* ctl, int limit) {
  if (threadIdx.x == 0 && blockIdx.x == 0) ctl->stall_limit = limit;
}

// Mixed precision, refinement step: r64 = b - A x (FP64, the caller's K and M),
// then the complex64 recursion restarts from it: r = demote(r64), p = 0,
// x32 = 0.  Partials of rho = r^T z (z = r / diag, complex64 values) and |r64|^2.
template <int CT, int TMAX>
__global__ void __launch_bounds__(kBS) k_refine(int64_t n, int R,
                                                const int32_t* __restrict__ rowptr,
                                                const int32_t* __restrict__ cols,
                                                const double* __restrict__ K,
                                                const double* __restrict__ M,
                                                const double* __restrict__ dK,
                                                const double* __restrict__ dM,
                                                const double* __restrict__ omu,
                                                const double2* __restrict__ X,
                                                const double2* __restrict__ B,
                                                float2* __restrict__ Rr,
                                                float2* __restrict__ P, float2* __restrict__ X32,
                                                double* part) {
  constexpr int RPB = kBS / CT;
  const int c = threadIdx.x % CT, rs = threadIdx.x / CT;
  const int T = R / CT;
  const Sweep sw = xcd_sweep(n, RPB);
  double v[TMAX][3];
  double om[TMAX];
#pragma unroll
  for (int t = 0; t < TMAX; ++t) {
    v[t][0] = v[t][1] = v[t][2] = 0.0;
    om[t] = t < T ? omu[c + t * CT] : 0.0;
  }
  for (int64_t row = sw.first + rs; row < sw.end; row += sw.stride) {
    double2 acc[TMAX];
#pragma unroll
    for (int t = 0; t < TMAX; ++t) acc[t] = make_double2(0.0, 0.0);
    for (int k = rowptr[row]; k < rowptr[row + 1]; ++k) {
      const double kv = K[k], mv = M[k];
      const int64_t cj = (int64_t)cols[k] * R;
#pragma unroll
      for (int t = 0; t < TMAX; ++t)
        if (t < T) {
          const double2 x = X[cj + c + t * CT];
          const double ai = -om[t] * mv;
          acc[t].x += kv * x.x - ai * x.y;
          acc[t].y += kv * x.y + ai * x.x;
        }
    }
    const double kd = dK[row], md = dM[row];
#pragma unroll
    for (int t = 0; t < TMAX; ++t)
      if (t < T) {
        const int64_t idx = row * R + c + t * CT;
        const double2 bv = B[idx];
        const double2 r64 = make_double2(bv.x - acc[t].x, bv.y - acc[t].y);
        const float2 rv = from_d<float2>(r64);
        Rr[idx] = rv;
        const float2 zv = jacobi_z(rv, kd, md, om[t]);
        P[idx] = make_float2(0.f, 0.f);
        X32[idx] = make_float2(0.f, 0.f);
        const double2 rd = to_d(rv), zd = to_d(zv);
        v[t][0] += rd.x * zd.x - rd.y * zd.y;
        v[t][1] += rd.x * zd.y + rd.y * zd.x;
        v[t][2] += r64.x * r64.x + r64.y * r64.y;
      }
  }
  column_partials<CT, 3, TMAX>(R, v, part);
}

// finish a refinement step: first time |b|^2 (x = 0); a column whose FP64
// residual meets the tolerance (or that stopped) is finished for good, the
// others restart the complex64 recursion with target |r|^2 <= res64 * delta2
__global__ void __launch_bounds__(kBS) k_fin_refine(const Ctl* ctl, Col* col, int R,
                                                    const double* part, int np, int first,
                                                    double delta2) {
  double tot[3];
  column_totals<3>(part, R, np, tot);
  if (threadIdx.x != 0) return;
  Col& cc = col[blockIdx.x];
  if (first) {
    cc = Col{};
    cc.bnorm2 = tot[2];
  }
  // a round that stagnated after cutting the true residual at least tenfold
  // is followed by another complex64 round (iterative refinement limited by
  // complex64 accuracy); without that progress the column finishes in FP64
  if (!first && cc.stalled && tot[2] <= 1e-2 * 1e-2 * cc.res64) cc.stalled = 0;
  cc.res64 = tot[2];
  cc.res2 = tot[2];
  cc.rho = make_double2(tot[0], tot[1]);
  cc.beta = make_double2(0.0, 0.0);
  cc.alpha = make_double2(0.0, 0.0);
  cc.ref2 = tot[2] * delta2;
  cc.best2 = tot[2];
  cc.last_imp = cc.iter;
  cc.outer = (cc.status != 0 || cc.stalled || cc.iter >= ctl->maxit ||
              tot[2] <= ctl->tol2 * cc.bnorm2) ? 1 : 0;
  cc.done = cc.outer;
}

// x64 += promote(x32) for every column
__global__ void k_accum(int64_t n, double2* __restrict__ X, const float2* __restrict__ X32) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float2 v = X32[i];
  X[i] = make_double2(X[i].x + (double)v.x, X[i].y + (double)v.y);
}

// ---------------------------------------------------------------------------
// Entity-block Jacobi for the batch: per mesh entity (block_ptr range of d <= 6
// DOFs) and per distinct frequency f, the inverse of A_f restricted to the
// entity's DOFs, A_f = K - i omega_f mu0 M; column r uses frequency fidx[r].
// ---------------------------------------------------------------------------
constexpr int kBlk = 6;

// one thread per (entity, frequency): gather the d x d block from the CSR K, M
// and invert it (Gauss-Jordan, partial pivoting); row-major at
// binv[f * nb2 + boff[e]] (boff: prefix sums of d^2, nb2 = their total)
__global__ void k_batch_binv(int64_t nent, int nf, int64_t nb2, const int32_t* __restrict__ bptr,
                             const int32_t* __restrict__ boff,
                             const int32_t* __restrict__ rowptr,
                             const int32_t* __restrict__ cols, const double* __restrict__ K,
                             const double* __restrict__ M, const double* __restrict__ omf,
                             double2* __restrict__ binv, int32_t* __restrict__ bad) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nent * nf) return;
  const int64_t e = i % nent;
  const int f = (int)(i / nent);
  const int i0 = bptr[e], d = bptr[e + 1] - i0;
  double2 a[kBlk][kBlk], inv[kBlk][kBlk];
  for (int t = 0; t < d; ++t) {
    const int row = i0 + t;
    int lo = rowptr[row], hi = rowptr[row + 1] - 1, pos = -1;
    while (lo <= hi) {
      const int mid = (lo + hi) >> 1;
      const int c = cols[mid];
      if (c == i0) {
        pos = mid;
        break;
      }
      if (c < i0) lo = mid + 1;
      else hi = mid - 1;
    }
    if (pos < 0 || pos + d > rowptr[row + 1] || cols[pos + d - 1] != i0 + d - 1) {
      bad[0] = 1;
      return;
    }
    for (int u = 0; u < d; ++u) {
      a[t][u] = make_double2(K[pos + u], -omf[f] * M[pos + u]);
      inv[t][u] = make_double2(t == u ? 1.0 : 0.0, 0.0);
    }
  }
  for (int c = 0; c < d; ++c) {
    int piv = c;
    double best = hypot(a[c][c].x, a[c][c].y);
    for (int t = c + 1; t < d; ++t) {
      const double v = hypot(a[t][c].x, a[t][c].y);
      if (v > best) {
        best = v;
        piv = t;
      }
    }
    if (!(best > 0.0)) {
      bad[0] = 1;
      return;
    }
    if (piv != c)
      for (int u = 0; u < d; ++u) {
        const double2 t1 = a[c][u], t2 = inv[c][u];
        a[c][u] = a[piv][u];
        inv[c][u] = inv[piv][u];
        a[piv][u] = t1;
        inv[piv][u] = t2;
      }
    const double2 ip = zdiv(make_double2(1.0, 0.0), a[c][c]);
    for (int u = 0; u < d; ++u) {
      a[c][u] = zmul(a[c][u], ip);
      inv[c][u] = zmul(inv[c][u], ip);
    }
    for (int t = 0; t < d; ++t) {
      if (t == c) continue;
      const double2 fct = a[t][c];
      for (int u = 0; u < d; ++u) {
        a[t][u] = zsub(a[t][u], zmul(fct, a[c][u]));
        inv[t][u] = zsub(inv[t][u], zmul(fct, inv[c][u]));
      }
    }
  }
  double2* out = binv + (int64_t)f * nb2 + boff[e];
  for (int t = 0; t < d; ++t)
    for (int u = 0; u < d; ++u) out[t * d + u] = inv[t][u];
}

// entities of this workgroup (the same chunking as the row kernels, over entities)
__device__ inline void ents_of_block(int64_t nent, int64_t* lo, int64_t* hi) {
  rows_of_block(nent, lo, hi);
}

// z_k = (B_e r)_k over an entity's d DOFs (B_e row-major d x d), as every
// kernel that forms the preconditioned residual computes it
template <typename VT>
__device__ inline VT block_z(const double2* Bm, const double2* rl, int d, int k) {
  double2 zk = make_double2(0.0, 0.0);
#pragma unroll
  for (int u = 0; u < kBlk; ++u)
    if (u < d) zk = zadd(zk, zmul(Bm[k * d + u], rl[u]));
  return from_d<VT>(zk);
}

// r -= alpha q; partials of r^T z and |r|^2 with z = B_e r (entity blocks).
// As in k_update, z is not stored and x += alpha p is left to k_pupdate_blk.
template <int CT, int TMAX, typename VT>
__global__ void __launch_bounds__(kBS) k_update_blk(const Ctl* ctl, const Col* col, int64_t nent,
                                                    int R, const int32_t* __restrict__ bptr,
                                                    const int32_t* __restrict__ boff, int64_t nb2,
                                                    const double2* __restrict__ binv,
                                                    const int32_t* __restrict__ fidx,
                                                    VT* __restrict__ Rr, const VT* __restrict__ Q,
                                                    double* part) {
  if (ctl->all_done) return;
  constexpr int EPB = kBS / CT;
  const int c = threadIdx.x % CT, es = threadIdx.x / CT;
  const int T = R / CT;
  int64_t lo, hi;
  ents_of_block(nent, &lo, &hi);
  double v[TMAX][3];
  VT al[TMAX];
  int fi[TMAX];
#pragma unroll
  for (int t = 0; t < TMAX; ++t) {
    v[t][0] = v[t][1] = v[t][2] = 0.0;
    al[t] = from_d<VT>(t < T ? col[c + t * CT].alpha : make_double2(0.0, 0.0));
    fi[t] = t < T ? fidx[c + t * CT] : 0;
  }
  for (int64_t e = lo + es; e < hi; e += EPB) {
    const int i0 = bptr[e], d = bptr[e + 1] - i0;
#pragma unroll
    for (int t = 0; t < TMAX; ++t)
      if (t < T) {
        const int cc = c + t * CT;
        double2 rl[kBlk];
#pragma unroll
        for (int k = 0; k < kBlk; ++k)
          if (k < d) {
            const int64_t idx = (int64_t)(i0 + k) * R + cc;
            const VT qv = Q[idx];
            VT rv = Rr[idx];
            rv.x -= al[t].x * qv.x - al[t].y * qv.y;
            rv.y -= al[t].x * qv.y + al[t].y * qv.x;
            Rr[idx] = rv;
            rl[k] = to_d(rv);
          }
        const double2* Bm = binv + (int64_t)fi[t] * nb2 + boff[e];
#pragma unroll
        for (int k = 0; k < kBlk; ++k)
          if (k < d) {
            const double2 zd = to_d(block_z<VT>(Bm, rl, d, k));
            v[t][0] += rl[k].x * zd.x - rl[k].y * zd.y;
            v[t][1] += rl[k].x * zd.y + rl[k].y * zd.x;
            v[t][2] += rl[k].x * rl[k].x + rl[k].y * rl[k].y;
          }
      }
  }
  column_partials<CT, 3, TMAX>(R, v, part);
}

// x += alpha p (the previous iteration's step), z = B_e r, p = z + beta p
template <int CT, int TMAX, typename VT>
__global__ void __launch_bounds__(kBS) k_pupdate_blk(const Ctl* ctl, const Col* col, int64_t nent,
                                                     int R, const int32_t* __restrict__ bptr,
                                                     const int32_t* __restrict__ boff, int64_t nb2,
                                                     const double2* __restrict__ binv,
                                                     const int32_t* __restrict__ fidx,
                                                     VT* __restrict__ X, const VT* __restrict__ Rr,
                                                     VT* __restrict__ P) {
  if (ctl->all_done) return;
  constexpr int EPB = kBS / CT;
  const int c = threadIdx.x % CT, es = threadIdx.x / CT;
  const int T = R / CT;
  int64_t lo, hi;
  ents_of_block(nent, &lo, &hi);
  VT al[TMAX], be[TMAX];
  int fi[TMAX];
#pragma unroll
  for (int t = 0; t < TMAX; ++t) {
    al[t] = from_d<VT>(t < T ? col[c + t * CT].alpha : make_double2(0.0, 0.0));
    be[t] = from_d<VT>(t < T ? col[c + t * CT].beta : make_double2(0.0, 0.0));
    fi[t] = t < T ? fidx[c + t * CT] : 0;
  }
  for (int64_t e = lo + es; e < hi; e += EPB) {
    const int i0 = bptr[e], d = bptr[e + 1] - i0;
#pragma unroll
    for (int t = 0; t < TMAX; ++t)
      if (t < T) {
        const int cc = c + t * CT;
        double2 rl[kBlk];
#pragma unroll
        for (int k = 0; k < kBlk; ++k)
          if (k < d) rl[k] = to_d(Rr[(int64_t)(i0 + k) * R + cc]);
        const double2* Bm = binv + (int64_t)fi[t] * nb2 + boff[e];
#pragma unroll
        for (int k = 0; k < kBlk; ++k)
          if (k < d)
            step_and_direction(X, P, (int64_t)(i0 + k) * R + cc, al[t], be[t],
                               block_z<VT>(Bm, rl, d, k));
      }
  }
}

// the r^T z partials (z = B_e r) for every entity and column after a
// (re)start that formed r; overwrites slots 0, 1 of the kernel before it
template <int CT, int TMAX, typename VT>
__global__ void __launch_bounds__(kBS) k_bapply(int64_t nent, int R,
                                                const int32_t* __restrict__ bptr,
                                                const int32_t* __restrict__ boff, int64_t nb2,
                                                const double2* __restrict__ binv,
                                                const int32_t* __restrict__ fidx,
                                                const VT* __restrict__ Rr, double* part) {
  constexpr int EPB = kBS / CT;
  const int c = threadIdx.x % CT, es = threadIdx.x / CT;
  const int T = R / CT;
  int64_t lo, hi;
  ents_of_block(nent, &lo, &hi);
  double v[TMAX][2];
#pragma unroll
  for (int t = 0; t < TMAX; ++t) v[t][0] = v[t][1] = 0.0;
  for (int64_t e = lo + es; e < hi; e += EPB) {
    const int i0 = bptr[e], d = bptr[e + 1] - i0;
#pragma unroll
    for (int t = 0; t < TMAX; ++t)
      if (t < T) {
        const int cc = c + t * CT;
        double2 rl[kBlk];
#pragma unroll
        for (int k = 0; k < kBlk; ++k)
          if (k < d) rl[k] = to_d(Rr[(int64_t)(i0 + k) * R + cc]);
        const double2* Bm = binv + (int64_t)fidx[cc] * nb2 + boff[e];
#pragma unroll
        for (int k = 0; k < kBlk; ++k)
          if (k < d) {
            const double2 zd = to_d(block_z<VT>(Bm, rl, d, k));
            v[t][0] += rl[k].x * zd.x - rl[k].y * zd.y;
            v[t][1] += rl[k].x * zd.y + rl[k].y * zd.x;
          }
      }
  }
  column_partials<CT, 2, TMAX>(R, v, part);
}

__global__ void k_diag(int64_t n, const int32_t* __restrict__ rowptr,
                       const int32_t* __restrict__ cols, const double* __restrict__ K,
                       const double* __restrict__ M, double* __restrict__ dK,
                       double* __restrict__ dM, int32_t* bad) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  int lo = rowptr[r], hi = rowptr[r + 1] - 1;
  double kd = 0.0, md = 0.0;
  while (lo <= hi) {
    const int mid = (lo + hi) >> 1;
    const int c = cols[mid];
    if (c == r) {
      kd = K[mid];
      md = M[mid];
      break;
    }
    if (c < r) lo = mid + 1;
    else hi = mid - 1;
  }
  if (kd == 0.0 && md == 0.0) {
    bad[0] = 1;
    kd = 1.0;
  }
  dK[r] = kd;
  dM[r] = md;
}

__global__ void k_ctl_init(Ctl* ctl, int maxit, double tol2, int R) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  Ctl c{};
  c.tol2 = tol2;
  c.maxit = maxit;
  c.ncols = R;
  *ctl = c;
}

struct Batch {
  int64_t n = 0, nnz = 0;
  int R = 0, CT = 64, TM = 1;
  int grid = 1;
  const int32_t *rowptr = nullptr, *cols = nullptr;
  const double *K = nullptr, *M = nullptr;
  double *omu = nullptr, *dK = nullptr, *dM = nullptr, *part = nullptr;
  double2 *Rr = nullptr, *P = nullptr, *Q = nullptr;
  Ctl* ctl = nullptr;
  Col* col = nullptr;
  Ctl* hctl = nullptr;
  Col* hcol = nullptr;
  int32_t* bad = nullptr;
  double rtol = 1e-8;
  int maxit = 100000, check = 25;
  // entity rows (optional): schedule descriptors {r0, d | L << 8, cbase, vbase},
  // each scheduled entity's CSR slot, the scheduled column lists and K, M
  int64_t nent = 0;
  int4* em = nullptr;
  int32_t *esrc = nullptr, *cols_s = nullptr;
  double *Ks = nullptr, *Ms = nullptr;
  int dmax = 0, egrid = 1;
  // mixed precision (optional): complex64 recursion vectors, float K, M in
  // schedule order, residual drop per refinement
  bool mixed = false;
  double delta2 = 1e-8;
  int stall = 400;
  // entity-block Jacobi (optional): entity DOF ranges, d^2 prefix sums, block
  // inverses per distinct frequency, frequency index per column
  int32_t *bptr = nullptr, *boff = nullptr, *fidx = nullptr;
  int64_t nb2 = 0;
  int nfreq = 0;
  double2* binv = nullptr;
  std::vector<double> hom;  // omega * mu0 per column (host)
  float2 *R32 = nullptr, *P32 = nullptr, *Q32 = nullptr, *X32 = nullptr;
  float *Ks32 = nullptr, *Ms32 = nullptr;
  // element-by-element products (optional, pg_batch_set_element_operator):
  // the operator, its summation schedule, the local results per precision and
  // the summation grid
  const ElemOp* eop = nullptr;
  EopSchedule esched{};
  double2* ye64 = nullptr;
  float2* ye32 = nullptr;
  int sgrid = 1;
  double* lpart = nullptr;  // the local products' x^T y partials (nlpart per column)
  int nlpart = 0;
  std::vector<void*> owned;
  ~Batch() {  // release errors cannot be reported from here
    for (void* v : owned) (void)hipFree(v);
    if (hctl) (void)hipHostFree(hctl);
    if (hcol) (void)hipHostFree(hcol);
  }
  template <typename T>
  int alloc(T** p, size_t count) {
    void* v = nullptr;
    hipError_t e = hipMalloc(&v, std::max<size_t>(count, 1) * sizeof(T));
    if (e != hipSuccess) PG_FAIL(PG_ERR_HIP, "hipMalloc(%zu): %s", count * sizeof(T), hipGetErrorString(e));
    owned.push_back(v);
    *p = static_cast<T*>(v);
    return PG_OK;
  }
};

template <int CT, int TM, typename VT, typename MT>
static void launch_iter_t(Batch* B, hipStream_t st, VT* X, VT* Rr, VT* P, VT* Q, const MT* Ks,
                          const MT* Ms) {
  const dim3 g(B->grid), b(kBS), gr(B->R);
  if (B->binv)
    hipLaunchKernelGGL((k_pupdate_blk<CT, TM, VT>), g, b, 0, st, B->ctl, B->col, B->nent, B->R,
                       B->bptr, B->boff, B->nb2, B->binv, B->fidx, X, Rr, P);
  else
    hipLaunchKernelGGL((k_pupdate<CT, TM, VT>), g, b, 0, st, B->ctl, B->col, B->n, B->R, B->dK,
                       B->dM, B->omu, X, Rr, P);
  int np = B->grid;
  if (B->eop) {
    const int32_t* done = &B->ctl->all_done;
    VT* ye = nullptr;
    if constexpr (sizeof(VT) == sizeof(float2)) ye = B->ye32;
    else ye = B->ye64;
    eop_multi_local(B->eop, B->R, done, B->omu, P, ye, B->lpart, st);
    const EopSchedule& es = B->esched;
    np = B->sgrid;
    if (es.D <= 1)
      hipLaunchKernelGGL((k_meop_sum<CT, TM, 1, VT>), dim3(B->sgrid), b, 0, st, B->ctl, es.nent,
                         B->R, es.gm, es.gslot, ye, P, Q, B->part);
    else
      hipLaunchKernelGGL((k_meop_sum<CT, TM, 2, VT>), dim3(B->sgrid), b, 0, st, B->ctl, es.nent,
                         B->R, es.gm, es.gslot, ye, P, Q, B->part);
  } else if (B->em) {
    const dim3 ge(B->egrid);
    np = B->egrid;
    if (B->dmax <= 1)
      hipLaunchKernelGGL((k_spmm_er<CT, TM, 1, VT, MT>), ge, b, 0, st, B->ctl, B->nent, B->R,
                         B->em, B->cols_s, Ks, Ms, B->omu, P, Q, B->part);
    else if (B->dmax <= 2)
      hipLaunchKernelGGL((k_spmm_er<CT, TM, 2, VT, MT>), ge, b, 0, st, B->ctl, B->nent, B->R,
                         B->em, B->cols_s, Ks, Ms, B->omu, P, Q, B->part);
    else
      hipLaunchKernelGGL((k_spmm_er<CT, TM, 6, VT, MT>), ge, b, 0, st, B->ctl, B->nent, B->R,
                         B->em, B->cols_s, Ks, Ms, B->omu, P, Q, B->part);
  } else {
    hipLaunchKernelGGL((k_spmm<CT, TM>), g, b, 0, st, B->ctl, B->n, B->R, B->rowptr, B->cols,
                       B->K, B->M, B->omu, (const double2*)P, (double2*)Q, B->part);
  }
  hipLaunchKernelGGL(k_fin_mu, gr, b, 0, st, B->ctl, B->col, B->R, B->part, np,
                     B->eop ? B->lpart : nullptr, B->nlpart);
  if (B->binv)
    hipLaunchKernelGGL((k_update_blk<CT, TM, VT>), g, b, 0, st, B->ctl, B->col, B->nent, B->R,
                       B->bptr, B->boff, B->nb2, B->binv, B->fidx, Rr, Q, B->part);
  else
    hipLaunchKernelGGL((k_update<CT, TM, VT>), g, b, 0, st, B->ctl, B->col, B->n, B->R, B->dK,
                       B->dM, B->omu, Rr, Q, B->part);
  hipLaunchKernelGGL(k_fin_update, gr, b, 0, st, B->ctl, B->col, B->R, B->part, B->grid);
  hipLaunchKernelGGL(k_fin_all, dim3(1), b, 0, st, B->ctl, B->col, B->R);
}

template <int CT, int TM>
static void launch_iter(Batch* B, hipStream_t st, double2* X) {
  if (B->mixed)
    launch_iter_t<CT, TM, float2, float>(B, st, B->X32, B->R32, B->P32, B->Q32, B->Ks32,
                                         B->Ms32);
  else
    launch_iter_t<CT, TM, double2, double>(B, st, X, B->Rr, B->P, B->Q, B->Ks, B->Ms);
}

// the last iteration's step x += alpha p once every column has paused
static void launch_xstep(Batch* B, hipStream_t st, double2* X) {
  const int64_t nr = B->n * (int64_t)B->R;
  const dim3 g((unsigned)((nr + kBS - 1) / kBS)), b(kBS);
  if (B->mixed)
    hipLaunchKernelGGL(k_bxstep<float2>, g, b, 0, st, B->col, nr, B->R, B->P32, B->X32);
  else
    hipLaunchKernelGGL(k_bxstep<double2>, g, b, 0, st, B->col, nr, B->R, B->P, X);
}

template <int CT, int TM>
static void launch_refine(Batch* B, hipStream_t st, const double2* Bv, const double2* X,
                          int first) {
  hipLaunchKernelGGL((k_refine<CT, TM>), dim3(B->grid), dim3(kBS), 0, st, B->n, B->R, B->rowptr,
                     B->cols, B->K, B->M, B->dK, B->dM, B->omu, X, Bv, B->R32, B->P32,
                     B->X32, B->part);
  if (B->binv)
    hipLaunchKernelGGL((k_bapply<CT, TM, float2>), dim3(B->grid), dim3(kBS), 0, st, B->nent, B->R,
                       B->bptr, B->boff, B->nb2, B->binv, B->fidx, B->R32, B->part);
  hipLaunchKernelGGL(k_fin_refine, dim3(B->R), dim3(kBS), 0, st, B->ctl, B->col, B->R, B->part,
                     B->grid, first, B->delta2);
  hipLaunchKernelGGL(k_ctl_resume, dim3(1), dim3(64), 0, st, B->ctl);
  hipLaunchKernelGGL(k_fin_all, dim3(1), dim3(kBS), 0, st, B->ctl, B->col, B->R);
}

template <int CT, int TM>
static void launch_init(Batch* B, hipStream_t st, const double2* Bv, double2* X) {
  hipLaunchKernelGGL((k_init<CT, TM>), dim3(B->grid), dim3(kBS), 0, st, B->n, B->R, B->dK, B->dM,
                     B->omu, Bv, X, B->Rr, B->P, B->part);
  if (B->binv)
    hipLaunchKernelGGL((k_bapply<CT, TM, double2>), dim3(B->grid), dim3(kBS), 0, st, B->nent, B->R,
                       B->bptr, B->boff, B->nb2, B->binv, B->fidx, B->Rr, B->part);
  hipLaunchKernelGGL(k_fin_init, dim3(B->R), dim3(kBS), 0, st, B->col, B->R, B->part, B->grid);
}

template <int CT, int TM>
static void launch_true(Batch* B, hipStream_t st, const double2* Bv, const double2* X) {
  hipLaunchKernelGGL((k_true_res<CT, TM>), dim3(B->grid), dim3(kBS), 0, st, B->n, B->R, B->rowptr,
                     B->cols, B->K, B->M, B->omu, X, Bv, B->part);
  hipLaunchKernelGGL(k_fin_true, dim3(B->R), dim3(kBS), 0, st, B->col, B->R, B->part, B->grid);
}

// Jacobi diagonals K_ii, M_ii from the caller's K and M, and the schedule-order
// copies of K and M (at creation and after the caller has reassembled them in place)
static int refresh_diag(Batch* B, hipStream_t st) {
  if (B->em) {
    hipLaunchKernelGGL((k_sched_km<double>), dim3((32 * B->nent + kBS - 1) / kBS), dim3(kBS), 0,
                       st, B->nent, B->em, B->esrc, B->K, B->M, B->Ks, B->Ms);
    if (B->mixed)
      hipLaunchKernelGGL((k_sched_km<float>), dim3((32 * B->nent + kBS - 1) / kBS), dim3(kBS), 0,
                         st, B->nent, B->em, B->esrc, B->K, B->M, B->Ks32, B->Ms32);
  }
  PG_HIP(hipMemsetAsync(B->bad, 0, sizeof(int32_t), st));
  hipLaunchKernelGGL(k_diag, dim3((B->n + kBS - 1) / kBS), dim3(kBS), 0, st, B->n, B->rowptr,
                     B->cols, B->K, B->M, B->dK, B->dM, B->bad);
  if (B->binv) {
    std::vector<double> fo;  // distinct frequencies, in order of first appearance
    for (double v : B->hom)
      if (std::find(fo.begin(), fo.end(), v) == fo.end()) fo.push_back(v);
    double* dfo = reinterpret_cast<double*>(B->part);  // staging: partials are free here
    PG_HIP(hipMemcpyAsync(dfo, fo.data(), fo.size() * sizeof(double), hipMemcpyHostToDevice, st));
    const int64_t work = B->nent * (int64_t)fo.size();
    hipLaunchKernelGGL(k_batch_binv, dim3((unsigned)((work + kBS - 1) / kBS)), dim3(kBS), 0, st,
                       B->nent, (int)fo.size(), B->nb2, B->bptr, B->boff, B->rowptr, B->cols, B->K,
                       B->M, dfo, B->binv, B->bad);
  }
  PG_CHECK_LAUNCH();
  PG_HIP(hipMemcpyAsync(&B->hctl->maxit, B->bad, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  PG_HIP(hipStreamSynchronize(st));
  if (B->hctl->maxit != 0)
    PG_FAIL(PG_ERR_ARG, "zero diagonal entry or singular block: preconditioner undefined");
  return PG_OK;
}

template <int CT, int TM>
static void launch_reinit(Batch* B, hipStream_t st, const double2* Bv, const double2* X) {
  hipLaunchKernelGGL((k_reinit<CT, TM>), dim3(B->grid), dim3(kBS), 0, st, B->n, B->R, B->rowptr,
                     B->cols, B->K, B->M, B->dK, B->dM, B->omu, X, Bv, B->Rr, B->P,
                     B->part);
  if (B->binv)
    hipLaunchKernelGGL((k_bapply<CT, TM, double2>), dim3(B->grid), dim3(kBS), 0, st, B->nent, B->R,
                       B->bptr, B->boff, B->nb2, B->binv, B->fidx, B->Rr, B->part);
  hipLaunchKernelGGL(k_fin_reinit, dim3(B->R), dim3(kBS), 0, st, B->ctl, B->col, B->R, B->part,
                     B->grid);
  hipLaunchKernelGGL(k_ctl_resume, dim3(1), dim3(64), 0, st, B->ctl);
  hipLaunchKernelGGL(k_fin_all, dim3(1), dim3(kBS), 0, st, B->ctl, B->col, B->R);
}

// (CT, TM) configurations: R = CT * T with T <= TM
#define PG_BATCH_DISPATCH(FN, ...)                                \
  do {                                                            \
    if (B->CT == 64 && B->TM == 4) FN<64, 4>(__VA_ARGS__);        \
    else if (B->CT == 64 && B->TM == 1) FN<64, 1>(__VA_ARGS__);   \
    else if (B->CT == 32) FN<32, 1>(__VA_ARGS__);                 \
    else if (B->CT == 16) FN<16, 1>(__VA_ARGS__);                 \
    else if (B->CT == 8) FN<8, 1>(__VA_ARGS__);                   \
    else if (B->CT == 4) FN<4, 1>(__VA_ARGS__);                   \
    else if (B->CT == 2) FN<2, 1>(__VA_ARGS__);                   \
    else FN<1, 1>(__VA_ARGS__);                                   \
  } while (0)

}  // namespace batch
}  // namespace pg

using namespace pg;
using namespace pg::batch;

extern "C" {

int pg_batch_create(int64_t n, int64_t nnz, const int32_t* rowptr, const int32_t* cols,
                    const double* K, const double* M, int nrhs, const double* omega,
                    double rtol, int maxit, void** out, void* stream) {
  if (!out || n <= 0 || nnz <= 0 || !rowptr || !cols || !K || !M || !omega || nrhs <= 0)
    PG_FAIL(PG_ERR_ARG, "bad arguments");
  const int R = nrhs;
  int CT, TM;
  if (R >= 64) {
    if (R % 64 != 0 || R > 256)
      PG_FAIL(PG_ERR_ARG, "nrhs >= 64 must be a multiple of 64 and <= 256 (got %d)", R);
    CT = 64;
    TM = R == 64 ? 1 : 4;
  } else {
    if (R & (R - 1)) PG_FAIL(PG_ERR_ARG, "nrhs < 64 must be a power of two (got %d)", R);
    CT = R;
    TM = 1;
  }
  hipStream_t st = as_stream(stream);
  Batch* B = new Batch();
  B->n = n;
  B->nnz = nnz;
  B->R = R;
  B->CT = CT;
  B->TM = TM;
  B->rowptr = rowptr;
  B->cols = cols;
  B->K = K;
  B->M = M;
  B->rtol = rtol > 0 ? rtol : 1e-8;
  B->maxit = maxit > 0 ? maxit : 100000;
  const int64_t rpb = kBS / CT;
  B->grid = (int)std::max<int64_t>(1, std::min<int64_t>(kMaxGrid, (n + rpb - 1) / rpb));
  if (B->grid >= 64) B->grid = (B->grid / 8) * 8;
  int rc;
#define PG_BA(p, c)                                \
  if ((rc = B->alloc(&(p), (c))) != PG_OK) {       \
    delete B;                                      \
    return rc;                                     \
  }
  PG_BA(B->omu, R);
  PG_BA(B->dK, n);
  PG_BA(B->dM, n);
  PG_BA(B->part, (size_t)3 * R * kMaxGrid);  // partials of any launch shape
  PG_BA(B->Rr, (size_t)n * R);
  PG_BA(B->P, (size_t)n * R);
  PG_BA(B->Q, (size_t)n * R);
  PG_BA(B->ctl, 1);
  PG_BA(B->col, R);
  PG_BA(B->bad, 1);
#undef PG_BA
  if (hipHostMalloc(reinterpret_cast<void**>(&B->hctl), sizeof(Ctl), 0) != hipSuccess ||
      hipHostMalloc(reinterpret_cast<void**>(&B->hcol), R * sizeof(Col), 0) != hipSuccess) {
    delete B;
    PG_FAIL(PG_ERR_HIP, "hipHostMalloc failed");
  }
  std::vector<double> om(R);
  for (int r = 0; r < R; ++r) om[r] = omega[r] * 4.0 * 3.14159265358979323846 * 1e-7;
  B->hom = om;
  if (hipMemcpy(B->omu, om.data(), R * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
    delete B;
    PG_FAIL(PG_ERR_HIP, "frequency upload failed");
  }
  const int rc_diag = refresh_diag(B, st);
  if (rc_diag != PG_OK) {
    delete B;
    return rc_diag;
  }
  *out = B;
  return PG_OK;
}

int pg_batch_solve(void* batch, const double* b, double* x, int32_t* iterations, double* relres,
                   int32_t* status, void* stream) {
  if (!batch || !b || !x) PG_FAIL(PG_ERR_ARG, "bad arguments");
  Batch* B = static_cast<Batch*>(batch);
  hipStream_t st = as_stream(stream);
  const double2* Bv = reinterpret_cast<const double2*>(b);
  double2* X = reinterpret_cast<double2*>(x);
  hipLaunchKernelGGL(k_ctl_init, dim3(1), dim3(64), 0, st, B->ctl, B->maxit, B->rtol * B->rtol,
                     B->R);
  bool fp64_loop = !B->mixed;
  if (B->mixed) {
    // iterative refinement: complex64 COCG rounds between FP64 residuals
    hipLaunchKernelGGL(k_ctl_stall, dim3(1), dim3(64), 0, st, B->ctl, B->stall);
    PG_HIP(hipMemsetAsync(X, 0, (size_t)B->n * B->R * sizeof(double2), st));
    PG_BATCH_DISPATCH(launch_refine, B, st, Bv, X, 1);
    PG_CHECK_LAUNCH();
    for (int round = 0;; ++round) {
      PG_HIP(hipMemcpyAsync(B->hctl, B->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, st));
      PG_HIP(hipStreamSynchronize(st));
      if (B->hctl->all_done || round >= 200) break;
      while (true) {
        for (int k = 0; k < B->check; ++k) PG_BATCH_DISPATCH(launch_iter, B, st, X);
        PG_CHECK_LAUNCH();
        PG_HIP(hipMemcpyAsync(B->hctl, B->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, st));
        PG_HIP(hipStreamSynchronize(st));
        if (B->hctl->all_done) break;
      }
      launch_xstep(B, st, X);
      hipLaunchKernelGGL(k_accum, dim3((unsigned)(((int64_t)B->n * B->R + kBS - 1) / kBS)),
                         dim3(kBS), 0, st, (int64_t)B->n * B->R, X, B->X32);
      PG_BATCH_DISPATCH(launch_refine, B, st, Bv, X, 0);
      PG_CHECK_LAUNCH();
    }
    PG_HIP(hipMemcpyAsync(B->hcol, B->col, B->R * sizeof(Col), hipMemcpyDeviceToHost, st));
    PG_HIP(hipStreamSynchronize(st));
    // columns whose complex64 rounds stagnated continue in FP64 from their iterate
    for (int r = 0; r < B->R; ++r) {
      const Col& c = B->hcol[r];
      B->hcol[r].res2 = c.res64;
      if (c.status == 0 && c.iter < B->maxit && c.res64 > B->rtol * B->rtol * c.bnorm2)
        fp64_loop = true;
    }
    if (fp64_loop) {
      hipLaunchKernelGGL(k_ctl_stall, dim3(1), dim3(64), 0, st, B->ctl, 0);
      B->mixed = false;  // the iteration kernels below run in FP64
      PG_BATCH_DISPATCH(launch_reinit, B, st, Bv, X);
      PG_CHECK_LAUNCH();
    }
  } else {
    PG_BATCH_DISPATCH(launch_init, B, st, Bv, X);
    PG_CHECK_LAUNCH();
  }
  if (fp64_loop) {
    // iterate until every column's recursion has converged, then check the true
    // residuals; columns above the tolerance restart from their current iterate
    for (int restart = 0;; ++restart) {
      while (true) {
        for (int k = 0; k < B->check; ++k) PG_BATCH_DISPATCH(launch_iter, B, st, X);
        PG_CHECK_LAUNCH();
        PG_HIP(hipMemcpyAsync(B->hctl, B->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, st));
        PG_HIP(hipStreamSynchronize(st));
        if (B->hctl->all_done) break;
      }
      launch_xstep(B, st, X);
      PG_BATCH_DISPATCH(launch_true, B, st, Bv, X);
      PG_CHECK_LAUNCH();
      PG_HIP(hipMemcpyAsync(B->hcol, B->col, B->R * sizeof(Col), hipMemcpyDeviceToHost, st));
      PG_HIP(hipStreamSynchronize(st));
      bool again = false;
      for (int r = 0; r < B->R; ++r) {
        const Col& c = B->hcol[r];
        if (c.status == 0 && c.iter < B->maxit && c.res2 > B->rtol * B->rtol * c.bnorm2)
          again = true;
      }
      if (!again || restart >= 20) break;
      PG_BATCH_DISPATCH(launch_reinit, B, st, Bv, X);
      PG_CHECK_LAUNCH();
    }
    B->mixed = B->R32 != nullptr;  // restore the batch's precision for the next solve
  }
  for (int r = 0; r < B->R; ++r) {
    const Col& c = B->hcol[r];
    const double rel = c.bnorm2 > 0 ? std::sqrt(c.res2 / c.bnorm2) : 0.0;
    if (iterations) iterations[r] = c.iter;
    if (relres) relres[r] = rel;
    if (status) status[r] = c.status != 0 ? c.status : (rel <= B->rtol * 1.0000001 ? 0 : -1);
  }
  return PG_OK;
}

int pg_batch_set_entities(void* batch, const int32_t* block_ptr, int64_t nblocks,
                          const int32_t* order, void* stream) {
  if (!batch || !block_ptr || nblocks <= 0) PG_FAIL(PG_ERR_ARG, "bad arguments");
  Batch* B = static_cast<Batch*>(batch);
  if (B->em) PG_FAIL(PG_ERR_ARG, "entities already set");
  if (block_ptr[0] != 0 || block_ptr[nblocks] != B->n)
    PG_FAIL(PG_ERR_ARG, "block_ptr must start at 0 and end at n");
  int dmax = 0;
  for (int64_t k = 0; k < nblocks; ++k) {
    const int d = block_ptr[k + 1] - block_ptr[k];
    if (d < 1 || d > 6) PG_FAIL(PG_ERR_ARG, "entity sizes must be in 1..6");
    dmax = std::max(dmax, d);
  }
  if (order) {
    std::vector<char> seen(nblocks, 0);
    for (int64_t i = 0; i < nblocks; ++i) {
      if (order[i] < 0 || order[i] >= nblocks || seen[order[i]])
        PG_FAIL(PG_ERR_ARG, "order must be a permutation of 0..nblocks-1");
      seen[order[i]] = 1;
    }
  }
  hipStream_t st = as_stream(stream);
  int32_t* bp = nullptr;
  int rc;
  if ((rc = B->alloc(&bp, nblocks + 1)) != PG_OK) return rc;
  PG_HIP(hipMemcpyAsync(bp, block_ptr, (nblocks + 1) * sizeof(int32_t), hipMemcpyHostToDevice, st));
  PG_HIP(hipMemsetAsync(B->bad, 0, sizeof(int32_t), st));
  hipLaunchKernelGGL(k_check_entities, dim3((nblocks + kBS - 1) / kBS), dim3(kBS), 0, st, nblocks,
                     bp, B->rowptr, B->cols, B->bad);
  PG_CHECK_LAUNCH();
  PG_HIP(hipMemcpyAsync(&B->hctl->maxit, B->bad, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  PG_HIP(hipStreamSynchronize(st));
  if (B->hctl->maxit != 0)
    PG_FAIL(PG_ERR_ARG, "the rows of an entity must be consecutive with one shared column list");
  // schedule descriptors on the host
  std::vector<int32_t> rp(B->n + 1);
  PG_HIP(hipMemcpy(rp.data(), B->rowptr, (B->n + 1) * sizeof(int32_t), hipMemcpyDeviceToHost));
  std::vector<int4> em(nblocks);
  std::vector<int32_t> src(nblocks);
  int64_t cb = 0, vb = 0;
  for (int64_t i = 0; i < nblocks; ++i) {
    const int e = order ? order[i] : (int)i;
    const int r0 = block_ptr[e], d = block_ptr[e + 1] - r0;
    const int base = rp[r0], L = rp[r0 + 1] - base;
    if (L >= (1 << 23)) PG_FAIL(PG_ERR_ARG, "entity row too long");
    em[i] = make_int4(r0, d | (L << 8), (int)cb, (int)vb);
    src[i] = base;
    cb += L;
    vb += (int64_t)d * L;
    if (cb >= (1LL << 31) || vb >= (1LL << 31)) PG_FAIL(PG_ERR_ARG, "matrix too large");
  }
  if (vb != B->nnz) PG_FAIL(PG_ERR_ARG, "the entity rows do not cover the matrix");
  int4* dem = nullptr;
  if ((rc = B->alloc(&dem, nblocks)) != PG_OK || (rc = B->alloc(&B->esrc, nblocks)) != PG_OK ||
      (rc = B->alloc(&B->cols_s, cb)) != PG_OK || (rc = B->alloc(&B->Ks, B->nnz)) != PG_OK ||
      (rc = B->alloc(&B->Ms, B->nnz)) != PG_OK)
    return rc;
  PG_HIP(hipMemcpy(dem, em.data(), nblocks * sizeof(int4), hipMemcpyHostToDevice));
  PG_HIP(hipMemcpy(B->esrc, src.data(), nblocks * sizeof(int32_t), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_sched_cols, dim3((32 * nblocks + kBS - 1) / kBS), dim3(kBS), 0, st,
                     nblocks, dem, B->esrc, B->cols, B->cols_s);
  PG_CHECK_LAUNCH();
  // partials are sized for kMaxGrid workgroups
  const int64_t epb = kBS / B->CT;
  int64_t g = std::min<int64_t>(kMaxGrid, (nblocks + epb - 1) / epb);
  if (g >= 64) g = (g / 8) * 8;
  B->egrid = (int)std::max<int64_t>(1, g);
  B->nent = nblocks;
  B->dmax = dmax;
  B->em = dem;
  B->bptr = bp;
  return refresh_diag(B, st);
}

int pg_batch_set_mixed(void* batch, double inner_rtol, int stall_iters, void* stream) {
  if (!batch) PG_FAIL(PG_ERR_ARG, "batch is NULL");
  Batch* B = static_cast<Batch*>(batch);
  if (!B->em) PG_FAIL(PG_ERR_ARG, "mixed precision needs the entities (pg_batch_set_entities)");
  if (B->mixed) return PG_OK;
  const size_t nr = (size_t)B->n * B->R;
  int rc;
  if ((rc = B->alloc(&B->R32, nr)) != PG_OK || (rc = B->alloc(&B->P32, nr)) != PG_OK ||
      (rc = B->alloc(&B->Q32, nr)) != PG_OK ||
      (rc = B->alloc(&B->X32, nr)) != PG_OK || (rc = B->alloc(&B->Ks32, B->nnz)) != PG_OK ||
      (rc = B->alloc(&B->Ms32, B->nnz)) != PG_OK)
    return rc;
  if (B->eop &&
      (rc = B->alloc(&B->ye32, (size_t)B->esched.nw * kWave * B->esched.N * B->R)) != PG_OK)
    return rc;
  const double d = (inner_rtol > 0.0 && inner_rtol < 1.0) ? inner_rtol : 1e-4;
  B->delta2 = d * d;
  B->stall = stall_iters > 0 ? stall_iters : 400;
  B->mixed = true;
  return refresh_diag(B, as_stream(stream));
}

int pg_batch_set_block_jacobi(void* batch, void* stream) {
  if (!batch) PG_FAIL(PG_ERR_ARG, "batch is NULL");
  Batch* B = static_cast<Batch*>(batch);
  if (!B->em) PG_FAIL(PG_ERR_ARG, "block Jacobi needs the entities (pg_batch_set_entities)");
  if (B->binv) return PG_OK;
  std::vector<int32_t> bp(B->nent + 1), off(B->nent + 1);
  PG_HIP(hipMemcpy(bp.data(), B->bptr, (B->nent + 1) * sizeof(int32_t), hipMemcpyDeviceToHost));
  off[0] = 0;
  for (int64_t e = 0; e < B->nent; ++e) {
    const int d = bp[e + 1] - bp[e];
    off[e + 1] = off[e] + d * d;
  }
  std::vector<double> fo;
  std::vector<int32_t> fi(B->R);
  for (int r = 0; r < B->R; ++r) {
    auto it = std::find(fo.begin(), fo.end(), B->hom[r]);
    fi[r] = (int32_t)(it - fo.begin());
    if (it == fo.end()) fo.push_back(B->hom[r]);
  }
  if ((int64_t)fo.size() > kMaxGrid) PG_FAIL(PG_ERR_ARG, "too many distinct frequencies");
  int rc;
  if ((rc = B->alloc(&B->boff, B->nent + 1)) != PG_OK || (rc = B->alloc(&B->fidx, B->R)) != PG_OK ||
      (rc = B->alloc(&B->binv, (size_t)fo.size() * off[B->nent])) != PG_OK)
    return rc;
  PG_HIP(hipMemcpy(B->boff, off.data(), (B->nent + 1) * sizeof(int32_t), hipMemcpyHostToDevice));
  PG_HIP(hipMemcpy(B->fidx, fi.data(), B->R * sizeof(int32_t), hipMemcpyHostToDevice));
  B->nb2 = off[B->nent];
  B->nfreq = (int)fo.size();
  const int rc2 = refresh_diag(B, as_stream(stream));
  if (rc2 != PG_OK) B->binv = nullptr;  // stays allocated; the batch keeps point Jacobi
  return rc2;
}

int pg_batch_set_element_operator(void* batch, void* op, const void* workspace, void* stream) {
  if (!batch || !op || !workspace) PG_FAIL(PG_ERR_ARG, "bad arguments");
  Batch* B = static_cast<Batch*>(batch);
  ElemOp* E = static_cast<ElemOp*>(op);
  if (B->eop) PG_FAIL(PG_ERR_ARG, "element operator already set");
  if (B->R < 64 || B->R % 64 != 0)
    PG_FAIL(PG_ERR_UNSUPPORTED, "element products in a batch need 64, 128, 192 or 256 columns "
                                "(got %d)", B->R);
  if (eop_ndof(E) != B->n) PG_FAIL(PG_ERR_ARG, "element operator of another system");
  hipStream_t st = as_stream(stream);
  int rc = eop_multi_prepare(E, workspace, st);
  if (rc != PG_OK) return rc;
  const EopSchedule es = eop_schedule(E);
  const size_t nye = (size_t)es.nw * kWave * es.N * B->R;
  if ((rc = B->alloc(&B->ye64, nye)) != PG_OK) return rc;
  if (B->mixed && (rc = B->alloc(&B->ye32, nye)) != PG_OK) return rc;
  const int64_t nlp = eop_multi_parts(E, B->R);
  if (nlp >= (1LL << 31)) PG_FAIL(PG_ERR_UNSUPPORTED, "too many element slot groups");
  B->nlpart = (int)nlp;
  if ((rc = B->alloc(&B->lpart, (size_t)2 * B->R * nlp)) != PG_OK) return rc;
  const int64_t epb = kBS / B->CT;
  int64_t g = std::min<int64_t>(kMaxGrid, (es.nent + epb - 1) / epb);
  if (g >= 64) g = (g / 8) * 8;
  B->sgrid = (int)std::max<int64_t>(1, g);
  B->esched = es;
  B->eop = E;
  PG_HIP(hipStreamSynchronize(st));
  return PG_OK;
}

int pg_batch_refresh(void* batch, void* stream) {
  if (!batch) PG_FAIL(PG_ERR_ARG, "batch is NULL");
  return refresh_diag(static_cast<Batch*>(batch), as_stream(stream));
}

int pg_batch_destroy(void* batch) {
  delete static_cast<Batch*>(batch);
  return PG_OK;
}

}  // extern "C"
