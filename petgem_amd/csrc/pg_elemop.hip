// Element-by-element ("matrix-free") application of the assembled operator
//     A = sum_e P_e^T (K_e - i omega mu0 M_e) P_e,   PEC rows and columns -> identity,
// for gfx950, as an alternative to the CSR SpMV inside the Krylov solver.
//
// The element matrices are never stored.  With the reference-tensor form used
// by the assembly (K_e, M_e linear in 6 + 6 geometric coefficients, tables T
// per orientation variant), one element's product costs about 4100 FMAs
// (symmetric, upper triangle) on 288 bytes of input and output per element
// in complex64, against about 2800 bytes per element for streaming the CSR
// rows the element contributes to.  The product is split in two kernels:
//
//  k_eop_local  one lane per element, 64 elements of one orientation
//               combination per wavefront, so every table entry is a
//               wave-uniform (scalar) load; writes the element's n local
//               results y_e = A_e x_e (slot layout [wave][i][lane]).
//  k_eop_sum    one thread per mesh entity (edge, face), in the entity
//               schedule: sums the entity's local results over its elements
//               in a fixed order (atomic-free, bitwise reproducible), and
//               fuses the COCG dot product or the residual.
//
// Orders p = 1, 2 and 3 (n = 6, 20, 45), complex128 and complex64; the p = 3
// complex128 products run on the FP64 matrix cores (k_eop_mfma3).  The
// summation order differs from the assembled CSR's, so results agree with the
// CSR SpMV to rounding.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <numeric>
#include <vector>

#include "pg_basis.h"
#include "pg_common.h"
#include "pg_eop.h"
#include "pg_eop_mask.h"

namespace pg {

namespace eop {

constexpr int kBS = 256;
constexpr int kMaxGrid = 2048;
// workgroup size of the element products: one wavefront in complex64
// (measured 10 % faster than four: finer dispatch of the last, partial round
// of waves), four in complex128 (4 % faster than one at two waves per SIMD)
template <typename R>
constexpr int local_bs() {
  return sizeof(R) == 4 ? 64 : 256;
}
// waves per SIMD the element products are compiled for: five in complex64 at
// p <= 2 (at most 102 VGPRs; the register allocator otherwise settles on 104
// and four); at p = 3 the 45 values of p and of the products need about 190
template <int P, typename R>
constexpr int min_waves() {
  return sizeof(R) == 4 && P <= 2 ? 5 : 1;
}
constexpr double kMu0 = 4.0 * 3.14159265358979323846 * 1e-7;  // petgem/solver.py:175

template <typename R>
struct Cx;
template <>
struct Cx<double> {
  using T = double2;
  using T2 = double4;  // two consecutive complex values
};
template <>
struct Cx<float> {
  using T = float2;
  using T2 = float4;
};

// local entities carrying DOFs (edges; faces from p = 2; the volume at p = 3),
// their first local DOF and their number of DOFs
template <int P>
constexpr int n_local_entities() {
  return P == 1 ? 6 : (P == 2 ? 10 : 11);
}
template <int P>
constexpr int loc_off(int la) {
  return la < 6 ? la * P : (la < 10 ? 6 * P + (la - 6) * P * (P - 1) : 6 * P + 4 * P * (P - 1));
}
template <int P>
constexpr int loc_nd(int la) {
  return la < 6 ? P : (la < 10 ? P * (P - 1) : P * (P - 1) * (P - 2) / 2);
}

// table components kept per orientation combination: those that are not
// structural zeros (zero_mask); the kernel skips the zeros at compile time
constexpr int popcount12(unsigned m) {
  int n = 0;
  for (int c = 0; c < 12; ++c) n += (m >> c) & 1;
  return n;
}
template <int P>
constexpr int n_table_entries() {
  constexpr int N = P * (P + 2) * (P + 3) / 2;
  int n = 0;
  for (int k = 0; k < N * (N + 1) / 2; ++k) n += 12 - popcount12(zero_mask<P>(k));
  return n;
}
// per-combination table stride: the entries zero-padded to a multiple of 32
// values, so that the element products load whole batches unconditionally
constexpr int kTabPad = 32;
template <int P>
constexpr int table_stride() {
  return (n_table_entries<P>() + kTabPad - 1) / kTabPad * kTabPad;
}

__device__ inline double2 to_d(double2 v) { return v; }
__device__ inline double2 to_d(float2 v) { return make_double2(v.x, v.y); }
__device__ inline void store(double2* p, double2 v) { *p = v; }
__device__ inline void store(float2* p, double2 v) { *p = make_float2((float)v.x, (float)v.y); }

// Table value `off` of a wavefront's element products, streamed through
// scalar registers in batches of KB values, double buffered: at the start of
// batch b the wave waits for the scalar loads in flight (batch b, issued one
// batch earlier), issues the loads of batch b + 1 and then computes with batch
// b, so that every scalar load has one batch of arithmetic to arrive.  Left to
// itself the compiler issues each load just before its first use and waits for
// it at once; measured at C2 (p = 2) the batches take the complex128 product
// from 0.205 to 0.153 ms and the complex64 product from 0.086 to 0.069 ms.
// Scalar loads can return out of order, so the wait is for all of them
// (lgkmcnt(0)).  `off` is a compile-time constant after unrolling; the tables
// are zero-padded to a multiple of KB, so whole batches load unconditionally
// (conditional loads here cost the compiler minutes).
constexpr unsigned kWaitScalarLoads = 0xC07F;  // s_waitcnt lgkmcnt(0), other counters untouched
template <int KB, int NTS, typename R>
__device__ __forceinline__ R tab_next(const R* __restrict__ t, R (&tb)[2][KB], int off) {
  if (off % KB == 0) {
    const int b = off / KB;
    __builtin_amdgcn_s_waitcnt(kWaitScalarLoads);
    __builtin_amdgcn_sched_barrier(0);
    if ((b + 1) * KB < NTS) {
#pragma unroll
      for (int k = 0; k < KB; ++k) tb[(b + 1) & 1][k] = t[(b + 1) * KB + k];
    }
    __builtin_amdgcn_sched_barrier(0);
  }
  return tb[(off / KB) & 1][off % KB];
}

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

// ---------------------------------------------------------------------------
// local products.  coef[w][c][lane]: c < 6 the mass coefficients times
// -omega mu0, c >= 6 the curl-curl coefficients; tab[combo][...] the table
// components of the pairs k = (i, j >= i) (row-major upper triangle; per pair
// c < 6 mass, c >= 6 curl-curl), structural zeros left out.  edofs[w][la][lane]
// = first global DOF of local entity la (its P DOFs at p = 1 and 2 are
// consecutive: 1 at p = 1, 2 at p = 2), -1 for a PEC entity or an empty lane
// (its x is taken as 0).  PAIR: at p = 2 both DOFs of an entity are loaded as
// one 16-byte (complex64) or 32-byte (complex128) value; the first DOF of
// every entity is even, so this needs x aligned to that size.
// ---------------------------------------------------------------------------
template <int P, typename R, bool PAIR>
__global__ void __launch_bounds__(local_bs<R>(), (min_waves<P, R>())) k_eop_local(int64_t nw, const int32_t* __restrict__ done,
                                                   const int32_t* __restrict__ wseg, int dstride,
                                                   const int32_t* __restrict__ wcombo,
                                                   const int32_t* __restrict__ edofs,
                                                   const R* __restrict__ coef,
                                                   const R* __restrict__ tab,
                                                   const typename Cx<R>::T* __restrict__ x,
                                                   typename Cx<R>::T* __restrict__ ye) {
  using C = typename Cx<R>::T;
  using C2 = typename Cx<R>::T2;
  constexpr int N = P * (P + 2) * (P + 3) / 2;
  constexpr int NLE = n_local_entities<P>();
  static_assert(N == loc_off<P>(NLE - 1) + loc_nd<P>(NLE - 1), "entity DOFs");
  if (done && !wseg && *done) return;
  // XCD-contiguous ranges of waves (workgroup b runs on XCD b % 8)
  const int64_t nb = gridDim.x;
  int64_t b = blockIdx.x;
  if (nb % 8 == 0) b = (b % 8) * (nb / 8) + b / 8;
  const int64_t w =
      b * (local_bs<R>() / kWave) + __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  if (w >= nw) return;
  // segmented operator: the wave's mesh has converged
  if (wseg && done[(int64_t)__builtin_amdgcn_readfirstlane(wseg[w]) * dstride]) return;
  const int lane = threadIdx.x % kWave;
  const int combo = __builtin_amdgcn_readfirstlane(wcombo[w]);
  constexpr int NTS = table_stride<P>();  // constant-evaluated, not a loop at run time
  const R* __restrict__ t = tab + (int64_t)combo * NTS;
  // first batch of table values (scalar loads) in flight during the gathers;
  // p = 3 (1035 entry pairs, a 17 600-instruction body) keeps the compiler's
  // own schedule
  constexpr bool kPipe = P <= 2;
  constexpr int KB = sizeof(R) == 8 ? 16 : 32;
  R tb[2][KB];
  if constexpr (kPipe) {
#pragma unroll
    for (int k = 0; k < KB; ++k) tb[0][k] = t[k];
    __builtin_amdgcn_sched_barrier(0);
  }
  // all index and coefficient loads first, then all gathers, unconditionally
  // (a masked entity reads x[0] and is zeroed after), so that the wave waits
  // for memory twice rather than once per DOF
  int fv[NLE];
#pragma unroll
  for (int la = 0; la < NLE; ++la) fv[la] = edofs[(w * NLE + la) * kWave + lane];
  R cf[12];
#pragma unroll
  for (int c = 0; c < 12; ++c) cf[c] = coef[(w * 12 + c) * kWave + lane];
  C xv[N];
#pragma unroll
  for (int la = 0; la < NLE; ++la) {
    const int f = fv[la] >= 0 ? fv[la] : 0;
    if (P == 2 && PAIR) {
      const C2 v = *reinterpret_cast<const C2*>(x + f);
      xv[2 * la] = C{v.x, v.y};
      xv[2 * la + 1] = C{v.z, v.w};
    } else {
#pragma unroll
      for (int k = 0; k < loc_nd<P>(la); ++k) xv[loc_off<P>(la) + k] = x[f + k];
    }
  }
#pragma unroll
  for (int la = 0; la < NLE; ++la)
    if (fv[la] < 0)
#pragma unroll
      for (int k = 0; k < loc_nd<P>(la); ++k) xv[loc_off<P>(la) + k] = C{0, 0};
  C y[N];
#pragma unroll
  for (int i = 0; i < N; ++i) y[i] = C{0, 0};
  // fully unrolled: the masks, the presence flags and the table offset are
  // compile-time constants, so only the nonzero terms are emitted
  int off = 0;
#pragma unroll
  for (int i = 0; i < N; ++i) {
#pragma unroll
    for (int j = i; j < N; ++j) {
      const unsigned zm = zero_mask<P>(i * N - (i * (i - 1)) / 2 + (j - i));
      R im = 0, re = 0;
      bool has_im = false, has_re = false;
#pragma unroll
      for (int c = 0; c < 6; ++c)
        if (!((zm >> c) & 1)) {
          const R tv = kPipe ? tab_next<KB, NTS>(t, tb, off) : t[off];
          im = has_im ? fma(cf[c], tv, im) : cf[c] * tv;
          has_im = true;
          ++off;
        }
#pragma unroll
      for (int c = 6; c < 12; ++c)
        if (!((zm >> c) & 1)) {
          const R tv = kPipe ? tab_next<KB, NTS>(t, tb, off) : t[off];
          re = has_re ? fma(cf[c], tv, re) : cf[c] * tv;
          has_re = true;
          ++off;
        }
      if (has_re) {
        y[i].x = fma(re, xv[j].x, y[i].x);
        y[i].y = fma(re, xv[j].y, y[i].y);
        if (j != i) {
          y[j].x = fma(re, xv[i].x, y[j].x);
          y[j].y = fma(re, xv[i].y, y[j].y);
        }
      }
      if (has_im) {
        y[i].x = fma(-im, xv[j].y, y[i].x);
        y[i].y = fma(im, xv[j].x, y[i].y);
        if (j != i) {
          y[j].x = fma(-im, xv[i].y, y[j].x);
          y[j].y = fma(im, xv[i].x, y[j].y);
        }
      }
    }
  }
#pragma unroll
  for (int i = 0; i < N; ++i) ye[(w * N + i) * kWave + lane] = y[i];
}

// ---------------------------------------------------------------------------
// p = 3 complex128 element products on the FP64 matrix cores.  With 45 DOFs
// the values of x and of the products take 360 registers per lane in
// complex128, so the lane-per-element kernel above does not fit; here one
// wave takes one slot group w (64 elements of one combination) and forms
//     Y (48 x 128) = Tf (48 x K) * B (K x 128)
// with v_mfma_f64_16x16x4f64: column n < 64 of Y is element n's real part,
// column 64 + n its imaginary part (rows 45..47 padding); column k = (c, j)
// of Tf is table component c of the pairs (i, j) (both triangles), and
//     B[(c, j)][n]      = cf_n[c] re x_n[j]   (c curl-curl) or -cf_n[c] im x_n[j] (c mass)
//     B[(c, j)][64 + n] = cf_n[c] im x_n[j]   (c curl-curl) or  cf_n[c] re x_n[j] (c mass)
// (mass coefficients carry -omega mu0, as in k_eop_local).  Only the columns
// (c, j) that are not zero for every orientation are kept (kcol, padded to a
// multiple of 4 with -1).  The element's x and coefficients are staged in LDS
// (52 KB per wave); the fragments of Tf are per combination in fragment order
// afrag[combo][ks][m-tile][lane] (L2-resident).  Operand layout of the
// instruction (measured): A lane l = (row l % 16, k l / 16), B lane l =
// (k l / 16, column l % 16), D lane l, value r = (row 4 r + l / 16, column l % 16).
// ---------------------------------------------------------------------------
typedef double mfma_d4 __attribute__((ext_vector_type(4)));
constexpr int kMfmaXStride = 46;  // complex values per element row in LDS (45 + 1 pad)

__global__ void __launch_bounds__(kWave) k_eop_mfma3(int64_t nw, const int32_t* __restrict__ done,
                                                     const int32_t* __restrict__ wseg, int dstride,
                                                     const int32_t* __restrict__ wcombo,
                                                     const int32_t* __restrict__ edofs,
                                                     const double* __restrict__ coef,
                                                     const double* __restrict__ afrag,
                                                     const int32_t* __restrict__ kcol, int nks,
                                                     const double2* __restrict__ x,
                                                     double2* __restrict__ ye) {
  constexpr int P = 3, N = 45, NLE = n_local_entities<P>();
  __shared__ double2 xs[kWave * kMfmaXStride];
  __shared__ double cfs[kWave * 12];
  if (done && !wseg && *done) return;
  const int64_t nb = gridDim.x;
  int64_t w = blockIdx.x;
  if (nb % 8 == 0) w = (w % 8) * (nb / 8) + w / 8;  // XCD-contiguous ranges of slot groups
  if (w >= nw) return;
  if (wseg && done[(int64_t)wseg[w] * dstride]) return;  // segmented: the mesh has converged
  const int lane = threadIdx.x;
  const int combo = __builtin_amdgcn_readfirstlane(wcombo[w]);
  // stage this lane's element: coefficients and x (zero for PEC / empty)
#pragma unroll
  for (int c = 0; c < 12; ++c) cfs[lane * 12 + c] = coef[(w * 12 + c) * kWave + lane];
#pragma unroll
  for (int la = 0; la < NLE; ++la) {
    const int f = edofs[(w * NLE + la) * kWave + lane];
#pragma unroll
    for (int k = 0; k < loc_nd<P>(la); ++k)
      xs[lane * kMfmaXStride + loc_off<P>(la) + k] = f >= 0 ? x[f + k] : double2{0.0, 0.0};
  }
  __syncthreads();
  mfma_d4 acc[3][8];
#pragma unroll
  for (int mt = 0; mt < 3; ++mt)
#pragma unroll
    for (int nt = 0; nt < 8; ++nt) acc[mt][nt] = mfma_d4{0.0, 0.0, 0.0, 0.0};
  const double* __restrict__ af = afrag + (int64_t)combo * nks * 3 * kWave;
  const int kr = lane / 16, en = lane % 16;
  // the fragments of Tf and the column of step ks + 1 are loaded during step ks
  double na0 = af[lane], na1 = af[kWave + lane], na2 = af[2 * kWave + lane];
  int nkc = kcol[kr];
  for (int ks = 0; ks < nks; ++ks) {
    const double a0 = na0, a1 = na1, a2 = na2;
    const int kc = nkc;
    if (ks + 1 < nks) {
      na0 = af[((ks + 1) * 3 + 0) * kWave + lane];
      na1 = af[((ks + 1) * 3 + 1) * kWave + lane];
      na2 = af[((ks + 1) * 3 + 2) * kWave + lane];
      nkc = kcol[(ks + 1) * 4 + kr];
    }
    double bre[4], bim[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) bre[q] = bim[q] = 0.0;
    if (kc >= 0) {
      const int c = kc >> 8, j = kc & 255;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int e = q * 16 + en;
        const double2 xv = xs[e * kMfmaXStride + j];
        const double cv = cfs[e * 12 + c];
        if (c >= 6) {
          bre[q] = cv * xv.x;
          bim[q] = cv * xv.y;
        } else {
          bre[q] = -(cv * xv.y);
          bim[q] = cv * xv.x;
        }
      }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      acc[0][q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, bre[q], acc[0][q], 0, 0, 0);
      acc[1][q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, bre[q], acc[1][q], 0, 0, 0);
      acc[2][q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a2, bre[q], acc[2][q], 0, 0, 0);
      acc[0][q + 4] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, bim[q], acc[0][q + 4], 0, 0, 0);
      acc[1][q + 4] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, bim[q], acc[1][q + 4], 0, 0, 0);
      acc[2][q + 4] = __builtin_amdgcn_mfma_f64_16x16x4f64(a2, bim[q], acc[2][q + 4], 0, 0, 0);
    }
  }
#pragma unroll
  for (int mt = 0; mt < 3; ++mt)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = mt * 16 + 4 * r + kr;
      if (i < N)
#pragma unroll
        for (int q = 0; q < 4; ++q)
          ye[((int64_t)w * N + i) * kWave + q * 16 + en] = double2{acc[mt][q][r], acc[mt][q + 4][r]};
    }
}

// ---------------------------------------------------------------------------
// per-entity sums.  gm[i] = {r0, d, s0, s1} for the i-th scheduled entity:
// rows r0 .. r0+d, local results at gslot[s0 .. s1) (+ t * 64 for DOF t);
// s0 == s1 marks a PEC entity (identity rows).  MODE as in the CSR SpMV.
// XCD x (workgroup b runs on XCD b % 8) sweeps the x-th eighth of the schedule.
// ---------------------------------------------------------------------------
template <int D, int MODE, typename C>
__global__ void __launch_bounds__(kBS) k_eop_sum(int64_t nent, const int32_t* __restrict__ done,
                                                 const int4* __restrict__ gm,
                                                 const int32_t* __restrict__ gslot,
                                                 const C* __restrict__ ye, const C* __restrict__ x,
                                                 C* __restrict__ y, const double2* __restrict__ b,
                                                 double* part) {
  if (done && *done) return;
  const int64_t blk = blockIdx.x, nb = gridDim.x;
  int64_t first = blk * kBS, end = nent, stride = nb * kBS;
  if (nb % 8 == 0) {
    const int64_t xc = blk % 8, nbx = nb / 8;
    first = nent * xc / 8 + (blk / 8) * kBS;
    end = nent * (xc + 1) / 8;
    stride = nbx * kBS;
  }
  double red[2] = {0.0, 0.0};
  for (int64_t ent = first + threadIdx.x; ent < end; ent += stride) {
    const int4 m = gm[ent];
    const int r0 = m.x, d = m.y, s0 = m.z, s1 = m.w;
    // x at the entity's rows (identity rows; the MODE 1 dot product), loaded
    // before the sums so that its latency overlaps theirs
    C xr[D];
    const bool need_x = MODE == 1 || s0 == s1;
#pragma unroll
    for (int t = 0; t < D; ++t) xr[t] = (need_x && t < d) ? x[r0 + t] : C{0, 0};
    C acc[D];
#pragma unroll
    for (int t = 0; t < D; ++t) acc[t] = C{0, 0};
    if (s0 == s1) {
#pragma unroll
      for (int t = 0; t < D; ++t) acc[t] = xr[t];
    } else {
      // two incidences per step (their loads in flight together), summed in
      // ascending slot order
      int k = s0;
      for (; k + 2 <= s1; k += 2) {
        const int b0 = gslot[k], b1 = gslot[k + 1];
        C v0[D], v1[D];
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
            const C v = ye[b0 + t * kWave];
            acc[t].x += v.x;
            acc[t].y += v.y;
          }
      }
    }
#pragma unroll
    for (int t = 0; t < D; ++t)
      if (t < d) {
        const int row = r0 + t;
        if (MODE == 2) {
          const double2 bb = b[row];
          const double2 rr = make_double2(bb.x - (double)acc[t].x, bb.y - (double)acc[t].y);
          store(y + row, rr);
          red[0] += rr.x * rr.x + rr.y * rr.y;
        } else {
          y[row] = acc[t];
          if (MODE == 1) {
            const double2 pv = to_d(xr[t]), qv = to_d(acc[t]);
            red[0] += pv.x * qv.x - pv.y * qv.y;
            red[1] += pv.x * qv.y + pv.y * qv.x;
          }
        }
      }
  }
  if (MODE == 1) write_partials<2>(red, part);
  if (MODE == 2) {
    double r1[1] = {red[0]};
    write_partials<1>(r1, part);
  }
}

// ---------------------------------------------------------------------------
// multi-right-hand-side local products for the batched solver: column r of the
// (ndof, R) row-major block x gets y_e,r = (K_e - i omega_r mu0 M_e) x_e,r.
// One workgroup of R threads (thread = column; R = 64 .. 256) per part of a
// slot group w of the single-RHS layout (64 elements of one orientation
// combination, split into 256 / R parts), which takes its elements one after
// another.  Per element the threads first form
// the n(n+1)/2 entry pairs (K_ij, M_ij) from the raw coefficients and the
// combination's table (pdesc[k] = {table offset, zero mask} of pair k) into
// LDS, double buffered so that one barrier per element suffices; then every
// thread gathers its column of x_e (R contiguous values per DOF) and applies
// the symmetric element matrix, each pair read once from LDS as a broadcast
// and applied to rows i and j.  The entries are formed once per element for
// all R columns.  Local results ye[((w N + i) 64 + l) R + r]: the single-RHS
// slot numbering times R, so the per-entity summation reuses the slot lists.
// With `dpart` each thread also sums x_e,r^T y_e,r over its elements (FP64):
// dpart[(k R + r) gridDim + block], k = 0 real, 1 imaginary part, the element
// share of x_r^T A_r x_r (PEC DOFs are zero in x_e and add nothing here).
// ---------------------------------------------------------------------------
template <int P, typename RT>
__global__ void __launch_bounds__(256) k_meop_local(
    int64_t nw, int R, int split, const int32_t* __restrict__ done,
    const int32_t* __restrict__ wcombo,
    const int32_t* __restrict__ edofs, const double* __restrict__ coefm,
    const double* __restrict__ tab, const int2* __restrict__ pdesc,
    const double* __restrict__ omu, const typename Cx<RT>::T* __restrict__ x,
    typename Cx<RT>::T* __restrict__ ye, double* __restrict__ dpart) {
  using C = typename Cx<RT>::T;
  constexpr int N = P * (P + 2) * (P + 3) / 2;
  constexpr int NP = N * (N + 1) / 2;
  constexpr int NLE = n_local_entities<P>();
  if (done && *done) return;
  constexpr int NTS = table_stride<P>();
  __shared__ C ent[2][NP];
  __shared__ double tl[NTS];  // the combination's table, staged once per workgroup
  // workgroup b: part b % split of slot group b / split (`split` workgroups per
  // group, so that narrow batches still launch several waves per SIMD)
  const int64_t nb = gridDim.x;
  int64_t b = blockIdx.x;
  if (nb % 8 == 0) b = (b % 8) * (nb / 8) + b / 8;  // XCD-contiguous ranges of slot groups
  const int64_t w = b / split;
  const int part = (int)(b - w * split);
  if (w >= nw) return;
  const int r = threadIdx.x;
  const RT om = (RT)omu[r];
  const int combo = wcombo[w];
  const double* __restrict__ t = tab + (int64_t)combo * NTS;
  for (int k = threadIdx.x; k < NTS; k += blockDim.x) tl[k] = t[k];
  __syncthreads();
  const int l0 = part * (kWave / split), l1 = l0 + kWave / split;
  double dre = 0.0, dim = 0.0;  // x^T y over this thread's elements
  for (int l = l0; l < l1; ++l) {
    int fv[NLE];
    bool any = false;
#pragma unroll
    for (int la = 0; la < NLE; ++la) {
      fv[la] = edofs[(w * NLE + la) * kWave + l];
      any |= fv[la] >= 0;
    }
    if (!any) continue;  // empty slot (or every entity PEC): no products; uniform branch
    C* e = ent[l & 1];
    double cfl[12];  // the element's raw coefficients (uniform: scalar loads)
#pragma unroll
    for (int c = 0; c < 12; ++c) cfl[c] = coefm[(w * 12 + c) * kWave + l];
    for (int k = threadIdx.x; k < NP; k += blockDim.x) {
      const int2 pd = pdesc[k];
      int o = pd.x;
      double m = 0.0, kk = 0.0;
      for (int c = 0; c < 6; ++c)
        if (!((pd.y >> c) & 1)) m = fma(cfl[c], tl[o++], m);
      for (int c = 6; c < 12; ++c)
        if (!((pd.y >> c) & 1)) kk = fma(cfl[c], tl[o++], kk);
      e[k] = C{(RT)kk, (RT)m};
    }
    C xv[N];
#pragma unroll
    for (int la = 0; la < NLE; ++la)
#pragma unroll
      for (int k = 0; k < loc_nd<P>(la); ++k)
        xv[loc_off<P>(la) + k] = fv[la] >= 0 ? x[(int64_t)(fv[la] + k) * R + r] : C{0, 0};
    __syncthreads();
    C y[N];
#pragma unroll
    for (int i = 0; i < N; ++i) y[i] = C{0, 0};
#pragma unroll
    for (int i = 0; i < N; ++i) {
#pragma unroll
      for (int j = i; j < N; ++j) {
        const int k = i * N - (i * (i - 1)) / 2 + (j - i);
        const unsigned zm = zero_mask<P>(k);
        const bool hasM = (zm & 0x3Fu) != 0x3Fu, hasK = (zm & 0xFC0u) != 0xFC0u;
        if (!hasM && !hasK) continue;
        const C a = e[k];  // (K_ij, M_ij)
        if (hasK) {
          y[i].x = fma(a.x, xv[j].x, y[i].x);
          y[i].y = fma(a.x, xv[j].y, y[i].y);
          if (j != i) {
            y[j].x = fma(a.x, xv[i].x, y[j].x);
            y[j].y = fma(a.x, xv[i].y, y[j].y);
          }
        }
        if (hasM) {  // -i omega mu0 M_ij
          const RT wm = om * a.y;
          y[i].x = fma(wm, xv[j].y, y[i].x);
          y[i].y = fma(-wm, xv[j].x, y[i].y);
          if (j != i) {
            y[j].x = fma(wm, xv[i].y, y[j].x);
            y[j].y = fma(-wm, xv[i].x, y[j].y);
          }
        }
      }
    }
#pragma unroll
    for (int i = 0; i < N; ++i) {
      ye[(((int64_t)w * N + i) * kWave + l) * R + r] = y[i];
      const double xr = xv[i].x, xi = xv[i].y, yr = y[i].x, yi = y[i].y;
      dre += xr * yr - xi * yi;
      dim += xr * yi + xi * yr;
    }
  }
  if (dpart) {
    dpart[(int64_t)r * gridDim.x + blockIdx.x] = dre;
    dpart[((int64_t)R + r) * gridDim.x + blockIdx.x] = dim;
  }
}

// coefficients in slot order: mass part times -omega mu0, then curl-curl part
// (multi-mesh operator: omega mu0 of the wave's segment, womu[wseg[w]])
template <typename R>
__global__ void k_eop_coef(int64_t nslot, const int32_t* __restrict__ perm,
                           const double* __restrict__ wscoef, double omega_mu,
                           const int32_t* __restrict__ wseg, const double* __restrict__ womu,
                           R* __restrict__ coef) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nslot) return;
  const int64_t w = s / kWave, lane = s % kWave;
  const int64_t e = perm[s];
  const double om = womu ? womu[wseg[w]] : omega_mu;
#pragma unroll
  for (int c = 0; c < 12; ++c) {
    double v = e >= 0 ? wscoef[12 * e + c] : 0.0;
    if (c < 6) v = -(om * v);
    coef[(w * 12 + c) * kWave + lane] = (R)v;
  }
}

}  // namespace eop

struct ElemOp {
  int p = 0, N = 0, D = 1;
  bool even_entities = false;  // every entity's first DOF even: paired gathers allowed
  int64_t E = 0, ndof = 0, nent = 0, nw = 0;
  int ncombo = 0, grid = 1;
  int32_t* wcombo = nullptr;
  int32_t* edofs = nullptr;
  int32_t* perm = nullptr;
  int4* gm = nullptr;
  int32_t* gslot = nullptr;
  double* coef64 = nullptr;
  float* coef32 = nullptr;
  double* tab64 = nullptr;
  float* tab32 = nullptr;
  double2* ye64 = nullptr;
  float2* ye32 = nullptr;
  // multi-right-hand-side products (eop_multi_prepare): raw coefficients in
  // slot order (mass part without -omega mu0) and per entry pair {offset in
  // the combination's table, zero mask}
  double* coefm = nullptr;
  int2* pdesc = nullptr;
  // p = 3 complex128 products on the matrix cores (k_eop_mfma3): Tf fragments
  // per combination and the kept (c, j) columns
  double* afrag = nullptr;
  int32_t* kcol = nullptr;
  int nks = 0;
  // multi-mesh operator (pg_eop_create_multi): segment of every wave, omega mu0
  // of every segment, first schedule position of every segment (nseg + 1)
  int nseg = 0;
  int32_t* wseg = nullptr;
  double* womu = nullptr;
  std::vector<int64_t> seg_pos, seg_wave;  // + first wave of every segment (nseg + 1)
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
  ~ElemOp() {  // release errors cannot be reported from here
    for (void* v : owned) (void)hipFree(v);
  }
};

int eop_grid(const ElemOp* op) { return op->grid; }
bool eop_supports(const ElemOp* op, bool complex128) {
  return !(complex128 && op->p == 3 && !op->afrag);
}
int64_t eop_ndof(const ElemOp* op) { return op->ndof; }

// element products of waves [w0, w0 + nwr) (all waves: 0, op->nw)
template <typename R>
static void launch_local(const ElemOp* op, const int32_t* done, const int32_t* wseg, int dstride,
                         const typename eop::Cx<R>::T* x, typename eop::Cx<R>::T* ye,
                         const R* coef, const R* tab, hipStream_t st, int64_t w0 = 0,
                         int64_t nwr = -1) {
  using C = typename eop::Cx<R>::T;
  if (nwr < 0) nwr = op->nw;
  if (nwr <= 0) return;
  const int nle = op->p == 1 ? 6 : (op->p == 2 ? 10 : 11);
  const int32_t* wc = op->wcombo + w0;
  const int32_t* ed = op->edofs + w0 * nle * kWave;
  const int32_t* ws = wseg ? wseg + w0 : nullptr;
  coef += w0 * 12 * kWave;
  ye += w0 * op->N * kWave;
  const int wpb = eop::local_bs<R>() / kWave;
  int64_t nbl = (nwr + wpb - 1) / wpb;
  if (nbl >= 64) nbl = (nbl + 7) / 8 * 8;
  const dim3 gl((unsigned)nbl), bll(wpb * kWave);
  const bool pair_ok =
      op->even_entities && reinterpret_cast<uintptr_t>(x) % (2 * sizeof(C)) == 0;
  if (op->p == 1) {
    hipLaunchKernelGGL((eop::k_eop_local<1, R, false>), gl, bll, 0, st, nwr, done, ws, dstride,
                       wc, ed, coef, tab, x, ye);
  } else if (op->p == 3) {
    // complex64: one lane per element; complex128: the matrix-core kernel (the
    // values of x and of the products do not fit one lane)
    if constexpr (sizeof(R) == 4) {
      hipLaunchKernelGGL((eop::k_eop_local<3, R, false>), gl, bll, 0, st, nwr, done, ws,
                         dstride, wc, ed, coef, tab, x, ye);
    } else {
      int64_t nbm = nwr;
      if (nbm >= 64) nbm = (nbm + 7) / 8 * 8;
      hipLaunchKernelGGL(eop::k_eop_mfma3, dim3((unsigned)nbm), dim3(kWave), 0, st, nwr, done,
                         ws, dstride, wc, ed, coef, op->afrag, op->kcol, op->nks, x, ye);
    }
  } else if (pair_ok) {
    hipLaunchKernelGGL((eop::k_eop_local<2, R, true>), gl, bll, 0, st, nwr, done, ws, dstride,
                       wc, ed, coef, tab, x, ye);
  } else {
    hipLaunchKernelGGL((eop::k_eop_local<2, R, false>), gl, bll, 0, st, nwr, done, ws, dstride,
                       wc, ed, coef, tab, x, ye);
  }
}

template <typename R>
static void launch_impl(const ElemOp* op, int mode, const int32_t* done,
                        const typename eop::Cx<R>::T* x, typename eop::Cx<R>::T* y,
                        const double2* b, double* part, hipStream_t st, hipEvent_t mid,
                        typename eop::Cx<R>::T* ye, const R* coef, const R* tab) {
  using C = typename eop::Cx<R>::T;
  launch_local<R>(op, done, nullptr, 0, x, ye, coef, tab, st);
  // a failed record only loses this timing sample (the elapsed-time query fails)
  if (mid) (void)hipEventRecord(mid, st);
  const dim3 g(op->grid), bl(eop::kBS);
#define PG_SUM(DD, MM)                                                                        \
  hipLaunchKernelGGL((eop::k_eop_sum<DD, MM, C>), g, bl, 0, st, op->nent, done, op->gm,       \
                     op->gslot, (const C*)ye, x, y, b, part)
  if (op->D == 1) {
    if (mode == 0) PG_SUM(1, 0);
    else if (mode == 1) PG_SUM(1, 1);
    else PG_SUM(1, 2);
  } else if (op->D == 2) {
    if (mode == 0) PG_SUM(2, 0);
    else if (mode == 1) PG_SUM(2, 1);
    else PG_SUM(2, 2);
  } else {
    if (mode == 0) PG_SUM(6, 0);
    else if (mode == 1) PG_SUM(6, 1);
    else PG_SUM(6, 2);
  }
#undef PG_SUM
}

void eop_launch(const ElemOp* op, int mode, const int32_t* done, const double2* x, double2* y,
                const double2* b, double* part, hipStream_t st, hipEvent_t mid) {
  launch_impl<double>(op, mode, done, x, y, b, part, st, mid, op->ye64, op->coef64, op->tab64);
}
void eop_launch(const ElemOp* op, int mode, const int32_t* done, const float2* x, float2* y,
                const double2* b, double* part, hipStream_t st, hipEvent_t mid) {
  launch_impl<float>(op, mode, done, x, y, b, part, st, mid, op->ye32, op->coef32, op->tab32);
}

static int eop_refresh(ElemOp* op, const void* ws, double omega, hipStream_t st) {
  const double* wscoef = reinterpret_cast<const double*>(ws);
  const int64_t nslot = op->nw * kWave;
  const double om = omega * eop::kMu0;
  const dim3 g((unsigned)((nslot + 255) / 256)), b(256);
  hipLaunchKernelGGL(eop::k_eop_coef<double>, g, b, 0, st, nslot, op->perm, wscoef, om,
                     op->wseg, op->womu, op->coef64);
  hipLaunchKernelGGL(eop::k_eop_coef<float>, g, b, 0, st, nslot, op->perm, wscoef, om,
                     op->wseg, op->womu, op->coef32);
  // raw coefficients: -(-1 * v) = v exactly
  if (op->coefm)
    hipLaunchKernelGGL(eop::k_eop_coef<double>, g, b, 0, st, nslot, op->perm, wscoef, -1.0,
                       nullptr, nullptr, op->coefm);
  PG_CHECK_LAUNCH();
  return PG_OK;
}

int eop_multi_prepare(ElemOp* op, const void* ws, hipStream_t st) {
  if (op->nseg > 0)
    PG_FAIL(PG_ERR_UNSUPPORTED, "multi-right-hand-side products of a multi-mesh operator");
  if (op->p > 2)
    PG_FAIL(PG_ERR_UNSUPPORTED, "multi-right-hand-side element products: p <= 2 only");
  if (op->coefm) return PG_OK;
  const int N = op->N, NP = N * (N + 1) / 2;
  std::vector<int2> pd(NP);
  int o = 0;
  for (int k = 0; k < NP; ++k) {
    const unsigned zm = op->p == 1 ? eop::zero_mask<1>(k) : eop::zero_mask<2>(k);
    pd[k] = make_int2(o, (int)zm);
    o += 12 - eop::popcount12(zm);
  }
  int rc;
  if ((rc = op->alloc(&op->pdesc, NP)) != PG_OK || (rc = op->alloc(&op->coefm, op->nw * kWave * 12)) != PG_OK)
    return rc;
  PG_HIP(hipMemcpyAsync(op->pdesc, pd.data(), NP * sizeof(int2), hipMemcpyHostToDevice, st));
  PG_HIP(hipStreamSynchronize(st));
  const int64_t nslot = op->nw * kWave;
  const dim3 g((unsigned)((nslot + 255) / 256)), b(256);
  hipLaunchKernelGGL(eop::k_eop_coef<double>, g, b, 0, st, nslot, op->perm,
                     reinterpret_cast<const double*>(ws), -1.0, nullptr, nullptr, op->coefm);
  PG_CHECK_LAUNCH();
  return PG_OK;
}

EopSchedule eop_schedule(const ElemOp* op) {
  return EopSchedule{op->gm, op->gslot, op->nent, op->nw, op->p, op->N, op->D};
}

int eop_segments(const ElemOp* op, const int64_t** seg_pos, const int64_t** seg_wave) {
  if (seg_pos) *seg_pos = op->seg_pos.data();
  if (seg_wave) *seg_wave = op->seg_wave.data();
  return op->nseg;
}

const double2* eop_local_segments(const ElemOp* op, const int32_t* done, int dstride,
                                  const double2* x, hipStream_t st, int64_t w0, int64_t w1) {
  if (w1 < 0) w1 = op->nw;
  launch_local<double>(op, done, op->wseg, dstride, x, op->ye64, op->coef64, op->tab64, st, w0,
                       w1 - w0);
  return op->ye64;
}
const float2* eop_local_segments(const ElemOp* op, const int32_t* done, int dstride,
                                 const float2* x, hipStream_t st, int64_t w0, int64_t w1) {
  if (w1 < 0) w1 = op->nw;
  launch_local<float>(op, done, op->wseg, dstride, x, op->ye32, op->coef32, op->tab32, st, w0,
                      w1 - w0);
  return op->ye32;
}

// workgroups of R threads; a slot group split over 256 / R workgroups (at
// most 4), about 4 waves per slot group at every width
static int multi_split(int R) { return std::max(1, std::min(4, 256 / R)); }
int64_t eop_multi_parts(const ElemOp* op, int R) { return op->nw * multi_split(R); }

template <typename RT>
static void multi_local(const ElemOp* op, int R, const int32_t* done, const double* omu,
                        const typename eop::Cx<RT>::T* x, typename eop::Cx<RT>::T* ye,
                        double* part, hipStream_t st) {
  const int split = multi_split(R);
  const dim3 g((unsigned)(op->nw * split)), b(R);
  const double* tab = op->tab64;
  if (op->p == 1)
    hipLaunchKernelGGL((eop::k_meop_local<1, RT>), g, b, 0, st, op->nw, R, split, done,
                       op->wcombo, op->edofs, op->coefm, tab, op->pdesc, omu, x, ye, part);
  else
    hipLaunchKernelGGL((eop::k_meop_local<2, RT>), g, b, 0, st, op->nw, R, split, done,
                       op->wcombo, op->edofs, op->coefm, tab, op->pdesc, omu, x, ye, part);
}
void eop_multi_local(const ElemOp* op, int R, const int32_t* done, const double* omu,
                     const double2* x, double2* ye, double* part, hipStream_t st) {
  multi_local<double>(op, R, done, omu, x, ye, part, st);
}
void eop_multi_local(const ElemOp* op, int R, const int32_t* done, const double* omu,
                     const float2* x, float2* ye, double* part, hipStream_t st) {
  multi_local<float>(op, R, done, omu, x, ye, part, st);
}

// Operator set-up from host copies of the element DOFs (hd, E x n), element
// orientation codes (hor), entity DOF ranges (bp), schedule (order, may be
// null) and PEC entities (hbc); ws = the device geometry coefficients.  With
// seg (a multi-mesh operator) the elements are grouped by segment first, so
// that no wave of 64 elements spans two segments, and each segment gets its
// own frequency.
struct SegInfo {
  int nseg = 0;
  std::vector<int64_t> pos;  // first schedule position of each segment (nseg + 1)
  std::vector<double> omega;
};

// host checks of the entity DOF ranges (before any device work)
static int check_block_ptr(int p, const int32_t* bp, int64_t nblocks, int64_t ndof) {
  const int maxd = p == 1 ? 1 : (p == 2 ? 2 : 6);
  if (bp[0] != 0 || bp[nblocks] != ndof)
    PG_FAIL(PG_ERR_ARG, "block_ptr must start at 0 and end at ndof");
  for (int64_t g = 0; g < nblocks; ++g)
    if (bp[g + 1] - bp[g] < 1 || bp[g + 1] - bp[g] > maxd)
      PG_FAIL(PG_ERR_ARG, "entity sizes must be in 1..%d at p=%d", maxd, p);
  return PG_OK;
}

static int eop_build(int p, int64_t E, const void* ws, int64_t ndof, const int32_t* bp,
                     int64_t nblocks, const int32_t* order, const std::vector<int32_t>& hd,
                     const std::vector<int32_t>& hor, const std::vector<uint8_t>& hbc,
                     double omega, const SegInfo* seg, ElemOp** out, hipStream_t st) {
  const int N = ndof_element(p);
  const int nle = p == 1 ? 6 : (p == 2 ? 10 : 11);  // local entities carrying DOFs
  const int maxd = p == 1 ? 1 : (p == 2 ? 2 : 6);
  int rc0 = check_block_ptr(p, bp, nblocks, ndof);
  if (rc0 != PG_OK) return rc0;
  std::vector<int32_t> ent_of(ndof);
  for (int64_t g = 0; g < nblocks; ++g)
    for (int d = bp[g]; d < bp[g + 1]; ++d) ent_of[d] = (int32_t)g;
  // each local entity's global entity; its DOFs must be that entity's range in order
  std::vector<int32_t> eg((size_t)E * nle);
  std::vector<int64_t> cnt(nblocks, 0);
  for (int64_t e = 0; e < E; ++e)
    for (int la = 0; la < nle; ++la) {
      const int off = entity_local_offset(la, p), nd = entity_ndof(la, p);
      const int d0 = hd[e * N + off];
      if (d0 < 0 || d0 >= ndof) PG_FAIL(PG_ERR_ARG, "element %lld: DOF out of range", (long long)e);
      const int g = ent_of[d0];
      if (bp[g + 1] - bp[g] != nd || bp[g] != d0)
        PG_FAIL(PG_ERR_ARG, "element %lld: local entity %d does not match a block_ptr range",
                (long long)e, la);
      for (int k = 1; k < nd; ++k)
        if (hd[e * N + off + k] != d0 + k)
          PG_FAIL(PG_ERR_ARG, "element %lld: entity DOFs out of order", (long long)e);
      eg[e * nle + la] = g;
      if (!hbc[g]) cnt[g]++;
    }
  for (int64_t g = 0; g < nblocks; ++g)
    if (!hbc[g] && cnt[g] == 0)
      PG_FAIL(PG_ERR_ARG, "entity %lld belongs to no element", (long long)g);
  std::vector<int32_t> pos(nblocks);
  if (order) {
    std::vector<char> seen(nblocks, 0);
    for (int64_t i = 0; i < nblocks; ++i) {
      if (order[i] < 0 || order[i] >= nblocks || seen[order[i]])
        PG_FAIL(PG_ERR_ARG, "entity_order must be a permutation of 0..nblocks-1");
      seen[order[i]] = 1;
      pos[order[i]] = (int32_t)i;
    }
  } else {
    std::iota(pos.begin(), pos.end(), 0);
  }
  // orientation combinations
  std::map<int, int> cid;
  std::vector<int> combos;
  std::vector<int32_t> ec(E);
  for (int64_t e = 0; e < E; ++e) {
    auto it = cid.find(hor[e]);
    if (it == cid.end()) {
      it = cid.emplace(hor[e], (int)combos.size()).first;
      combos.push_back(hor[e]);
    }
    ec[e] = it->second;
  }
  // element order: by segment (multi-mesh operator), by combination, then by
  // the earliest schedule position of the element's entities (spatial order
  // when the schedule is)
  std::vector<int64_t> key(E);
  std::vector<int32_t> es(E, 0);
  for (int64_t e = 0; e < E; ++e) {
    int32_t m = pos[eg[e * nle]], mx = m;
    for (int la = 1; la < nle; ++la) {
      m = std::min(m, pos[eg[e * nle + la]]);
      mx = std::max(mx, pos[eg[e * nle + la]]);
    }
    key[e] = ((int64_t)ec[e] << 32) | (uint32_t)m;
    if (seg) {
      const auto& sp = seg->pos;
      const int sm = (int)(std::upper_bound(sp.begin(), sp.end(), (int64_t)m) - sp.begin()) - 1;
      if (sm < 0 || sm >= seg->nseg || mx >= sp[sm + 1])
        PG_FAIL(PG_ERR_ARG, "element %lld has entities of two meshes", (long long)e);
      es[e] = sm;
    }
  }
  std::vector<int64_t> sorted(E);
  std::iota(sorted.begin(), sorted.end(), 0);
  std::stable_sort(sorted.begin(), sorted.end(), [&](int64_t a, int64_t b) {
    return es[a] != es[b] ? es[a] < es[b] : key[a] < key[b];
  });
  // waves of 64 elements of one combination (and one segment)
  std::vector<int32_t> wcombo, perm, wsg;
  for (int64_t k = 0; k < E;) {
    const int c = ec[sorted[k]], sg = es[sorted[k]];
    int64_t k1 = k;
    while (k1 < E && ec[sorted[k1]] == c && es[sorted[k1]] == sg) ++k1;
    for (int64_t a = k; a < k1; a += kWave) {
      wcombo.push_back(c);
      wsg.push_back(sg);
      for (int l = 0; l < kWave; ++l) perm.push_back(a + l < k1 ? (int32_t)sorted[a + l] : -1);
    }
    k = k1;
  }
  const int64_t nw = (int64_t)wcombo.size(), nslot = nw * kWave;
  if (nslot * N >= (1LL << 31)) PG_FAIL(PG_ERR_UNSUPPORTED, "too many elements");
  // per slot and local entity: the entity's first DOF (-1: PEC or empty)
  std::vector<int32_t> hed((size_t)nslot * nle, -1);
  for (int64_t s = 0; s < nslot; ++s) {
    const int64_t e = perm[s];
    if (e < 0) continue;
    const int64_t w = s / kWave, l = s % kWave;
    for (int la = 0; la < nle; ++la) {
      const int g = eg[e * nle + la];
      hed[(w * nle + la) * kWave + l] = hbc[g] ? -1 : bp[g];
    }
  }
  // per-entity lists of local-result slots, in schedule order; within an
  // entity in ascending slot order (a fixed summation order)
  std::vector<int64_t> start(nblocks + 1, 0);
  for (int64_t i = 0; i < nblocks; ++i) {
    const int64_t g = order ? order[i] : i;
    start[i + 1] = start[i] + (hbc[g] ? 0 : cnt[g]);
  }
  if (start[nblocks] >= (1LL << 31)) PG_FAIL(PG_ERR_UNSUPPORTED, "too many incidences");
  std::vector<int32_t> gslot(start[nblocks]);
  std::vector<int64_t> cur(nblocks);
  for (int64_t g = 0; g < nblocks; ++g) cur[g] = start[pos[g]];
  for (int64_t s = 0; s < nslot; ++s) {
    const int64_t e = perm[s];
    if (e < 0) continue;
    const int64_t w = s / kWave, l = s % kWave;
    for (int la = 0; la < nle; ++la) {
      const int g = eg[e * nle + la];
      if (hbc[g]) continue;
      gslot[cur[g]++] = (int32_t)((w * N + entity_local_offset(la, p)) * kWave + l);
    }
  }
  std::vector<int4> gm(nblocks);
  for (int64_t i = 0; i < nblocks; ++i) {
    const int64_t g = order ? order[i] : i;
    gm[i] = make_int4(bp[g], bp[g + 1] - bp[g], (int)start[i], (int)start[i + 1]);
  }
  // per-combination tables of the upper triangle without the structural zeros
  // (which must vanish here too: a stale mask would drop real terms)
  const int V = n_variants(p), nc = (int)combos.size();
  const int NT = p == 1   ? eop::n_table_entries<1>()
                 : p == 2 ? eop::n_table_entries<2>()
                          : eop::n_table_entries<3>();
  const int NTS = p == 1   ? eop::table_stride<1>()
                  : p == 2 ? eop::table_stride<2>()
                           : eop::table_stride<3>();
  std::vector<double> T((size_t)V * V * 12);
  int rc = pg_tables_host(p, T.data(), (int64_t)T.size());
  if (rc != PG_OK) return rc;
  double tmax = 0.0;
  for (double v : T) tmax = std::max(tmax, std::fabs(v));
  std::vector<double> tab((size_t)nc * NTS, 0.0);
  std::vector<int> var(N);
  for (int c = 0; c < nc; ++c) {
    for (int la = 0; la < nle; ++la)
      for (int k = 0; k < entity_ndof(la, p); ++k)
        var[entity_local_offset(la, p) + k] =
            entity_variant_base(la, orient_of(combos[c], la), p) + k;
    int k = 0;
    size_t o = (size_t)c * NTS;
    for (int i = 0; i < N; ++i)
      for (int j = i; j < N; ++j, ++k) {
        const unsigned zm = p == 1   ? eop::zero_mask<1>(k)
                            : p == 2 ? eop::zero_mask<2>(k)
                                     : eop::zero_mask<3>(k);
        for (int q = 0; q < 12; ++q) {
          const double v = T[((size_t)var[i] * V + var[j]) * 12 + q];
          if ((zm >> q) & 1) {
            if (std::fabs(v) > 1e-15 * tmax)
              PG_FAIL(PG_ERR_UNSUPPORTED, "element table entry (%d, %d, %d) is not zero", i, j, q);
          } else {
            tab[o++] = v;
          }
        }
      }
    if (o != (size_t)c * NTS + NT) PG_FAIL(PG_ERR_UNSUPPORTED, "element table size mismatch");
  }
  std::vector<float> tab32(tab.begin(), tab.end());
  // p = 3: matrix-core fragments.  kcol = the (c, j) columns of Tf that are not
  // zero for every orientation (c << 8 | j), padded to a multiple of 4 with -1;
  // afrag[combo][ks][mt][lane] = Tf[mt * 16 + lane % 16][kcol[4 ks + lane / 16]]
  std::vector<int32_t> kcol;
  std::vector<double> afrag;
  int nks = 0;
  if (p == 3) {
    const int NPp = N * (N + 1) / 2;
    std::vector<int> poff(NPp + 1, 0);
    for (int k = 0; k < NPp; ++k) poff[k + 1] = poff[k] + 12 - eop::popcount12(eop::zero_mask<3>(k));
    auto pair_of = [&](int i, int j) {
      const int a = std::min(i, j), b2 = std::max(i, j);
      return a * N - (a * (a - 1)) / 2 + (b2 - a);
    };
    for (int c = 0; c < 12; ++c)
      for (int j = 0; j < N; ++j) {
        bool any = false;
        for (int i = 0; i < N && !any; ++i) any = !((eop::zero_mask<3>(pair_of(i, j)) >> c) & 1);
        if (any) kcol.push_back((c << 8) | j);
      }
    while (kcol.size() % 4) kcol.push_back(-1);
    nks = (int)kcol.size() / 4;
    afrag.assign((size_t)nc * nks * 3 * kWave, 0.0);
    for (int cb = 0; cb < nc; ++cb)
      for (int ks = 0; ks < nks; ++ks)
        for (int mt = 0; mt < 3; ++mt)
          for (int l = 0; l < kWave; ++l) {
            const int i = mt * 16 + l % 16, kc = kcol[ks * 4 + l / 16];
            if (i >= N || kc < 0) continue;
            const int c = kc >> 8, j = kc & 255, k = pair_of(i, j);
            const unsigned zm = eop::zero_mask<3>(k);
            if ((zm >> c) & 1) continue;
            int rank = 0;  // position of component c among the pair's kept components
            for (int q = 0; q < c; ++q) rank += !((zm >> q) & 1);
            afrag[(((size_t)cb * nks + ks) * 3 + mt) * kWave + l] =
                tab[(size_t)cb * NTS + poff[k] + rank];
          }
  }

  ElemOp* op = new ElemOp();
  op->p = p;
  op->N = N;
  op->D = maxd;
  op->even_entities = true;
  for (int64_t g = 0; g < nblocks; ++g)
    if (bp[g] % 2) op->even_entities = false;
  op->E = E;
  op->ndof = ndof;
  op->nent = nblocks;
  op->nw = nw;
  op->ncombo = nc;
  {
    int64_t g = (nblocks + eop::kBS - 1) / eop::kBS;
    g = std::min<int64_t>(g, eop::kMaxGrid);
    if (g >= 64) g = (g / 8) * 8;
    op->grid = (int)std::max<int64_t>(g, 1);
  }
#define PG_A(ptr, cnt)                            \
  if ((rc = op->alloc(&(ptr), (cnt))) != PG_OK) { \
    delete op;                                    \
    return rc;                                    \
  }
  PG_A(op->wcombo, nw);
  PG_A(op->edofs, hed.size());
  PG_A(op->perm, nslot);
  PG_A(op->gm, nblocks);
  PG_A(op->gslot, gslot.size());
  PG_A(op->coef64, nslot * 12);
  PG_A(op->coef32, nslot * 12);
  PG_A(op->tab64, tab.size());
  PG_A(op->tab32, tab.size());
  PG_A(op->ye64, nslot * N);
  PG_A(op->ye32, nslot * N);
  if (p == 3) {
    PG_A(op->afrag, afrag.size());
    PG_A(op->kcol, kcol.size());
    op->nks = nks;
  }
  std::vector<double> womu;
  if (seg) {
    op->nseg = seg->nseg;
    op->seg_pos = seg->pos;
    op->seg_wave.assign(seg->nseg + 1, nw);
    for (int64_t w = nw - 1; w >= 0; --w) op->seg_wave[wsg[w]] = w;
    for (int sg = seg->nseg - 1; sg >= 0; --sg)  // a segment without waves starts where the next does
      op->seg_wave[sg] = std::min(op->seg_wave[sg], op->seg_wave[sg + 1]);
    for (double om : seg->omega) womu.push_back(om * eop::kMu0);
    PG_A(op->wseg, nw);
    PG_A(op->womu, womu.size());
  }
#undef PG_A
  const struct {
    void* dst;
    const void* src;
    size_t bytes;
  } up[] = {{op->wcombo, wcombo.data(), wcombo.size() * sizeof(int32_t)},
            {op->edofs, hed.data(), hed.size() * sizeof(int32_t)},
            {op->perm, perm.data(), perm.size() * sizeof(int32_t)},
            {op->gm, gm.data(), gm.size() * sizeof(int4)},
            {op->gslot, gslot.data(), gslot.size() * sizeof(int32_t)},
            {op->tab64, tab.data(), tab.size() * sizeof(double)},
            {op->tab32, tab32.data(), tab32.size() * sizeof(float)},
            {op->afrag, afrag.data(), afrag.size() * sizeof(double)},
            {op->kcol, kcol.data(), kcol.size() * sizeof(int32_t)},
            {op->wseg, wsg.data(), op->wseg ? wsg.size() * sizeof(int32_t) : 0},
            {op->womu, womu.data(), womu.size() * sizeof(double)}};
  for (const auto& u : up)
    if (u.bytes && hipMemcpy(u.dst, u.src, u.bytes, hipMemcpyHostToDevice) != hipSuccess) {
      delete op;
      PG_FAIL(PG_ERR_HIP, "element operator upload failed");
    }
  if ((rc = eop_refresh(op, ws, omega, st)) != PG_OK) {
    delete op;
    return rc;
  }
  if (hipStreamSynchronize(st) != hipSuccess) {
    delete op;
    PG_FAIL(PG_ERR_HIP, "element operator setup failed");
  }
  *out = op;
  return PG_OK;
}

}  // namespace pg

using namespace pg;

extern "C" {

int pg_eop_create(int p, int64_t E, const void* ws, const int32_t* dofs, int64_t ndof,
                  const int32_t* block_ptr, int64_t nblocks, const int32_t* order,
                  const uint8_t* ent_bc, double omega, void** out, void* stream) {
  if (!out || !ws || !dofs || !block_ptr || E <= 0 || ndof <= 0 || nblocks <= 0)
    PG_FAIL(PG_ERR_ARG, "bad arguments");
  if (p < 1 || p > 3)
    PG_FAIL(PG_ERR_UNSUPPORTED, "element-by-element operator: order p=%d not supported (1..3)", p);
  const int N = ndof_element(p);
  hipStream_t st = as_stream(stream);
  int rc0 = check_block_ptr(p, block_ptr, nblocks, ndof);
  if (rc0 != PG_OK) return rc0;
  // host copies of the element DOFs, orientations and PEC entities
  std::vector<int32_t> hd((size_t)E * N), hor(E);
  std::vector<uint8_t> hbc(nblocks, 0);
  const int32_t* dor = reinterpret_cast<const int32_t*>(reinterpret_cast<const double*>(ws) + 12 * E);
  PG_HIP(hipMemcpyAsync(hd.data(), dofs, hd.size() * sizeof(int32_t), hipMemcpyDeviceToHost, st));
  PG_HIP(hipMemcpyAsync(hor.data(), dor, E * sizeof(int32_t), hipMemcpyDeviceToHost, st));
  if (ent_bc)
    PG_HIP(hipMemcpyAsync(hbc.data(), ent_bc, nblocks, hipMemcpyDeviceToHost, st));
  PG_HIP(hipStreamSynchronize(st));
  ElemOp* op = nullptr;
  const int rc = eop_build(p, E, ws, ndof, block_ptr, nblocks, order, hd, hor, hbc, omega,
                           nullptr, &op, st);
  if (rc != PG_OK) return rc;
  *out = op;
  return PG_OK;
}

int pg_eop_create_multi(int p, int nseg, const int64_t* n_elems, const void* const* ws,
                        const int32_t* const* dofs, const int64_t* ndof,
                        const int32_t* const* block_ptr, const int64_t* nblocks,
                        const int32_t* const* order, const uint8_t* const* ent_bc,
                        const double* omega, void** out, void* stream) {
  if (!out || nseg <= 0 || !n_elems || !ws || !dofs || !ndof || !block_ptr || !nblocks || !omega)
    PG_FAIL(PG_ERR_ARG, "bad arguments");
  if (p < 1 || p > 3)
    PG_FAIL(PG_ERR_UNSUPPORTED, "element-by-element operator: order p=%d not supported (1..3)", p);
  const int N = ndof_element(p);
  hipStream_t st = as_stream(stream);
  // the union: mesh s's elements, DOFs and entities follow those of meshes < s
  int64_t E = 0, nd = 0, nb = 0;
  SegInfo seg;
  seg.nseg = nseg;
  seg.pos.push_back(0);
  for (int s = 0; s < nseg; ++s) {
    if (n_elems[s] <= 0 || ndof[s] <= 0 || nblocks[s] <= 0 || !ws[s] || !dofs[s] ||
        !block_ptr[s])
      PG_FAIL(PG_ERR_ARG, "mesh %d: bad arguments", s);
    int rc0 = check_block_ptr(p, block_ptr[s], nblocks[s], ndof[s]);
    if (rc0 != PG_OK) return rc0;
    E += n_elems[s];
    nd += ndof[s];
    nb += nblocks[s];
    seg.pos.push_back(nb);
    seg.omega.push_back(omega[s]);
  }
  if (nd >= (1LL << 31) || nb >= (1LL << 31) || E * N >= (1LL << 31))
    PG_FAIL(PG_ERR_UNSUPPORTED, "the meshes together are too large for 32-bit indices");
  std::vector<int32_t> hd((size_t)E * N), hor(E), bp(nb + 1), ord(nb);
  std::vector<uint8_t> hbc(nb, 0);
  double* wsu = nullptr;
  if (hipMalloc(&wsu, (size_t)E * 12 * sizeof(double)) != hipSuccess)
    PG_FAIL(PG_ERR_HIP, "hipMalloc(%lld) failed", (long long)(E * 12 * sizeof(double)));
  int64_t e0 = 0, d0 = 0, b0 = 0;
  int rc = PG_OK;
  for (int s = 0; s < nseg && rc == PG_OK; ++s) {
    const int64_t Es = n_elems[s];
    const double* wss = reinterpret_cast<const double*>(ws[s]);
    if (hipMemcpyAsync(wsu + 12 * e0, wss, Es * 12 * sizeof(double), hipMemcpyDeviceToDevice,
                       st) != hipSuccess ||
        hipMemcpyAsync(hd.data() + e0 * N, dofs[s], Es * N * sizeof(int32_t),
                       hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(hor.data() + e0, reinterpret_cast<const int32_t*>(wss + 12 * Es),
                       Es * sizeof(int32_t), hipMemcpyDeviceToHost, st) != hipSuccess ||
        (ent_bc && ent_bc[s] &&
         hipMemcpyAsync(hbc.data() + b0, ent_bc[s], nblocks[s], hipMemcpyDeviceToHost, st) !=
             hipSuccess) ||
        hipStreamSynchronize(st) != hipSuccess) {
      rc = PG_ERR_HIP;
      break;
    }
    for (int64_t k = e0 * N; k < (e0 + Es) * N; ++k) hd[k] += (int32_t)d0;
    for (int64_t g = 0; g < nblocks[s]; ++g) {
      bp[b0 + g] = (int32_t)(block_ptr[s][g] + d0);
      const int32_t o = order && order[s] ? order[s][g] : (int32_t)g;
      if (o < 0 || o >= nblocks[s]) {
        rc = PG_ERR_ARG;
        break;
      }
      ord[b0 + g] = (int32_t)(o + b0);
    }
    e0 += Es;
    d0 += ndof[s];
    b0 += nblocks[s];
  }
  bp[nb] = (int32_t)nd;
  ElemOp* op = nullptr;
  if (rc == PG_OK)
    rc = eop_build(p, E, wsu, nd, bp.data(), nb, ord.data(), hd, hor, hbc, 0.0, &seg, &op, st);
  else
    set_error("multi-mesh element operator: copies failed or an entity order out of range");
  if (rc != PG_OK) {
    (void)hipFree(wsu);
    return rc;
  }
  op->owned.push_back(wsu);  // the union's coefficients, read by no refresh
  *out = op;
  return PG_OK;
}

int pg_eop_refresh(void* op, const void* ws, double omega, void* stream) {
  if (!op || !ws) PG_FAIL(PG_ERR_ARG, "bad arguments");
  if (static_cast<ElemOp*>(op)->nseg > 0)
    PG_FAIL(PG_ERR_UNSUPPORTED, "a multi-mesh element operator keeps its creation coefficients");
  return eop_refresh(static_cast<ElemOp*>(op), ws, omega, as_stream(stream));
}

int pg_eop_apply_z(const void* op, const double* x, double* y, void* stream) {
  if (!op || !x || !y) PG_FAIL(PG_ERR_ARG, "bad arguments");
  if (!eop_supports(static_cast<const ElemOp*>(op), true))
    PG_FAIL(PG_ERR_UNSUPPORTED, "element-by-element operator without a complex128 product");
  eop_launch(static_cast<const ElemOp*>(op), 0, nullptr, reinterpret_cast<const double2*>(x),
             reinterpret_cast<double2*>(y), nullptr, nullptr, as_stream(stream));
  PG_CHECK_LAUNCH();
  return PG_OK;
}

int pg_eop_apply_c(const void* op, const float* x, float* y, void* stream) {
  if (!op || !x || !y) PG_FAIL(PG_ERR_ARG, "bad arguments");
  eop_launch(static_cast<const ElemOp*>(op), 0, nullptr, reinterpret_cast<const float2*>(x),
             reinterpret_cast<float2*>(y), nullptr, nullptr, as_stream(stream));
  PG_CHECK_LAUNCH();
  return PG_OK;
}

int pg_eop_destroy(void* op) {
  delete static_cast<ElemOp*>(op);
  return PG_OK;
}

}  // extern "C"
