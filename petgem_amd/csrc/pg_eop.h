// Element-by-element operator: the interface the Krylov solver uses to apply
// A = sum_e A_e without the assembled matrix (kernels in pg_elemop.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace pg {

struct ElemOp;

// Workgroups of the per-entity summation kernel: the number of reduction
// partials (per quantity) that MODE 1 and MODE 2 leave in `part`.
int eop_grid(const ElemOp* op);
int64_t eop_ndof(const ElemOp* op);
// whether the operator has a complex128 (true) or complex64 (false) product
// (at p = 3 only complex64)
bool eop_supports(const ElemOp* op, bool complex128);

// MODE 0: y = A x.  MODE 1: y = A x and partials of x^T y (two per workgroup,
// part[k * grid + b]).  MODE 2: y = b - A x and partials of |y|^2.  With `done`
// non-null both kernels return at once while *done is set.  A non-null `mid`
// is recorded on `st` between the two kernels (timing of the element products).
void eop_launch(const ElemOp* op, int mode, const int32_t* done, const double2* x, double2* y,
                const double2* b, double* part, hipStream_t st, hipEvent_t mid = nullptr);
void eop_launch(const ElemOp* op, int mode, const int32_t* done, const float2* x, float2* y,
                const double2* b, double* part, hipStream_t st, hipEvent_t mid = nullptr);

// Multi-right-hand-side products for the batched solver (p = 1, 2): column r
// of a row-major (ndof, R) block gets its own frequency.  eop_multi_prepare
// (once; `ws` = the geometry workspace the operator was created from) keeps
// raw coefficients that pg_eop_refresh updates from then on.
// eop_multi_local writes the local results ye (nw * 64 * N * R values) of
// every element and column and, with part non-null, per workgroup and column
// the FP64 sums of x_e^T y_e over its elements (part[(k R + r) np + b],
// k = 0 real, 1 imaginary, np = eop_multi_parts partials per column: the
// elements' share of x^T A x); the batch sums them per entity with the schedule
// below (local result of slot g, entity DOF t, column r at ye[(g + 64 t) R + r];
// s0 == s1 marks a PEC entity, an identity row).  R = 64, 128, 192 or 256.
struct EopSchedule {
  const int4* gm;        // per scheduled entity {first DOF, d, s0, s1}
  const int32_t* gslot;  // local-result slots, ascending per entity
  int64_t nent, nw;
  int p, N, D;
};
int eop_multi_prepare(ElemOp* op, const void* ws, hipStream_t st);
EopSchedule eop_schedule(const ElemOp* op);
int64_t eop_multi_parts(const ElemOp* op, int R);
void eop_multi_local(const ElemOp* op, int R, const int32_t* done, const double* omu,
                     const double2* x, double2* ye, double* part, hipStream_t st);
void eop_multi_local(const ElemOp* op, int R, const int32_t* done, const double* omu,
                     const float2* x, float2* ye, double* part, hipStream_t st);

// Multi-mesh operators (pg_eop_create_multi): the meshes' systems side by side
// (block-diagonal), every wave of 64 elements within one mesh ("segment").
// eop_segments returns the number of segments (0 for a one-mesh operator) and
// points seg_pos at the first schedule position of each and seg_wave at the
// first wave of each (nseg + 1 values each; schedule and waves are segment
// after segment).  eop_local_segments writes the local results (complex128
// or complex64, as x) of the waves [w0, w1) (w1 < 0: to the last) whose
// segment s is not done (done[s * dstride] == 0) into the operator's buffer
// and returns it; the caller sums them per entity with eop_schedule.
int eop_segments(const ElemOp* op, const int64_t** seg_pos, const int64_t** seg_wave);
const double2* eop_local_segments(const ElemOp* op, const int32_t* done, int dstride,
                                  const double2* x, hipStream_t st, int64_t w0 = 0,
                                  int64_t w1 = -1);
const float2* eop_local_segments(const ElemOp* op, const int32_t* done, int dstride,
                                 const float2* x, hipStream_t st, int64_t w0 = 0,
                                 int64_t w1 = -1);

}  // namespace pg
