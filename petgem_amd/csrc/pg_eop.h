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

}  // namespace pg
