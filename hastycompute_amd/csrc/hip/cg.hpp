// Device-side batched conjugate gradient, see cg.hip.
#pragma once

#include "common.hpp"

namespace hasty_hip {

size_t cg_workspace_doubles(int64_t P, int64_t n);
void cg_init(int64_t P, int64_t n, const float2* b, const float2* Ax, float2* r, float2* p, double* state, double* ws,
             hipStream_t s);
// shift > 0: the iteration for A + shift I, with Ap = A p from the caller
void cg_iterate(int64_t P, int64_t n, float2* x, float2* r, float2* p, const float2* Ap, double* state, double* ws,
                double tol, hipStream_t s, double shift = 0.0);
void cg_init_precond(int64_t P, int64_t n, const float2* r, const float2* z, float2* p, double* state, double* ws,
                     hipStream_t s);
void cg_iterate_xr(int64_t P, int64_t n, float2* x, float2* r, const float2* p, const float2* Ap, double* state,
                   double* ws, double tol, hipStream_t s);
void cg_iterate_pz(int64_t P, int64_t n, const float2* r, const float2* z, float2* p, double* state, double* ws,
                   hipStream_t s);

}  // namespace hasty_hip
