// Host launchers of the column operators and small utility kernels.
#pragma once

#include "qg_common.hpp"

namespace qg {

void launch_thomas(const Geo& g, hipStream_t st, const cplx<double>* rhs, cplx<double>* out,
                   cplx<double>* Cout, double* cscr, const double* a, int kdiv, int* err,
                   unsigned long long* warn);
void launch_omega(const Geo& g, hipStream_t st, const cplx<double>* rhsw, cplx<double>* wk, double* sd,
                  double* sdu, double* sfill, const double* cz, const double* fN2, int* err);
void launch_q_from_psi(const Geo& g, hipStream_t st, const cplx<double>* psi, cplx<double>* q,
                       const double* a, int barotropic);
template <typename T>
void launch_transpose(const T* in, T* out, int nx, int ny, int nz, bool to_planar, hipStream_t st);
template <typename From, typename To>
void launch_convert(const From* in, To* out, size_t n, hipStream_t st);
void launch_nonfinite(const cplx<double>* a, size_t n, int* flag, hipStream_t st);
void launch_zero_bytes(void* p, size_t bytes, hipStream_t st);
void launch_zero(cplx<double>* p, size_t n, hipStream_t st);

}  // namespace qg
