// Host launchers of the column operators and small utility kernels.
#pragma once

#include "qg_common.hpp"

namespace qg {

// LDS budget of one column-solver tile (TC columns x nz levels, FP64 complex,
// plus the per-level coefficients); two tiles fit in a CU's 160 KB.
constexpr size_t kThomasLdsBytes = 72 * 1024;
int thomas_tile_columns(int nz);
bool thomas_lds_setup();
// Thomas elimination tables (pivot, its reciprocal, c') of every column,
// stored for the quarter of wavenumbers |ix| <= nx/2, |iy| <= ny/2
// (kdiv 1: q->psi, 4: B->A); thomas_table_size() entries each.
size_t thomas_table_size(const Geo& g);
void launch_thomas_factor(const Geo& g, hipStream_t st, const double* a, int kdiv, double* den, double* inv,
                          double* cp, int* err);
// Thomas solves of all columns; C (= dA/dz) is written when Cout is set.
void launch_thomas_lds(const Geo& g, hipStream_t st, const cplx<double>* rhs, cplx<double>* out,
                       cplx<double>* Cout, const double* den, const double* inv, const double* cp,
                       const double* a, int kdiv, unsigned long long* warn);
// QG omega-equation interior solves: elimination recorded once per model,
// then applied to every step's right-hand sides.
void launch_omega_factor(const Geo& g, hipStream_t st, const double* cz, double* fact, double* dp, double* dinv,
                         unsigned char* pivoted, int* err);
bool omega_lds_setup();
void launch_omega_tile(const Geo& g, hipStream_t st, const cplx<double>* rhsw, cplx<double>* wk,
                       const double* fact, const double* dp, const double* dinv, const unsigned char* pivoted,
                       const double* cz, const double* fN2, double* sd, double* sdu, double* sfill);
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
