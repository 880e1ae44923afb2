// Host launchers of the column operators and small utility kernels.
#pragma once

#include "qg_common.hpp"

namespace qg {

// Column solves run as partitioned column kernels (qg_colscan.hpp); depths up
// to 1024 levels are supported.  `shape` selects a compiled launch shape
// (an index into kScanShapes; -1: the default for the depth).
bool column_depth_supported(int nz);
// Thomas elimination tables (1/pivot, c') of the quarter columns
// |ix| <= nx/2, |iy| <= ny/2, level-indexed [k][rc] (kdiv 1: q->psi,
// 4: B->A); thomas_table_size() entries each.
size_t thomas_table_size(const Geo& g);
void launch_thomas_factor(const Geo& g, hipStream_t st, const double* a, int kdiv, double* inv, double* cp,
                          int* err);
bool thomas_setup();
// `al` is a shifted by one level (al[k] = a[k-1], al[0] = 0), zero-padded to
// thomas_shifted_a_size(nzc) entries per member.  Per-member arrays (a, N2
// derived) are concatenated in member order.
int thomas_shifted_a_size(int nzc);
// Thomas solves of all columns; C (= dA/dz) is written when Cout is set.
// Field arguments (rhs, out, Cout, and below psi, q, B, A, C) are cplx<float>
// when f32 is set and cplx<double> otherwise; the solves run in FP64.
void launch_thomas(const Geo& g, hipStream_t st, const void* rhs, void* out, void* Cout, bool f32,
                   const double* inv, const double* cp, const double* al, int kdiv, unsigned long long* warn,
                   int shape = -1);
// The q -> psi solve (tables inv_q, cp_q; warn as for kdiv 1) and the
// B -> A solve (inv_b, cp_b; C when Cout is set) in one launch, both at
// launch shape `shape`: the same values as the two launch_thomas calls.
void launch_thomas_pair(const Geo& g, hipStream_t st, bool f32, const double* al, int shape, const void* q,
                        void* psi, const double* inv_q, const double* cp_q, unsigned long long* warn,
                        const void* B, void* A, void* Cout, const double* inv_b, const double* cp_b);
// QG omega-equation interior solves: elimination recorded once per model
// (tables of thomas_table_size() entries, pivoted flags per quarter column,
// err[3] set when any column interchanged rows), then applied to every
// step's right-hand sides.  cu: upper diagonal by level, zero-padded to
// thomas_shifted_a_size(nz) entries.  Columns with interchanges are solved
// by a second kernel when any_pivoted is set.
void launch_omega_factor(const Geo& g, hipStream_t st, const double* cz, double* M, double* D,
                         unsigned char* pivoted, int* err);
bool omega_setup();
void launch_omega(const Geo& g, hipStream_t st, const void* rhsw, void* wk, bool f32, const double* M,
                  const double* D, const unsigned char* pivoted, bool any_pivoted, const double* cz,
                  const double* cu, const double* fN2, double* sd, double* sdu, double* sfill, int shape = -1);
// Initial conditions on the device (qg_init.hip): the seeded random
// streamfunction of members [m0, m0 + nm) and the spectral surface wave
// (profile: nzc values per member, the amplitude times the vertical profile).
void launch_random_psi(const Geo& g, hipStream_t st, int m0, int nm, unsigned long long seed, double amplitude,
                       double slope, void* psi, bool f32);
void launch_surface_wave(const Geo& g, hipStream_t st, int m0, int nm, const double* profile, void* B, bool f32);
// Energy diagnostics (qg_diag.hip) of members [m0, m0 + nmem): out receives
// 5 values per member (wave KE, PE, CE, mean-flow KE, PE); partial holds
// energy_partial_size(g, nmem) doubles.
size_t energy_partial_size(const Geo& g, int nmem);
// max hypot(u, v) over n points of two real fields (FP32 or FP64), folded into
// *out (the bit pattern of a non-negative double; the caller clears it).
void launch_max_hypot(const void* u, const void* v, bool f32, size_t n, unsigned long long* out, hipStream_t st);
void launch_energies(const Geo& g, hipStream_t st, int m0, int nmem, const void* psi, const void* A,
                     const void* C, bool f32, const double* a, double* partial, double* out);
// psi_f64: psi is cplx<double> while q is stored per f32 (set-up from a
// psi still held in FP64).
void launch_q_from_psi(const Geo& g, hipStream_t st, const void* psi, void* q, bool f32, const double* a,
                       int barotropic, bool psi_f64 = false);
template <typename T>
void launch_transpose(const T* in, T* out, int nx, int ny, int nz, bool to_planar, hipStream_t st);
template <typename From, typename To>
void launch_convert(const From* in, To* out, size_t n, hipStream_t st);
void launch_zero_bytes(void* p, size_t bytes, hipStream_t st);
// Zeroes a spectral field (cplx<float> when f32, else cplx<double>) outside
// the dealiased disk of a pruned geometry, on all its g.nz levels.
void launch_zero_outside_disk(const Geo& g, hipStream_t st, void* field, bool f32);
void launch_zero(cplx<double>* p, size_t n, hipStream_t st);

}  // namespace qg
