// Host runtime of the QG-YBJ+ library: model handle, HBM buffers, the
// ETD-RK2 step orchestration and the C ABI declared in include/qgybj.h.
#include <algorithm>
#include <array>
#include <cmath>
#include <limits>
#include <map>
#include <mutex>
#include <tuple>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/qgybj.h"
#include "qg_colscan.hpp"
#include "qg_columns.hpp"
#include "qg_ops.hpp"

using qg::cplx;
using D = cplx<double>;

namespace {

thread_local std::string g_create_error;

struct QgError : std::runtime_error {
  int code;
  QgError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPCHECK(expr)                                                                 \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      throw QgError(QG_ERR_DEVICE, std::string(#expr " failed: ") + hipGetErrorString(e_)); \
  } while (0)

}  // namespace

namespace qg {
// Tunable passes of the step: the fused FFT passes (tile variant 0..2) and
// the column operators (launch shape index into kScanShapes).
enum Pass : int {
  P_YA_QW, P_XA_QW, P_YA2_QEFF, P_YB_ADV, P_X_C2R, P_XB_Q, P_XB_B, P_YB2_Q, P_YB2_B, P_YF_OMEGA, P_XF,
  P_Y_R2C, P_Y_HERM, P_FFT_END,
  P_THOMAS_PSI = P_FFT_END, P_THOMAS_A, P_THOMAS_AC, P_OMEGA, P_THOMAS_PAIR, P_END
};
}  // namespace qg

struct qg_handle {
  // nz counts all stacked levels: members * nzc (see qg::Geo)
  int nx = 0, ny = 0, nz = 0, nzc = 0, members = 1, precision = 0;
  // members [0, active) are stepped; g describes them (g.members == active)
  int active = 1;
  double f = 0, dz = 1, nu_z = 0;
  bool fixed_flow = false, feedback = false, passive = false, inviscid = false, linear = false,
       no_disp = false;
  std::vector<double> kx, ky, N2, a, fc, wc;
  std::vector<int> fo, wo;
  hipStream_t st = nullptr;
  // g: every pass unpruned; gp: the same grid pruned to the dealiased disk
  // (qg::Geo).  A step runs pruned when the model can be (prunable: q and psi
  // are stepped, so every spectral field it reads has been through the mask)
  // and `clean` holds: q, B of both parities, psi, A and C are zero outside
  // the disk.  Any set-up or operator call clears `clean`; the next step then
  // runs unpruned and zeroes the outside of the buffers it leaves behind.
  qg::Geo g{}, gp{};
  bool prunable = false, prune_enabled = true, clean = false;
  // points a pass touches (the algorithmic bytes of profiles): spectral
  // points inside the disk per plane, for gp
  size_t disk_plane = 0, disk_quarter = 0;
  size_t plane = 0, npts = 0, nhalf = 0;
  std::vector<void*> allocs;
  // device
  double* d_prof = nullptr;  // a vertical profile per stacked level (set-up input)
  double *d_epart = nullptr, *d_energy = nullptr;  // energy reduction partials, results
  double *d_kx = nullptr, *d_ky = nullptr, *d_a = nullptr, *d_al = nullptr, *d_cz = nullptr, *d_cu = nullptr,
         *d_fN2 = nullptr;
  unsigned char* d_mask = nullptr;
  qg::EtdParams* d_etd = nullptr;  // dt and closure terms of the update passes
  double etd_dt = std::nan("");
  // Complex fields in the mode's storage precision: cplx<double> in FP64
  // mode, cplx<float> in mixed mode (f32()).  stg and stg2 are FP64 staging
  // for host exchange and plain transforms (stg2 aliases qeff in FP64 mode).
  void *q[2] = {nullptr, nullptr}, *B[2] = {nullptr, nullptr};
  int cur = 0;
  void *psi = nullptr, *A = nullptr, *C = nullptr, *qs = nullptr, *Bs = nullptr, *psis = nullptr,
       *As = nullptr, *rq0 = nullptr, *rB0 = nullptr, *qeff = nullptr;
  D *stg = nullptr, *stg2 = nullptr;
  // mixed mode: members [psi64_m0, psi64_m0 + psi64_nm) have the FP64 psi-hat
  // of their last set-up in stg2 (at their own offsets), so q is derived from
  // it rather than from the FP32-rounded psi; any other use of stg2 clears it
  int psi64_m0 = 0, psi64_nm = 0;
  double *s0 = nullptr, *s1 = nullptr, *s2 = nullptr;
  // Thomas elimination factors of the q->psi and B->A operators, and where
  // (if anywhere) their pivots fall below sqrt(eps): {code, level, column}
  double *th_inv[2] = {nullptr, nullptr}, *th_cp[2] = {nullptr, nullptr};
  int sing_psi[3] = {0, 0, 0}, sing_A[3] = {0, 0, 0};
  // recorded dgtsv elimination of the omega-equation columns
  double *om_M = nullptr, *om_D = nullptr;
  unsigned char* om_piv = nullptr;
  bool om_any_piv = false;  // some column's elimination interchanged rows
  int sing_omega[3] = {0, 0, 0};
  void *u = nullptr, *v = nullptr, *w = nullptr;
  void *T1 = nullptr, *T3 = nullptr, *G3 = nullptr, *G4 = nullptr;
  void *Qa = nullptr, *Qb = nullptr, *Pu = nullptr, *Pv = nullptr, *Pq = nullptr, *Pz = nullptr,
       *G1 = nullptr, *G2 = nullptr;
  void *twx = nullptr, *twy = nullptr;  // intermediate precision
  D *tw64x = nullptr, *tw64y = nullptr;
  int* d_err = nullptr;
  int* d_ferr = nullptr;  // error flags of member re-factorisations
  unsigned long long* d_warn = nullptr;
  unsigned long long* d_absmax = nullptr;
  unsigned long long* d_pmax = nullptr;  // progress maxima: wave speed, flow speed
  int* d_flag = nullptr;
  // per stacked level: the bound of the termination check's max |ifft2(psi)|
  // that a bound-checked step accumulates (yf_omega) and clears (xf)
  double* d_lsum = nullptr;
  bool diag_valid = false, t1_valid = false;
  // hipGraphs of two steady-state steps starting from each parity of the
  // double buffer (after two swaps the buffers are back in place)
  bool graphs = true;
  hipGraphExec_t graph[2] = {nullptr, nullptr};
  // the same two steps, each followed by the termination check (qg_run's
  // replay after a failing chunk), or by its bound (qg_run's chunks)
  hipGraphExec_t graph_chk[2] = {nullptr, nullptr};
  hipGraphExec_t graph_bnd[2] = {nullptr, nullptr};
  // the same with kLongGraphSteps steps each (fewer graph launches, and
  // fewer steps whose w passes must finish before their graph ends)
  hipGraphExec_t graph_l[2] = {nullptr, nullptr};
  hipGraphExec_t graph_chk_l[2] = {nullptr, nullptr};
  hipGraphExec_t graph_bnd_l[2] = {nullptr, nullptr};
  // the state a step starts from (q, B and the diagnostics psi, A, T1, u, v
  // it reuses) at the start of a qg_run chunk; allocated by the first qg_run
  void* snap = nullptr;
  // device address of the file-scope snapshot scratch on this handle's
  // device (qg_run_noalloc), looked up by qg_create
  void* noalloc_snap = nullptr;
  std::string err;
  // profiling: bit 1 = events around every kernel, bit 2 = one event per step
  struct ProfRec {
    const char* name;
    int pass;
    double bytes;
    hipEvent_t a, b;
  };
  // tuned choice per pass (-1: the default variant / shape)
  int tune[qg::P_END];
  int prof_mode = 0;
  std::vector<ProfRec> prof;
  qg_handle() { std::fill(tune, tune + qg::P_END, -1); }
  std::vector<hipEvent_t> step_events;
  std::vector<int> step_weights;  // steps covered by each interval of step_events
  std::vector<hipEvent_t> pool;
  size_t pool_used = 0;
  // side stream of the step's independent passes (default priority, as the
  // main stream) and the events that order the two; reused in turn (a wait
  // refers to the record issued just before it)
  hipStream_t st2 = nullptr;
  bool concurrent = true;
  // stream of the finalize's w passes, which run beside the next step (their
  // buffers Pw, Pw2, Wr, Ww are theirs alone); w_pending: issued there and
  // not yet ordered before the main stream
  hipStream_t st3 = nullptr;
  bool w_pending = false;
  void *Pw = nullptr, *Pw2 = nullptr, *Wr = nullptr, *Ww = nullptr;
  hipEvent_t sync_ev[8] = {};
  int sync_next = 0;
  hipEvent_t last_sync = nullptr;
  hipEvent_t sync_event() {
    last_sync = sync_ev[sync_next];
    sync_next = (sync_next + 1) % 8;
    return last_sync;
  }
  hipEvent_t event() {
    if (pool_used == pool.size()) {
      hipEvent_t e;
      HIPCHECK(hipEventCreate(&e));
      pool.push_back(e);
    }
    return pool[pool_used++];
  }

  template <typename T>
  T* alloc(size_t count) {
    void* p = nullptr;
    HIPCHECK(hipMalloc(&p, count * sizeof(T) + 256));
    allocs.push_back(p);
    return static_cast<T*>(p);
  }
  size_t rsize() const { return precision == QG_PRECISION_MIXED ? 4 : 8; }
  bool f32() const { return precision == QG_PRECISION_MIXED; }
  size_t fsize() const { return 2 * rsize(); }  // bytes of one complex field value
  // field pointer advanced by `elems` complex values
  void* at(void* p, size_t elems) const { return p ? static_cast<char*>(p) + elems * fsize() : nullptr; }
  void* falloc(size_t count) { return ralloc(2 * count); }
  void* ralloc(size_t count) {  // count of R scalars
    void* p = nullptr;
    HIPCHECK(hipMalloc(&p, count * rsize() + 256));
    allocs.push_back(p);
    return p;
  }
  void drop_graphs() {
    for (hipGraphExec_t* set : {graph, graph_chk, graph_bnd, graph_l, graph_chk_l, graph_bnd_l})
      for (int i = 0; i < 2; ++i)
        if (set[i]) {
          (void)hipGraphExecDestroy(set[i]);
          set[i] = nullptr;
        }
  }
  ~qg_handle() {
    if (st) (void)hipStreamSynchronize(st);
    if (st2) (void)hipStreamSynchronize(st2);
    if (st3) (void)hipStreamSynchronize(st3);
    drop_graphs();
    for (hipEvent_t e : pool) (void)hipEventDestroy(e);
    for (hipEvent_t e : sync_ev)
      if (e) (void)hipEventDestroy(e);
    for (void* p : allocs) (void)hipFree(p);
    if (st2) (void)hipStreamDestroy(st2);
    if (st3) (void)hipStreamDestroy(st3);
    if (st) (void)hipStreamDestroy(st);
  }
};

namespace {

// exp(-2*pi*i*m/n) correctly rounded from long double.
template <typename R>
std::vector<cplx<R>> twiddles(int n) {
  std::vector<cplx<R>> t(n);
  const long double pi = 3.141592653589793238462643383279502884L;
  for (int m = 0; m < n; ++m) {
    long double ang = -2.0L * pi * (long double)m / (long double)n;
    t[m].re = (R)cosl(ang);
    t[m].im = (R)sinl(ang);
  }
  return t;
}

// The parameters the update passes compute int_factor (src/nonlinear.jl:
// 91-97, 1160-1183) and _etd_coefficients (src/timestep.jl:177-190) from, per
// point (qg::Etd): dt and the closure terms, written in place (graphs
// captured earlier read the new values).
void build_etd(qg_handle& h, double dt) {
  if (h.etd_dt == dt) return;
  qg::EtdParams P{};
  P.dt = dt;
  const std::vector<double>* c[2] = {&h.fc, &h.wc};
  const std::vector<int>* o[2] = {&h.fo, &h.wo};
  for (int w = 0; w < 2; ++w) {
    P.n[w] = h.inviscid ? 0 : (int)c[w]->size();
    for (int i = 0; i < P.n[w]; ++i) {
      P.coef[w][i] = (*c[w])[i];
      P.half[w][i] = (*o[w])[i] / 2;
    }
  }
  HIPCHECK(hipMemcpyAsync(h.d_etd, &P, sizeof P, hipMemcpyHostToDevice, h.st));
  HIPCHECK(hipStreamSynchronize(h.st));
  h.etd_dt = dt;
}

bool is_dealiased(int i, int j, int nx, int ny) {  // src/physics.jl:36-43 (0-based)
  int kmax = std::min(nx, ny) / 3;
  int ix = i <= nx / 2 ? i : i - nx;
  int jy = j <= ny / 2 ? j : j - ny;
  return ix * ix + jy * jy <= kmax * kmax;
}

void check_status(qg_handle& h) {
  int e[3] = {0, 0, 0};
  HIPCHECK(hipMemcpyAsync(e, h.d_err, sizeof(e), hipMemcpyDeviceToHost, h.st));
  HIPCHECK(hipStreamSynchronize(h.st));
  if (e[0] != 0) {
    HIPCHECK(hipMemsetAsync(h.d_err, 0, 3 * sizeof(int), h.st));
    HIPCHECK(hipStreamSynchronize(h.st));
    h.diag_valid = false;
    h.clean = false;
    h.t1_valid = false;
    char buf[256];
    if (e[0] == 1) {
      snprintf(buf, sizeof buf,
               "Singular matrix in Thomas solver at level %d of column %d (|pivot| < sqrt(eps)). "
               "This may indicate an ill-conditioned elliptic problem from N^2~0 or a degenerate wavenumber.",
               e[1] + 1, e[2]);
      throw QgError(QG_ERR_SINGULAR, buf);
    }
    snprintf(buf, sizeof buf, "LAPACK-style gtsv failed for vertical velocity at row %d of column %d",
             e[1] + 1, e[2]);
    throw QgError(QG_ERR_GTSV, buf);
  }
}

[[noreturn]] void raise_singular(const int* s) {
  char buf[256];
  snprintf(buf, sizeof buf,
           "Singular matrix in Thomas solver at level %d of column %d (|pivot| < sqrt(eps)). "
           "This may indicate an ill-conditioned elliptic problem from N^2~0 or a degenerate wavenumber.",
           s[1] + 1, s[2]);
  throw QgError(QG_ERR_SINGULAR, buf);
}

// invert_q_to_psi! (src/elliptic.jl:189-314) on device arrays (the columns of
// geometry g: all, or the disk's when pruned).
void solve_psi(qg_handle& h, const qg::Geo& g, hipStream_t st, const void* q, void* psi) {
  if (h.sing_psi[0]) raise_singular(h.sing_psi);
  qg::launch_thomas(g, st, q, psi, nullptr, h.f32(), h.th_inv[0], h.th_cp[0], h.d_al, 1, h.d_warn,
                    h.tune[qg::P_THOMAS_PSI]);
}

// invert_B_to_A! (src/elliptic.jl:836-990) on device arrays; C optional.
void solve_A(qg_handle& h, const qg::Geo& g, hipStream_t st, const void* B, void* A, void* C) {
  if (h.sing_A[0]) raise_singular(h.sing_A);
  qg::launch_thomas(g, st, B, A, C, h.f32(), h.th_inv[1], h.th_cp[1], h.d_al, 4, nullptr,
                    h.tune[C ? qg::P_THOMAS_AC : qg::P_THOMAS_A]);
}

// solve_psi and solve_A (with C when C is set) in one launch, at the pair's
// own launch shape: the same values, side by side without a second stream.
void solve_psi_A(qg_handle& h, const qg::Geo& g, hipStream_t st, const void* q, void* psi, const void* B, void* A,
                 void* C) {
  if (h.sing_psi[0]) raise_singular(h.sing_psi);
  if (h.sing_A[0]) raise_singular(h.sing_A);
  qg::launch_thomas_pair(g, st, h.f32(), h.d_al, h.tune[qg::P_THOMAS_PAIR], q, psi, h.th_inv[0], h.th_cp[0],
                         h.d_warn, B, A, C, h.th_inv[1], h.th_cp[1]);
}

template <typename R>
struct Engine {
  using C = cplx<R>;
  using O = qg::Ops<R>;
  qg_handle& h;
  // algorithmic bytes per point: complex R, real R, complex field value in
  // storage (the fields are stored at R)
  const double cR = 2.0 * sizeof(R), rR = sizeof(R), cS = 2.0 * sizeof(R);
  // The geometry the passes run on (h.g, or h.gp for a pruned step) and the
  // points one field of each kind has there: P physical (every point), Ps
  // spectral (the disk's points when pruned), Ph half-width and Pf full-width
  // x-transformed intermediates (the columns carried), and Tq the two
  // per-level tables of a column solve (8 bytes per quarter column solved).
  const qg::Geo* G = nullptr;
  double P = 0, Ps = 0, Ph = 0, Pf = 0, Tq = 0;
  // The stream launches go to: the handle's stream, or its side stream for
  // passes that run beside the main chain (side()).
  hipStream_t S;
  bool joined = true;
  explicit Engine(qg_handle& hh) : h(hh), S(hh.st) { use(hh.g); }

  // Passes independent of the main chain run on the side stream, ordered
  // after everything issued so far on the main stream; join() orders the main
  // stream after them.  Under hipGraph capture the two streams become parallel
  // branches of the graph.  With per-launch profiling (or the side stream
  // switched off) everything runs on the main stream in issue order.
  bool concurrent() const { return h.concurrent && h.st2 && !(h.prof_mode & 1); }
  template <class F>
  void side(F&& fn, hipStream_t to = nullptr) {
    if (!to) to = h.st2;
    if (!concurrent()) {
      fn();
      return;
    }
    HIPCHECK(hipEventRecord(h.sync_event(), h.st));
    HIPCHECK(hipStreamWaitEvent(to, h.last_sync, 0));
    S = to;
    fn();
    S = h.st;
    if (to == h.st2) joined = false;
  }
  void join() {
    if (joined) return;
    HIPCHECK(hipEventRecord(h.sync_event(), h.st2));
    HIPCHECK(hipStreamWaitEvent(h.st, h.last_sync, 0));
    joined = true;
  }
  // orders the main stream after the w passes left running by a step
  void join_w() {
    if (!h.w_pending) return;
    HIPCHECK(hipEventRecord(h.sync_event(), h.st3));
    HIPCHECK(hipStreamWaitEvent(h.st, h.last_sync, 0));
    h.w_pending = false;
  }
  void use(const qg::Geo& g) {
    G = &g;
    const double lv = g.nz;
    P = (double)h.plane * lv;
    if (g.pruned()) {
      Ps = (double)h.disk_plane * lv;
      Ph = (double)g.Kx * h.ny * lv;
      Pf = (double)(2 * g.Kx - 1) * h.ny * lv;
      Tq = 2.0 * 8.0 * (double)h.disk_quarter * lv;
    } else {
      Ps = P;
      Ph = (double)g.W * h.ny * lv;
      Pf = P;
      Tq = 2.0 * 8.0 * (h.nx / 2 + 1) * (h.ny / 2 + 1) * lv;
    }
  }
  C* c(void* p) const { return static_cast<C*>(p); }
  R* r(void* p) const { return static_cast<R*>(p); }
  const C* twx() const { return static_cast<const C*>(h.twx); }
  const C* twy() const { return static_cast<const C*>(h.twy); }
  qg::EtdView etd() const { return qg::EtdView{h.d_etd}; }

  template <class F>
  void K(const char* name, int pass, double bytes, F&& fn) {
    if (h.prof_mode & 1) {
      hipEvent_t a = h.event(), b = h.event();
      HIPCHECK(hipEventRecord(a, S));
      fn();
      HIPCHECK(hipEventRecord(b, S));
      h.prof.push_back({name, pass, bytes, a, b});
    } else {
      fn();
    }
  }

  // _check_termination_conditions! (src/simulation.jl:25-44) on the levels of
  // members [m0, m0 + nm): non-finite q, B, psi -> d_flag, max |ifft2(psi)| ->
  // d_absmax, both accumulated (the caller clears them).  G3 is free between
  // steps and holds the y-transformed psi.
  // The passes are those of the check the step fuses (yf_omega's psi output,
  // xf's max), at the same tile variants, so both give the same values.  Pz
  // is free between steps and holds psi's y-transformed half rows.
  void check(int m0, int nm) {
    qg::Geo g = h.g;
    g.nz = nm * h.nzc;
    g.members = nm;
    const size_t off = (size_t)m0 * h.nzc * h.plane;
    C* P = c(h.Pz) + (size_t)m0 * h.nzc * h.ny * h.g.Wp;
    const double pts = (double)h.plane * g.nz, hpts = (double)h.g.W * h.ny * g.nz;
    K("check_y", -1, pts * 3 * cS + hpts * cR, [&] {
      O::check_y(h.tune[qg::P_YF_OMEGA], g, S, static_cast<const C*>(h.at(h.psi, off)),
                 static_cast<const C*>(h.at(h.q[h.cur], off)), static_cast<const C*>(h.at(h.B[h.cur], off)), P,
                 h.d_flag, twy());
    });
    K("check_x", -1, hpts * cR, [&] { O::check_x(h.tune[qg::P_XF], g, S, P, h.d_absmax, twx()); });
  }

  // _progress_maxima (src/simulation.jl:46-81): max |ifft2(LA)| with LA =
  // B + kh2 A / 4 (YBJ+) into d_pmax[0], max hypot(u, v) into d_pmax[1].
  // The formulations built here are YBJ+ and passive (normal YBJ is
  // rejected), so LA always carries kh2 A / 4; A is zero when passive or
  // without dispersion, as the reference fills it.
  void progress_maxima() {
    O::la_max(1, h.g, S, static_cast<const C*>(h.B[h.cur]), static_cast<const C*>(h.A), 1, c(h.G3), h.d_pmax,
              twx(), twy());
    qg::launch_max_hypot(h.u, h.v, h.f32(), (size_t)h.plane * h.g.nz, h.d_pmax + 1, S);
  }

  void t1_only(const C* Bsrc) {
    K("ya_T1", qg::P_YA_QW, Ps * cS + Pf * cR, [&] { O::ya_qw(h.tune[qg::P_YA_QW], *G, S, Bsrc, c(h.T1), nullptr, twy()); });
  }

  // _diagnose_flow! (src/timestep.jl:192-215) minus the velocities: q_eff with
  // wave feedback, then q -> psi.  Leaves T1 = IFFTy(mask B) when need_t1.
  // t_ready: T1 (and T3 when q_w is needed) already hold the inverse y
  // transforms of Bsrc, made by the yb2_B pass that produced Bsrc.
  // Bw: also the wave inversion Bw -> Aw (and Cw when set), independent of
  // the chain: with concurrency on it is solved in the psi solve's launch
  // (on the side stream when there is no psi solve); otherwise first.
  void diag(const C* qsrc, const C* Bsrc, C* psi_out, bool need_t1, bool t_ready = false, const C* Bw = nullptr,
            C* Aw = nullptr, C* Cw = nullptr) {
    const bool pair = Bw && concurrent() && !h.fixed_flow && !(h.passive || h.no_disp);
    if (Bw && !pair) side([&] { wave_inv(Bw, Aw, Cw); });
    bool t1 = t_ready;
    if (!h.fixed_flow) {
      const C* src = qsrc;
      if (h.feedback) {
        if (!t_ready)
          K("ya_qw", qg::P_YA_QW, Ps * cS + Pf * 2 * cR, [&] { O::ya_qw(h.tune[qg::P_YA_QW], *G, S, Bsrc, c(h.T1), c(h.T3), twy()); });
        K("xa_qw", qg::P_XA_QW, Pf * 2 * cR + Ph * 2 * cR,
          [&] { O::xa_qw(h.tune[qg::P_XA_QW], *G, S, c(h.T1), c(h.T3), c(h.Qa), c(h.Qb), twx()); });
        K("ya2_qeff", qg::P_YA2_QEFF, Ph * 2 * cR + Ps * 2 * cS,
          [&] { O::ya2_qeff(h.tune[qg::P_YA2_QEFF], *G, S, c(h.Qa), c(h.Qb), qsrc, c(h.qeff), nullptr, 1.0 / h.f, twy()); });
        src = c(h.qeff);
        t1 = true;
      }
      if (pair)
        K(Cw ? "thomas_psi_AC" : "thomas_psi_A", qg::P_THOMAS_PSI, Ps * 2 * cS + Ps * (Cw ? 3 : 2) * cS + 2 * Tq,
          [&] { solve_psi_A(h, *G, S, src, psi_out, Bw, Aw, Cw); });
      else
        K("thomas_psi", qg::P_THOMAS_PSI, Ps * 2 * cS + Tq,
          [&] { solve_psi(h, *G, S, src, psi_out); });
    }
    if (need_t1 && !t1) t1_only(Bsrc);
    if (Bw && !pair) join();
  }

  void wave_inv(const C* Bsrc, C* Aout, C* Cout) {
    if (h.passive || h.no_disp) {
      qg::launch_zero_bytes(Aout, h.plane * G->nz * h.fsize(), S);
      if (Cout) qg::launch_zero_bytes(Cout, h.plane * G->nz * h.fsize(), S);
      return;
    }
    K(Cout ? "thomas_AC" : "thomas_A", Cout ? qg::P_THOMAS_AC : qg::P_THOMAS_A, Ps * (Cout ? 3 : 2) * cS + Tq,
      [&] { solve_A(h, *G, S, Bsrc, Aout, Cout); });
  }

  void velocities(const C* psi) {
    K("yb_uv", qg::P_YB_ADV, Ps * cS + Ph * 2 * cR, [&] {
      O::yb_adv(h.tune[qg::P_YB_ADV], *G, S, qg::OUT_U | qg::OUT_V, psi, nullptr, c(h.Pu), c(h.Pv), nullptr, nullptr, twy());
    });
    K("x_c2r_uv", qg::P_X_C2R, Ph * 2 * cR + P * 2 * rR,
      [&] { O::x_c2r(h.tune[qg::P_X_C2R], *G, S, 2, c(h.Pu), c(h.Pv), r(h.u), r(h.v), twx()); });
  }

  // u, v and the QG omega-equation w (compute_velocities!(compute_w=true)).
  // One Y-pass over psi makes the half rows of u and v with the four of the
  // omega right-hand side.  chk = 1: the same passes also make the
  // termination check of psi (non-finite values; max |ifft2(psi)| from psi's
  // own half rows, in Pz until y_herm_w reuses it); chk = 2: the non-finite
  // test and, in place of the maximum, its bound (the largest over the levels
  // of sum |psi's Hermitian part| / (nx ny)) into d_absmax.
  // defer: the w passes run on their own stream, beside whatever the main
  // stream does next (the next step), until join_w().
  void velocities_w(const C* psi, int chk = 0, bool defer = false) {
    C* Ppsi = chk == 1 ? c(h.Pz) : nullptr;
    double* lsum = chk == 2 ? h.d_lsum : nullptr;
    join_w();  // the previous step's w passes read Pw, which xf rewrites
    K("yf_omega", qg::P_YF_OMEGA, Ps * cS + Ph * (chk == 1 ? 7 : 6) * cR, [&] {
      O::yf_omega(h.tune[qg::P_YF_OMEGA], *G, S, psi, c(h.Qa), c(h.Qb), c(h.G1), c(h.G2), c(h.Pu), c(h.Pv),
                  Ppsi, h.d_flag, lsum, twy());
    });
    K("xf", qg::P_XF, Ph * (chk == 1 ? 8 : 7) * cR + P * 2 * rR, [&] {
      O::xf(h.tune[qg::P_XF], *G, S, 1, c(h.Pu), c(h.Pv), c(h.Qa), c(h.Qb), c(h.G1), c(h.G2), r(h.u), r(h.v), c(h.Pw),
            Ppsi, h.d_absmax, lsum, twx());
    });
    if (h.sing_omega[0]) {
      char buf[160];
      snprintf(buf, sizeof buf, "LAPACK-style gtsv failed for vertical velocity at row %d of column %d",
               h.sing_omega[1] + 1, h.sing_omega[2]);
      throw QgError(QG_ERR_GTSV, buf);
    }
    // w from the omega right-hand side: buffers of its own (Pw, Wr, Ww, Pw2,
    // the pivoted solve's scratch), so nothing the next step does touches them
    auto w_passes = [&] {
      K("y_r2c_omega", qg::P_Y_R2C, Ph * cR + Ps * cS, [&] { O::y_r2c_mask(h.tune[qg::P_Y_R2C], *G, S, c(h.Pw), c(h.Wr), twy()); });
      K("omega_gtsv", qg::P_OMEGA, Ps * 2 * cS + Tq, [&] {
        qg::launch_omega(*G, S, c(h.Wr), c(h.Ww), h.f32(), h.om_M, h.om_D, h.om_piv, h.om_any_piv, h.d_cz, h.d_cu,
                         h.d_fN2, h.s0, h.s1, h.s2, h.tune[qg::P_OMEGA]);
      });
      K("y_herm_w", qg::P_Y_HERM, Ps * cS + Ph * cR, [&] { O::y_herm(h.tune[qg::P_Y_HERM], *G, S, c(h.Ww), c(h.Pw2), twy()); });
      K("x_c2r_w", qg::P_X_C2R, Ph * cR + P * rR, [&] { O::x_c2r(h.tune[qg::P_X_C2R], *G, S, 1, c(h.Pw2), nullptr, r(h.w), nullptr, twx()); });
    };
    if (defer && concurrent()) {
      side(w_passes, h.st3);
      h.w_pending = true;
    } else {
      w_passes();
    }
  }

  // One RHS evaluation at (qS, psiS, AS, T1, u, v) fused with the ETD update
  // from the base state (qn, Bn): _compute_etdrk2_rhs! (src/timestep.jl:230-298)
  // and the stage/final loops (src/timestep.jl:362-404).
  // uv: u and v are not yet made from psiS; the Y-pass over psiS and q_S
  // makes their half rows too, and the q-advection X-pass turns them into u
  // and v (written for the B pass).  Needs the advection of q.
  // flag: the new q and B are tested for non-finite values (the step's fused
  // termination check; final update only).
  void rhs_update(int mode, const C* qS, const C* psiS, const C* AS, const C* qn, const C* Bn, C* qout,
                  C* Bout, bool uv = false, int* flag = nullptr) {
    const bool adv = !h.linear;
    const bool refr = !h.passive;
    const bool qdyn = !h.fixed_flow;
    int outs = 0;
    if (adv && qdyn) outs |= qg::OUT_Q;
    if (refr) outs |= qg::OUT_Z;
    if (uv) {
      if (!(adv && qdyn)) throw QgError(QG_ERR_ARGUMENT, "u, v from the RHS passes need the advection of q");
      outs |= qg::OUT_U | qg::OUT_V;
    }
    const int nouts = ((outs & qg::OUT_Q) ? 1 : 0) + ((outs & qg::OUT_Z) ? 1 : 0) + (uv ? 2 : 0);
    if (outs)
      K(uv ? "yb_adv_uv" : "yb_adv", qg::P_YB_ADV,
        // reads q for OUT_Q and psi for the others
        Ps * (((outs & qg::OUT_Q) ? cS : 0) + ((outs & ~qg::OUT_Q) ? cS : 0)) + Ph * nouts * cR, [&] {
          O::yb_adv(h.tune[qg::P_YB_ADV], *G, S, outs, psiS, qS, uv ? c(h.Pu) : nullptr, uv ? c(h.Pv) : nullptr,
                    c(h.Pq), c(h.Pz), twy());
        });
    if (adv && qdyn && refr) {
      // both X-passes in one launch: Pq, Pz, T1 and (u, v | Pu, Pv) in;
      // G1, G2 (half width), G3, G4 out
      K(uv ? "xb_all_uv" : "xb_all", qg::P_XB_B,
        Ph * 4 * cR + Pf * cR + (uv ? Ph * 2 * cR : P * 2 * rR) + Pf * 2 * cR, [&] {
          O::xb_all(h.tune[qg::P_XB_B], *G, S, c(h.Pq), c(h.Pz), c(h.T1), uv ? c(h.Pu) : nullptr,
                    uv ? c(h.Pv) : nullptr, r(h.u), r(h.v), c(h.G1), c(h.G2), c(h.G3), c(h.G4), twx());
        });
    } else if (adv && qdyn) {
      K(uv ? "xb_q_uv" : "xb_q", qg::P_XB_Q, Ph * (uv ? 3 : 1) * cR + Ph * 2 * cR + P * 2 * rR, [&] {
        O::xb_q(h.tune[qg::P_XB_Q], *G, S, c(h.Pq), uv ? c(h.Pu) : nullptr, uv ? c(h.Pv) : nullptr, r(h.u),
                r(h.v), c(h.G1), c(h.G2), twx());
      });
    }
    int xbf = (adv ? qg::XB_ADV : 0) | (refr ? qg::XB_REFR : 0);
    if (xbf && !(adv && qdyn && refr))
      K("xb_B", qg::P_XB_B, Pf * cR + (refr ? Ph * cR : 0) + (adv ? P * 2 * rR : 0) + Pf * cR * (1 + (adv ? 1 : 0)),
        [&] { O::xb_B(h.tune[qg::P_XB_B], *G, S, xbf, c(h.T1), c(h.Pz), r(h.u), r(h.v), c(h.G3), c(h.G4), twx()); });
    const bool diff = h.nu_z != 0.0 && !h.inviscid && h.nz > 1;
    int qf = ((adv && qdyn) ? qg::YQ_ADV : 0) | (h.fixed_flow ? qg::YQ_FIXED : 0) | (diff ? qg::YQ_DIFF : 0);
    const double inv_dz2 = h.nz > 1 ? 1.0 / (h.dz * h.dz) : 0.0;
    // stage: q_n, the kept bracket and q_S; final: the bracket and q_{n+1}
    const double qbytes = ((qf & qg::YQ_ADV) ? Ph * 2 * cR : 0) +
                          Ps * cS * ((mode == qg::MODE_STAGE || h.fixed_flow ? 3 : 2) + (diff ? 1 : 0));
    const bool disp = !(h.passive || h.no_disp);
    int bf = (adv ? qg::YB_ADV : 0) | (refr ? qg::YB_REFR : 0) | (disp ? qg::YB_DISP : 0);
    // the new B's inverse y transforms for the next diagnosis: T1 always
    // (xb_B), T3 when q_w is computed
    C* T3 = (h.feedback && !h.fixed_flow) ? c(h.T3) : nullptr;
    const double bbytes = Pf * cR * (((adv || refr) ? 1 : 0) + (adv ? 1 : 0) + 1 + (T3 ? 1 : 0)) +
                          Ps * cS * ((mode == qg::MODE_STAGE ? 3 : 2) + (disp ? 1 : 0));
    // the q update beside the B update (disjoint inputs and outputs): one
    // launch of both (the evolving flow without vertical diffusion), or the
    // q update on the side stream
    if (concurrent() && qf == qg::YQ_ADV) {
      K(mode == qg::MODE_STAGE ? "yb2_qB_stage" : "yb2_qB_final", qg::P_YB2_B, qbytes + bbytes, [&] {
        O::yb2_qB(h.tune[qg::P_YB2_B], *G, S, mode, c(h.G1), c(h.G2), qn, c(h.rq0), qout, bf, c(h.G3), c(h.G4), AS,
                  Bn, c(h.rB0), Bout, c(h.T1), T3, etd(), h.f / 2, flag, twy());
      });
      return;
    }
    side([&] {
      K(mode == qg::MODE_STAGE ? "yb2_q_stage" : "yb2_q_final", qg::P_YB2_Q, qbytes,
        [&] { O::yb2_q(h.tune[qg::P_YB2_Q], *G, S, mode, qf, c(h.G1), c(h.G2), qS, qn, c(h.rq0), qout, etd(), h.nu_z, inv_dz2, flag, twy()); });
    });
    K(mode == qg::MODE_STAGE ? "yb2_B_stage" : "yb2_B_final", qg::P_YB2_B, bbytes,
      [&] { O::yb2_B(h.tune[qg::P_YB2_B], *G, S, mode, bf, c(h.G3), c(h.G4), AS, Bn, c(h.rB0), Bout, c(h.T1), T3, etd(), h.f / 2, flag, twy()); });
    join();
  }

  // step!(model, ExponentialRungeKutta2) -> _advance_etdrk2! (src/timestep.jl:335-415).
  // chk = 1: the step's last passes also make _check_termination_conditions!
  // (src/simulation.jl:25-44) of the state they leave, accumulated into
  // d_flag and d_absmax: non-finite q and B where the final update writes
  // them, non-finite psi and max |ifft2(psi)| in the finalize's velocity and
  // omega passes, which read psi anyway.  The values are those of check().
  // chk = 2: the same with an upper bound of max |ifft2(psi)| in d_absmax
  // (velocities_w).
  void step(int chk = 0) {
    const bool can = h.prunable && h.prune_enabled, prune = can && h.clean;
    use(prune ? h.gp : h.g);
    C* qn = c(h.q[h.cur]);
    C* Bn = c(h.B[h.cur]);
    C* q1 = c(h.q[1 - h.cur]);
    C* B1 = c(h.B[1 - h.cur]);
    const bool need_uv = !h.linear;
    // RHS at S_n.  psi, A, u, v and T1 left by the previous finalize are the
    // values this evaluation would recompute from (q_n, B_n).
    if (!h.diag_valid) {
      diag(qn, Bn, c(h.psi), true);
      wave_inv(Bn, c(h.A), nullptr);
      if (need_uv) velocities(c(h.psi));
    } else if (!h.t1_valid) {
      t1_only(Bn);
    }
    rhs_update(qg::MODE_STAGE, qn, c(h.psi), c(h.A), qn, Bn, c(h.qs), c(h.Bs));
    // RHS at the stage (yb2_B left T1, T3 of B_S)
    C* psiS = h.fixed_flow ? c(h.psi) : c(h.psis);
    // A_S = L+^-1 B_S beside the q_w -> q_eff -> psi_S chain (independent)
    diag(c(h.qs), c(h.Bs), psiS, true, true, c(h.Bs), c(h.As));
    // the stage's u, v come out of the RHS passes (they are needed nowhere else)
    rhs_update(qg::MODE_FINAL, c(h.qs), psiS, c(h.As), qn, Bn, q1, B1, need_uv && !h.fixed_flow,
               chk ? h.d_flag : nullptr);
    // finalize (src/timestep.jl:300-332); yb2_B left T1, T3 of B_{n+1}.  A
    // and C beside the diagnosis of psi, which reads neither.
    diag(q1, B1, c(h.psi), false, true, B1, c(h.A), c(h.C));
    velocities_w(c(h.psi), chk, true);
    if (can && !prune) {
      // The step's outputs (q, B, psi, A, C) are masked and written at every
      // point; the buffers it read from are the next step's outputs, which a
      // pruned step writes inside the disk only: zero their outside.
      qg::launch_zero_outside_disk(h.gp, S, qn, h.f32());
      qg::launch_zero_outside_disk(h.gp, S, Bn, h.f32());
      h.clean = true;
    }
    h.cur = 1 - h.cur;
    h.diag_valid = true;
    h.t1_valid = true;
    use(h.g);
  }

  void compute_qw(C* out) {
    O::ya_qw(h.tune[qg::P_YA_QW], h.g, S, c(h.B[h.cur]), c(h.T1), c(h.T3), twy());
    O::xa_qw(h.tune[qg::P_XA_QW], h.g, S, c(h.T1), c(h.T3), c(h.Qa), c(h.Qb), twx());
    O::ya2_qeff(h.tune[qg::P_YA2_QEFF], h.g, S, c(h.Qa), c(h.Qb), c(h.q[h.cur]), nullptr, out, 1.0 / h.f, twy());
    h.t1_valid = false;
  }

  void compute_velocities(bool compute_w) {
    if (compute_w) {
      velocities_w(c(h.psi));
    } else {
      velocities(c(h.psi));
      qg::launch_zero_bytes(h.w, h.plane * h.g.nz * h.rsize(), S);
    }
  }
};

template <typename F>
int guarded(qg_handle* h, F&& fn) {
  if (!h) return QG_ERR_ARGUMENT;
  try {
    fn();
    h->err.clear();
    return QG_OK;
  } catch (const QgError& e) {
    h->err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    h->err = e.what();
    return QG_ERR_UNSUPPORTED;
  }
}

template <typename Fn>
void with_engine(qg_handle& h, Fn&& fn) {
  if (h.precision == QG_PRECISION_MIXED) {
    Engine<float> e(h);
    fn(e);
  } else {
    Engine<double> e(h);
    fn(e);
  }
}

void* complex_field(qg_handle& h, int field) {
  switch (field) {
    case QG_FIELD_Q: return h.q[h.cur];
    case QG_FIELD_B: return h.B[h.cur];
    case QG_FIELD_PSI: return h.psi;
    case QG_FIELD_A: return h.A;
    case QG_FIELD_C: return h.C;
    default: return nullptr;
  }
}
void* real_field(qg_handle& h, int field) {
  switch (field) {
    case QG_FIELD_U: return h.u;
    case QG_FIELD_V: return h.v;
    case QG_FIELD_W: return h.w;
    default: return nullptr;
  }
}

// planar double field -> device planar field in place: forward FFT.
// 2-D transforms of `levels` planes starting at f (all stacked levels by
// default).
void fft2_inplace(qg_handle& h, D* f, int dir, D* out = nullptr, int levels = -1) {
  qg::Geo g = h.g;
  if (levels >= 0) g.nz = levels;
  // y first for the inverse (spectral -> physical), x first for the forward
  if (dir < 0) {
    qg::Ops<double>::x_c2c(1, g, h.st, -1, f, f, h.tw64x);
    qg::Ops<double>::y_c2c(1, g, h.st, -1, f, out ? out : f, h.tw64y);
  } else {
    D* o = out ? out : f;
    qg::Ops<double>::y_c2c(1, g, h.st, +1, f, o, h.tw64y);
    qg::Ops<double>::x_c2c(1, g, h.st, +1, o, o, h.tw64x);
  }
}

}  // namespace

// Steps per long graph (even, so that a graph leaves the double buffer as it
// found it).  Each graph launch costs a few microseconds of idle GPU, and the
// last step of a graph waits for its w passes instead of overlapping them
// with the next step.
constexpr int kLongGraphSteps = 16;

// `nsteps` steps of the engine: steady-state runs of steps (diagnostics
// valid, no per-kernel events) replay captured graphs of kLongGraphSteps or
// two steps, the others run eagerly.  With chk = 1 every step is followed by
// the termination check of the active members, inside the graph too; chk = 2
// by its bound (Engine::step).  mark(n) follows every n steps.
template <class E, class Mark>
void advance(qg_handle& H, E& e, int nsteps, int chk, Mark&& mark) {
  const bool use_graphs = H.graphs && !(H.prof_mode & 1);
  auto one = [&] { e.step(chk); };  // the check is fused into the step's last passes
  int i = 0;
  while (i < nsteps) {
    // an unclean state's step runs unpruned and eagerly (it makes the state clean)
    if (use_graphs && H.diag_valid && (H.clean || !(H.prunable && H.prune_enabled)) && nsteps - i >= 2) {
      const int c = H.cur;
      const int gs = nsteps - i >= kLongGraphSteps ? kLongGraphSteps : 2;
      hipGraphExec_t* set = gs == 2 ? (chk == 2 ? H.graph_bnd : chk == 1 ? H.graph_chk : H.graph)
                                    : (chk == 2 ? H.graph_bnd_l : chk == 1 ? H.graph_chk_l : H.graph_l);
      hipGraphExec_t& gx = set[c];
      // w passes left running by an eager step finish before the graph's
      // steps reuse their buffers
      e.join_w();
      if (!gx) {
        hipGraph_t g = nullptr;
        HIPCHECK(hipStreamBeginCapture(H.st, hipStreamCaptureModeThreadLocal));
        try {
          for (int s = 0; s < gs; ++s) one();
          e.join_w();  // every branch rejoins the captured stream
        } catch (...) {
          H.w_pending = false;
          (void)hipStreamEndCapture(H.st, &g);
          if (g) (void)hipGraphDestroy(g);
          throw;
        }
        HIPCHECK(hipStreamEndCapture(H.st, &g));
        const hipError_t ie = hipGraphInstantiate(&gx, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ie != hipSuccess) {
          gx = nullptr;
          throw QgError(QG_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
        }
      }
      HIPCHECK(hipGraphLaunch(gx, H.st));
      i += gs;
      mark(gs);
    } else {
      one();
      ++i;
      mark(1);
    }
  }
  e.join_w();  // the caller may read w or synchronize the main stream
}

// Steps per qg_run chunk: one host round trip per chunk.
constexpr int kRunChunk = 32;

// Snapshot space of qg_run_noalloc: the state a chunk starts from (12 scalars
// of the storage precision per point of the active members), held in a
// file-scope device array so the entry point allocates nothing.  Every GPU the
// library runs on holds one copy, at its own address, until the process ends;
// calls that overlap on one GPU share it and race.
constexpr size_t kNoAllocSnapBytes = QG_RUN_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kNoAllocSnapBytes];

namespace {
std::mutex g_scratch_mu;
std::map<int, void*> g_scratch_addr;  // device id -> address of `scratch` there

// Address of `scratch` on the current device (looked up once per device).
// In host code the array's name is only a stand-in; this address is what
// kernels and copies use.
void* device_scratch() {
  int dev = -1;
  HIPCHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  auto it = g_scratch_addr.find(dev);
  if (it != g_scratch_addr.end()) return it->second;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    throw QgError(QG_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch_addr[dev] = ptr;
  return ptr;
}
}  // namespace


// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

const char* qg_last_error(const qg_handle* h) { return h ? h->err.c_str() : g_create_error.c_str(); }

int qg_create(const qg_config* cfg, qg_handle** out) {
  g_create_error.clear();
  if (!cfg || !out) {
    g_create_error = "qg_create: null argument";
    return QG_ERR_ARGUMENT;
  }
  *out = nullptr;
  qg_handle* h = nullptr;
  try {
    if (cfg->nx <= 0 || cfg->ny <= 0 || cfg->nz <= 0)
      throw QgError(QG_ERR_ARGUMENT, "grid dimensions must be positive");
    if (!qg::fft_size_supported(cfg->nx) || !qg::fft_size_supported(cfg->ny))
      throw QgError(QG_ERR_UNSUPPORTED,
                    "horizontal sizes must be one of " + qg::fft_sizes_text() + " (got " +
                        std::to_string(cfg->nx) + "x" + std::to_string(cfg->ny) + ")");
    if (!qg::column_depth_supported(cfg->nz))
      throw QgError(QG_ERR_UNSUPPORTED, "nz must be between 1 and 1024 for the column solvers");
    const int members = cfg->members > 0 ? cfg->members : 1;
    if ((long long)members * cfg->nz > (1 << 20))
      throw QgError(QG_ERR_ARGUMENT, "members * nz must not exceed 2^20 stacked levels");
    if (cfg->formulation == QG_YBJ)
      throw QgError(QG_ERR_UNSUPPORTED, "the normal YBJ formulation is not part of this library; use YBJ+");
    if (cfg->formulation != QG_YBJ_PLUS && cfg->formulation != QG_PASSIVE)
      throw QgError(QG_ERR_ARGUMENT, "unknown wave formulation");
    if (cfg->precision != QG_PRECISION_FP64 && cfg->precision != QG_PRECISION_MIXED)
      throw QgError(QG_ERR_ARGUMENT, "unknown precision mode");
    if (!cfg->kx || !cfg->ky || !cfg->N2 || !cfg->a_ell)
      throw QgError(QG_ERR_ARGUMENT, "kx, ky, N2 and a_ell are required");
    if (!(std::isfinite(cfg->dz) && cfg->dz > 0)) throw QgError(QG_ERR_ARGUMENT, "dz must be finite and positive");
    if (!(std::isfinite(cfg->nu_z) && cfg->nu_z >= 0))
      throw QgError(QG_ERR_ARGUMENT, "vertical diffusivity must be finite and non-negative");
    const bool feedback = cfg->wave_feedback && !cfg->fixed_flow;
    if (feedback && !(std::isfinite(cfg->f) && cfg->f != 0))
      throw QgError(QG_ERR_ARGUMENT, "wave feedback requires a finite, nonzero Coriolis frequency");
    if (cfg->n_flow_terms > qg::kEtdTerms || cfg->n_wave_terms > qg::kEtdTerms)
      throw QgError(QG_ERR_UNSUPPORTED, "at most " + std::to_string(qg::kEtdTerms) +
                                            " hyperdiffusivity terms per closure are supported");
    for (int i = 0; i < cfg->n_flow_terms; ++i)
      if (cfg->flow_orders[i] <= 0 || cfg->flow_orders[i] % 2)
        throw QgError(QG_ERR_ARGUMENT, "flow hyperdiffusivity orders must be positive and even");
    for (int i = 0; i < cfg->n_wave_terms; ++i)
      if (cfg->wave_orders[i] <= 0 || cfg->wave_orders[i] % 2)
        throw QgError(QG_ERR_ARGUMENT, "wave hyperdiffusivity orders must be positive and even");

    h = new qg_handle();
    h->nx = cfg->nx;
    h->ny = cfg->ny;
    h->nzc = cfg->nz;
    h->members = members;
    h->active = members;
    h->nz = members * cfg->nz;
    h->precision = cfg->precision;
    h->f = cfg->f;
    h->dz = cfg->nz > 1 ? cfg->dz : 1.0;
    h->nu_z = cfg->nu_z;
    h->fixed_flow = cfg->fixed_flow != 0;
    h->feedback = feedback;
    h->passive = cfg->formulation == QG_PASSIVE;
    h->inviscid = cfg->inviscid != 0;
    h->linear = cfg->linear != 0;
    h->no_disp = cfg->no_dispersion != 0;
    h->kx.assign(cfg->kx, cfg->kx + cfg->nx);
    h->ky.assign(cfg->ky, cfg->ky + cfg->ny);
    h->N2.assign(cfg->N2, cfg->N2 + h->nz);  // per member, concatenated
    h->a.assign(cfg->a_ell, cfg->a_ell + h->nz);
    if (cfg->n_flow_terms > 0) {
      h->fc.assign(cfg->flow_coefficients, cfg->flow_coefficients + cfg->n_flow_terms);
      h->fo.assign(cfg->flow_orders, cfg->flow_orders + cfg->n_flow_terms);
    }
    if (cfg->n_wave_terms > 0) {
      h->wc.assign(cfg->wave_coefficients, cfg->wave_coefficients + cfg->n_wave_terms);
      h->wo.assign(cfg->wave_orders, cfg->wave_orders + cfg->n_wave_terms);
    }
    HIPCHECK(hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking));
    HIPCHECK(hipStreamCreateWithFlags(&h->st2, hipStreamNonBlocking));
    HIPCHECK(hipStreamCreateWithFlags(&h->st3, hipStreamNonBlocking));
    for (hipEvent_t& e : h->sync_ev) HIPCHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    const int nx = h->nx, ny = h->ny, nz = h->nz;
    h->plane = (size_t)nx * ny;
    h->npts = h->plane * nz;
    const int W = nx / 2 + 1;
    const int Wp = (W + 15) / 16 * 16;
    h->nhalf = (size_t)Wp * ny * nz;

    h->d_kx = h->alloc<double>(nx);
    h->d_ky = h->alloc<double>(ny);
    h->d_a = h->alloc<double>(nz);
    h->d_cz = h->alloc<double>(nz);
    h->d_fN2 = h->alloc<double>(nz);
    h->d_prof = h->alloc<double>(nz);
    h->d_mask = h->alloc<unsigned char>(h->plane);
    h->d_etd = h->alloc<qg::EtdParams>(1);
    std::vector<unsigned char> mask(h->plane);
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) mask[(size_t)y * nx + x] = is_dealiased(x, y, nx, ny) ? 1 : 0;
    std::vector<double> cz(nz), fN2(nz);
    const double f2 = h->f * h->f;
    for (int k = 0; k < nz; ++k) {
      cz[k] = (f2 / h->N2[k]) / (h->dz * h->dz);  // src/operators.jl:428
      fN2[k] = h->f / h->N2[k];                   // src/operators.jl:439
    }
    HIPCHECK(hipMemcpy(h->d_kx, h->kx.data(), nx * sizeof(double), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(h->d_ky, h->ky.data(), ny * sizeof(double), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(h->d_a, h->a.data(), nz * sizeof(double), hipMemcpyHostToDevice));
    const int nzc = h->nzc, kpad = qg::thomas_shifted_a_size(nzc);
    {
      std::vector<double> al((size_t)members * kpad, 0.0);
      for (int m = 0; m < members; ++m)
        for (int k = 1; k < nzc; ++k) al[(size_t)m * kpad + k] = h->a[(size_t)m * nzc + k - 1];
      h->d_al = h->alloc<double>(al.size());
      HIPCHECK(hipMemcpy(h->d_al, al.data(), al.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    HIPCHECK(hipMemcpy(h->d_cz, cz.data(), nz * sizeof(double), hipMemcpyHostToDevice));
    {
      // upper diagonal of the omega interior row k-1, by level
      std::vector<double> cu((size_t)members * kpad, 0.0);
      for (int m = 0; m < members; ++m)
        for (int k = 1; k <= nzc - 3; ++k) cu[(size_t)m * kpad + k] = cz[(size_t)m * nzc + k];
      h->d_cu = h->alloc<double>(cu.size());
      HIPCHECK(hipMemcpy(h->d_cu, cu.data(), cu.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    HIPCHECK(hipMemcpy(h->d_fN2, fN2.data(), nz * sizeof(double), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(h->d_mask, mask.data(), h->plane, hipMemcpyHostToDevice));

    for (int i = 0; i < 2; ++i) {
      h->q[i] = h->falloc(h->npts);
      h->B[i] = h->falloc(h->npts);
    }
    for (void** p : {&h->psi, &h->A, &h->C, &h->qs, &h->Bs, &h->psis, &h->As, &h->rq0, &h->rB0, &h->qeff})
      *p = h->falloc(h->npts);
    h->stg = h->alloc<D>(h->npts);
    h->stg2 = h->f32() ? h->alloc<D>(h->npts) : static_cast<D*>(h->qeff);
    h->s0 = h->alloc<double>(h->npts);
    h->s1 = h->alloc<double>(h->npts);
    h->s2 = h->alloc<double>(h->npts);
    h->u = h->ralloc(h->npts);
    h->v = h->ralloc(h->npts);
    h->w = h->ralloc(h->npts);
    for (void** p : {&h->T1, &h->T3, &h->G3, &h->G4}) *p = h->ralloc(2 * h->npts);
    for (void** p : {&h->Qa, &h->Qb, &h->Pu, &h->Pv, &h->Pq, &h->Pz, &h->G1, &h->G2, &h->Pw, &h->Pw2})
      *p = h->ralloc(2 * h->nhalf);
    h->Wr = h->falloc(h->npts);
    h->Ww = h->falloc(h->npts);
    h->d_err = h->alloc<int>(4);
    h->d_ferr = h->alloc<int>(8);
    h->d_warn = h->alloc<unsigned long long>(1);
    h->d_absmax = h->alloc<unsigned long long>(1);
    h->d_pmax = h->alloc<unsigned long long>(2);
    h->d_flag = h->alloc<int>(1);
    h->d_lsum = h->alloc<double>(h->nz);
    HIPCHECK(hipMemset(h->d_lsum, 0, h->nz * sizeof(double)));
    HIPCHECK(hipMemset(h->d_err, 0, 4 * sizeof(int)));
    HIPCHECK(hipMemset(h->d_warn, 0, sizeof(unsigned long long)));
    for (void* p : {h->q[0], h->q[1], h->B[0], h->B[1], h->psi, h->A, h->C})
      HIPCHECK(hipMemset(p, 0, h->npts * h->fsize()));
    for (void* p : {h->u, h->v, h->w}) HIPCHECK(hipMemset(p, 0, h->npts * h->rsize()));

    h->tw64x = h->alloc<D>(nx);
    h->tw64y = h->alloc<D>(ny);
    {
      auto tx = twiddles<double>(nx), ty = twiddles<double>(ny);
      HIPCHECK(hipMemcpy(h->tw64x, tx.data(), nx * sizeof(D), hipMemcpyHostToDevice));
      HIPCHECK(hipMemcpy(h->tw64y, ty.data(), ny * sizeof(D), hipMemcpyHostToDevice));
    }
    if (h->precision == QG_PRECISION_MIXED) {
      auto tx = twiddles<float>(nx), ty = twiddles<float>(ny);
      h->twx = h->alloc<cplx<float>>(nx);
      h->twy = h->alloc<cplx<float>>(ny);
      HIPCHECK(hipMemcpy(h->twx, tx.data(), nx * sizeof(cplx<float>), hipMemcpyHostToDevice));
      HIPCHECK(hipMemcpy(h->twy, ty.data(), ny * sizeof(cplx<float>), hipMemcpyHostToDevice));
    } else {
      h->twx = h->tw64x;
      h->twy = h->tw64y;
    }

    qg::Geo& g = h->g;
    g.nx = nx;
    g.ny = ny;
    g.nz = nz;
    g.W = W;
    g.Wp = Wp;
    g.nzc = h->nzc;
    g.members = h->members;
    g.kx = h->d_kx;
    g.ky = h->d_ky;
    g.mask = h->d_mask;
    g.inv_n = 1.0 / ((double)nx * (double)ny);
    g.dz = h->dz;
    g.Kx = W;
    g.kd2 = -1;
    g.km2 = (std::min(nx, ny) / 3) * (std::min(nx, ny) / 3);  // is_dealiased (src/physics.jl:36-43)
    {
      const int kmax = std::min(nx, ny) / 3;
      h->gp = g;
      h->gp.Kx = kmax + 1;
      h->gp.kd2 = kmax * kmax;
      h->prunable = !h->fixed_flow && kmax + 1 < W;
      for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) h->disk_plane += mask[(size_t)y * nx + x];
      for (int y = 0; y <= ny / 2; ++y)
        for (int x = 0; x <= nx / 2; ++x) h->disk_quarter += (x * x + y * y <= kmax * kmax) ? 1 : 0;
    }
    // Thomas factors, computed once: they depend only on (kh2, a, dz)
    if (!qg::thomas_setup() || !qg::omega_setup())
      throw QgError(QG_ERR_DEVICE, "cannot reserve the LDS of the column solvers");
    {
      const size_t tsz = qg::thomas_table_size(g);
      for (int o = 0; o < 2; ++o) {
        h->th_inv[o] = h->alloc<double>(tsz);
        h->th_cp[o] = h->alloc<double>(tsz);
      }
      int* e = h->alloc<int>(8);
      // the error words are cleared on the stream the factor kernels run on:
      // a plain hipMemset (null stream) is not ordered before work on this
      // non-blocking stream and could clear an error they recorded
      HIPCHECK(hipMemsetAsync(e, 0, 8 * sizeof(int), h->st));
      qg::launch_thomas_factor(g, h->st, h->d_a, 1, h->th_inv[0], h->th_cp[0], e);
      qg::launch_thomas_factor(g, h->st, h->d_a, 4, h->th_inv[1], h->th_cp[1], e + 4);
      int he[8];
      HIPCHECK(hipMemcpyAsync(he, e, sizeof he, hipMemcpyDeviceToHost, h->st));
      HIPCHECK(hipStreamSynchronize(h->st));
      std::memcpy(h->sing_psi, he, 3 * sizeof(int));
      std::memcpy(h->sing_A, he + 4, 3 * sizeof(int));
      h->om_M = h->alloc<double>(tsz);
      h->om_D = h->alloc<double>(tsz);
      h->om_piv = h->alloc<unsigned char>(tsz / nzc);
      HIPCHECK(hipMemsetAsync(e, 0, 8 * sizeof(int), h->st));
      qg::launch_omega_factor(g, h->st, h->d_cz, h->om_M, h->om_D, h->om_piv, e);
      HIPCHECK(hipMemcpyAsync(he, e, sizeof he, hipMemcpyDeviceToHost, h->st));
      HIPCHECK(hipStreamSynchronize(h->st));
      std::memcpy(h->sing_omega, he, 3 * sizeof(int));
      h->om_any_piv = he[3] != 0;
    }
    h->noalloc_snap = device_scratch();
    HIPCHECK(hipDeviceSynchronize());
    *out = h;
    return QG_OK;
  } catch (const QgError& e) {
    g_create_error = e.what();
    delete h;
    return e.code;
  } catch (const std::exception& e) {
    g_create_error = e.what();
    delete h;
    return QG_ERR_DEVICE;
  }
}

int qg_memory_required(const qg_config* cfg, size_t* bytes) {
  if (!cfg || !bytes || cfg->nx <= 0 || cfg->ny <= 0 || cfg->nz <= 0) return QG_ERR_ARGUMENT;
  const size_t members = cfg->members > 0 ? (size_t)cfg->members : 1;
  const size_t nx = cfg->nx, ny = cfg->ny, nz = members * cfg->nz, plane = nx * ny, npts = plane * nz;
  const size_t nhalf = ((nx / 2 + 1 + 15) / 16 * 16) * ny * nz;
  const size_t r = cfg->precision == QG_PRECISION_MIXED ? 4 : 8;
  const size_t quarter = (nx / 2 + 1) * (ny / 2 + 1);
  const size_t kpad = (size_t)qg::thomas_shifted_a_size(cfg->nz);
  size_t b = 16 * 2 * r * npts    // complex fields and stage/work buffers (storage precision)
             + (r == 4 ? 2 : 1) * 16 * npts  // FP64 staging (the second aliases qeff in FP64 mode)
             + 3 * 8 * npts       // omega scratch
             + 3 * r * npts       // u, v, w
             + 4 * 2 * r * npts   // full-width transform intermediates
             + 10 * 2 * r * nhalf  // half-width transform intermediates
             + 6 * 8 * quarter * nz + quarter * members  // Thomas and omega tables
             + plane + 4 * 16 * (nx + ny) + (4 * 8 * nz) + 2 * 8 * members * kpad + sizeof(qg::EtdParams);
  b += 64 * 256;  // allocation padding
  *bytes = b;
  return QG_OK;
}

int qg_device_memory(size_t* free_bytes, size_t* total_bytes) {
  size_t f = 0, t = 0;
  if (hipMemGetInfo(&f, &t) != hipSuccess) return QG_ERR_DEVICE;
  if (free_bytes) *free_bytes = f;
  if (total_bytes) *total_bytes = t;
  return QG_OK;
}

int qg_destroy(qg_handle* h) {
  delete h;
  return QG_OK;
}

}  // extern "C"

namespace {

// n values of a complex field from value `off` on, as FP64: the field itself
// in FP64 mode, else converted into the FP64 staging stg2.
D* as_fp64(qg_handle& H, void* field, size_t off, size_t n) {
  if (!H.f32()) return static_cast<D*>(field) + off;
  H.psi64_nm = 0;
  qg::launch_convert<float, double>(static_cast<const float*>(H.at(field, off)), reinterpret_cast<double*>(H.stg2),
                                    2 * n, H.st);
  return H.stg2;
}

// Field exchange for members [m0, m0 + nm): a (z, x, y) array of nm * nzc
// levels (member-major along z).
void set_field_range(qg_handle& H, int field, const double* host, int space, int m0, int nm) {
  if (!host) throw QgError(QG_ERR_ARGUMENT, "null host buffer");
  const int lv = nm * H.nzc;
  const size_t off = (size_t)m0 * H.nzc * H.plane, n = (size_t)lv * H.plane;
  if (void* dst = complex_field(H, field)) {
    if (space != QG_SPACE_PHYSICAL && space != QG_SPACE_SPECTRAL)
      throw QgError(QG_ERR_ARGUMENT, "space must be spectral or physical");
    // FP64 planar values (transformed in FP64), then stored at the field's
    // precision
    D* pl = H.f32() ? H.stg2 + off : static_cast<D*>(dst) + off;
    if (H.f32()) H.psi64_nm = 0;
    HIPCHECK(hipMemcpyAsync(H.stg, host, n * sizeof(D), hipMemcpyHostToDevice, H.st));
    qg::launch_transpose<D>(H.stg, pl, H.nx, H.ny, lv, true, H.st);
    if (space == QG_SPACE_PHYSICAL) fft2_inplace(H, pl, -1, nullptr, lv);
    if (H.f32()) {
      qg::launch_convert<double, float>(reinterpret_cast<const double*>(pl),
                                        static_cast<float*>(H.at(dst, off)), 2 * n, H.st);
      if (field == QG_FIELD_PSI) {
        H.psi64_m0 = m0;
        H.psi64_nm = nm;
      }
    }
  } else if (void* rdst = real_field(H, field)) {
    double* tmp = reinterpret_cast<double*>(H.stg);
    double* pl = reinterpret_cast<double*>(H.stg2);
    H.psi64_nm = 0;
    HIPCHECK(hipMemcpyAsync(tmp, host, n * sizeof(double), hipMemcpyHostToDevice, H.st));
    if (H.precision == QG_PRECISION_MIXED) {
      qg::launch_transpose<double>(tmp, pl, H.nx, H.ny, lv, true, H.st);
      qg::launch_convert<double, float>(pl, static_cast<float*>(rdst) + off, n, H.st);
    } else {
      qg::launch_transpose<double>(tmp, static_cast<double*>(rdst) + off, H.nx, H.ny, lv, true, H.st);
    }
  } else {
    throw QgError(QG_ERR_ARGUMENT, "unknown field id " + std::to_string(field));
  }
  HIPCHECK(hipStreamSynchronize(H.st));
  H.diag_valid = false;
  H.clean = false;
  H.t1_valid = false;
}

void get_field_range(qg_handle& H, int field, double* host, int m0, int nm) {
  if (!host) throw QgError(QG_ERR_ARGUMENT, "null host buffer");
  const int lv = nm * H.nzc;
  const size_t off = (size_t)m0 * H.nzc * H.plane, n = (size_t)lv * H.plane;
  if (void* src = complex_field(H, field)) {
    qg::launch_transpose<D>(as_fp64(H, src, off, n), H.stg, H.nx, H.ny, lv, false, H.st);
    HIPCHECK(hipMemcpyAsync(host, H.stg, n * sizeof(D), hipMemcpyDeviceToHost, H.st));
  } else if (void* rsrc = real_field(H, field)) {
    double* tmp = reinterpret_cast<double*>(H.stg);
    const double* pl = static_cast<const double*>(rsrc) + off;
    if (H.precision == QG_PRECISION_MIXED) {
      double* conv = reinterpret_cast<double*>(H.stg2);
      H.psi64_nm = 0;
      qg::launch_convert<float, double>(static_cast<const float*>(rsrc) + off, conv, n, H.st);
      pl = conv;
    }
    qg::launch_transpose<double>(pl, tmp, H.nx, H.ny, lv, false, H.st);
    HIPCHECK(hipMemcpyAsync(host, tmp, n * sizeof(double), hipMemcpyDeviceToHost, H.st));
  } else {
    throw QgError(QG_ERR_ARGUMENT, "unknown field id " + std::to_string(field));
  }
  HIPCHECK(hipStreamSynchronize(H.st));
}

// Reads (and waits for) the check accumulators.
void read_check(qg_handle& H, int& flag, double& m) {
  unsigned long long bits = 0;
  HIPCHECK(hipMemcpyAsync(&flag, H.d_flag, sizeof(int), hipMemcpyDeviceToHost, H.st));
  HIPCHECK(hipMemcpyAsync(&bits, H.d_absmax, sizeof(bits), hipMemcpyDeviceToHost, H.st));
  HIPCHECK(hipStreamSynchronize(H.st));
  std::memcpy(&m, &bits, sizeof m);
}

void check_range(qg_handle& H, int m0, int nm, int* nonfinite, double* max_abs_psi) {
  HIPCHECK(hipMemsetAsync(H.d_flag, 0, sizeof(int), H.st));
  HIPCHECK(hipMemsetAsync(H.d_absmax, 0, sizeof(unsigned long long), H.st));
  with_engine(H, [&](auto& e) { e.check(m0, nm); });
  int flag = 0;
  double m = 0.0;
  read_check(H, flag, m);
  if (nonfinite) *nonfinite = flag;
  if (max_abs_psi) *max_abs_psi = m;
}

// The bound a chunk's steps must stay under to pass without the exact check:
// 1e10 less a margin that covers the transform's rounding of the exact
// maximum (relative 1e-6 at FP32).
constexpr double kBoundPass = 1e10 * (1.0 - 1e-4);

// run!'s stepping loop with the check on the device, shared by qg_run and
// qg_run_noalloc: `snap` holds 12 R scalars per point of the active members;
// graphs = false launches every step eagerly (no graph is instantiated).
// A chunk's steps are checked against the bound of max |ifft2(psi)| (the
// non-finite tests in full): when every step's bound is under kBoundPass
// the exact maximum is too, and the chunk passes.  Otherwise the chunk is
// stepped again from its snapshot with the exact check after every step, up
// to the first failing step, so the stop, the failure kind and the state at
// the stop are those of the exact check.  After a run without failure the
// state it leaves is checked once more, exactly, for max_abs_psi.
void run_checked(qg_handle& H, double dt, int nsteps, void* snap, bool graphs, int& d, int& fail,
                 double& mx) {
  d = 0;
  fail = 0;
  mx = 0.0;
  if (H.active <= 0 || nsteps <= 0) return;
  H.psi64_nm = 0;
  build_etd(H, dt);
  // everything the next step reads: q, B (current buffers), and the
  // diagnostics a steady-state step reuses (psi, A, T1, u, v); R scalars
  // per point of each
  const size_t pts = (size_t)H.active * H.nzc * H.plane, r = H.rsize();
  struct Part {
    void* field;
    size_t per_point;  // R scalars
  };
  auto parts = [&](int c) {
    return std::array<Part, 7>{Part{H.q[c], 2}, Part{H.B[c], 2}, Part{H.psi, 2}, Part{H.A, 2},
                               Part{H.T1, 2}, Part{H.u, 1}, Part{H.v, 1}};
  };
  auto copy_parts = [&](int c, bool save) {
    char* at = static_cast<char*>(snap);
    for (const Part& p : parts(c)) {
      const size_t bytes = pts * p.per_point * r;
      HIPCHECK(hipMemcpyAsync(save ? at : p.field, save ? p.field : at, bytes, hipMemcpyDeviceToDevice, H.st));
      at += bytes;
    }
  };
  auto clear = [&] {
    HIPCHECK(hipMemsetAsync(H.d_flag, 0, sizeof(int), H.st));
    HIPCHECK(hipMemsetAsync(H.d_absmax, 0, sizeof(unsigned long long), H.st));
    HIPCHECK(hipMemsetAsync(H.d_lsum, 0, H.nz * sizeof(double), H.st));
  };
  auto none = [](int) {};
  const bool graphs_were = H.graphs;
  H.graphs = graphs_were && graphs;
  struct Restore {
    qg_handle& H;
    bool g;
    ~Restore() { H.graphs = g; }
  } restore{H, graphs_were};
  with_engine(H, [&](auto& e) {
    while (d < nsteps && fail == 0) {
      const int k = std::min(kRunChunk, nsteps - d);
      const int c0 = H.cur;
      const bool dv = H.diag_valid, tv = H.t1_valid, cl = H.clean;
      copy_parts(c0, true);
      clear();
      advance(H, e, k, 2, none);
      int flag = 0;
      double m = 0.0;
      read_check(H, flag, m);
      check_status(H);
      if (flag == 0 && m <= kBoundPass) {
        d += k;
        continue;
      }
      // A step of this chunk failed or its bound is too high: restore the
      // chunk's starting state and step again, one step and one exact check
      // at a time, stopping right after a failing step.  The same kernels on
      // the same state give the same values, so the state at the stop is the
      // one the failing step produced.
      copy_parts(c0, false);
      H.cur = c0;
      H.diag_valid = dv;
      H.t1_valid = tv;
      H.clean = cl;
      for (int j = 0; j < k && fail == 0; ++j) {
        clear();
        advance(H, e, 1, 1, none);
        read_check(H, flag, m);
        check_status(H);
        ++d;
        if (flag != 0 || !(m <= 1e10)) {
          fail = flag != 0 ? 1 : 2;
          mx = m;
        }
      }
    }
    if (fail == 0 && d > 0) {  // the exact maximum of the state left
      clear();
      e.check(0, H.active);
      int flag = 0;
      read_check(H, flag, mx);
    }
  });
  HIPCHECK(hipGetLastError());
}

// One member seen as a single model: its Geo and the offsets of its levels,
// its Thomas tables and its shifted coefficient array.
struct MemberView {
  qg::Geo g;
  size_t off, toff, aoff;
  MemberView(const qg_handle& H, int m) : g(H.g) {
    g.nz = H.nzc;
    g.members = 1;
    off = (size_t)m * H.nzc * H.plane;
    toff = (size_t)m * qg::thomas_table_size(g);
    aoff = (size_t)m * qg::thomas_shifted_a_size(H.nzc);
  }
};

void check_member_index(const qg_handle& H, int member) {
  if (member < 0 || member >= H.members)
    throw QgError(QG_ERR_ARGUMENT, "member " + std::to_string(member) + " out of range [0, " +
                                       std::to_string(H.members) + ")");
}

}  // namespace

extern "C" {

int qg_set_field(qg_handle* h, int field, const double* host, int space) {
  return guarded(h, [&] { set_field_range(*h, field, host, space, 0, h->members); });
}

int qg_get_field(qg_handle* h, int field, double* host) {
  return guarded(h, [&] { get_field_range(*h, field, host, 0, h->members); });
}

int qg_set_field_member(qg_handle* h, int member, int field, const double* host, int space) {
  return guarded(h, [&] {
    check_member_index(*h, member);
    set_field_range(*h, field, host, space, member, 1);
  });
}

int qg_get_field_member(qg_handle* h, int member, int field, double* host) {
  return guarded(h, [&] {
    check_member_index(*h, member);
    get_field_range(*h, field, host, member, 1);
  });
}

int qg_step(qg_handle* h, double dt, int nsteps) {
  return guarded(h, [&] {
    if (!(std::isfinite(dt) && dt > 0)) throw QgError(QG_ERR_ARGUMENT, "Δt must be finite and positive");
    if (nsteps < 0) throw QgError(QG_ERR_ARGUMENT, "nsteps must be non-negative");
    qg_handle& H = *h;
    if (H.active == 0) return;
    H.psi64_nm = 0;    // psi is rediagnosed from q
    build_etd(H, dt);  // rewrites the ETD tables in place: captured graphs stay valid
    auto mark = [&](int weight) {
      if (!(H.prof_mode & 2)) return;
      if (H.step_events.empty()) {
        hipEvent_t ev = H.event();
        HIPCHECK(hipEventRecord(ev, H.st));
        H.step_events.push_back(ev);
        if (weight == 0) return;
      }
      hipEvent_t ev = H.event();
      HIPCHECK(hipEventRecord(ev, H.st));
      H.step_events.push_back(ev);
      H.step_weights.push_back(weight);
    };
    with_engine(H, [&](auto& e) {
      if (H.prof_mode & 2) mark(0);
      advance(H, e, nsteps, 0, mark);
    });
    HIPCHECK(hipGetLastError());
    check_status(H);
  });
}

int qg_run(qg_handle* h, double dt, int nsteps, int* done, int* failure, double* max_abs_psi) {
  return guarded(h, [&] {
    if (!(std::isfinite(dt) && dt > 0)) throw QgError(QG_ERR_ARGUMENT, "Δt must be finite and positive");
    if (nsteps < 0) throw QgError(QG_ERR_ARGUMENT, "nsteps must be non-negative");
    qg_handle& H = *h;
    int d = 0, fail = 0;
    double mx = 0.0;
    if (H.active > 0 && nsteps > 0 && !H.snap) H.snap = H.ralloc(12 * H.npts);
    run_checked(H, dt, nsteps, H.snap, true, d, fail, mx);
    if (done) *done = d;
    if (failure) *failure = fail;
    if (max_abs_psi) *max_abs_psi = mx;
  });
}

int qg_run_noalloc(qg_handle* h, double dt, int nsteps, int* done, int* failure, double* max_abs_psi) {
  return guarded(h, [&] {
    if (!(std::isfinite(dt) && dt > 0)) throw QgError(QG_ERR_ARGUMENT, "Δt must be finite and positive");
    if (nsteps < 0) throw QgError(QG_ERR_ARGUMENT, "nsteps must be non-negative");
    qg_handle& H = *h;
    if (!H.noalloc_snap) throw QgError(QG_ERR_DEVICE, "qg_run_noalloc: no scratch address for this handle");
    if (H.prof_mode)
      throw QgError(QG_ERR_ARGUMENT, "qg_run_noalloc: not while a profile is recorded (profiling creates events)");
    const size_t need = (size_t)12 * H.rsize() * H.active * H.nzc * H.plane;
    if (need > kNoAllocSnapBytes) {
      char buf[256];
      snprintf(buf, sizeof buf,
               "qg_run_noalloc: the snapshot of the active members needs %zu bytes, more than the %zu-byte "
               "static scratch (12 scalars per grid point); use qg_run",
               need, kNoAllocSnapBytes);
      throw QgError(QG_ERR_ARGUMENT, buf);
    }
    int d = 0, fail = 0;
    double mx = 0.0;
    run_checked(H, dt, nsteps, H.noalloc_snap, false, d, fail, mx);
    if (done) *done = d;
    if (failure) *failure = fail;
    if (max_abs_psi) *max_abs_psi = mx;
  });
}

// Members addressed by a set-up call: one member, or all (member == -1).
static void member_span(const qg_handle& H, int member, int& m0, int& nm) {
  if (member == -1) {
    m0 = 0;
    nm = H.members;
    return;
  }
  check_member_index(H, member);
  m0 = member;
  nm = 1;
}

int qg_set_random_psi(qg_handle* h, int member, double amplitude, double slope, long long seed) {
  return guarded(h, [&] {
    qg_handle& H = *h;
    int m0, nm;
    member_span(H, member, m0, nm);
    if (!(std::isfinite(amplitude) && std::isfinite(slope)))
      throw QgError(QG_ERR_ARGUMENT, "random streamfunction parameters must be finite");
    if (amplitude < 0) throw QgError(QG_ERR_ARGUMENT, "amplitude must be non-negative");
    if (seed < 0) throw QgError(QG_ERR_ARGUMENT, "seed must be non-negative (UInt64(seed))");
    qg::Geo g = H.g;
    g.nz = H.nz;
    g.members = H.members;
    if (H.f32()) {  // FP64 values kept in stg2 for the q of the set-up
      const size_t off = (size_t)m0 * H.nzc * H.plane, n = (size_t)nm * H.nzc * H.plane;
      qg::launch_random_psi(g, H.st, m0, nm, (unsigned long long)seed, amplitude, slope, H.stg2, false);
      qg::launch_convert<double, float>(reinterpret_cast<const double*>(H.stg2 + off),
                                        static_cast<float*>(H.at(H.psi, off)), 2 * n, H.st);
      H.psi64_m0 = m0;
      H.psi64_nm = nm;
    } else {
      qg::launch_random_psi(g, H.st, m0, nm, (unsigned long long)seed, amplitude, slope, H.psi, false);
    }
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(H.st));
    H.diag_valid = false;
    H.clean = false;
    H.t1_valid = false;
  });
}

int qg_set_surface_wave(qg_handle* h, int member, const double* profile) {
  return guarded(h, [&] {
    qg_handle& H = *h;
    int m0, nm;
    member_span(H, member, m0, nm);
    if (!profile) throw QgError(QG_ERR_ARGUMENT, "profile is required");
    const size_t n = (size_t)nm * H.nzc;
    for (size_t k = 0; k < n; ++k)
      if (!std::isfinite(profile[k])) throw QgError(QG_ERR_ARGUMENT, "wave profile must be finite");
    HIPCHECK(hipMemcpyAsync(H.d_prof, profile, n * sizeof(double), hipMemcpyHostToDevice, H.st));
    qg::Geo g = H.g;
    g.nz = H.nz;
    g.members = H.members;
    qg::launch_surface_wave(g, H.st, m0, nm, H.d_prof, H.B[H.cur], H.f32());
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(H.st));
    H.diag_valid = false;
    H.clean = false;
    H.t1_valid = false;
  });
}

int qg_energies(qg_handle* h, int member, double* out) {
  return guarded(h, [&] {
    qg_handle& H = *h;
    int m0, nm;
    member_span(H, member, m0, nm);
    if (!out) throw QgError(QG_ERR_ARGUMENT, "out is required");
    if (H.sing_psi[0]) raise_singular(H.sing_psi);
    if (H.sing_A[0]) raise_singular(H.sing_A);
    qg::Geo all = H.g;
    all.nz = H.nz;
    all.members = H.members;
    if (!H.d_epart) {
      H.d_epart = H.alloc<double>(qg::energy_partial_size(all, H.members));
      H.d_energy = H.alloc<double>(5 * (size_t)H.members);
    }
    // _record_energy_diagnostics! (src/core/io.jl:405-419): psi from q
    // without the wave feedback, A and C from B, on scratch fields
    const MemberView mv(H, m0);
    qg::Geo g = mv.g;
    g.members = nm;
    g.nz = nm * H.nzc;
    qg::launch_thomas(g, H.st, H.at(H.q[H.cur], mv.off), H.at(H.psis, mv.off), nullptr, H.f32(), H.th_inv[0] + mv.toff,
                      H.th_cp[0] + mv.toff, H.d_al + mv.aoff, 1, nullptr, H.tune[qg::P_THOMAS_PSI]);
    qg::launch_thomas(g, H.st, H.at(H.B[H.cur], mv.off), H.at(H.As, mv.off), H.at(H.qeff, mv.off), H.f32(),
                      H.th_inv[1] + mv.toff,
                      H.th_cp[1] + mv.toff, H.d_al + mv.aoff, 4, nullptr, H.tune[qg::P_THOMAS_AC]);
    qg::launch_energies(all, H.st, m0, nm, H.psis, H.As, H.qeff, H.f32(), H.d_a, H.d_epart, H.d_energy);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemcpyAsync(out, H.d_energy, 5 * (size_t)nm * sizeof(double), hipMemcpyDeviceToHost, H.st));
    HIPCHECK(hipStreamSynchronize(H.st));
  });
}

int qg_set_member_stratification(qg_handle* h, int member, const double* N2, const double* a_ell) {
  return guarded(h, [&] {
    qg_handle& H = *h;
    check_member_index(H, member);
    if (!N2 || !a_ell) throw QgError(QG_ERR_ARGUMENT, "N2 and a_ell are required");
    const int nzc = H.nzc, kpad = qg::thomas_shifted_a_size(nzc);
    for (int k = 0; k < nzc; ++k)
      if (!(std::isfinite(N2[k]) && N2[k] > 0 && std::isfinite(a_ell[k])))
        throw QgError(QG_ERR_ARGUMENT, "N² must be finite and positive at every level");
    const size_t o = (size_t)member * nzc;
    std::vector<double> cz(nzc), fN2(nzc), al(kpad, 0.0), cu(kpad, 0.0);
    const double f2 = H.f * H.f;
    for (int k = 0; k < nzc; ++k) {
      cz[k] = (f2 / N2[k]) / (H.dz * H.dz);  // src/operators.jl:428
      fN2[k] = H.f / N2[k];                  // src/operators.jl:439
    }
    for (int k = 1; k < nzc; ++k) al[k] = a_ell[k - 1];
    for (int k = 1; k <= nzc - 3; ++k) cu[k] = cz[k];
    const MemberView mv(H, member);
    const size_t tq = qg::thomas_table_size(mv.g);  // table entries of one member
    const size_t quarter = tq / nzc;
    // The eliminations of the new profile are recorded into scratch tables
    // first: the member keeps its profile and tables (host and device alike)
    // unless every one of them succeeds.
    struct Scratch {
      hipStream_t st;
      void* p = nullptr;
      ~Scratch() {
        if (p) {
          (void)hipStreamSynchronize(st);
          (void)hipFree(p);
        }
      }
    } sc{H.st};
    const size_t ntab = 6 * tq + 2 * (size_t)nzc;
    HIPCHECK(hipMalloc(&sc.p, ntab * sizeof(double) + quarter + 256));
    double* s_a = static_cast<double*>(sc.p);
    double* s_cz = s_a + nzc;
    double* s_tab = s_cz + nzc;  // th_inv[0], th_cp[0], th_inv[1], th_cp[1], om_M, om_D
    unsigned char* s_piv = reinterpret_cast<unsigned char*>(s_a + ntab);
    HIPCHECK(hipMemcpyAsync(s_a, a_ell, nzc * sizeof(double), hipMemcpyHostToDevice, H.st));
    HIPCHECK(hipMemcpyAsync(s_cz, cz.data(), nzc * sizeof(double), hipMemcpyHostToDevice, H.st));
    HIPCHECK(hipMemsetAsync(H.d_ferr, 0, 8 * sizeof(int), H.st));
    qg::launch_thomas_factor(mv.g, H.st, s_a, 1, s_tab, s_tab + tq, H.d_ferr);
    qg::launch_thomas_factor(mv.g, H.st, s_a, 4, s_tab + 2 * tq, s_tab + 3 * tq, H.d_ferr + 4);
    int he[8];
    HIPCHECK(hipMemcpyAsync(he, H.d_ferr, sizeof he, hipMemcpyDeviceToHost, H.st));
    HIPCHECK(hipStreamSynchronize(H.st));
    if (he[0] || he[4]) raise_singular(he[0] ? he : he + 4);
    HIPCHECK(hipMemsetAsync(H.d_ferr, 0, 8 * sizeof(int), H.st));
    qg::launch_omega_factor(mv.g, H.st, s_cz, s_tab + 4 * tq, s_tab + 5 * tq, s_piv, H.d_ferr);
    HIPCHECK(hipMemcpyAsync(he, H.d_ferr, sizeof he, hipMemcpyDeviceToHost, H.st));
    HIPCHECK(hipStreamSynchronize(H.st));
    if (he[0]) {
      char buf[160];
      snprintf(buf, sizeof buf, "LAPACK-style gtsv failed for vertical velocity at row %d of column %d",
               he[1] + 1, he[2]);
      throw QgError(QG_ERR_GTSV, buf);
    }
    // commit: coefficients, tables and the host copies of the profile
    auto d2d = [&](void* dst, const void* src, size_t bytes) {
      HIPCHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, H.st));
    };
    double* tabs[6] = {H.th_inv[0], H.th_cp[0], H.th_inv[1], H.th_cp[1], H.om_M, H.om_D};
    for (int t = 0; t < 6; ++t) d2d(tabs[t] + mv.toff, s_tab + t * tq, tq * sizeof(double));
    d2d(H.om_piv + (size_t)member * quarter, s_piv, quarter);
    d2d(H.d_a + o, s_a, nzc * sizeof(double));
    d2d(H.d_cz + o, s_cz, nzc * sizeof(double));
    HIPCHECK(hipMemcpyAsync(H.d_fN2 + o, fN2.data(), nzc * sizeof(double), hipMemcpyHostToDevice, H.st));
    HIPCHECK(hipMemcpyAsync(H.d_al + mv.aoff, al.data(), kpad * sizeof(double), hipMemcpyHostToDevice, H.st));
    HIPCHECK(hipMemcpyAsync(H.d_cu + mv.aoff, cu.data(), kpad * sizeof(double), hipMemcpyHostToDevice, H.st));
    HIPCHECK(hipStreamSynchronize(H.st));
    std::copy(N2, N2 + nzc, H.N2.begin() + o);
    std::copy(a_ell, a_ell + nzc, H.a.begin() + o);
    if (he[3] && !H.om_any_piv) {
      // the step graphs were captured without the pivoted-column omega
      // kernel; it now has columns to solve (kept for the ensemble's lifetime)
      H.drop_graphs();
      H.om_any_piv = true;
    }
    H.diag_valid = false;
    H.clean = false;
    H.t1_valid = false;
  });
}

int qg_set_active_members(qg_handle* h, int active) {
  return guarded(h, [&] {
    qg_handle& H = *h;
    if (active < 0 || active > H.members)
      throw QgError(QG_ERR_ARGUMENT, "active members must be in [0, " + std::to_string(H.members) + "]");
    if (active == H.active) return;
    HIPCHECK(hipStreamSynchronize(H.st));
    H.drop_graphs();  // captured for the previous extent
    if (active > H.active) {  // newly stepped members have no diagnostics yet
      H.diag_valid = false;
      H.clean = false;
      H.t1_valid = false;
    }
    H.active = active;
    H.g.members = H.gp.members = active;
    H.g.nz = H.gp.nz = active * H.nzc;
  });
}

int qg_move_member(qg_handle* h, int src, int dst) {
  return guarded(h, [&] {
    qg_handle& H = *h;
    check_member_index(H, src);
    check_member_index(H, dst);
    if (src == dst) return;
    H.psi64_nm = 0;
    const size_t lv = (size_t)H.nzc * H.plane;
    auto mv = [&](void* base, size_t elem, size_t count) {
      char* b = static_cast<char*>(base);
      HIPCHECK(hipMemcpyAsync(b + dst * count * elem, b + src * count * elem, count * elem,
                              hipMemcpyDeviceToDevice, H.st));
    };
    // prognostic state and the diagnostics a caller may read back
    for (void* f : {H.q[H.cur], H.B[H.cur], H.psi, H.A, H.C}) mv(f, H.fsize(), lv);
    for (void* f : {H.u, H.v, H.w}) mv(f, H.rsize(), lv);
    // the member's coefficients and recorded eliminations
    const int kpad = qg::thomas_shifted_a_size(H.nzc);
    const size_t tq = (size_t)(H.nx / 2 + 1) * (H.ny / 2 + 1) * H.nzc;  // table entries per member
    for (double* f : {H.d_a, H.d_cz, H.d_fN2}) mv(f, sizeof(double), H.nzc);
    for (double* f : {H.d_al, H.d_cu}) mv(f, sizeof(double), kpad);
    for (double* f : {H.th_inv[0], H.th_cp[0], H.th_inv[1], H.th_cp[1], H.om_M, H.om_D}) mv(f, sizeof(double), tq);
    mv(H.om_piv, 1, tq / H.nzc);
    HIPCHECK(hipStreamSynchronize(H.st));
    std::copy(H.N2.begin() + (size_t)src * H.nzc, H.N2.begin() + (size_t)(src + 1) * H.nzc,
              H.N2.begin() + (size_t)dst * H.nzc);
    std::copy(H.a.begin() + (size_t)src * H.nzc, H.a.begin() + (size_t)(src + 1) * H.nzc,
              H.a.begin() + (size_t)dst * H.nzc);
    H.diag_valid = false;
    H.clean = false;
    H.t1_valid = false;
  });
}

int qg_use_graphs(qg_handle* h, int enable) {
  return guarded(h, [&] {
    h->graphs = enable != 0;
    if (!h->graphs) {
      HIPCHECK(hipStreamSynchronize(h->st));
      h->drop_graphs();
    }
  });
}

int qg_set_pruning(qg_handle* h, int enable) {
  return guarded(h, [&] {
    if (h->prune_enabled == (enable != 0)) return;
    HIPCHECK(hipStreamSynchronize(h->st));
    h->drop_graphs();  // captured with the other geometry
    h->prune_enabled = enable != 0;
    h->clean = false;  // the outside of the disk is zeroed again by the next step
  });
}

int qg_synchronize(qg_handle* h) {
  return guarded(h, [&] { HIPCHECK(hipStreamSynchronize(h->st)); });
}

int qg_check(qg_handle* h, int* nonfinite, double* max_abs_psi) {
  return guarded(h, [&] { check_range(*h, 0, h->members, nonfinite, max_abs_psi); });
}

int qg_check_member(qg_handle* h, int member, int* nonfinite, double* max_abs_psi) {
  return guarded(h, [&] {
    check_member_index(*h, member);
    check_range(*h, member, 1, nonfinite, max_abs_psi);
  });
}

int qg_progress_maxima(qg_handle* h, double* wave_speed, double* flow_speed) {
  return guarded(h, [&] {
    qg_handle& H = *h;
    HIPCHECK(hipMemsetAsync(H.d_pmax, 0, 2 * sizeof(unsigned long long), H.st));
    with_engine(H, [&](auto& e) { e.progress_maxima(); });
    unsigned long long bits[2] = {0, 0};
    HIPCHECK(hipMemcpyAsync(bits, H.d_pmax, sizeof(bits), hipMemcpyDeviceToHost, H.st));
    HIPCHECK(hipStreamSynchronize(H.st));
    double m[2];
    std::memcpy(m, bits, sizeof m);
    if (wave_speed) *wave_speed = m[0];
    if (flow_speed) *flow_speed = m[1];
  });
}

int qg_invert_q_to_psi(qg_handle* h) {
  return guarded(h, [&] {
    h->psi64_nm = 0;
    solve_psi(*h, h->g, h->st, h->q[h->cur], h->psi);
    h->diag_valid = false;
    h->clean = false;
    check_status(*h);
  });
}

int qg_invert_B_to_A(qg_handle* h) {
  return guarded(h, [&] {
    solve_A(*h, h->g, h->st, h->B[h->cur], h->A, h->C);
    h->diag_valid = false;
    h->clean = false;
    check_status(*h);
  });
}

int qg_invert_B_to_A_member(qg_handle* h, int member) {
  return guarded(h, [&] {
    qg_handle& H = *h;
    check_member_index(H, member);
    if (H.sing_A[0]) raise_singular(H.sing_A);
    const MemberView mv(H, member);
    qg::launch_thomas(mv.g, H.st, H.at(H.B[H.cur], mv.off), H.at(H.A, mv.off), H.at(H.C, mv.off), H.f32(),
                      H.th_inv[1] + mv.toff,
                      H.th_cp[1] + mv.toff, H.d_al + mv.aoff, 4, nullptr, H.tune[qg::P_THOMAS_AC]);
    H.diag_valid = false;
    H.clean = false;
    check_status(H);
  });
}

int qg_compute_q_from_psi_member(qg_handle* h, int member, int method) {
  return guarded(h, [&] {
    qg_handle& H = *h;
    check_member_index(H, member);
    if (method != 0 && method != 1) throw QgError(QG_ERR_ARGUMENT, "pv method must be 0 (qg) or 1 (barotropic)");
    const MemberView mv(H, member);
    const bool staged = H.f32() && member >= H.psi64_m0 && member < H.psi64_m0 + H.psi64_nm;
    qg::launch_q_from_psi(mv.g, H.st, staged ? static_cast<void*>(H.stg2 + mv.off) : H.at(H.psi, mv.off),
                          H.at(H.q[H.cur], mv.off), H.f32(), H.d_a + (size_t)member * H.nzc, method, staged);
    H.diag_valid = false;
    H.clean = false;
    HIPCHECK(hipStreamSynchronize(H.st));
  });
}

int qg_compute_velocities(qg_handle* h, int compute_w) {
  return guarded(h, [&] {
    with_engine(*h, [&](auto& e) { e.compute_velocities(compute_w != 0); });
    h->diag_valid = false;
    h->clean = false;
    check_status(*h);
  });
}

int qg_compute_q_from_psi(qg_handle* h, int method) {
  return guarded(h, [&] {
    if (method != 0 && method != 1) throw QgError(QG_ERR_ARGUMENT, "pv method must be 0 (qg) or 1 (barotropic)");
    const bool staged = h->f32() && h->psi64_m0 == 0 && h->psi64_nm >= h->g.members;
    qg::launch_q_from_psi(h->g, h->st, staged ? static_cast<void*>(h->stg2) : h->psi, h->q[h->cur], h->f32(),
                          h->d_a, method, staged);
    h->diag_valid = false;
    h->clean = false;
    HIPCHECK(hipStreamSynchronize(h->st));
  });
}

int qg_compute_qw(qg_handle* h, double* host_out) {
  return guarded(h, [&] {
    if (!host_out) throw QgError(QG_ERR_ARGUMENT, "null host buffer");
    if (!(std::isfinite(h->f) && h->f != 0))
      throw QgError(QG_ERR_ARGUMENT, "wave feedback requires a finite, nonzero Coriolis frequency");
    with_engine(*h, [&](auto& e) { e.compute_qw(e.c(h->qeff)); });
    qg::launch_transpose<D>(as_fp64(*h, h->qeff, 0, h->npts), h->stg, h->nx, h->ny, h->nz, false, h->st);
    HIPCHECK(hipMemcpyAsync(host_out, h->stg, h->npts * sizeof(D), hipMemcpyDeviceToHost, h->st));
    HIPCHECK(hipStreamSynchronize(h->st));
  });
}

int qg_fft(qg_handle* h, const double* host_in, double* host_out, int direction) {
  return guarded(h, [&] {
    if (!host_in || !host_out) throw QgError(QG_ERR_ARGUMENT, "null host buffer");
    if (direction != -1 && direction != 1) throw QgError(QG_ERR_ARGUMENT, "direction must be -1 or +1");
    qg_handle& H = *h;
    HIPCHECK(hipMemcpyAsync(H.stg, host_in, H.npts * sizeof(D), hipMemcpyHostToDevice, H.st));
    H.psi64_nm = 0;
    qg::launch_transpose<D>(H.stg, H.stg2, H.nx, H.ny, H.nz, true, H.st);
    fft2_inplace(H, H.stg2, direction);
    qg::launch_transpose<D>(H.stg2, H.stg, H.nx, H.ny, H.nz, false, H.st);
    HIPCHECK(hipMemcpyAsync(host_out, H.stg, H.npts * sizeof(D), hipMemcpyDeviceToHost, H.st));
    HIPCHECK(hipStreamSynchronize(H.st));
  });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Autotuner: tile variants of the FFT passes and launch shapes of the column
// solves, chosen per grid class (nx, ny, nzc, precision) by timing whole
// eager steps on the model's own state with HIP events around every launch.
// Candidate c runs every FFT pass at tile variant c (c < 3) and every column
// operator at the c-th launch shape that fits the depth.  A pass's time for a
// candidate is the median of its launch durations; a pass leaves its default
// only for a candidate at least 1 % faster than the default (timing noise
// between equal variants is about that size).  With QG_TUNE_IN_STEP the
// passes that take most of the step are then re-chosen one at a time from
// whole graph-replayed steps.  The state is restored after every candidate
// and at the end, so tuning leaves the model where it was.
// ---------------------------------------------------------------------------
namespace {

std::mutex g_tune_mu;
std::map<std::tuple<int, int, int, int, int>, std::array<int, qg::P_END>> g_tune_cache;

// the grid class, and whether the choice was refined on whole steps
std::tuple<int, int, int, int, int> tune_key(const qg_handle& h, bool in_step) {
  return {h.nx, h.ny, h.nzc, h.precision, (int)in_step};
}

std::vector<int> column_candidates(int nzc) {
  std::vector<int> c;
  for (int i = 0; i < qg::kScanShapeCount; ++i)
    if (qg::scan_shape_fits(qg::scan_shape_of(i, nzc))) c.push_back(i);
  return c;
}

// FFT tile variants other than the base are compiled for the lengths of
// fft_variants_supported only.
bool fft_variants_ok(const qg_handle& h) {
  return qg::fft_variants_supported(h.nx) && qg::fft_variants_supported(h.ny);
}

bool valid_choice(const qg_handle& h, int pass, int v) {
  if (v == -1) return true;
  if (pass < qg::P_FFT_END) return v >= 0 && v < qg::kTileVariants && (v == 1 || fft_variants_ok(h));
  return v >= 0 && v < qg::kScanShapeCount && qg::scan_shape_fits(qg::scan_shape_of(v, h.nzc));
}

void autotune(qg_handle& H, double dt, int steps, bool in_step) {
  if (H.prof_mode) throw QgError(QG_ERR_ARGUMENT, "autotune while a profile is being recorded");
  H.psi64_nm = 0;
  HIPCHECK(hipStreamSynchronize(H.st));
  // Every field a caller can read is saved and put back afterwards: the
  // prognostic q, B (the candidates step from them) and the diagnostics psi,
  // A, C, u, v, w of the state as the caller left it.
  const size_t n = (size_t)H.nzc * H.plane * H.active;
  const size_t fb = n * H.fsize(), rb = n * H.rsize();  // bytes of one complex / real field
  struct Saved {
    void* field;
    size_t off, bytes;
  };
  std::vector<Saved> kept;
  size_t total = 0;
  for (void* f : {H.q[H.cur], H.B[H.cur], H.psi, H.A, H.C}) {
    kept.push_back({f, total, fb});
    total += fb;
  }
  for (void* f : {H.u, H.v, H.w}) {
    kept.push_back({f, total, rb});
    total += rb;
  }
  char* save = nullptr;
  HIPCHECK(hipMalloc(&save, total));
  struct Cleanup {
    qg_handle& H;
    char* save;
    const std::vector<Saved>& kept;
    int cur;
    ~Cleanup() {
      (void)hipStreamSynchronize(H.st);
      if (H.cur != cur) H.cur = cur;  // kept[0], kept[1] are the buffers of parity cur
      for (const Saved& k : kept)
        (void)hipMemcpyAsync(k.field, save + k.off, k.bytes, hipMemcpyDeviceToDevice, H.st);
      (void)hipStreamSynchronize(H.st);
      (void)hipFree(save);
      H.prof_mode = 0;
      H.prof.clear();
      H.pool_used = 0;
      H.diag_valid = false;
      H.clean = false;
      H.t1_valid = false;
      H.drop_graphs();
    }
  } cleanup{H, save, kept, H.cur};
  for (const Saved& k : kept)
    HIPCHECK(hipMemcpyAsync(save + k.off, k.field, k.bytes, hipMemcpyDeviceToDevice, H.st));
  build_etd(H, dt);
  const std::vector<int> shapes = column_candidates(H.nzc);
  const int ncfg = std::max<int>(qg::kTileVariants, (int)shapes.size());
  double best[qg::P_END], dflt[qg::P_END];
  int choice[qg::P_END];
  std::fill(best, best + qg::P_END, std::numeric_limits<double>::infinity());
  std::fill(dflt, dflt + qg::P_END, std::numeric_limits<double>::infinity());
  std::fill(choice, choice + qg::P_END, -1);
  for (int c = 0; c < ncfg; ++c) {
    const int fv = (c < qg::kTileVariants && fft_variants_ok(H)) ? c : 1;
    const int cs = c < (int)shapes.size() ? shapes[c] : qg::scan_default_shape(H.nzc);
    for (int p = 0; p < qg::P_FFT_END; ++p) H.tune[p] = fv;
    for (int p = qg::P_FFT_END; p < qg::P_END; ++p) H.tune[p] = cs;
    HIPCHECK(hipStreamSynchronize(H.st));
    H.cur = cleanup.cur;
    HIPCHECK(hipMemcpyAsync(H.q[H.cur], save + kept[0].off, fb, hipMemcpyDeviceToDevice, H.st));
    HIPCHECK(hipMemcpyAsync(H.B[H.cur], save + kept[1].off, fb, hipMemcpyDeviceToDevice, H.st));
    H.diag_valid = false;
    H.clean = false;
    H.t1_valid = false;
    std::vector<std::vector<float>> ms(qg::P_END);
    with_engine(H, [&](auto& e) {
      e.step();  // warm-up; recomputes the diagnostics of the restored state
      H.prof.clear();
      H.pool_used = 0;
      H.prof_mode = 1;
      for (int i = 0; i < steps; ++i) e.step();
      // The paired psi and A solve (the step's launch when concurrency is
      // on; per-launch profiling runs the two solves apart) at this
      // candidate's shape, timed on its own from the current q and B into
      // the stage buffers (restored state is not touched; no mean warning).
      if (!H.fixed_flow && !(H.passive || H.no_disp) && H.nzc > 1) {
        const qg::Geo& g = (H.prunable && H.prune_enabled && H.clean) ? H.gp : H.g;
        for (int i = 0; i < 8; ++i) {
          hipEvent_t a = H.event(), b = H.event();
          HIPCHECK(hipEventRecord(a, H.st));
          qg::launch_thomas_pair(g, H.st, H.f32(), H.d_al, H.tune[qg::P_THOMAS_PAIR], H.q[H.cur], H.psis,
                                 H.th_inv[0], H.th_cp[0], nullptr, H.B[H.cur], H.As, nullptr, H.th_inv[1],
                                 H.th_cp[1]);
          HIPCHECK(hipEventRecord(b, H.st));
          H.prof.push_back({"thomas_pair", qg::P_THOMAS_PAIR, 0.0, a, b});
        }
      }
      H.prof_mode = 0;
    });
    HIPCHECK(hipStreamSynchronize(H.st));
    check_status(H);
    for (auto& r : H.prof) {
      float t = 0;
      HIPCHECK(hipEventElapsedTime(&t, r.a, r.b));
      if (r.pass >= 0) ms[r.pass].push_back(t);
    }
    H.prof.clear();
    H.pool_used = 0;
    for (int p = 0; p < qg::P_END; ++p) {
      const bool attributable =
          p < qg::P_FFT_END ? (c < qg::kTileVariants && fft_variants_ok(H)) : c < (int)shapes.size();
      if (!attributable || ms[p].empty()) continue;
      std::vector<float>& v = ms[p];
      std::nth_element(v.begin(), v.begin() + v.size() / 2, v.end());
      const double med = v[v.size() / 2];
      const int def = p < qg::P_FFT_END              ? 1
                      : p == qg::P_THOMAS_PAIR ? qg::scan_pair_default_shape(H.nzc)
                                               : qg::scan_default_shape(H.nzc);
      if (H.tune[p] == def) dflt[p] = med;
      if (med < best[p]) {
        best[p] = med;
        choice[p] = H.tune[p];
      }
    }
  }
  for (int p = 0; p < qg::P_END; ++p)  // keep the default unless clearly beaten
    if (choice[p] >= 0 && dflt[p] < std::numeric_limits<double>::infinity() && best[p] > 0.99 * dflt[p])
      choice[p] = -1;
  std::copy(choice, choice + qg::P_END, H.tune);
  if (!in_step || !H.graphs) return;

  // In-step refinement.  The choices above come from launches timed one at
  // a time; in the captured step some passes overlap the w passes' stream,
  // which can favour another variant.  For the passes that take most of the
  // step, and then for the w passes (which run beside the next step's
  // passes), each candidate is timed as whole graph-replayed steps (the other
  // passes fixed), and one that beats the current choice by 0.3 % replaces it.
  const int key_passes[] = {qg::P_YB2_B,    qg::P_XB_B,     qg::P_YB_ADV,  qg::P_XA_QW,
                            qg::P_XF,       qg::P_YF_OMEGA, qg::P_YA2_QEFF, qg::P_THOMAS_PAIR,
                            qg::P_OMEGA,    qg::P_Y_R2C,    qg::P_Y_HERM,  qg::P_X_C2R};
  hipEvent_t ea = H.event(), eb = H.event();
  with_engine(H, [&](auto& e) {
    auto none = [](int) {};
    auto step_ms = [&]() -> double {  // every candidate steps from the caller's q and B
      HIPCHECK(hipStreamSynchronize(H.st));
      H.cur = cleanup.cur;
      HIPCHECK(hipMemcpyAsync(H.q[H.cur], save + kept[0].off, fb, hipMemcpyDeviceToDevice, H.st));
      HIPCHECK(hipMemcpyAsync(H.B[H.cur], save + kept[1].off, fb, hipMemcpyDeviceToDevice, H.st));
      H.diag_valid = false;
      H.clean = false;
      H.t1_valid = false;
      H.drop_graphs();
      advance(H, e, 1 + 2 * kLongGraphSteps, 0, none);  // an eager step, then the graphs' capture
      HIPCHECK(hipEventRecord(ea, H.st));
      advance(H, e, 2 * kLongGraphSteps, 0, none);
      HIPCHECK(hipEventRecord(eb, H.st));
      HIPCHECK(hipEventSynchronize(eb));
      check_status(H);
      float ms = 0;
      HIPCHECK(hipEventElapsedTime(&ms, ea, eb));
      return ms;
    };
    for (int p : key_passes) {
      std::vector<int> cands;
      if (p < qg::P_FFT_END) {
        if (!fft_variants_ok(H)) continue;
        for (int v = 0; v < qg::kTileVariants; ++v) cands.push_back(v);
      } else if (p == qg::P_THOMAS_PAIR) {
        if (H.fixed_flow || H.passive || H.no_disp || H.nzc <= 1) continue;
        cands = shapes;
      } else {
        if (H.nzc <= 2) continue;  // no interior level: no omega solve
        cands = shapes;
      }
      const int def = p < qg::P_FFT_END              ? 1
                      : p == qg::P_THOMAS_PAIR ? qg::scan_pair_default_shape(H.nzc)
                                               : qg::scan_default_shape(H.nzc);
      int cur = H.tune[p];
      double best_ms = step_ms();
      for (int c : cands) {
        if (c == cur || (cur < 0 && c == def)) continue;
        H.tune[p] = c;
        const double t = step_ms();
        if (t < 0.997 * best_ms) {
          best_ms = t;
          cur = c;
        }
      }
      H.tune[p] = cur;
    }
  });
  std::copy(H.tune, H.tune + qg::P_END, choice);
  H.pool_used = 0;
  std::copy(choice, choice + qg::P_END, H.tune);
}

}  // namespace

extern "C" {

int qg_autotune(qg_handle* h, double dt, int steps, int flags) {
  return guarded(h, [&] {
    if (!(std::isfinite(dt) && dt > 0)) throw QgError(QG_ERR_ARGUMENT, "Δt must be finite and positive");
    if (steps <= 0) steps = 4;
    qg_handle& H = *h;
    if (H.active == 0) return;
    const bool in_step = (flags & QG_TUNE_IN_STEP) != 0;
    const auto key = tune_key(H, in_step);
    if (!(flags & QG_TUNE_FORCE)) {
      std::lock_guard<std::mutex> lk(g_tune_mu);
      auto it = g_tune_cache.find(key);
      if (it != g_tune_cache.end()) {
        HIPCHECK(hipStreamSynchronize(H.st));
        std::copy(it->second.begin(), it->second.end(), H.tune);
        H.drop_graphs();
        return;
      }
    }
    autotune(H, dt, steps, in_step);
    std::array<int, qg::P_END> t;
    std::copy(H.tune, H.tune + qg::P_END, t.begin());
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_cache[key] = t;
  });
}

int qg_get_tuning(qg_handle* h, int* choices, int n) {
  return guarded(h, [&] {
    if (!choices || n < (int)qg::P_END) throw QgError(QG_ERR_ARGUMENT, "need QG_TUNE_PASSES entries");
    std::copy(h->tune, h->tune + qg::P_END, choices);
  });
}

int qg_set_tuning(qg_handle* h, const int* choices, int n) {
  return guarded(h, [&] {
    if (!choices || n < (int)qg::P_END) throw QgError(QG_ERR_ARGUMENT, "need QG_TUNE_PASSES entries");
    for (int p = 0; p < qg::P_END; ++p)
      if (!valid_choice(*h, p, choices[p]))
        throw QgError(QG_ERR_ARGUMENT, "tuning choice " + std::to_string(choices[p]) + " is not valid for pass " +
                                           std::to_string(p));
    HIPCHECK(hipStreamSynchronize(h->st));
    std::copy(choices, choices + qg::P_END, h->tune);
    h->drop_graphs();  // captured with the previous kernels
  });
}

int qg_profile_begin(qg_handle* h, int mode) {
  return guarded(h, [&] {
    HIPCHECK(hipStreamSynchronize(h->st));
    h->prof.clear();
    h->step_events.clear();
    h->step_weights.clear();
    h->pool_used = 0;
    h->prof_mode = mode & 3;
  });
}

int qg_profile_end(qg_handle* h, char* buf, size_t len, size_t* needed) {
  return guarded(h, [&] {
    HIPCHECK(hipStreamSynchronize(h->st));
    struct Agg {
      std::string name;
      int count = 0;
      double ms = 0, bytes = 0;
    };
    std::vector<Agg> agg;
    for (auto& r : h->prof) {
      float ms = 0;
      HIPCHECK(hipEventElapsedTime(&ms, r.a, r.b));
      Agg* a = nullptr;
      for (auto& x : agg)
        if (x.name == r.name) a = &x;
      if (!a) {
        agg.push_back(Agg{r.name});
        a = &agg.back();
      }
      a->count += 1;
      a->ms += ms;
      a->bytes += r.bytes;
    }
    std::string js = "{\"kernels\": {";
    char tmp[256];
    for (size_t i = 0; i < agg.size(); ++i) {
      snprintf(tmp, sizeof tmp, "%s\"%s\": {\"count\": %d, \"ms\": %.6f, \"bytes\": %.1f}", i ? ", " : "",
               agg[i].name.c_str(), agg[i].count, agg[i].ms, agg[i].bytes);
      js += tmp;
    }
    // an interval covering w steps (a replayed two-step graph) counts as w
    // steps of equal length; intervals of weight 0 are gaps between calls
    js += "}, \"steps_ms\": [";
    bool first = true;
    for (size_t i = 1; i < h->step_events.size() && i - 1 < h->step_weights.size(); ++i) {
      const int w = h->step_weights[i - 1];
      if (w <= 0) continue;
      float ms = 0;
      HIPCHECK(hipEventElapsedTime(&ms, h->step_events[i - 1], h->step_events[i]));
      for (int k = 0; k < w; ++k) {
        snprintf(tmp, sizeof tmp, "%s%.6f", first ? "" : ", ", ms / w);
        js += tmp;
        first = false;
      }
    }
    js += "]}";
    h->prof_mode = 0;
    h->prof.clear();
    h->step_events.clear();
    h->step_weights.clear();
    h->pool_used = 0;
    if (needed) *needed = js.size() + 1;
    if (buf && len > 0) {
      size_t n = std::min(len - 1, js.size());
      std::memcpy(buf, js.data(), n);
      buf[n] = 0;
      if (js.size() + 1 > len) throw QgError(QG_ERR_ARGUMENT, "profile buffer too small");
    }
  });
}

int qg_take_mean_q_warning(qg_handle* h, double* max_abs_mean_q) {
  return guarded(h, [&] {
    unsigned long long bits = 0;
    HIPCHECK(hipMemcpyAsync(&bits, h->d_warn, sizeof bits, hipMemcpyDeviceToHost, h->st));
    HIPCHECK(hipMemsetAsync(h->d_warn, 0, sizeof bits, h->st));
    HIPCHECK(hipStreamSynchronize(h->st));
    double m;
    std::memcpy(&m, &bits, sizeof m);
    if (max_abs_mean_q) *max_abs_mean_q = m;
  });
}

}  // extern "C"
