/*
 * qgybj.h - C ABI of the MI355X (gfx950) QG-YBJ+ ETD-RK2 library.
 *
 * One handle is one model: the whole grid is resident in HBM and every entry
 * point is serialised on the handle's HIP stream.  A handle is not
 * thread-safe.  No exceptions cross the ABI: every function returns a status
 * code (QG_OK = 0) and qg_last_error() returns the message of the last
 * failure on that handle.
 *
 * Field exchange uses the reference's array layout: Julia column-major
 * (z, x, y) with z fastest, i.e. element (k, i, j) at k + nz*(i + nx*j).
 * Complex fields (q, B, psi, A, C) are interleaved (re, im) doubles; real
 * fields (u, v, w) are doubles.  Spectral fields follow FFTW ordering of
 * (kx, ky).  Inside the library fields are stored planar [z][y][x].
 *
 * Citations name the reference (subhk/QGYBJplus.jl v1.0.27) function each
 * entry point replaces.
 */
#ifndef QGYBJ_H
#define QGYBJ_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define QG_OK 0
#define QG_ERR_ARGUMENT 1     /* ArgumentError / DimensionMismatch in the reference */
#define QG_ERR_DEVICE 2       /* HIP runtime failure */
#define QG_ERR_SINGULAR 3     /* "Singular matrix in Thomas solver" (src/elliptic.jl:1177-1193) */
#define QG_ERR_UNSUPPORTED 4  /* configuration outside the built path (e.g. normal YBJ) */
#define QG_ERR_GTSV 5         /* "LAPACK gtsv failed" (src/operators.jl:447-464) */

/* field ids: ModelFields (src/core/fields.jl:11-24) */
#define QG_FIELD_Q 0
#define QG_FIELD_B 1
#define QG_FIELD_PSI 2
#define QG_FIELD_A 3
#define QG_FIELD_C 4
#define QG_FIELD_U 5
#define QG_FIELD_V 6
#define QG_FIELD_W 7

/* wave formulations: src/core/components.jl:59-67 */
#define QG_YBJ_PLUS 0
#define QG_PASSIVE 1
#define QG_YBJ 2 /* normal YBJ: rejected with QG_ERR_UNSUPPORTED */

/* precision modes */
#define QG_PRECISION_FP64 0  /* every transform, product and solve in FP64 */
#define QG_PRECISION_MIXED 1 /* FP32 storage of every field (state included), FP32 transforms and
                                products; FP64 column-solve recurrences and ETD coefficients */

/* space of a field handed to qg_set_field */
#define QG_SPACE_SPECTRAL 0
#define QG_SPACE_PHYSICAL 1

typedef struct qg_handle qg_handle;

/*
 * Model configuration.  The host computes kx, ky, N2 and a_ell with the
 * reference's own code (src/core/grid.jl:122-125, src/stratification.jl:308-318,
 * src/physics.jl:2-25) so the coefficients are bit-identical; dz is the
 * solver spacing G.z[2]-G.z[1] (1.0 when nz == 1).
 * Closures: HorizontalHyperdiffusivity fills both term lists,
 * FlowHyperdiffusivity only the flow list, WaveHyperdiffusivity only the wave
 * list (src/core/components.jl:105-208, src/nonlinear.jl:1160-1183).
 */
typedef struct qg_config {
  int nx, ny, nz;
  double f;                 /* Coriolis parameter of FPlane */
  double dz;                /* G.z[2]-G.z[1] */
  const double* kx;         /* [nx] */
  const double* ky;         /* [ny] */
  const double* N2;         /* [nz] runtime N^2 profile (OperatorCoefficients.N²) */
  const double* a_ell;      /* [nz] f^2/max(N^2, sqrt(eps)) */
  int n_flow_terms;          /* at most 8 per closure (QG_ERR_UNSUPPORTED beyond) */
  const double* flow_coefficients;
  const int* flow_orders;
  int n_wave_terms;
  const double* wave_coefficients;
  const int* wave_orders;
  double nu_z;              /* VerticalDiffusivity coefficient */
  int fixed_flow;           /* FixedFlow() */
  int wave_feedback;        /* WaveMeanFeedback() (NoFeedback/NoWaveFeedback = 0) */
  int formulation;          /* QG_YBJ_PLUS | QG_PASSIVE | QG_YBJ */
  int inviscid;             /* Inviscid() */
  int linear;               /* LinearDynamics() */
  int no_dispersion;        /* NoDispersion() */
  int precision;            /* QG_PRECISION_FP64 | QG_PRECISION_MIXED */
  /* Ensemble of models of one grid class (0 or 1: a single model).  Member m
   * occupies stacked levels [m*nz, (m+1)*nz); N2 and a_ell then hold
   * members*nz entries (member-major), every other setting is shared.  Each
   * pass of the step covers all members in one launch. */
  int members;
} qg_config;

/* Replaces build_runtime + allocate_fields (src/core/runtime.jl:34-83,
 * src/core/model.jl:79-121): allocates the model on the current device. */
int qg_create(const qg_config* cfg, qg_handle** out);

/* Device bytes qg_create would allocate for cfg (no allocation; used to admit
 * runs by free HBM), and the device's free and total memory.  No reference
 * counterpart. */
int qg_memory_required(const qg_config* cfg, size_t* bytes);
int qg_device_memory(size_t* free_bytes, size_t* total_bytes);

/* Replaces finalize_model! (src/core/model.jl:124). Safe on NULL. */
int qg_destroy(qg_handle* h);

/* Message of the last failing call on h (or of the last failed qg_create when
 * h is NULL). */
const char* qg_last_error(const qg_handle* h);

/* Replaces _assign_initial_field! / FieldArray upload (src/simulation.jl:405-427):
 * copies one field from host memory.  For complex fields `space` selects
 * spectral values (copied) or physical values (forward FFT on the device, as
 * fft_forward!).  Real fields (u, v, w) are physical. */
int qg_set_field(qg_handle* h, int field, const double* host, int space);

/* Copies one field to host memory (spectral for q, B, psi, A, C).  With
 * members > 1 both calls exchange the whole stack, a (members*nz, nx, ny)
 * array. */
int qg_get_field(qg_handle* h, int field, double* host);

/* The same for one ensemble member: a (nz, nx, ny) array.  No reference
 * counterpart (the reference has no ensembles); used by batched runs and the
 * serving loop to set up and read back one run. */
int qg_set_field_member(qg_handle* h, int member, int field, const double* host, int space);
int qg_get_field_member(qg_handle* h, int member, int field, double* host);

/* Replaces step!(model, ::ExponentialRungeKutta2) (src/core/runtime.jl:239-263)
 * -> _advance_etdrk2! (src/timestep.jl:335-415), nsteps times with step size
 * dt, including the double-buffer swap of the prognostic fields. */
int qg_step(qg_handle* h, double dt, int nsteps);

/* Replaces _check_termination_conditions! (src/simulation.jl:25-44): any
 * non-finite value in q, B or psi, and max |ifft(psi)| over the grid (the
 * complex inverse transform, computed at the mode's transform precision). */
int qg_check(qg_handle* h, int* nonfinite, double* max_abs_psi);

/* Replaces the stepping loop of run!(simulation) (src/simulation.jl:621-625):
 * up to nsteps steps of dt, each followed by _check_termination_conditions!,
 * with the checks made on the device (one host round trip per 32 steps).
 * Stops right after the first step whose check fails, as run! does:
 *   *done        steps taken, the failing step included;
 *   *failure     0 none, 1 non-finite q, B or psi ("non-finite value
 *                detected in the model state"), 2 max |ifft(psi)| > 1e10
 *                ("solution appears to be blowing up");
 *   *max_abs_psi the failing step's max |ifft(psi)|, else that of the state
 *                after the last step (the value qg_check gives).
 * While the steps stay far from the limit the maximum is bounded rather than
 * computed (sum of |psi| over each level's spectrum / (nx ny) under
 * 1e10 * (1 - 1e-4)); a stretch of 32 steps that does not meet the bound is
 * stepped again with the exact check, so stops and failures are those of the
 * exact check after every step.
 * The model state after a failure is the state right after the failing step.
 * The first call allocates five complex and two real fields of scratch: the
 * state every chunk of steps starts from (q, B, psi, A, the y-transformed B,
 * u, v), kept to step again up to a failing step. */
int qg_run(qg_handle* h, double dt, int nsteps, int* done, int* failure, double* max_abs_psi);
/* No counterpart in the reference: qg_run for callers that cannot allocate
 * device memory (tight serving loops, harnesses that forbid hipMalloc).  The
 * same loop, checks, failure semantics and results as qg_run, bit for bit,
 * with every step launched eagerly (no graph is instantiated) and the chunk
 * snapshot kept in a static device scratch array of the library instead of
 * the buffer qg_run allocates on its first call; nothing is allocated.
 * Limit: the snapshot takes 12 scalars of the storage precision per grid
 * point of the active members and must fit the 128 MiB scratch
 * (QG_RUN_NOALLOC_SCRATCH_BYTES); larger models get QG_ERR_ARGUMENT and
 * nothing runs.  Every call on one GPU shares that GPU's scratch: calls that
 * overlap on one GPU (two handles on two host threads) race.  Not while a
 * profile is recorded (QG_ERR_ARGUMENT).  Not capturable into a graph: it
 * reads the check back once per chunk. */
#define QG_RUN_NOALLOC_SCRATCH_BYTES ((size_t)128 << 20)
int qg_run_noalloc(qg_handle* h, double dt, int nsteps, int* done, int* failure, double* max_abs_psi);
/* Replaces _progress_maxima (src/simulation.jl:46-81): the maximum physical
 * wave velocity |ifft2(B + kh2 A / 4)| and the maximum flow speed
 * hypot(u, v) over the grid (all members), at the transform precision. */
int qg_progress_maxima(qg_handle* h, double* wave_speed, double* flow_speed);
/* The check of one ensemble member. */
int qg_check_member(qg_handle* h, int member, int* nonfinite, double* max_abs_psi);

/* Model-level operators (src/core/runtime.jl:169-206). */
int qg_invert_q_to_psi(qg_handle* h);                 /* invert_q_to_psi!(model) */
int qg_invert_B_to_A(qg_handle* h);                   /* invert_B_to_A!(model) */
int qg_compute_velocities(qg_handle* h, int compute_w); /* compute_velocities!(model) */
/* q from psi: method 0 = compute_q_from_psi! (src/initialization.jl:597-644),
 * 1 = compute_barotropic_q_from_psi! (src/initialization.jl:653-668). */
int qg_compute_q_from_psi(qg_handle* h, int method);
/* The same two set-up operators for one ensemble member (no reference
 * counterpart; a run admitted into an ensemble is set up without touching the
 * other members). */
int qg_compute_q_from_psi_member(qg_handle* h, int member, int method);
int qg_invert_B_to_A_member(qg_handle* h, int member);
/* Seeded random streamfunction on the device: replaces init_mpi_random_psi!
 * (src/parallel_mpi.jl:1159-1216) as called by set_mean_flow!(method=:random)
 * for RandomStreamfunction (src/simulation.jl:268-275, src/core/components.jl:
 * 218-222,264-272).  Writes psi-hat of one member (member == -1: every member,
 * each with the same seed) with the reference's counter hash over the
 * 1-based (level, x, y) indices of one model, energy amplitude*|k|^slope, and
 * the 2/3 mask.  q, u, v are derived by the caller's set-up operators
 * (qg_compute_q_from_psi[_member], qg_compute_velocities).  A negative seed
 * is an ArgumentError (the reference's UInt64(seed) throws). */
int qg_set_random_psi(qg_handle* h, int member, double amplitude, double slope, long long seed);

/* SurfaceWave on the device: replaces the field fill and fft_forward! of
 * set_surface_waves! (src/simulation.jl:292-329).  profile[k] is the
 * amplitude times the vertical profile of level k (nz entries for one member,
 * members*nz for member == -1); B-hat becomes nx*ny*profile[k] on the kh = 0
 * mode and zero elsewhere.  A follows from qg_invert_B_to_A[_member]. */
int qg_set_surface_wave(qg_handle* h, int member, const double* profile);

/* Energy diagnostics: replaces _local_energy_components /
 * _record_energy_diagnostics! (src/core/io.jl:331-419) on the device.  psi is
 * recomputed from q (no wave feedback) and A, C from B, on scratch fields (the
 * model's diagnostics are not touched), then out receives per member
 * {wave_KE, wave_PE, wave_CE, mean_flow_KE, mean_flow_PE}: 5 values for one
 * member, 5*members for member == -1.  The sums are reduced in a fixed order,
 * so repeated calls give identical values. */
int qg_energies(qg_handle* h, int member, double* out);

/* Replaces one member's N2 and a_ell profiles (nz entries each, computed as
 * for qg_config) and re-records its column eliminations, so a free ensemble
 * slot can take a run with another stratification.  Fails with
 * QG_ERR_SINGULAR / QG_ERR_GTSV, leaving the other members untouched, when
 * the new profile gives a singular column.  No reference counterpart. */
int qg_set_member_stratification(qg_handle* h, int member, const double* N2, const double* a_ell);
/* compute_qw_complex!(qw, B, ...) (src/nonlinear.jl:859-944) on the current
 * B; result copied to host (complex, (z,x,y)). */
int qg_compute_qw(qg_handle* h, double* host_out);

/* fft_forward! (direction -1) / fft_backward! (+1) of one complex (z,x,y)
 * host array (src/transforms.jl:19-23, src/parallel_mpi.jl:735-797). */
int qg_fft(qg_handle* h, const double* host_in, double* host_out, int direction);

/* Ensemble extent: qg_step advances members [0, active) only (all members
 * after qg_create; 0 makes qg_step a no-op), and qg_move_member copies member
 * src's state (q, B, psi, A, C, u, v, w) and coefficients over member dst, so a
 * serving loop keeps its resident runs in the leading slots.  The moved run's
 * next step recomputes its diagnostics, exactly as the previous step left
 * them.  No reference counterpart. */
int qg_set_active_members(qg_handle* h, int active);
int qg_move_member(qg_handle* h, int src, int dst);

/* Steady-state steps of qg_step (diagnostics left valid by the previous step,
 * no per-kernel profiling) replay a hipGraph of two captured steps; enable = 0
 * runs every step eagerly.  On by default.  No reference counterpart. */
int qg_use_graphs(qg_handle* h, int enable);

/* Dealiased-disk pruning of the step.  The 2/3 mask (src/physics.jl:28-43)
 * keeps the disk ix^2 + iy^2 <= kmax^2 of the spectral plane, and after one
 * masked update every spectral field of the state is zero outside it.  A
 * pruned step then reads and writes spectral data inside the disk only, runs
 * its y transforms over the columns |kx| <= kmax only, and solves the columns
 * inside the disk only; its results are those of the unpruned step (the work
 * left out is transforms of zeros and values the mask discards).  The first
 * step after any set-up or operator call runs unpruned.  Models whose flow is
 * fixed (psi is never re-diagnosed from q) always step unpruned.  On by
 * default; enable = 0 steps every step unpruned.  No reference counterpart. */
int qg_set_pruning(qg_handle* h, int enable);

/* Waits for the handle's stream. */
int qg_synchronize(qg_handle* h);

/* Autotuner (no reference counterpart: FFTW plans are fixed by PencilFFTs,
 * src/parallel_mpi.jl:676-720, and the Thomas loops have no launch shape).
 * The tunable passes are the 13 fused FFT passes (tile variant 0, 1 or 2:
 * half, base or twice the values per block, bitwise the same result; 3: the
 * alternative FFT plan where one is defined (radix 16 x 16 at N = 256) in
 * half tiles, equal to round-off) and the 4 column operators q->psi, B->A,
 * B->A with C, omega
 * (launch shape: partition depth and tile width; shapes differ in the
 * association of the partition carries, a few ulp).  qg_autotune times
 * `steps` eager steps of size dt per candidate on the model's own state,
 * restores the state, and keeps per pass the candidate with the least median
 * launch time when it beats the default by at least 1 %; the choice is
 * cached per grid class (nx, ny, nz per member, precision) and
 * QG_TUNE_IN_STEP flag in the process and reused by later calls unless flags
 * has QG_TUNE_FORCE.  -1 is the default
 * choice of a pass.  choices[] has QG_TUNE_PASSES entries in the order
 * ya_qw, xa_qw, ya2_qeff, yb_adv, x_c2r, xb_q, xb_B, yb2_q, yb2_B, yf_omega,
 * xf, y_r2c, y_herm, thomas_psi, thomas_A, thomas_AC, omega, thomas_pair
 * (xb_B also sets the X-pass that advects q and B together when both are
 * advected, yb2_B the launch of the q and B updates together; thomas_pair is
 * the launch that solves q -> psi and B -> A (and C) together). */
#define QG_TUNE_PASSES 18
#define QG_TUNE_FORCE 1
/* After the per-launch search, re-choose the passes that take most of the
 * step (yb2_B, xb_B, yb_adv, xa_qw, xf, yf_omega, ya2_qeff, thomas_pair) and
 * the w passes that run beside the next step (omega, y_r2c, y_herm, x_c2r)
 * one at a time from whole graph-replayed steps, 32 per candidate from the
 * caller's state; a candidate replaces the choice when 0.3 % faster (about
 * three seconds more at 256x256x128).  Without graphs (qg_use_graphs off)
 * the per-launch choice stands. */
#define QG_TUNE_IN_STEP 2
int qg_autotune(qg_handle* h, double dt, int steps, int flags);
int qg_get_tuning(qg_handle* h, int* choices, int n);
int qg_set_tuning(qg_handle* h, const int* choices, int n);

/* Measurement hooks (no reference counterpart: the reference has no
 * profiler, docs/src/advanced/performance.md:47-51).  mode bit 1 records HIP
 * events around every kernel launch on the handle's stream, bit 2 one event
 * per step.  qg_profile_end writes a JSON summary
 * {"kernels": {name: {"count", "ms", "bytes"}}, "steps_ms": [...]} where
 * "bytes" is the algorithmic HBM traffic of those launches; *needed receives
 * the buffer size the summary requires. */
int qg_profile_begin(qg_handle* h, int mode);
int qg_profile_end(qg_handle* h, char* buf, size_t len, size_t* needed);

/* Largest |q| met on the kh=0 column by invert_q_to_psi since the last call
 * (the reference warns once when it exceeds 1e-10, src/elliptic.jl:257-265);
 * resets the record. */
int qg_take_mean_q_warning(qg_handle* h, double* max_abs_mean_q);

#ifdef __cplusplus
}
#endif

#endif /* QGYBJ_H */
