/*
 * maxwell_fdtd.h -- C ABI of the MI355X (gfx950) FDTD runtime.
 *
 * Plain pointers, sizes and scalars only; every device pointer is a float32
 * (or where stated float64/int32) buffer owned by the caller, laid out like the
 * reference's torch tensors: [x][y][z], z fastest, component shapes
 *   Ex (Nx-1,Ny,Nz)  Ey (Nx,Ny-1,Nz)  Ez (Nx,Ny,Nz-1)
 *   Hx (Nx,Ny-1,Nz-1) Hy (Nx-1,Ny,Nz-1) Hz (Nx-1,Ny-1,Nz)
 * (reference: witwin/maxwell/fdtd/runtime/stepping.py:1625-1630).  `stream` is a
 * hipStream_t passed as void*; nothing here synchronises it.  Every entry point
 * returns 0 on success or a negative MW_ERR_* code, with a message available
 * from mw_last_error() (mirrors the reference's STD_TORCH_CHECK RuntimeErrors,
 * witwin/maxwell/fdtd/cuda/tensors.h:10-55).
 *
 * The reference binds its kernels as torch ops
 * (witwin/maxwell/fdtd/cuda/extension.cpp:2059-2212) called through
 * NativeFDTDModule.<Name>(**kw).launchRaw() (fdtd/cuda/backend.py:81-115); each
 * declaration below names the reference op family it replaces.
 */
#ifndef MAXWELL_FDTD_H
#define MAXWELL_FDTD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MW_OK 0
#define MW_ERR_ARGUMENT -1
#define MW_ERR_LAUNCH -2
#define MW_ERR_RUNTIME -3

/* boundary face codes, witwin/maxwell/fdtd/cuda/kernels/common.cuh:45-50 */
#define MW_BC_NONE 0
#define MW_BC_PML 1
#define MW_BC_PERIODIC 2
#define MW_BC_PEC 4
#define MW_BC_PMC 5

#define MW_MAX_POLES 4
#define MW_MAX_TERMS 16
#define MW_MAX_PLANES 16

/* Yee grid: node counts, per-face boundary codes and spacing reciprocals
 * (dual spacing for E, length N; primal for H, length N-1;
 * reference fdtd/runtime/initialization.py:131-139). */
typedef struct mw_grid {
  int32_t nx, ny, nz;
  int32_t bc_lo[3], bc_hi[3];
  const float* inv_e[3];
  const float* inv_h[3];
} mw_grid;

/* Update coefficients per component, order Ex Ey Ez Hx Hy Hz.  A NULL array
 * means the coefficient is uniform and *_value is used
 * (reference runtime/materials.py:1550-1574 uniform fast path). */
typedef struct mw_coeffs {
  const float* decay[6];
  const float* curl[6];
  float decay_value[6];
  float curl_value[6];
} mw_coeffs;

/* CPML: 1-D profiles per axis and 12 slab-stored psi arrays, order
 * ex_y ex_z ey_x ey_z ez_x ez_y hx_y hx_z hy_x hy_z hz_x hz_y.  psi[q] spans
 * psi_low[q] + psi_high_len[q] samples along its derivative axis
 * (reference boundary/cpml.py:218-257, kernels/common.cuh:65-96). */
typedef struct mw_cpml {
  int32_t enabled;
  int32_t pad;
  const float* inv_kappa_e[3];
  const float* b_e[3];
  const float* c_e[3];
  const float* inv_kappa_h[3];
  const float* b_h[3];
  const float* c_h[3];
  float* psi[12];
  int32_t psi_low[12];
  int32_t psi_high_start[12];
  int32_t psi_high_len[12];
} mw_cpml;

typedef struct mw_fields {
  float* e[3];
  float* h[3];
} mw_fields;

/* One Debye (0) / Drude (1) / Lorentz (2) pole on one E component; drive and the
 * state arrays have the component's shape (reference kernels/dispersive.cu:18-70).
 * drive_scale: the pole's drive on a fully occupied edge, used with mw_ade.weight. */
typedef struct mw_pole {
  int32_t kind;
  float decay;
  float restoring;
  float drive_scale;
  const float* drive;
  float* polarization; /* unused (may be NULL) for Drude */
  float* current;
} mw_pole;

typedef struct mw_ade {
  int32_t count[3];
  float dt;
  const float* inv_eps[3];
  mw_pole poles[3][MW_MAX_POLES];
  /* Optional (NULL: every pole's drive is read at every sample): per-sample byte of
   * the component's shape, bit p set where pole p's drive is non-zero.  A pole whose
   * bit is clear is skipped there -- its state stays zero and it adds nothing -- so
   * the drive arrays are read only where some pole acts.  No reference counterpart
   * (the reference streams every pole array at every sample). */
  const uint8_t* active[3];
  /* Optional (NULL: the drive arrays are read): per-sample byte of the component's
   * shape, two bits per pole (pole p in bits 2p..2p+1): 0 = the pole does not act
   * there, 1 = half of an edge inside the medium (drive = drive_scale / 2), 2 = the
   * whole edge (drive = drive_scale).  Node occupancy is 0 or 1, so every drive is
   * one of those three values; with the codes the E passes read one byte instead
   * of one float per pole.  On a sample whose E decay is 1 (no conductivity) the
   * subtraction uses dt / eps_inf = the E curl coefficient and reads no inv_eps.
   * No reference counterpart. */
  const uint8_t* weight[3];
} mw_ade;

/* H <- H*decay - curl*(curl E) with CPML psi for the components in comp_mask
 * (bit c = Hx,Hy,Hz).  Replaces updateMagneticField{Hx,Hy,Hz}{3D,CpmlCompressed3D,
 * Standard3D} (kernels/magnetic.cu:20-212, launchers :368-921). */
int mw_update_magnetic(const mw_grid* grid, const mw_coeffs* coeffs, const mw_cpml* cpml, const mw_fields* fields,
                       int32_t comp_mask, void* stream);

/* E <- E*decay + curl*(curl H) with boundary modes, CPML psi and (optional, may
 * be NULL) fused ADE advance + polarisation-current subtraction.  Replaces
 * updateElectricField{Ex,Ey,Ez}{Cpml3D,CpmlCompressed3D,Standard3D} plus
 * update{Debye,Drude,Lorentz}Current3D / applyPolarizationCurrent3D
 * (kernels/electric.cu:242-941, kernels/dispersive.cu:18-83). */
int mw_update_electric(const mw_grid* grid, const mw_coeffs* coeffs, const mw_cpml* cpml, const mw_ade* ade,
                       const mw_fields* fields, int32_t comp_mask, void* stream);

/* field[o + p] += signal * patch[p].  Replaces addSourcePatch3D
 * (kernels/sources.cu:135-154). */
int mw_add_source_patch(float* field, int32_t fnx, int32_t fny, int32_t fnz, const float* patch, int32_t pnx,
                        int32_t pny, int32_t pnz, int32_t oi, int32_t oj, int32_t ok, float signal, void* stream);

/* field[o + p] += patch[p] * s(time - delay[p]) with the CW (0) / Gaussian pulse (1)
 * / Ricker (2) waveform evaluated in double precision; delay is float64.
 * Replaces addTimeShiftedSourcePatch3D (kernels/sources.cu:179-261). */
int mw_add_time_shifted_patch(float* field, int32_t fnx, int32_t fny, int32_t fnz, const float* patch,
                              const double* delay, int32_t pnx, int32_t pny, int32_t pnz, int32_t oi, int32_t oj,
                              int32_t ok, double time, int32_t kind, double frequency, double fwidth,
                              double amplitude, double phase, double delay0, void* stream);

/* re[f][l] += v[l]*wcos[f]; im[f][l] += v[l]*wsin[f] for f < nf, weights in device
 * memory.  Replaces accumulateDft* (kernels/spectral.cu:9-297). */
int mw_accumulate_dft(const float* field, int64_t count, int32_t nf, const float* wcos, const float* wsin,
                      float* re, float* im, void* stream);

/* Plane observer: accumulate one axis-normal plane of a component
 * (axis 0/1/2, plane index) into re/im of shape (nf, plane).  Replaces
 * accumulatePlaneObserver3D (kernels/observers.cu:35-78). */
int mw_accumulate_plane(const float* field, int32_t nx, int32_t ny, int32_t nz, int32_t axis, int32_t index,
                        int32_t nf, const float* wcos, const float* wsin, float* re, float* im, void* stream);

/* Point observers: accumulate count samples field[idx[p]] (linear indices) into
 * re/im of shape (nf, count).  Replaces accumulatePointObservers3D
 * (kernels/observers.cu:7-33). */
int mw_accumulate_points(const float* field, const int32_t* linear_index, int32_t count, int32_t nf,
                         const float* wcos, const float* wsin, float* re, float* im, void* stream);

/* One pole's current on one E component over `count` samples: kind 0 Debye
 * (P' = decay P + drive E, J = (P' - P)/dt), 1 Drude (J' = decay J + drive E;
 * polarization unused, may be NULL), 2 Lorentz (J' = decay J - restoring P +
 * drive E, P' = P + dt J').  Replaces update{Debye,Drude,Lorentz}Current3D
 * (kernels/dispersive.cu:18-70); the step kernels fuse it into the E update. */
int mw_update_pole_current(int32_t kind, const float* e, const float* drive, float* polarization, float* current,
                           int64_t count, float decay, float restoring, float dt, void* stream);

/* e -= dt * current * inv_eps over `count` samples.  Replaces
 * applyPolarizationCurrent3D (kernels/dispersive.cu:72-83). */
int mw_apply_polarization_current(float* e, const float* current, const float* inv_eps, int64_t count, float dt,
                                  void* stream);

/* TFSF auxiliary line, one half step each: h[m] = decay[m] h[m] - curl[m] (e[m+1] -
 * e[m]) for m < n_h; and e[n_e-1] = 0, e[source_index] = source_value, e[m] =
 * decay[m] e[m] - curl[m] (h[m] - h[m-1]) for the other 0 < m < n_e - 1.  Replace
 * updateAuxiliary{Magnetic,Electric}1D (kernels/sources.cu:803-839); mw_run_steps
 * advances the line itself. */
int mw_update_aux_magnetic(float* h, const float* e, const float* decay, const float* curl, int32_t n_h,
                           void* stream);
int mw_update_aux_electric(float* e, const float* h, const float* decay, const float* curl, int32_t n_e,
                           int32_t source_index, float source_value, void* stream);

/* ---- multi-step program: the whole per-step sequence with device-side step
 * tables, launched from native code (reference hot loop
 * fdtd/runtime/stepping.py:2289-2357 and its graph runners :1818-2024). ---- */

#define MW_TERM_ADDITIVE 0 /* value = signal_table[step] * patch            */
#define MW_TERM_SHIFTED 1  /* value = patch * s(step*dt + t_offset - delay) */
#define MW_TERM_LINE 2     /* value = patch * line[line_base + line_step * p_axis] (TFSF face) */

typedef struct mw_term {
  int32_t kind;
  int32_t component; /* 0..5 = Ex..Hz */
  int32_t offset[3];
  int32_t shape[3];
  const float* patch;
  const double* delay;        /* SHIFTED */
  const float* signal_table;  /* ADDITIVE: one value per step */
  int32_t wave_kind;          /* SHIFTED: 0 CW, 1 Gaussian, 2 Ricker */
  int32_t pad;
  double t_offset, frequency, fwidth, amplitude, phase, delay0;
  float* values;              /* SHIFTED, single-pass mode: device buffer of
                                 patch-size floats the step fills with
                                 patch * s(t - delay) before the fused kernel */
  const float* line;          /* LINE: the auxiliary line's E (H-side terms) or H
                                 (E-side terms); p_axis is the patch-local
                                 coordinate along line_axis (line_axis < 0: one
                                 sample for the whole patch) */
  int32_t line_axis, line_base, line_step, line_len;
} mw_term;

/* TFSF auxiliary 1-D line (reference fdtd/excitation/spatial.py:311-469,
 * kernels/sources.cu:803-839): advanced once per step, its H before the field
 * update and its E (hard source e[0] = source[step], e[n_e-1] = 0) after it. */
typedef struct mw_aux_line {
  int32_t n_e; /* E samples; 0 = no line */
  int32_t pad;
  float* e;    /* n_e */
  float* h;    /* n_e - 1 */
  const float* decay_e;
  const float* curl_e;
  const float* decay_h;
  const float* curl_h;
  const float* source; /* total_steps values */
} mw_aux_line;

typedef struct mw_plane_obs {
  int32_t component, axis, index, pad;
  float* re; /* (nf, plane) */
  float* im;
} mw_plane_obs;

typedef struct mw_program {
  mw_grid grid;
  mw_coeffs coeffs;
  mw_cpml cpml;
  mw_ade ade;
  mw_fields fields;
  double dt;
  int32_t total_steps;
  int32_t has_ade;
  /* H-side (applied after the H update) and E-side terms; HOST arrays of at
   * most MW_MAX_TERMS descriptors each (they travel as kernel arguments) */
  int32_t n_h_terms, n_e_terms;
  const mw_term* h_terms;
  const mw_term* e_terms;
  /* observers: weight tables are (total_steps, nf) float32 in device memory */
  int32_t nf_obs, n_planes, n_points;
  /* single-pass tiling of the z ends: 0 or 2 = interior tiles aligned to the z
   * PML slabs with the z ends in the per-cell shell, 1 = z-edge tiles beside the
   * interior tiles, 3 = interior tiles aligned as for 2 and both z ends of two
   * rows in one wavefront (when each end and a halo lane fit 32 lanes and the z
   * faces are not periodic; otherwise as 2), 4 = as 3 with the z-end launch on a
   * second stream beside the interior launch (a stream and two events per
   * concurrent caller, created on first use and kept by the library; the z ends
   * stay on the caller's stream while it is being captured) */
  int32_t z_tiling;
  const float* obs_e_cos;
  const float* obs_e_sin;
  const float* obs_h_cos;
  const float* obs_h_sin;
  const mw_plane_obs* planes; /* HOST array of at most MW_MAX_PLANES */
  const int32_t* point_component; /* device arrays, n_points */
  const int32_t* point_linear;
  float* point_re; /* (nf, n_points) */
  float* point_im;
  /* full-field DFT of E */
  int32_t nf_dft;
  /* single-pass tiling: x planes per block of the PML-shell launches (0 = automatic) */
  int32_t shell_chunk;
  const float* dft_cos;
  const float* dft_sin;
  float* dft_re[3]; /* (nf, component numel) */
  float* dft_im[3];
  /* Single-pass mode (fused = 1): E, H and the H-side psi are double-buffered.
   * At step n the state lives in `fields` / cpml.psi[6..11] when n is even and in
   * `fields_alt` / psi_h_alt when n is odd.  Not available with periodic faces.
   * h_terms_dev / e_terms_dev are device copies of h_terms / e_terms.
   * xchunk: x planes per block of the interior launch (0 = automatic). */
  int32_t fused, xchunk;
  mw_fields fields_alt;
  float* psi_h_alt[6];
  const mw_term* h_terms_dev;
  const mw_term* e_terms_dev;
  mw_aux_line aux;
  /* Staged z-normal plane monitors (NULL = every plane added to its sums after
   * each step).  After step n the samples of every plane with axis 2 are copied
   * into stage row n % plane_stage_depth; when a block of plane_stage_depth steps
   * completes, and at the end of each mw_run_steps call, one kernel adds the
   * block's rows to the planes' sums in step order (the roundings of per-step
   * accumulation).  A row holds each distinct (component, index) of the axis-2
   * planes once, in `planes` order, in the sums' layout, so plane_stage needs
   * plane_stage_depth x (sum of their cells) floats.  A call
   * that ends inside a block folds that block's rows so far and zeroes them, so a
   * later call that continues the block adds each step once.  The stage must be
   * zero before the first call and whenever the next call does not start where
   * the last one ended (a restart or a restored state). */
  float* plane_stage;
  int64_t plane_stage_floats; /* capacity of plane_stage in floats */
  int32_t plane_stage_depth;  /* 1..64 steps per block */
  int32_t pad_stage;
} mw_program;

/* Advance `count` steps starting at step index `first` (t = n*dt).  Replaces the
 * reference's per-step Python loop with its graph-captured field/tail blocks
 * (fdtd/runtime/stepping.py:2289-2357, :1818-2024). */
int mw_run_steps(const mw_program* program, int32_t first, int32_t count, void* stream);

/* ---- adjoint (reverse-mode) ----
 * Design window: the node box [lo, hi) over which dL/d(decay_E) and dL/d(curl_E)
 * are accumulated, each as three z-fastest (hi-lo) boxes, one per E component.
 * The box must satisfy 1 <= lo, hi <= N-1 and lie outside every E-side CPML
 * slab, so every window sample is an updated, psi-free sample.  A capture
 * stores E^n on [lo, hi) and H^{n+1/2} on [lo-1, hi) per replayed step
 * (replaces the reference's segment replay that keeps whole states,
 * fdtd/adjoint/bridge.py:728-853, fdtd/checkpoint.py:365-440). */
typedef struct mw_window {
  int32_t lo[3], hi[3];
  int32_t capacity; /* steps the stores hold */
  int32_t pad;
  float* e_store;   /* (capacity + 1) x 3 x W,  W  = prod(hi - lo)     */
  float* h_store;   /* capacity x 3 x Wh,       Wh = prod(hi - lo + 1) */
  float* g_decay[3]; /* W each, accumulated */
  float* g_curl[3];
} mw_window;

/* Adjoint state (same shapes as the forward state; psi in the program's slab
 * layout) and the gradients of the observer sums that seed it (NULL = zero;
 * layouts of point_re/im, planes[q].re/im and dft_re/im). */
typedef struct mw_adjoint {
  mw_fields fields;
  float* psi[12];
  float* polarization[3][MW_MAX_POLES];
  float* current[3][MW_MAX_POLES];
  const float* g_point_re;
  const float* g_point_im;
  const float* g_plane_re[MW_MAX_PLANES];
  const float* g_plane_im[MW_MAX_PLANES];
  const float* g_dft_re[3];
  const float* g_dft_im[3];
} mw_adjoint;

/* Replay steps [first, first+count) of a program (observers are not
 * accumulated) and capture the design window: E^first..E^{first+count} into
 * e_store[0..count], H^{n+1/2} into h_store[n-first].  The program's state must
 * hold step `first` on entry.  Replaces _replay_segment_states / _step_state
 * (fdtd/adjoint/core.py:3495-3565). */
int mw_capture_steps(const mw_program* program, int32_t first, int32_t count, const mw_window* window,
                     void* stream);

/* Reverse steps first+count-1 down to first: seeds of each step's observers,
 * then the transposed H/E/CPML/ADE updates; with a window (may be NULL) also the
 * coefficient gradients, reading the capture of [first, first+count).  The
 * adjoint must hold the adjoint of the state after step first+count-1 on entry.
 * Replaces reverse_phase{1,2,3} / reverse_cpml_phase_* / reverse_*_current and
 * seed_inject_* (fdtd/adjoint/native.py:48-727, kernels/adjoint.cu:149-420,
 * :898-979, :1224-1452, :1737-1870). */
int mw_reverse_steps(const mw_program* program, const mw_adjoint* adjoint, const mw_window* window, int32_t first,
                     int32_t count, void* stream);

/* ---- scene batches ----
 * Many independent scenes (uneven sizes, single-pass programs) advance in lockstep:
 * every launch of a step covers the tiles of all scenes, and the step index lives
 * in a device counter, so a run of `graph_steps` steps is captured once as a
 * hipGraph and replayed.  The handle owns its small device tables, a private
 * stream (default priority) and the graph; the scenes' buffers stay the caller's
 * and must outlive the handle.  `xchunk` > 0 forces the x-planes per tile
 * (0 = automatic).  No counterpart in the reference, which runs one scene per
 * solver (its multi-scene ensemble, execution/ensemble.py:166-180, threads whole
 * solvers across GPUs). */
typedef struct mw_batch mw_batch;
int mw_batch_create(const mw_program* const* programs, int32_t count, int32_t xchunk, mw_batch** out);
/* Advance steps [first, first+steps) of every scene; graph_steps <= 1 launches
 * directly.  Ordered after prior work on `stream` and before later work on it. */
int mw_batch_run(mw_batch* batch, int32_t first, int32_t steps, int32_t graph_steps, void* stream);
int mw_batch_destroy(mw_batch* batch);

/* No-allocation batch run.  No counterpart in the reference: for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Advances steps [first, first+steps)
 * of every scene like mw_batch_create + mw_batch_run (direct launches, no graph)
 * + mw_batch_destroy, with the same kernels, but allocates nothing: the batch's
 * descriptor tables and step counter live in a static device array of
 * MW_NOALLOC_SCRATCH_BYTES on each GPU the library runs on (held until the
 * process ends), rewritten by every call.  Tables that do not fit (about 150
 * scenes) are refused with MW_ERR_ARGUMENT.  The call returns once the tables
 * are on the device; the steps run on `stream`.  Calls that overlap on one GPU
 * (another stream, another thread before the first call's steps finished) share
 * that array and race. */
#define MW_NOALLOC_SCRATCH_BYTES (1 << 20)
int mw_batch_run_noalloc(const mw_program* const* programs, int32_t count, int32_t xchunk, int32_t first,
                         int32_t steps, void* stream);

/* Adjoint of a scene batch: one adjoint state and design window per scene
 * (the single-scene mw_reverse_steps / mw_capture_steps, batched: every launch
 * covers all scenes).  `forward` must be the batch the scenes advanced in; its
 * replay inside a capture accumulates no observers.  Full-field DFT seeds are not
 * batched.  Same reference counterparts as mw_reverse_steps / mw_capture_steps. */
typedef struct mw_adjoint_batch mw_adjoint_batch;
int mw_adjoint_batch_create(mw_batch* forward, const mw_program* const* programs, const mw_adjoint* const* adjoints,
                            const mw_window* const* windows, int32_t count, mw_adjoint_batch** out);
int mw_adjoint_batch_capture(mw_adjoint_batch* batch, int32_t first, int32_t count, void* stream);
int mw_adjoint_batch_reverse(mw_adjoint_batch* batch, int32_t first, int32_t count, void* stream);
int mw_adjoint_batch_destroy(mw_adjoint_batch* batch);

const char* mw_last_error(void);
int32_t mw_abi_version(void);
/* sizeof of mw_grid, mw_coeffs, mw_cpml, mw_fields, mw_pole, mw_ade, mw_term,
 * mw_plane_obs, mw_program (in that order) for binding layout checks; returns
 * the number of entries. */
int32_t mw_abi_struct_sizes(int64_t* out, int32_t capacity);

#ifdef __cplusplus
}
#endif

#endif /* MAXWELL_FDTD_H */
