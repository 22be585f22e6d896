/*
 * CPU restatement of the reference FDTD update equations -- TEST INFRASTRUCTURE.
 *
 * This file is the checker for the HIP runtime, never part of the product path:
 * only tests/, the smoke check and bench.py's CPU-baseline leg load it.  It is
 * compiled twice:
 *   - REAL_T=double          -> liboracle_f64.so  (the oracle for parity tests)
 *   - REAL_T=float, -fopenmp -> libcpu_f32.so     (the timed CPU baseline)
 *
 * Every routine restates a reference kernel, cited by file:line under
 * witwin/maxwell/fdtd/cuda/kernels/ unless stated otherwise:
 *   or_update_h        magnetic.cu:62-117   (H CPML update, dense psi)
 *   or_update_e        electric.cu:18-119 (boundary modes), :463-553 (E CPML update)
 *   or_ade_*           dispersive.cu:18-83  (Debye/Drude/Lorentz, polarisation current)
 *   or_add_patch       sources.cu:135-154   (additive source patch)
 *   or_add_shifted     sources.cu:71-93, :179-261 (time-shifted patch; waveform in double)
 *   or_accumulate      spectral.cu:9-46, observers.cu:7-78 (running DFT sums)
 * Reverse (adjoint) routines restate adjoint.cu as the exact transpose of the
 * forward ones (see or_reverse_*).
 *
 * Pinning (tests/test_oracle_*.py): discrete PEC-cavity eigenmodes against the
 * closed-form Yee dispersion relation, discrete energy conservation, CPML decay
 * and reflection, ADE susceptibilities against their closed-form discrete-time
 * transfer functions, DFT sums against direct numpy sums, and transpose
 * (dot-product) identities for the reverse routines.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#ifndef REAL_T
#define REAL_T double
#endif
typedef REAL_T real_t;

#define BC_NONE 0
#define BC_PML 1
#define BC_PERIODIC 2
#define BC_PEC 4
#define BC_PMC 5

typedef struct {
  int32_t nx, ny, nz;
  int32_t bc_lo[3], bc_hi[3];
  int32_t cpml;
  int32_t pad;
  const real_t* inv_e[3];
  const real_t* inv_h[3];
  const real_t* decay[6]; /* NULL: the component's coefficient is decay_value[c] everywhere */
  const real_t* curl[6];
  real_t decay_value[6];
  real_t curl_value[6];
  const real_t* kinv_e[3];
  const real_t* b_e[3];
  const real_t* c_e[3];
  const real_t* kinv_h[3];
  const real_t* b_h[3];
  const real_t* c_h[3];
} or_grid;

/* OpenMP only for large arrays: small grids (the pinning tests) run serially,
 * free of fork/join overhead */
#define PARALLEL(count) ((count) >= (1 << 18))

#define DECAY(c, l) (g->decay[c] ? g->decay[c][l] : g->decay_value[c])
#define CURL(c, l) (g->curl[c] ? g->curl[c][l] : g->curl_value[c])

/* CPML recursion psi' = b psi + c d.  Outside the absorbing layers b = 1 and
 * c = 0: psi starts at zero and stays exactly zero there, so it is neither read
 * nor written (the same numbers as the dense update, without its traffic). */
static inline real_t psi_step(real_t* p, real_t b, real_t c, real_t d) {
  if (b == (real_t)1 && c == (real_t)0) return 0;
  const real_t v = b * *p + c * d;
  *p = v;
  return v;
}

static inline int64_t off3(int64_t i, int64_t j, int64_t k, int64_t ny, int64_t nz) {
  return (i * ny + j) * nz + k;
}

/* magnetic.cu:62-117 */
void or_update_h(const or_grid* g, real_t* const* e, real_t* const* h, real_t* const* psi) {
  const int64_t nx = g->nx, ny = g->ny, nz = g->nz;
  const real_t *ex = e[0], *ey = e[1], *ez = e[2];
  /* Hx (nx, ny-1, nz-1): a = dEz/dy, b = dEy/dz */
#pragma omp parallel for schedule(static) if (PARALLEL(nx * ny * nz))
  for (int64_t i = 0; i < nx; ++i)
    for (int64_t j = 0; j < ny - 1; ++j)
      for (int64_t k = 0; k < nz - 1; ++k) {
        const int64_t l = off3(i, j, k, ny - 1, nz - 1);
        real_t da = (ez[off3(i, j + 1, k, ny, nz - 1)] - ez[off3(i, j, k, ny, nz - 1)]) * g->inv_h[1][j];
        real_t db = (ey[off3(i, j, k + 1, ny - 1, nz)] - ey[off3(i, j, k, ny - 1, nz)]) * g->inv_h[2][k];
        real_t curl;
        if (g->cpml) {
          real_t pa = psi_step(&psi[6][l], g->b_h[1][j], g->c_h[1][j], da);
          real_t pb = psi_step(&psi[7][l], g->b_h[2][k], g->c_h[2][k], db);
          curl = (da * g->kinv_h[1][j] + pa) - (db * g->kinv_h[2][k] + pb);
        } else {
          curl = da - db;
        }
        h[0][l] = h[0][l] * DECAY(3, l) - CURL(3, l) * curl;
      }
  /* Hy (nx-1, ny, nz-1): a = dEz/dx, b = dEx/dz, curl = b - a */
#pragma omp parallel for schedule(static) if (PARALLEL(nx * ny * nz))
  for (int64_t i = 0; i < nx - 1; ++i)
    for (int64_t j = 0; j < ny; ++j)
      for (int64_t k = 0; k < nz - 1; ++k) {
        const int64_t l = off3(i, j, k, ny, nz - 1);
        real_t da = (ez[off3(i + 1, j, k, ny, nz - 1)] - ez[off3(i, j, k, ny, nz - 1)]) * g->inv_h[0][i];
        real_t db = (ex[off3(i, j, k + 1, ny, nz)] - ex[off3(i, j, k, ny, nz)]) * g->inv_h[2][k];
        real_t curl;
        if (g->cpml) {
          real_t pa = psi_step(&psi[8][l], g->b_h[0][i], g->c_h[0][i], da);
          real_t pb = psi_step(&psi[9][l], g->b_h[2][k], g->c_h[2][k], db);
          curl = (db * g->kinv_h[2][k] + pb) - (da * g->kinv_h[0][i] + pa);
        } else {
          curl = db - da;
        }
        h[1][l] = h[1][l] * DECAY(4, l) - CURL(4, l) * curl;
      }
  /* Hz (nx-1, ny-1, nz): a = dEy/dx, b = dEx/dy */
#pragma omp parallel for schedule(static) if (PARALLEL(nx * ny * nz))
  for (int64_t i = 0; i < nx - 1; ++i)
    for (int64_t j = 0; j < ny - 1; ++j)
      for (int64_t k = 0; k < nz; ++k) {
        const int64_t l = off3(i, j, k, ny - 1, nz);
        real_t da = (ey[off3(i + 1, j, k, ny - 1, nz)] - ey[off3(i, j, k, ny - 1, nz)]) * g->inv_h[0][i];
        real_t db = (ex[off3(i, j + 1, k, ny, nz)] - ex[off3(i, j, k, ny, nz)]) * g->inv_h[1][j];
        real_t curl;
        if (g->cpml) {
          real_t pa = psi_step(&psi[10][l], g->b_h[0][i], g->c_h[0][i], da);
          real_t pb = psi_step(&psi[11][l], g->b_h[1][j], g->c_h[1][j], db);
          curl = (da * g->kinv_h[0][i] + pa) - (db * g->kinv_h[1][j] + pb);
        } else {
          curl = da - db;
        }
        h[2][l] = h[2][l] * DECAY(5, l) - CURL(5, l) * curl;
      }
}

/* electric.cu:23-51 */
static inline int is_pec(int64_t c, int64_t size, int lo, int hi) {
  return (c == 0 && lo == BC_PEC) || (c + 1 == size && hi == BC_PEC);
}
static inline int is_inactive(int64_t c, int64_t size, int lo, int hi) {
  return (c == 0 && (lo == BC_NONE || lo == BC_PML)) || (c + 1 == size && (hi == BC_NONE || hi == BC_PML));
}

/* Backward difference of `src` along one axis with the boundary modes of
 * electric.cu:23-87.  `cur`/`prev` are the current and previous samples,
 * `first`/`last` the samples at source index 0 and size-1 of that line. */
static inline real_t bdiff(int64_t c, int64_t size, int lo, int hi, real_t inv, real_t cur, real_t prev, real_t first,
                           real_t last) {
  if (c == 0) {
    if (lo == BC_PERIODIC) return (first - last) * inv;
    if (lo == BC_PMC) return 2.0f * first * inv;
    return 0;
  }
  if (c + 1 == size) {
    if (hi == BC_PERIODIC) return (first - last) * inv;
    if (hi == BC_PMC) return -2.0f * last * inv;
    return 0;
  }
  return (cur - prev) * inv;
}

/* electric.cu:463-553 */
void or_update_e(const or_grid* g, real_t* const* e, real_t* const* h, real_t* const* psi) {
  const int64_t nx = g->nx, ny = g->ny, nz = g->nz;
  const real_t *hx = h[0], *hy = h[1], *hz = h[2];
  const int *lo = g->bc_lo, *hi = g->bc_hi;
  /* Ex (nx-1, ny, nz): a = dHz/dy (Hz: nx-1, ny-1, nz), b = dHy/dz (Hy: nx-1, ny, nz-1) */
#pragma omp parallel for schedule(static) if (PARALLEL(nx * ny * nz))
  for (int64_t i = 0; i < nx - 1; ++i)
    for (int64_t j = 0; j < ny; ++j)
      for (int64_t k = 0; k < nz; ++k) {
        const int64_t l = off3(i, j, k, ny, nz);
        if (is_pec(j, ny, lo[1], hi[1]) || is_pec(k, nz, lo[2], hi[2])) { e[0][l] = 0; continue; }
        if (is_inactive(j, ny, lo[1], hi[1]) || is_inactive(k, nz, lo[2], hi[2])) continue;
        const int64_t jc = j < ny - 1 ? j : ny - 2, jp = j > 0 ? j - 1 : 0;
        const int64_t kc = k < nz - 1 ? k : nz - 2, kp = k > 0 ? k - 1 : 0;
        real_t da = bdiff(j, ny, lo[1], hi[1], g->inv_e[1][j], hz[off3(i, jc, k, ny - 1, nz)],
                          hz[off3(i, jp, k, ny - 1, nz)], hz[off3(i, 0, k, ny - 1, nz)],
                          hz[off3(i, ny - 2, k, ny - 1, nz)]);
        real_t db = bdiff(k, nz, lo[2], hi[2], g->inv_e[2][k], hy[off3(i, j, kc, ny, nz - 1)],
                          hy[off3(i, j, kp, ny, nz - 1)], hy[off3(i, j, 0, ny, nz - 1)],
                          hy[off3(i, j, nz - 2, ny, nz - 1)]);
        real_t curl;
        if (g->cpml) {
          real_t pa = psi_step(&psi[0][l], g->b_e[1][j], g->c_e[1][j], da);
          real_t pb = psi_step(&psi[1][l], g->b_e[2][k], g->c_e[2][k], db);
          curl = (da * g->kinv_e[1][j] + pa) - (db * g->kinv_e[2][k] + pb);
        } else {
          curl = da - db;
        }
        e[0][l] = e[0][l] * DECAY(0, l) + CURL(0, l) * curl;
      }
  /* Ey (nx, ny-1, nz): a = dHz/dx (Hz: nx-1, ny-1, nz), b = dHx/dz (Hx: nx, ny-1, nz-1), curl = b - a */
#pragma omp parallel for schedule(static) if (PARALLEL(nx * ny * nz))
  for (int64_t i = 0; i < nx; ++i)
    for (int64_t j = 0; j < ny - 1; ++j)
      for (int64_t k = 0; k < nz; ++k) {
        const int64_t l = off3(i, j, k, ny - 1, nz);
        if (is_pec(i, nx, lo[0], hi[0]) || is_pec(k, nz, lo[2], hi[2])) { e[1][l] = 0; continue; }
        if (is_inactive(i, nx, lo[0], hi[0]) || is_inactive(k, nz, lo[2], hi[2])) continue;
        const int64_t ic = i < nx - 1 ? i : nx - 2, ip = i > 0 ? i - 1 : 0;
        const int64_t kc = k < nz - 1 ? k : nz - 2, kp = k > 0 ? k - 1 : 0;
        real_t da = bdiff(i, nx, lo[0], hi[0], g->inv_e[0][i], hz[off3(ic, j, k, ny - 1, nz)],
                          hz[off3(ip, j, k, ny - 1, nz)], hz[off3(0, j, k, ny - 1, nz)],
                          hz[off3(nx - 2, j, k, ny - 1, nz)]);
        real_t db = bdiff(k, nz, lo[2], hi[2], g->inv_e[2][k], hx[off3(i, j, kc, ny - 1, nz - 1)],
                          hx[off3(i, j, kp, ny - 1, nz - 1)], hx[off3(i, j, 0, ny - 1, nz - 1)],
                          hx[off3(i, j, nz - 2, ny - 1, nz - 1)]);
        real_t curl;
        if (g->cpml) {
          real_t pa = psi_step(&psi[2][l], g->b_e[0][i], g->c_e[0][i], da);
          real_t pb = psi_step(&psi[3][l], g->b_e[2][k], g->c_e[2][k], db);
          curl = (db * g->kinv_e[2][k] + pb) - (da * g->kinv_e[0][i] + pa);
        } else {
          curl = db - da;
        }
        e[1][l] = e[1][l] * DECAY(1, l) + CURL(1, l) * curl;
      }
  /* Ez (nx, ny, nz-1): a = dHy/dx (Hy: nx-1, ny, nz-1), b = dHx/dy (Hx: nx, ny-1, nz-1) */
#pragma omp parallel for schedule(static) if (PARALLEL(nx * ny * nz))
  for (int64_t i = 0; i < nx; ++i)
    for (int64_t j = 0; j < ny; ++j)
      for (int64_t k = 0; k < nz - 1; ++k) {
        const int64_t l = off3(i, j, k, ny, nz - 1);
        if (is_pec(i, nx, lo[0], hi[0]) || is_pec(j, ny, lo[1], hi[1])) { e[2][l] = 0; continue; }
        if (is_inactive(i, nx, lo[0], hi[0]) || is_inactive(j, ny, lo[1], hi[1])) continue;
        const int64_t ic = i < nx - 1 ? i : nx - 2, ip = i > 0 ? i - 1 : 0;
        const int64_t jc = j < ny - 1 ? j : ny - 2, jp = j > 0 ? j - 1 : 0;
        real_t da = bdiff(i, nx, lo[0], hi[0], g->inv_e[0][i], hy[off3(ic, j, k, ny, nz - 1)],
                          hy[off3(ip, j, k, ny, nz - 1)], hy[off3(0, j, k, ny, nz - 1)],
                          hy[off3(nx - 2, j, k, ny, nz - 1)]);
        real_t db = bdiff(j, ny, lo[1], hi[1], g->inv_e[1][j], hx[off3(i, jc, k, ny - 1, nz - 1)],
                          hx[off3(i, jp, k, ny - 1, nz - 1)], hx[off3(i, 0, k, ny - 1, nz - 1)],
                          hx[off3(i, ny - 2, k, ny - 1, nz - 1)]);
        real_t curl;
        if (g->cpml) {
          real_t pa = psi_step(&psi[4][l], g->b_e[0][i], g->c_e[0][i], da);
          real_t pb = psi_step(&psi[5][l], g->b_e[1][j], g->c_e[1][j], db);
          curl = (da * g->kinv_e[0][i] + pa) - (db * g->kinv_e[1][j] + pb);
        } else {
          curl = da - db;
        }
        e[2][l] = e[2][l] * DECAY(2, l) + CURL(2, l) * curl;
      }
}

/* dispersive.cu:18-70; kind 0 = Debye, 1 = Drude, 2 = Lorentz */
void or_ade_advance(int kind, int64_t n, const real_t* e, const real_t* drive, double decay, double restoring,
                    double dt, real_t* pol, real_t* cur) {
#pragma omp parallel for schedule(static) if (PARALLEL(n))
  for (int64_t l = 0; l < n; ++l) {
    if (kind == 0) {
      real_t prev = pol[l];
      real_t next = (real_t)decay * prev + drive[l] * e[l];
      pol[l] = next;
      cur[l] = (next - prev) * (real_t)(1.0 / dt);
    } else if (kind == 1) {
      cur[l] = (real_t)decay * cur[l] + drive[l] * e[l];
    } else {
      real_t pc = cur[l], pp = pol[l];
      real_t nc = (real_t)decay * pc - (real_t)restoring * pp + drive[l] * e[l];
      cur[l] = nc;
      pol[l] = pp + (real_t)dt * nc;
    }
  }
}

/* or_ade_advance over the samples idx[0..count) only: where a pole's drive is
 * zero its state starts at zero and stays zero, so skipping those samples gives
 * the dense routine's numbers */
void or_ade_advance_idx(int kind, int64_t count, const int64_t* idx, const real_t* e, const real_t* drive,
                        double decay, double restoring, double dt, real_t* pol, real_t* cur) {
#pragma omp parallel for schedule(static) if (PARALLEL(count))
  for (int64_t q = 0; q < count; ++q) {
    const int64_t l = idx[q];
    if (kind == 0) {
      real_t prev = pol[l];
      real_t next = (real_t)decay * prev + drive[l] * e[l];
      pol[l] = next;
      cur[l] = (next - prev) * (real_t)(1.0 / dt);
    } else if (kind == 1) {
      cur[l] = (real_t)decay * cur[l] + drive[l] * e[l];
    } else {
      real_t pc = cur[l], pp = pol[l];
      real_t nc = (real_t)decay * pc - (real_t)restoring * pp + drive[l] * e[l];
      cur[l] = nc;
      pol[l] = pp + (real_t)dt * nc;
    }
  }
}

/* or_ade_apply over idx[0..count): elsewhere the pole's current is zero */
void or_ade_apply_idx(int64_t count, const int64_t* idx, const real_t* cur, const real_t* inv_eps, double dt,
                      real_t* e) {
#pragma omp parallel for schedule(static) if (PARALLEL(count))
  for (int64_t q = 0; q < count; ++q) {
    const int64_t l = idx[q];
    e[l] -= (real_t)dt * cur[l] * inv_eps[l];
  }
}

/* dispersive.cu:72-83 */
void or_ade_apply(int64_t n, const real_t* cur, const real_t* inv_eps, double dt, real_t* e) {
#pragma omp parallel for schedule(static) if (PARALLEL(n))
  for (int64_t l = 0; l < n; ++l) e[l] -= (real_t)dt * cur[l] * inv_eps[l];
}

/* sources.cu:135-154 */
void or_add_patch(real_t* f, int32_t fny, int32_t fnz, const real_t* patch, int32_t pnx, int32_t pny, int32_t pnz,
                  int32_t oi, int32_t oj, int32_t ok, double signal) {
  for (int64_t i = 0; i < pnx; ++i)
    for (int64_t j = 0; j < pny; ++j)
      for (int64_t k = 0; k < pnz; ++k)
        f[off3(oi + i, oj + j, ok + k, fny, fnz)] += (real_t)signal * patch[off3(i, j, k, pny, pnz)];
}

/* sources.cu:71-93 waveform, evaluated in double for the checker */
static double waveform(int kind, double t, double freq, double fwidth, double amp, double phase, double delay) {
  const double two_pi = 6.283185307179586476925286766559;
  if (kind == 0) return amp * cos(two_pi * freq * t + phase);
  if (kind == 1) {
    double tau = t - delay;
    double inv_sigma = two_pi * fwidth;
    if (inv_sigma < 1e-30) inv_sigma = 1e-30;
    double x = tau * inv_sigma;
    return amp * exp(-0.5 * x * x) * cos(two_pi * freq * tau + phase);
  }
  double tau = t - delay;
  double a = 3.14159265358979323846 * freq * tau;
  return amp * (1.0 - 2.0 * a * a) * exp(-a * a);
}

/* sources.cu:179-261 (no causal gate) */
void or_add_shifted(real_t* f, int32_t fny, int32_t fnz, const real_t* patch, const double* delay, int32_t pnx,
                    int32_t pny, int32_t pnz, int32_t oi, int32_t oj, int32_t ok, double time, int32_t kind,
                    double freq, double fwidth, double amp, double phase, double delay0) {
  for (int64_t i = 0; i < pnx; ++i)
    for (int64_t j = 0; j < pny; ++j)
      for (int64_t k = 0; k < pnz; ++k) {
        const int64_t p = off3(i, j, k, pny, pnz);
        double s = waveform(kind, time - delay[p], freq, fwidth, amp, phase, delay0);
        f[off3(oi + i, oj + j, ok + k, fny, fnz)] += (real_t)(s * (double)patch[p]);
      }
}

/* spectral.cu:9-46 / observers.cu:35-78: re[f][l] += v[l]*wcos[f], im += v*wsin[f] */
void or_accumulate(int64_t n, const real_t* v, int32_t nf, const double* wcos, const double* wsin, real_t* re,
                   real_t* im) {
  for (int32_t f = 0; f < nf; ++f) {
    const real_t c = (real_t)wcos[f], s = (real_t)wsin[f];
    if (c == 0 && s == 0) continue;
#pragma omp parallel for schedule(static) if (PARALLEL(n))
    for (int64_t l = 0; l < n; ++l) {
      re[f * n + l] += v[l] * c;
      im[f * n + l] += v[l] * s;
    }
  }
}

int32_t or_real_bytes(void) { return (int32_t)sizeof(real_t); }

/* max |v| over n samples (the running field peak the parity tests normalise by) */
double or_max_abs(int64_t n, const real_t* v) {
  double m = 0.0;
#pragma omp parallel for schedule(static) if (PARALLEL(n)) reduction(max : m)
  for (int64_t l = 0; l < n; ++l) {
    const double a = fabs((double)v[l]);
    if (a > m) m = a;
  }
  return m;
}

/* ------------------------------------------------------------------ reverse
 * Exact transpose of one forward step in the driver's order (H update, ADE
 * advance, E update, ADE subtraction), restating adjoint.cu:149-420 (standard),
 * :1224-1452 (CPML psi) and :898-979 (ADE) in scatter form: every forward read
 * x -> y becomes lambda_x += d y/d x * lambda_y.  Source terms are constants and
 * drop out.  On entry ae/ah/apsi/ade hold the adjoint of the state after step n
 * (E^{n+1}, H^{n+1/2}, psi^{n+1}, ADE state^{n+1}) with the step's observation
 * seeds already added; on exit they hold the adjoint of the state before it.
 *
 * When fe/fh/fpsi (E^n, H^{n+1/2}, psi_E^{n+1} of the forward run) are given, the
 * gradients of the loss with respect to the E decay and curl coefficients are
 * accumulated into g_decay/g_curl (E-shaped): dL/dcurl += lambda_post * curl^n,
 * dL/ddecay += lambda_post * E^n on every updated sample. */

typedef struct {
  int32_t count[3];
  int32_t kind[3][4];
  double decay[3][4];
  double restoring[3][4];
  const real_t* drive[3][4];
  real_t* apol[3][4];
  real_t* acur[3][4];
  const real_t* inv_eps[3];
  double dt;
} or_ade_adj;

/* transpose of bdiff: scatter w into the H line (same index rules) */
static inline void bdiff_t(int64_t c, int64_t size, int lo, int hi, real_t w, real_t* cur, real_t* prev, real_t* first,
                           real_t* last) {
  if (c == 0) {
    if (lo == BC_PERIODIC) { *first += w; *last -= w; }
    else if (lo == BC_PMC) *first += 2.0f * w;
    return;
  }
  if (c + 1 == size) {
    if (hi == BC_PERIODIC) { *first += w; *last -= w; }
    else if (hi == BC_PMC) *last -= 2.0f * w;
    return;
  }
  *cur += w;
  *prev -= w;
}

/* one derivative term of an updated sample: psi adjoint and the weight of d */
static inline real_t term_t(const or_grid* g, int cpml, int e_side, int axis, int64_t coord, real_t s, real_t mu,
                            real_t* apsi_l) {
  if (!cpml) return s * mu;
  const real_t kinv = e_side ? g->kinv_e[axis][coord] : g->kinv_h[axis][coord];
  const real_t b = e_side ? g->b_e[axis][coord] : g->b_h[axis][coord];
  const real_t c = e_side ? g->c_e[axis][coord] : g->c_h[axis][coord];
  const real_t tot = *apsi_l + s * mu;
  *apsi_l = b * tot;
  return s * mu * kinv + c * tot;
}

void or_reverse_step(const or_grid* g, real_t* const* ae, real_t* const* ah, real_t* const* apsi, real_t* const* tmp,
                     const or_ade_adj* ad, real_t* const* fe, real_t* const* fh, real_t* const* fpsi,
                     real_t* const* g_decay, real_t* const* g_curl) {
  const int64_t nx = g->nx, ny = g->ny, nz = g->nz;
  const int *lo = g->bc_lo, *hi = g->bc_hi;
  const int cp = g->cpml;
  real_t *hx = ah[0], *hy = ah[1], *hz = ah[2];
  /* ---- E side: ADE subtraction, E update and ADE advance, transposed ---- */
  for (int c = 0; c < 3; ++c) {
    const int64_t sx = c == 0 ? nx - 1 : nx, sy = c == 1 ? ny - 1 : ny, sz = c == 2 ? nz - 1 : nz;
    for (int64_t i = 0; i < sx; ++i)
      for (int64_t j = 0; j < sy; ++j)
        for (int64_t k = 0; k < sz; ++k) {
          const int64_t l = off3(i, j, k, sy, sz);
          const real_t lam = ae[c][l];
          real_t out = 0;
          /* transverse axes (a, b) of this component and the sign of each term */
          const int ax_a = c == 0 ? 1 : 0, ax_b = c == 2 ? 1 : 2;
          const int64_t ca = ax_a == 0 ? i : j, cb = ax_b == 1 ? j : k;
          const int64_t na = ax_a == 0 ? nx : ny, nb = ax_b == 1 ? ny : nz;
          const real_t sa = c == 1 ? -1.0f : 1.0f, sb = -sa;
          const int pec = is_pec(ca, na, lo[ax_a], hi[ax_a]) || is_pec(cb, nb, lo[ax_b], hi[ax_b]);
          const int inactive = is_inactive(ca, na, lo[ax_a], hi[ax_a]) || is_inactive(cb, nb, lo[ax_b], hi[ax_b]);
          if (pec) {
            out = 0;
          } else if (inactive) {
            out = lam;
          } else {
            out = DECAY(c, l) * lam;
            const real_t mu = CURL(c, l) * lam;
            const real_t wa = term_t(g, cp, 1, ax_a, ca, sa, mu, &apsi[2 * c][l]) * g->inv_e[ax_a][ca];
            const real_t wb = term_t(g, cp, 1, ax_b, cb, sb, mu, &apsi[2 * c + 1][l]) * g->inv_e[ax_b][cb];
            /* H lines read by the two differences: a = d(H_pa)/d(ax_a), b = d(H_pb)/d(ax_b) */
            if (c == 0) { /* Ex: a = dHz/dy, b = dHy/dz */
              const int64_t jc = j < ny - 1 ? j : ny - 2, jp = j > 0 ? j - 1 : 0;
              const int64_t kc = k < nz - 1 ? k : nz - 2, kp = k > 0 ? k - 1 : 0;
              bdiff_t(j, ny, lo[1], hi[1], wa, &hz[off3(i, jc, k, ny - 1, nz)], &hz[off3(i, jp, k, ny - 1, nz)],
                      &hz[off3(i, 0, k, ny - 1, nz)], &hz[off3(i, ny - 2, k, ny - 1, nz)]);
              bdiff_t(k, nz, lo[2], hi[2], wb, &hy[off3(i, j, kc, ny, nz - 1)], &hy[off3(i, j, kp, ny, nz - 1)],
                      &hy[off3(i, j, 0, ny, nz - 1)], &hy[off3(i, j, nz - 2, ny, nz - 1)]);
            } else if (c == 1) { /* Ey: a = dHz/dx, b = dHx/dz */
              const int64_t ic = i < nx - 1 ? i : nx - 2, ip = i > 0 ? i - 1 : 0;
              const int64_t kc = k < nz - 1 ? k : nz - 2, kp = k > 0 ? k - 1 : 0;
              bdiff_t(i, nx, lo[0], hi[0], wa, &hz[off3(ic, j, k, ny - 1, nz)], &hz[off3(ip, j, k, ny - 1, nz)],
                      &hz[off3(0, j, k, ny - 1, nz)], &hz[off3(nx - 2, j, k, ny - 1, nz)]);
              bdiff_t(k, nz, lo[2], hi[2], wb, &hx[off3(i, j, kc, ny - 1, nz - 1)],
                      &hx[off3(i, j, kp, ny - 1, nz - 1)], &hx[off3(i, j, 0, ny - 1, nz - 1)],
                      &hx[off3(i, j, nz - 2, ny - 1, nz - 1)]);
            } else { /* Ez: a = dHy/dx, b = dHx/dy */
              const int64_t ic = i < nx - 1 ? i : nx - 2, ip = i > 0 ? i - 1 : 0;
              const int64_t jc = j < ny - 1 ? j : ny - 2, jp = j > 0 ? j - 1 : 0;
              bdiff_t(i, nx, lo[0], hi[0], wa, &hy[off3(ic, j, k, ny, nz - 1)], &hy[off3(ip, j, k, ny, nz - 1)],
                      &hy[off3(0, j, k, ny, nz - 1)], &hy[off3(nx - 2, j, k, ny, nz - 1)]);
              bdiff_t(j, ny, lo[1], hi[1], wb, &hx[off3(i, jc, k, ny - 1, nz - 1)],
                      &hx[off3(i, jp, k, ny - 1, nz - 1)], &hx[off3(i, 0, k, ny - 1, nz - 1)],
                      &hx[off3(i, ny - 2, k, ny - 1, nz - 1)]);
            }
            if (fe && g_curl) {
              /* curl^n of the forward update at this sample, recomputed */
              const real_t* fa = c == 0 ? fh[2] : (c == 1 ? fh[2] : fh[1]);
              const real_t* fb = c == 0 ? fh[1] : (c == 1 ? fh[0] : fh[0]);
              real_t da, db;
              if (c == 0) {
                const int64_t jc = j < ny - 1 ? j : ny - 2, jp = j > 0 ? j - 1 : 0;
                const int64_t kc = k < nz - 1 ? k : nz - 2, kp = k > 0 ? k - 1 : 0;
                da = bdiff(j, ny, lo[1], hi[1], g->inv_e[1][j], fa[off3(i, jc, k, ny - 1, nz)],
                           fa[off3(i, jp, k, ny - 1, nz)], fa[off3(i, 0, k, ny - 1, nz)],
                           fa[off3(i, ny - 2, k, ny - 1, nz)]);
                db = bdiff(k, nz, lo[2], hi[2], g->inv_e[2][k], fb[off3(i, j, kc, ny, nz - 1)],
                           fb[off3(i, j, kp, ny, nz - 1)], fb[off3(i, j, 0, ny, nz - 1)],
                           fb[off3(i, j, nz - 2, ny, nz - 1)]);
              } else if (c == 1) {
                const int64_t ic = i < nx - 1 ? i : nx - 2, ip = i > 0 ? i - 1 : 0;
                const int64_t kc = k < nz - 1 ? k : nz - 2, kp = k > 0 ? k - 1 : 0;
                da = bdiff(i, nx, lo[0], hi[0], g->inv_e[0][i], fa[off3(ic, j, k, ny - 1, nz)],
                           fa[off3(ip, j, k, ny - 1, nz)], fa[off3(0, j, k, ny - 1, nz)],
                           fa[off3(nx - 2, j, k, ny - 1, nz)]);
                db = bdiff(k, nz, lo[2], hi[2], g->inv_e[2][k], fb[off3(i, j, kc, ny - 1, nz - 1)],
                           fb[off3(i, j, kp, ny - 1, nz - 1)], fb[off3(i, j, 0, ny - 1, nz - 1)],
                           fb[off3(i, j, nz - 2, ny - 1, nz - 1)]);
              } else {
                const int64_t ic = i < nx - 1 ? i : nx - 2, ip = i > 0 ? i - 1 : 0;
                const int64_t jc = j < ny - 1 ? j : ny - 2, jp = j > 0 ? j - 1 : 0;
                da = bdiff(i, nx, lo[0], hi[0], g->inv_e[0][i], fa[off3(ic, j, k, ny, nz - 1)],
                           fa[off3(ip, j, k, ny, nz - 1)], fa[off3(0, j, k, ny, nz - 1)],
                           fa[off3(nx - 2, j, k, ny, nz - 1)]);
                db = bdiff(j, ny, lo[1], hi[1], g->inv_e[1][j], fb[off3(i, jc, k, ny - 1, nz - 1)],
                           fb[off3(i, jp, k, ny - 1, nz - 1)], fb[off3(i, 0, k, ny - 1, nz - 1)],
                           fb[off3(i, ny - 2, k, ny - 1, nz - 1)]);
              }
              real_t ta = da, tb = db;
              if (cp) {
                ta = da * g->kinv_e[ax_a][ca] + fpsi[2 * c][l];
                tb = db * g->kinv_e[ax_b][cb] + fpsi[2 * c + 1][l];
              }
              const real_t curl = sa * (ta - tb);
              g_curl[c][l] += lam * curl;
              g_decay[c][l] += lam * fe[c][l];
            }
          }
          /* ADE: subtraction E -= dt*J'*inv_eps, then the advance from E^n */
          if (ad) {
            const real_t lj = -(real_t)ad->dt * ad->inv_eps[c][l] * lam;
            for (int p = 0; p < ad->count[c]; ++p) {
              const real_t d = (real_t)ad->decay[c][p], r = (real_t)ad->restoring[c][p];
              const real_t gdr = ad->drive[c][p][l];
              real_t* apol = ad->apol[c][p];
              real_t* acur = ad->acur[c][p];
              if (ad->kind[c][p] == 0) {
                /* J' = (P' - P)/dt feeds the subtraction; J^n itself is never read */
                const real_t ljt = acur[l] + lj;
                const real_t lp = apol[l] + ljt * (real_t)(1.0 / ad->dt);
                apol[l] = d * lp - ljt * (real_t)(1.0 / ad->dt);
                acur[l] = 0;
                out += gdr * lp;
              } else if (ad->kind[c][p] == 1) {
                const real_t lc = acur[l] + lj;
                acur[l] = d * lc;
                out += gdr * lc;
              } else {
                const real_t lpol = apol[l];
                const real_t lnc = acur[l] + lj + (real_t)ad->dt * lpol;
                acur[l] = d * lnc;
                apol[l] = lpol - r * lnc;
                out += gdr * lnc;
              }
            }
          }
          tmp[c][l] = out;
        }
  }
  /* ---- H side: H update transposed; scatter into the E adjoint ---- */
  real_t *tx = tmp[0], *ty = tmp[1], *tz = tmp[2];
  /* Hx (nx, ny-1, nz-1): a = dEz/dy (+), b = dEy/dz (-) */
  for (int64_t i = 0; i < nx; ++i)
    for (int64_t j = 0; j < ny - 1; ++j)
      for (int64_t k = 0; k < nz - 1; ++k) {
        const int64_t l = off3(i, j, k, ny - 1, nz - 1);
        const real_t lam = hx[l];
        const real_t nu = -CURL(3, l) * lam;
        const real_t wa = term_t(g, cp, 0, 1, j, 1.0f, nu, &apsi[6][l]) * g->inv_h[1][j];
        const real_t wb = term_t(g, cp, 0, 2, k, -1.0f, nu, &apsi[7][l]) * g->inv_h[2][k];
        tz[off3(i, j + 1, k, ny, nz - 1)] += wa;
        tz[off3(i, j, k, ny, nz - 1)] -= wa;
        ty[off3(i, j, k + 1, ny - 1, nz)] += wb;
        ty[off3(i, j, k, ny - 1, nz)] -= wb;
        hx[l] = DECAY(3, l) * lam;
      }
  /* Hy (nx-1, ny, nz-1): a = dEz/dx (-), b = dEx/dz (+) */
  for (int64_t i = 0; i < nx - 1; ++i)
    for (int64_t j = 0; j < ny; ++j)
      for (int64_t k = 0; k < nz - 1; ++k) {
        const int64_t l = off3(i, j, k, ny, nz - 1);
        const real_t lam = hy[l];
        const real_t nu = -CURL(4, l) * lam;
        const real_t wa = term_t(g, cp, 0, 0, i, -1.0f, nu, &apsi[8][l]) * g->inv_h[0][i];
        const real_t wb = term_t(g, cp, 0, 2, k, 1.0f, nu, &apsi[9][l]) * g->inv_h[2][k];
        tz[off3(i + 1, j, k, ny, nz - 1)] += wa;
        tz[off3(i, j, k, ny, nz - 1)] -= wa;
        tx[off3(i, j, k + 1, ny, nz)] += wb;
        tx[off3(i, j, k, ny, nz)] -= wb;
        hy[l] = DECAY(4, l) * lam;
      }
  /* Hz (nx-1, ny-1, nz): a = dEy/dx (+), b = dEx/dy (-) */
  for (int64_t i = 0; i < nx - 1; ++i)
    for (int64_t j = 0; j < ny - 1; ++j)
      for (int64_t k = 0; k < nz; ++k) {
        const int64_t l = off3(i, j, k, ny - 1, nz);
        const real_t lam = hz[l];
        const real_t nu = -CURL(5, l) * lam;
        const real_t wa = term_t(g, cp, 0, 0, i, 1.0f, nu, &apsi[10][l]) * g->inv_h[0][i];
        const real_t wb = term_t(g, cp, 0, 1, j, -1.0f, nu, &apsi[11][l]) * g->inv_h[1][j];
        ty[off3(i + 1, j, k, ny - 1, nz)] += wa;
        ty[off3(i, j, k, ny - 1, nz)] -= wa;
        tx[off3(i, j + 1, k, ny, nz)] += wb;
        tx[off3(i, j, k, ny, nz)] -= wb;
        hz[l] = DECAY(5, l) * lam;
      }
  for (int c = 0; c < 3; ++c) {
    const int64_t n = (int64_t)(c == 0 ? nx - 1 : nx) * (c == 1 ? ny - 1 : ny) * (c == 2 ? nz - 1 : nz);
    memcpy(ae[c], tmp[c], (size_t)n * sizeof(real_t));
  }
}
