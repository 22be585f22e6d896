// Source terms and running-DFT accumulation kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "maxwell_fdtd.h"

namespace mw {

// Waveforms of kernels/sources.cu:71-93, evaluated in double: the argument
// t - delay spans thousands of periods, where float32 loses ~1e-5 of phase.
__device__ __forceinline__ double waveform(int kind, double t, double freq, double fwidth, double amp, double phase,
                                           double delay0) {
  const double two_pi = 6.283185307179586476925286766559;
  if (kind == 0) return amp * cos(two_pi * freq * t + phase);
  const double tau = t - delay0;
  if (kind == 1) {
    const double inv_sigma = fmax(two_pi * fwidth, 1e-30);
    const double x = tau * inv_sigma;
    return amp * exp(-0.5 * x * x) * cos(two_pi * freq * tau + phase);
  }
  const double a = 3.14159265358979323846 * freq * tau;
  return amp * (1.0 - 2.0 * a * a) * exp(-a * a);
}

struct TermSet {
  int32_t count;
  int32_t pad;
  mw_term t[MW_MAX_TERMS];
};

struct FieldSet {
  float* ptr[6];
  int32_t shape[6][3];
};

// One launch applies every term of a set: blockIdx.y = term, x = patch cell.
__global__ __launch_bounds__(256) void apply_terms_kernel(TermSet set, FieldSet fs, int step, double dt) {
  const int ti = blockIdx.y;
  if (ti >= set.count) return;
  const mw_term& t = set.t[ti];
  const int64_t cells = (int64_t)t.shape[0] * t.shape[1] * t.shape[2];
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= cells) return;
  const int pz = (int)(p % t.shape[2]);
  const int py = (int)((p / t.shape[2]) % t.shape[1]);
  const int px = (int)(p / ((int64_t)t.shape[2] * t.shape[1]));
  const int c = t.component;
  const int fy = fs.shape[c][1], fz = fs.shape[c][2];
  const uint32_t o = ((uint32_t)(t.offset[0] + px) * fy + (t.offset[1] + py)) * fz + (t.offset[2] + pz);
  float value;
  if (t.kind == MW_TERM_ADDITIVE) {
    value = t.signal_table[step] * t.patch[p];
  } else if (t.kind == MW_TERM_LINE) {
    const int pa = t.line_axis == 0 ? px : (t.line_axis == 1 ? py : pz);
    value = t.patch[p] * t.line[t.line_axis < 0 ? t.line_base : t.line_base + t.line_step * pa];
  } else {
    const double time = (double)step * dt + t.t_offset - t.delay[p];
    value = (float)(waveform(t.wave_kind, time, t.frequency, t.fwidth, t.amplitude, t.phase, t.delay0) *
                    (double)t.patch[p]);
  }
  fs.ptr[c][o] += value;
}

// Per-step waveform * patch of every time-shifted term of a set, written to the
// term's value buffer (the fused step adds these instead of evaluating the
// double-precision waveform per sample).
__global__ __launch_bounds__(256) void shifted_values_kernel(TermSet set, int step, double dt) {
  const int ti = blockIdx.y;
  if (ti >= set.count) return;
  const mw_term& t = set.t[ti];
  if (t.kind != MW_TERM_SHIFTED) return;
  const int64_t cells = (int64_t)t.shape[0] * t.shape[1] * t.shape[2];
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= cells) return;
  const double time = (double)step * dt + t.t_offset - t.delay[p];
  t.values[p] = (float)(waveform(t.wave_kind, time, t.frequency, t.fwidth, t.amplitude, t.phase, t.delay0) *
                        (double)t.patch[p]);
}

__global__ __launch_bounds__(256) void add_patch_kernel(float* field, int fy, int fz, const float* patch, int pnx,
                                                         int pny, int pnz, int oi, int oj, int ok, float signal) {
  const int64_t cells = (int64_t)pnx * pny * pnz;
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= cells) return;
  const int pz = (int)(p % pnz);
  const int py = (int)((p / pnz) % pny);
  const int px = (int)(p / ((int64_t)pnz * pny));
  field[((uint32_t)(oi + px) * fy + (oj + py)) * fz + (ok + pz)] += signal * patch[p];
}

__global__ __launch_bounds__(256) void add_shifted_kernel(float* field, int fy, int fz, const float* patch,
                                                           const double* delay, int pnx, int pny, int pnz, int oi,
                                                           int oj, int ok, double time, int kind, double freq,
                                                           double fwidth, double amp, double phase, double delay0) {
  const int64_t cells = (int64_t)pnx * pny * pnz;
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= cells) return;
  const int pz = (int)(p % pnz);
  const int py = (int)((p / pnz) % pny);
  const int px = (int)(p / ((int64_t)pnz * pny));
  const double s = waveform(kind, time - delay[p], freq, fwidth, amp, phase, delay0);
  field[((uint32_t)(oi + px) * fy + (oj + py)) * fz + (ok + pz)] += (float)(s * (double)patch[p]);
}

// ------------------------------------------------------------ TFSF line
// One workgroup-free pass each (the line holds a few hundred samples).
__global__ __launch_bounds__(256) void aux_line_h_kernel(mw_aux_line a) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= a.n_e - 1) return;
  a.h[m] = a.decay_h[m] * a.h[m] - a.curl_h[m] * (a.e[m + 1] - a.e[m]);
}

__global__ __launch_bounds__(256) void aux_line_e_kernel(mw_aux_line a, int step) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= a.n_e) return;
  if (m == a.n_e - 1) { a.e[m] = 0.0f; return; }
  if (m == 0) { a.e[0] = a.source[step]; return; }
  a.e[m] = a.decay_e[m] * a.e[m] - a.curl_e[m] * (a.h[m] - a.h[m - 1]);
}

// Standalone auxiliary-line updates with the reference's per-op arguments
// (kernels/sources.cu:803-839): the hard source at `source_index`, the far end
// held at zero, sample 0 left alone unless it is the source.
__global__ __launch_bounds__(256) void aux_magnetic_kernel(float* __restrict__ h, const float* __restrict__ e,
                                                           const float* __restrict__ decay,
                                                           const float* __restrict__ curl, int n_h) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n_h) return;
  h[m] = decay[m] * h[m] - curl[m] * (e[m + 1] - e[m]);
}

__global__ __launch_bounds__(256) void aux_electric_kernel(float* __restrict__ e, const float* __restrict__ h,
                                                           const float* __restrict__ decay,
                                                           const float* __restrict__ curl, int n_e, int source_index,
                                                           float source_value) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n_e) return;
  if (m == n_e - 1) { e[m] = 0.0f; return; }
  if (m == source_index) { e[m] = source_value; return; }
  if (m > 0) e[m] = decay[m] * e[m] - curl[m] * (h[m] - h[m - 1]);
}

// Standalone pole update over every sample (kernels/dispersive.cu:18-70):
// Debye P' = d P + g E, J = (P' - P)/dt; Drude J' = d J + g E; Lorentz
// J' = d J - r P + g E, P' = P + dt J'.  The fused E passes use ade_advance.
__global__ __launch_bounds__(256) void pole_current_kernel(int kind, const float* __restrict__ e,
                                                           const float* __restrict__ drive, float* __restrict__ pol,
                                                           float* __restrict__ cur, int64_t n, float decay,
                                                           float restoring, float dt) {
  const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= n) return;
  const float g = drive[l] * e[l];
  if (kind == 0) {
    const float prev = pol[l];
    const float next = decay * prev + g;
    pol[l] = next;
    cur[l] = (next - prev) * (1.0f / dt);
  } else if (kind == 1) {
    cur[l] = decay * cur[l] + g;
  } else {
    const float nc = decay * cur[l] - restoring * pol[l] + g;
    cur[l] = nc;
    pol[l] = pol[l] + dt * nc;
  }
}

// E -= dt J / eps_inf (kernels/dispersive.cu:72-83).
__global__ __launch_bounds__(256) void apply_polarization_kernel(float* __restrict__ e, const float* __restrict__ cur,
                                                                 const float* __restrict__ inv_eps, int64_t n,
                                                                 float dt) {
  const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= n) return;
  e[l] -= dt * cur[l] * inv_eps[l];
}

// ------------------------------------------------------------ spectral sums
// re[f][l] += v[l]*wc[f]; the weights of one step are read once per block.
__global__ __launch_bounds__(256) void accumulate_kernel(const float* __restrict__ v, int64_t n, int nf,
                                                         const float* __restrict__ wc, const float* __restrict__ ws,
                                                         float* __restrict__ re, float* __restrict__ im) {
  const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= n) return;
  const float x = v[l];
  for (int f = 0; f < nf; ++f) {
    const float c = wc[f], s = ws[f];
    re[f * n + l] += x * c;
    im[f * n + l] += x * s;
  }
}

struct PlaneSet {
  int32_t count;
  int32_t pad;
  mw_plane_obs p[MW_MAX_PLANES];
};

// Running DFT of one observed sample x into its column l of (nf, n) re/im sums.
// The sums are restrict-qualified parameters, so all 2 nf loads issue before the
// stores instead of each waiting for the previous store.
__device__ __forceinline__ void dft_column(float x, int nf, int64_t n, int64_t l, const float* __restrict__ wc,
                                           const float* __restrict__ ws, float* __restrict__ re,
                                           float* __restrict__ im) {
  for (int f = 0; f < nf; ++f) {
    re[f * n + l] += x * wc[f];
    im[f * n + l] += x * ws[f];
  }
}

// All plane observers of one step in one launch: blockIdx.y = plane.
__global__ __launch_bounds__(256) void accumulate_planes_kernel(PlaneSet set, FieldSet fs, int nf,
                                                                const float* __restrict__ e_cos,
                                                                const float* __restrict__ e_sin,
                                                                const float* __restrict__ h_cos,
                                                                const float* __restrict__ h_sin) {
  const int qi = blockIdx.y;
  if (qi >= set.count) return;
  const mw_plane_obs& q = set.p[qi];
  const int c = q.component;
  const int sx = fs.shape[c][0], sy = fs.shape[c][1], sz = fs.shape[c][2];
  int a, b;
  if (q.axis == 0) { a = sy; b = sz; } else if (q.axis == 1) { a = sx; b = sz; } else { a = sx; b = sy; }
  const int64_t n = (int64_t)a * b;
  const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= n) return;
  const int ia = (int)(l / b), ib = (int)(l % b);
  uint32_t o;
  if (q.axis == 0) o = ((uint32_t)q.index * sy + ia) * sz + ib;
  else if (q.axis == 1) o = ((uint32_t)ia * sy + q.index) * sz + ib;
  else o = ((uint32_t)ia * sy + ib) * sz + q.index;
  const float x = fs.ptr[c][o];
  dft_column(x, nf, n, l, c < 3 ? e_cos : h_cos, c < 3 ? e_sin : h_sin, q.re, q.im);
}

// z-normal plane monitors, staged.  Sampling a z-normal plane reads one float per
// (x, y) column of a z-fastest field, a cache line each, and adding it to nf
// complex running sums per step reads and writes 16 nf bytes per sample.  Staged,
// each step only copies the plane's samples into a stage row (gather below), and
// once per block of steps one pass adds the block's rows to the sums (flush), so
// the sums move once per block instead of once per step.  Plane q's samples of
// stage row t sit at stage + t * row_floats + off, in the layout of its sums
// (nf, cells): sample (i, j) at i * sy + j.
constexpr int kMaxStageDepth = 64;

// The staged (component, column) pairs, each listed once; row = this step's row.
struct ZGatherSet {
  int32_t count;
  int32_t pad;
  float* row;
  int32_t comp[MW_MAX_PLANES];
  int32_t k[MW_MAX_PLANES];
  int64_t off[MW_MAX_PLANES];
  // samples in [own_lo, own_hi) (x, y, z node box) were stored by the interior
  // tiles of the step and are skipped (an empty box: every sample is gathered)
  int32_t own_lo[3], own_hi[3];
};

// row[off + i * sy + j] = component(i, j, k) for each staged pair.  Thread t
// takes pair t % count of sample t / count: the pairs of one sample sit in one
// wavefront, so pairs that share a cache line (one component at neighbouring
// columns, as a flux plane's H samples) fetch it once, when it is read, instead of
// once per pair in passes far apart.
__device__ __forceinline__ void gather_zplane_sample(const ZGatherSet& set, const FieldSet& fs, int64_t samples,
                                                     int64_t t) {
  if (t >= samples * set.count) return;
  const int q = (int)(t % set.count);
  const int64_t l = t / set.count;
  const int c = set.comp[q];
  if (l >= (int64_t)fs.shape[c][0] * fs.shape[c][1]) return;
  const int kq = set.k[q];
  if (kq >= set.own_lo[2] && kq < set.own_hi[2]) {
    const uint32_t lu = (uint32_t)l, rows = (uint32_t)fs.shape[c][1];  // a plane holds < 2^32 samples
    const int i = (int)(lu / rows), j = (int)(lu % rows);
    if (i >= set.own_lo[0] && i < set.own_hi[0] && j >= set.own_lo[1] && j < set.own_hi[1]) return;
  }
  set.row[set.off[q] + l] = fs.ptr[c][l * fs.shape[c][2] + kq];
}

__global__ __launch_bounds__(256) void gather_zplanes_kernel(ZGatherSet set, FieldSet fs, int64_t samples) {
  gather_zplane_sample(set, fs, samples, (int64_t)blockIdx.x * blockDim.x + threadIdx.x);
}

// One step's z-plane gather handed to a later launch: the stage row of that
// step, the fields it wrote and the samples per plane (count = 0: none).
struct ZGatherJob {
  ZGatherSet set;
  FieldSet fs;
  int64_t samples;
  int32_t blocks;  // 256-thread blocks that cover samples * set.count
  int32_t pad;
};

struct ZFlushPlane {
  int32_t component;
  int32_t pad;
  int64_t off;
  int64_t cells;
  float* re;
  float* im;
};
struct ZFlushSet {
  int32_t count;
  int32_t pad;
  int64_t row_floats;
  const float* stage;
  ZFlushPlane p[MW_MAX_PLANES];
};

// Fold stage rows [0, rows) -- the steps first .. first+rows-1, weights wc/ws
// already offset to step `first` -- into the running DFT sums: for each sample the
// rows are added in step order with one fused multiply-add per step, the same
// chain of roundings as adding each step as it is taken.  Planes that observe the
// same (component, column) read the same stage samples.  blockIdx.y = plane.
__global__ __launch_bounds__(256) void flush_zstage_kernel(ZFlushSet set, int rows, int nf,
                                                           const float* __restrict__ e_cos,
                                                           const float* __restrict__ e_sin,
                                                           const float* __restrict__ h_cos,
                                                           const float* __restrict__ h_sin) {
  const int qi = blockIdx.y;
  if (qi >= set.count) return;
  const ZFlushPlane& q = set.p[qi];
  const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= q.cells) return;
  const float* __restrict__ wc = q.component < 3 ? e_cos : h_cos;
  const float* __restrict__ ws = q.component < 3 ? e_sin : h_sin;
  const float* __restrict__ col = set.stage + q.off + l;
  float v[kMaxStageDepth];
#pragma unroll
  for (int t = 0; t < kMaxStageDepth; ++t) v[t] = t < rows ? col[t * set.row_floats] : 0.0f;
  for (int f = 0; f < nf; ++f) {
    float re = q.re[f * q.cells + l], im = q.im[f * q.cells + l];
#pragma unroll
    for (int t = 0; t < kMaxStageDepth; ++t) {
      if (t < rows) {
        re = __builtin_fmaf(v[t], wc[t * nf + f], re);
        im = __builtin_fmaf(v[t], ws[t * nf + f], im);
      }
    }
    q.re[f * q.cells + l] = re;
    q.im[f * q.cells + l] = im;
  }
}

// p[0, n) = 0 (folded stage rows of a block that a later call continues).
__global__ __launch_bounds__(256) void clear_floats_kernel(float* __restrict__ p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0f;
}

// The point observers of one step: the fields that step wrote and the step's
// weight rows.  Point p adds x * w[f] to its sums (one fused multiply-add each,
// wherever the sample is taken).
struct PointSet {
  const float* ptr[6];
  const int32_t* comp;
  const int32_t* lin;
  const float* e_cos;
  const float* e_sin;
  const float* h_cos;
  const float* h_sin;
  float* re;
  float* im;
  int32_t count, nf;
};

__device__ __forceinline__ void accumulate_point(const PointSet& ps, int p) {
  const int c = ps.comp[p];
  const float x = ps.ptr[c][ps.lin[p]];
  const float* wc = c < 3 ? ps.e_cos : ps.h_cos;
  const float* ws = c < 3 ? ps.e_sin : ps.h_sin;
  for (int f = 0; f < ps.nf; ++f) {
    const int o = f * ps.count + p;
    ps.re[o] = fmaf(x, wc[f], ps.re[o]);
    ps.im[o] = fmaf(x, ws[f], ps.im[o]);
  }
}

__global__ __launch_bounds__(64) void accumulate_points_kernel(PointSet ps) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < ps.count) accumulate_point(ps, p);
}

__global__ __launch_bounds__(64) void accumulate_points_single_kernel(const float* field, const int32_t* lin,
                                                                      int count, int nf, const float* wc,
                                                                      const float* ws, float* re, float* im) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= count) return;
  const float x = field[lin[p]];
  for (int f = 0; f < nf; ++f) {
    re[f * count + p] += x * wc[f];
    im[f * count + p] += x * ws[f];
  }
}

__global__ __launch_bounds__(256) void accumulate_plane_single_kernel(const float* field, int sx, int sy, int sz,
                                                                      int axis, int index, int nf, const float* wc,
                                                                      const float* ws, float* re, float* im) {
  int a, b;
  if (axis == 0) { a = sy; b = sz; } else if (axis == 1) { a = sx; b = sz; } else { a = sx; b = sy; }
  const int64_t n = (int64_t)a * b;
  const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= n) return;
  const int ia = (int)(l / b), ib = (int)(l % b);
  uint32_t o;
  if (axis == 0) o = ((uint32_t)index * sy + ia) * sz + ib;
  else if (axis == 1) o = ((uint32_t)ia * sy + index) * sz + ib;
  else o = ((uint32_t)ia * sy + ib) * sz + index;
  const float x = field[o];
  for (int f = 0; f < nf; ++f) {
    re[f * n + l] += x * wc[f];
    im[f * n + l] += x * ws[f];
  }
}

}  // namespace mw
