// Per-step helper kernels of a scene batch.  Every kernel reads the step index
// from a device counter (advanced by step_counter_kernel at the end of each
// step), so the whole per-step sequence can be captured once in a hipGraph and
// replayed with no host arguments.  Each kernel covers all scenes of the batch
// in one launch through a device table of per-scene descriptors.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "maxwell_fdtd.h"
#include "terms.h"
#include "yee.h"

namespace mw {

// The counter is written by lane 0 with an ordinary vector store.
__global__ __launch_bounds__(64) void step_counter_kernel(int* counter, int set, int value) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  counter[0] = set ? value : counter[0] + 1;
}

struct BatchTerm {
  mw_term t;  // a SHIFTED term of some scene (values buffer owned by that scene)
  double dt;
};

__global__ __launch_bounds__(256) void shifted_values_batch_kernel(const BatchTerm* __restrict__ terms, int count,
                                                                   const int* counter) {
  const int ti = blockIdx.y;
  if (ti >= count) return;
  const BatchTerm& b = terms[ti];
  const mw_term& t = b.t;
  const int64_t cells = (int64_t)t.shape[0] * t.shape[1] * t.shape[2];
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= cells) return;
  const int n = *counter;
  const double time = (double)n * b.dt + t.t_offset - G(t.delay)[p];
  G(t.values)[p] = (float)(waveform(t.wave_kind, time, t.frequency, t.fwidth, t.amplitude, t.phase, t.delay0) *
                           (double)G(t.patch)[p]);
}

__global__ __launch_bounds__(256) void aux_line_h_batch_kernel(const mw_aux_line* __restrict__ lines, int count) {
  const int s = blockIdx.y;
  if (s >= count) return;
  const mw_aux_line a = lines[s];
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= a.n_e - 1) return;
  G(a.h)[m] = G(a.decay_h)[m] * G(a.h)[m] - G(a.curl_h)[m] * (G(a.e)[m + 1] - G(a.e)[m]);
}

__global__ __launch_bounds__(256) void aux_line_e_batch_kernel(const mw_aux_line* __restrict__ lines, int count,
                                                               const int* counter) {
  const int s = blockIdx.y;
  if (s >= count) return;
  const mw_aux_line a = lines[s];
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= a.n_e) return;
  if (m == a.n_e - 1) { G(a.e)[m] = 0.0f; return; }
  if (m == 0) { G(a.e)[0] = G(a.source)[*counter]; return; }
  G(a.e)[m] = G(a.decay_e)[m] * G(a.e)[m] - G(a.curl_e)[m] * (G(a.h)[m] - G(a.h)[m - 1]);
}

// dft_column (terms.h) on global-memory views of table-held pointers.
__device__ __forceinline__ void plane_dft_column(float x, int nf, int64_t n, int64_t l,
                                                 gptr<const float> __restrict__ wc, gptr<const float> __restrict__ ws,
                                                 gptr<float> __restrict__ re, gptr<float> __restrict__ im) {
  for (int f = 0; f < nf; ++f) {
    re[f * n + l] += x * wc[f];
    im[f * n + l] += x * ws[f];
  }
}

// One observer plane of one scene.  field[p] is the component buffer that holds
// the state after a step of parity p (the single-pass step's output buffer).
struct BatchPlane {
  int32_t component, axis, index, nf;
  int32_t shape[3];
  int32_t pad;
  const float* field[2];
  float* re;
  float* im;
  const float* w_cos;  // (steps, nf) table of the component's family
  const float* w_sin;
};

__global__ __launch_bounds__(256) void accumulate_planes_batch_kernel(const BatchPlane* __restrict__ planes,
                                                                      int count, const int* counter) {
  const int qi = blockIdx.y;
  if (qi >= count) return;
  const BatchPlane& q = planes[qi];
  const int sx = q.shape[0], sy = q.shape[1], sz = q.shape[2];
  int na, nb;
  if (q.axis == 0) { na = sy; nb = sz; } else if (q.axis == 1) { na = sx; nb = sz; } else { na = sx; nb = sy; }
  const int64_t cells = (int64_t)na * nb;
  const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= cells) return;
  const int n = *counter;
  const int ia = (int)(l / nb), ib = (int)(l % nb);
  uint32_t o;
  if (q.axis == 0) o = ((uint32_t)q.index * sy + ia) * sz + ib;
  else if (q.axis == 1) o = ((uint32_t)ia * sy + q.index) * sz + ib;
  else o = ((uint32_t)ia * sy + ib) * sz + q.index;
  const float x = G(q.field[n & 1])[o];
  plane_dft_column(x, q.nf, cells, l, G(q.w_cos + (int64_t)n * q.nf), G(q.w_sin + (int64_t)n * q.nf), G(q.re),
                   G(q.im));
}

struct BatchPoint {
  int32_t linear, nf, stride, pad;  // stride = the scene's point count (re/im are (nf, stride))
  const float* field[2];
  float* re;  // already offset to this point's column
  float* im;
  const float* w_cos;
  const float* w_sin;
};

__global__ __launch_bounds__(64) void accumulate_points_batch_kernel(const BatchPoint* __restrict__ points, int count,
                                                                     const int* counter) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= count) return;
  const BatchPoint& q = points[p];
  const int n = *counter;
  const float x = G(q.field[n & 1])[q.linear];
  const auto wc = G(q.w_cos + (int64_t)n * q.nf);
  const auto ws = G(q.w_sin + (int64_t)n * q.nf);
  for (int f = 0; f < q.nf; ++f) {
    G(q.re)[(int64_t)f * q.stride] += x * wc[f];
    G(q.im)[(int64_t)f * q.stride] += x * ws[f];
  }
}

}  // namespace mw
