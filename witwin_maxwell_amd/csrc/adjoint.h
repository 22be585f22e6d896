// Reverse (adjoint) step kernels: the exact transpose of one forward step
// (H update, ADE advance, E update, ADE subtraction), in gather form so that no
// two threads write the same sample (no atomics on the field adjoints).
//
// Reference: fdtd/cuda/kernels/adjoint.cu:149-420 (standard reverse phases),
// :1224-1452 (CPML psi adjoints), :898-979 (Debye/Drude/Lorentz currents),
// :1737-1870 (seed injection); phase order fdtd/adjoint/native.py:48-727.
//
// One reverse step n, with lambda holding the adjoint of the state after step n:
//   1. E seeds of step n           lambda_E += sum_f gRe w_cos[n,f] + gIm w_sin[n,f]
//   2. reverse_h (P1), per H sample:
//        finish step n+1's H transpose (lambda_H <- decay_H lambda_H and the
//        H-side psi adjoints, deferred so P2 of step n+1 could read them), then
//        lambda_H += D_b^T q(lambda_E, lambda_psi_E)      (E update transposed)
//   3. H seeds of step n
//   4. coefficient gradients over the design window (optional)
//   5. reverse_e (P2), per E sample: boundary mask, decay, E-side psi adjoints,
//        ADE subtraction/advance transposes, lambda_E += D_f^T r(lambda_H, ...)
//        (H update transposed)
// Finishing the H transpose of step n+1 inside P1 of step n is exact because an
// all-zero adjoint finishes to zero, so the first reverse step needs no flag.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "maxwell_fdtd.h"
#include "terms.h"
#include "yee.h"

namespace mw {

struct RevArgs {
  mw_grid g;
  mw_coeffs cf;
  mw_cpml cp;   // CPML profiles of the forward run; psi[] are the adjoint psi slabs
  mw_fields a;  // adjoint fields
  mw_ade ad;    // forward pole constants; polarization/current are adjoint states
  int32_t cpml;
  int32_t ade;
  // per axis, the coordinates [in_lo, in_hi) whose reverse update involves no face,
  // no psi and no sample outside the grid (empty: lo >= hi)
  int32_t in_lo[3], in_hi[3];
  // the nodes outside that box as at most six boxes for the general launch: box b
  // covers [lo, hi), with 2^w_shift lanes along z per row, one plane per block
  int32_t nboxes, total;
  int32_t box_lo[6][3], box_hi[6][3], box_shift[6], box_nbz[6], box_nby[6], box_first[6];
};

__device__ __forceinline__ int axis_size(const mw_grid& g, int axis) {
  return axis == 0 ? g.nx : (axis == 1 ? g.ny : g.nz);
}

__device__ __forceinline__ void e_shape(const mw_grid& g, int c, int& sx, int& sy, int& sz) {
  sx = c == 0 ? g.nx - 1 : g.nx;
  sy = c == 1 ? g.ny - 1 : g.ny;
  sz = c == 2 ? g.nz - 1 : g.nz;
}
__device__ __forceinline__ void h_shape(const mw_grid& g, int h, int& sx, int& sy, int& sz) {
  sx = h == 0 ? g.nx : g.nx - 1;
  sy = h == 1 ? g.ny : g.ny - 1;
  sz = h == 2 ? g.nz : g.nz - 1;
}

// Derivative terms of the E update (kernels/electric.cu:463-553):
//   Ex: a = dHz/dy, b = dHy/dz, curl = a - b
//   Ey: a = dHz/dx, b = dHx/dz, curl = b - a
//   Ez: a = dHy/dx, b = dHx/dy, curl = a - b
__device__ __forceinline__ int e_axis(int c, int t) { return t == 0 ? (c == 0 ? 1 : 0) : (c == 2 ? 1 : 2); }
__device__ __forceinline__ float e_sign(int c, int t) { return ((c == 1) != (t == 1)) ? -1.0f : 1.0f; }
// Derivative terms of the H update (kernels/magnetic.cu:62-117):
//   Hx: a = dEz/dy, b = dEy/dz, curl = a - b
//   Hy: a = dEz/dx, b = dEx/dz, curl = b - a
//   Hz: a = dEy/dx, b = dEx/dy, curl = a - b
__device__ __forceinline__ int h_axis(int h, int t) { return t == 0 ? (h == 0 ? 1 : 0) : (h == 2 ? 1 : 2); }
__device__ __forceinline__ float h_sign(int h, int t) { return ((h == 1) != (t == 1)) ? -1.0f : 1.0f; }

__device__ __forceinline__ int pick(int axis, int i, int j, int k) { return axis == 0 ? i : (axis == 1 ? j : k); }

__device__ __forceinline__ float* psi_at(const mw_cpml& cp, int q, int axis, int coord, int i, int j, int k, int sy,
                                         int sz) {
  const int local = slab_local(coord, cp.psi_low[q], cp.psi_high_start[q], cp.psi_high_len[q]);
  if (local < 0) return nullptr;
  return cp.psi[q] + slab_offset(axis, i, j, k, local, sy, sz, cp.psi_low[q] + cp.psi_high_len[q]);
}

// Is E sample (c; i,j,k) held by its boundary mode (PEC -> 0, none/PML face ->
// unchanged)?  Returns 0 active, 1 PEC, 2 inactive (kernels/electric.cu:23-51).
__device__ __forceinline__ int e_mask(const mw_grid& g, int c, int i, int j, int k) {
  int m = 0;
  for (int t = 0; t < 2; ++t) {
    const int ax = e_axis(c, t);
    const int x = pick(ax, i, j, k), n = axis_size(g, ax), lo = g.bc_lo[ax], hi = g.bc_hi[ax];
    if ((x == 0 && lo == MW_BC_PEC) || (x == n - 1 && hi == MW_BC_PEC)) return 1;
    if ((x == 0 && (lo == MW_BC_NONE || lo == MW_BC_PML)) || (x == n - 1 && (hi == MW_BC_NONE || hi == MW_BC_PML)))
      m = 2;
  }
  return m;
}

// Adjoint weight of the backward difference of E sample (c; i,j,k), term t,
// times its spacing reciprocal: q = (s mu kinv + c_e (lambda_psi + s mu)) inv_e
// with mu = curl_E lambda_E; zero for held samples.
template <bool CPML>
__device__ __forceinline__ float q_e(const RevArgs& a, int c, int t, int i, int j, int k) {
  if (e_mask(a.g, c, i, j, k)) return 0.0f;
  int sx, sy, sz;
  e_shape(a.g, c, sx, sy, sz);
  const uint32_t o = ((uint32_t)i * sy + j) * sz + k;
  const float mu = coef(a.cf.curl[c], a.cf.curl_value[c], o) * a.a.e[c][o];
  const int ax = e_axis(c, t);
  const int x = pick(ax, i, j, k);
  const float s = e_sign(c, t);
  float w = s * mu;
  if (CPML) {
    const float* p = psi_at(a.cp, 2 * c + t, ax, x, i, j, k, sy, sz);
    const float lp = p ? *p : 0.0f;
    w = s * mu * a.cp.inv_kappa_e[ax][x] + a.cp.c_e[ax][x] * (lp + s * mu);
  }
  return w * a.g.inv_e[ax][x];
}

// Same for the forward difference of H sample (h; i,j,k), term t, with
// nu = -curl_H lambda_H (the H-side psi adjoint read before it is finished).
template <bool CPML>
__device__ __forceinline__ float r_h(const RevArgs& a, int h, int t, int i, int j, int k) {
  int sx, sy, sz;
  h_shape(a.g, h, sx, sy, sz);
  const uint32_t o = ((uint32_t)i * sy + j) * sz + k;
  const float nu = -coef(a.cf.curl[3 + h], a.cf.curl_value[3 + h], o) * a.a.h[h][o];
  const int ax = h_axis(h, t);
  const int x = pick(ax, i, j, k);
  const float s = h_sign(h, t);
  float w = s * nu;
  if (CPML) {
    const float* p = psi_at(a.cp, 6 + 2 * h + t, ax, x, i, j, k, sy, sz);
    const float lp = p ? *p : 0.0f;
    w = s * nu * a.cp.inv_kappa_h[ax][x] + a.cp.c_h[ax][x] * (lp + s * nu);
  }
  return w * a.g.inv_h[ax][x];
}

// Sum over the E samples whose backward difference (with the face rules of
// electric.cu:23-87) reads H at position m of this axis line.
template <bool CPML>
__device__ __forceinline__ float gather_e(const RevArgs& a, int c, int t, int i, int j, int k) {
  const int ax = e_axis(c, t);
  const int n = axis_size(a.g, ax), lo = a.g.bc_lo[ax], hi = a.g.bc_hi[ax];
  const int m = pick(ax, i, j, k);
  auto q = [&](int x) {
    return q_e<CPML>(a, c, t, ax == 0 ? x : i, ax == 1 ? x : j, ax == 2 ? x : k);
  };
  float s = 0.0f;
  if (m > 0) {
    s += q(m);
  } else {
    if (lo == MW_BC_PERIODIC) s += q(0);
    else if (lo == MW_BC_PMC) s += 2.0f * q(0);
    if (hi == MW_BC_PERIODIC) s += q(n - 1);
  }
  if (m + 1 < n - 1) {
    s -= q(m + 1);
  } else {
    if (hi == MW_BC_PERIODIC) s -= q(n - 1);
    else if (hi == MW_BC_PMC) s -= 2.0f * q(n - 1);
    if (lo == MW_BC_PERIODIC) s -= q(0);
  }
  return s;
}

// P1: lambda_H of every H sample at (i, j, k).
template <bool CPML>
__device__ __forceinline__ void reverse_h_at(const RevArgs& a, int i, int j, int k) {
  const int nx = a.g.nx, ny = a.g.ny, nz = a.g.nz;
  if (k >= nz || j >= ny || i >= nx) return;
  for (int h = 0; h < 3; ++h) {
    int sx, sy, sz;
    h_shape(a.g, h, sx, sy, sz);
    if (i >= sx || j >= sy || k >= sz) continue;
    const uint32_t o = ((uint32_t)i * sy + j) * sz + k;
    const float lam = a.a.h[h][o];
    // finish the previous reverse step's H transpose
    if (CPML) {
      const float nu = -coef(a.cf.curl[3 + h], a.cf.curl_value[3 + h], o) * lam;
      for (int t = 0; t < 2; ++t) {
        const int ax = h_axis(h, t);
        const int x = pick(ax, i, j, k);
        float* p = psi_at(a.cp, 6 + 2 * h + t, ax, x, i, j, k, sy, sz);
        if (p) *p = a.cp.b_h[ax][x] * (*p + h_sign(h, t) * nu);
      }
    }
    float v = coef(a.cf.decay[3 + h], a.cf.decay_value[3 + h], o) * lam;
    // E samples whose curl read this H sample
    if (h == 0) v += gather_e<CPML>(a, 2, 1, i, j, k) + gather_e<CPML>(a, 1, 1, i, j, k);       // Ez b, Ey b
    else if (h == 1) v += gather_e<CPML>(a, 0, 1, i, j, k) + gather_e<CPML>(a, 2, 0, i, j, k);  // Ex b, Ez a
    else v += gather_e<CPML>(a, 0, 0, i, j, k) + gather_e<CPML>(a, 1, 0, i, j, k);              // Ex a, Ey a
    a.a.h[h][o] = v;
  }
}

// Transpose of the fused ADE of E sample (c, o) (E -= dt J' inv_eps after the poles
// advance from E^n, kernels/dispersive.cu:18-83): updates the poles' adjoint state
// and returns what it adds to lambda_E.
__device__ __forceinline__ float reverse_ade(const RevArgs& a, int c, uint32_t o, float lam) {
  const int np = a.ad.count[c];
  if (np <= 0) return 0.0f;
  float out = 0.0f;
  const float dt = a.ad.dt;
  const float lj = -dt * G(a.ad.inv_eps[c])[o] * lam;
  for (int p = 0; p < np; ++p) {
    const mw_pole& pl = a.ad.poles[c][p];
    const float drive = G(pl.drive)[o];
    if (pl.kind == 0) {  // Debye: P' = d P + g E, J' = (P' - P)/dt; J^n unread
      const float ljt = G(pl.current)[o] + lj;
      const float lp = G(pl.polarization)[o] + ljt * (1.0f / dt);
      G(pl.polarization)[o] = pl.decay * lp - ljt * (1.0f / dt);
      G(pl.current)[o] = 0.0f;
      out += drive * lp;
    } else if (pl.kind == 1) {  // Drude: J' = d J + g E
      const float lc = G(pl.current)[o] + lj;
      G(pl.current)[o] = pl.decay * lc;
      out += drive * lc;
    } else {  // Lorentz: J' = d J - r P + g E, P' = P + dt J'
      const float lpol = G(pl.polarization)[o];
      const float lnc = G(pl.current)[o] + lj + dt * lpol;
      G(pl.current)[o] = pl.decay * lnc;
      G(pl.polarization)[o] = lpol - pl.restoring * lnc;
      out += drive * lnc;
    }
  }
  return out;
}

// P2: lambda_E of every E sample at (i, j, k).
template <bool CPML, bool ADE>
__device__ __forceinline__ void reverse_e_at(const RevArgs& a, int i, int j, int k) {
  const int nx = a.g.nx, ny = a.g.ny, nz = a.g.nz;
  if (k >= nz || j >= ny || i >= nx) return;
  for (int c = 0; c < 3; ++c) {
    int sx, sy, sz;
    e_shape(a.g, c, sx, sy, sz);
    if (i >= sx || j >= sy || k >= sz) continue;
    const uint32_t o = ((uint32_t)i * sy + j) * sz + k;
    const float lam = a.a.e[c][o];
    const int mask = e_mask(a.g, c, i, j, k);
    float out;
    if (mask == 1) {
      out = 0.0f;
    } else if (mask == 2) {
      out = lam;
    } else {
      out = coef(a.cf.decay[c], a.cf.decay_value[c], o) * lam;
      if (CPML) {
        const float mu = coef(a.cf.curl[c], a.cf.curl_value[c], o) * lam;
        for (int t = 0; t < 2; ++t) {
          const int ax = e_axis(c, t);
          const int x = pick(ax, i, j, k);
          float* p = psi_at(a.cp, 2 * c + t, ax, x, i, j, k, sy, sz);
          if (p) *p = a.cp.b_e[ax][x] * (*p + e_sign(c, t) * mu);
        }
      }
    }
    if (ADE) out += reverse_ade(a, c, o, lam);
    // H samples whose forward difference read this E sample: +r(m = x-1) - r(m = x)
    auto fwd = [&](int h, int t) {
      const int ax = h_axis(h, t);
      const int x = pick(ax, i, j, k), n = axis_size(a.g, ax);
      float s = 0.0f;
      if (x >= 1) s += r_h<CPML>(a, h, t, ax == 0 ? x - 1 : i, ax == 1 ? x - 1 : j, ax == 2 ? x - 1 : k);
      if (x <= n - 2) s -= r_h<CPML>(a, h, t, i, j, k);
      return s;
    };
    if (c == 0) out += fwd(1, 1) + fwd(2, 1);       // Hy b (dEx/dz), Hz b (dEx/dy)
    else if (c == 1) out += fwd(0, 1) + fwd(2, 0);  // Hx b (dEy/dz), Hz a (dEy/dx)
    else out += fwd(0, 0) + fwd(1, 0);              // Hx a (dEz/dy), Hy a (dEz/dx)
    a.a.e[c][o] = out;
  }
}

// ---------------------------------------------------------------- boundary boxes
// The same P1 / P2 sums as reverse_h_at / reverse_e_at, term by term in the same
// order, with every component and derivative axis fixed at compile time, every
// load of a node issued before its first store, and no load behind a run-time
// test: a sample that does not exist (or lies outside a psi slab) is read at
// offset 0, an array the scene does not have (a scalar coefficient, a psi slab of
// an axis without PML) through a stand-in buffer, and the value is then scaled by
// 0 (every value read is a finite field, coefficient or psi sample).  A select
// would let the compiler move the load behind a branch, and the argument table's
// pointers carry no aliasing information, so a load after a store waits for it:
// written the general way, one node cost several memory round trips.
template <int C, int T> struct EDer {  // term T of the E_C update
  static constexpr int ax = T == 0 ? (C == 0 ? 1 : 0) : (C == 2 ? 1 : 2);
  static constexpr int psi = 2 * C + T;
};
template <int H, int T> struct HDer {  // term T of the H_H update
  static constexpr int ax = T == 0 ? (H == 0 ? 1 : 0) : (H == 2 ? 1 : 2);
  static constexpr int psi = 6 + 2 * H + T;
};

__device__ __forceinline__ float keep_if(bool c) { return c ? 1.0f : 0.0f; }

// coefficient array value (or its scalar) with the load outside any branch
__device__ __forceinline__ float coef_nb(const float* arr, const float* stand_in, float value, uint32_t o) {
  const float v = G(arr ? arr : stand_in)[arr ? o : 0u];
  const float has = keep_if(arr != nullptr);
  return v * has + value * (1.0f - has);
}

// e_mask without early exits: 1 PEC, 2 held by a none/PML face, 0 active
template <int C>
__device__ __forceinline__ int e_mask_nb(const mw_grid& g, int i, int j, int k) {
  bool pec = false, held = false;
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    const int ax = t == 0 ? EDer<C, 0>::ax : EDer<C, 1>::ax;
    const int x = pick(ax, i, j, k), n = axis_size(g, ax), lo = g.bc_lo[ax], hi = g.bc_hi[ax];
    const bool at_lo = x == 0, at_hi = x == n - 1;
    pec = pec | (at_lo & (lo == MW_BC_PEC)) | (at_hi & (hi == MW_BC_PEC));
    held = held | (at_lo & ((lo == MW_BC_NONE) | (lo == MW_BC_PML))) | (at_hi & ((hi == MW_BC_NONE) | (hi == MW_BC_PML)));
  }
  return pec ? 1 : (held ? 2 : 0);
}

// q of E sample (C; i,j,k) for term T (q_e), 0 when `valid` is false
template <int C, int T, bool CPML>
__device__ __forceinline__ float q_nb(const RevArgs& a, int i, int j, int k, bool valid) {
  constexpr int ax = EDer<C, T>::ax, q = EDer<C, T>::psi;
  const float s = e_sign(C, T);
  int sx, sy, sz;
  e_shape(a.g, C, sx, sy, sz);
  const uint32_t o = valid ? ((uint32_t)i * sy + j) * sz + k : 0u;
  const int x = valid ? pick(ax, i, j, k) : 0;
  const float mu = coef_nb(a.cf.curl[C], a.a.e[C], a.cf.curl_value[C], o) * G(a.a.e[C])[o];
  float w = s * mu;
  if (CPML) {
    const int local = slab_local(x, a.cp.psi_low[q], a.cp.psi_high_start[q], a.cp.psi_high_len[q]);
    const bool in = valid && local >= 0;
    const uint32_t po = in ? slab_offset(ax, i, j, k, local, sy, sz, a.cp.psi_low[q] + a.cp.psi_high_len[q]) : 0u;
    const float lp = G(a.cp.psi[q] ? a.cp.psi[q] : a.a.e[C])[po] * keep_if(in);
    w = s * mu * G(a.cp.inv_kappa_e[ax])[x] + G(a.cp.c_e[ax])[x] * (lp + s * mu);
  }
  return w * G(a.g.inv_e[ax])[x] * keep_if(valid && e_mask_nb<C>(a.g, i, j, k) == 0);
}

// gather_e for E component C, term T, at H sample (i,j,k) (which exists iff `ok`);
// PER: the grid has periodic faces (the seam samples are then read as well)
template <int C, int T, bool CPML, bool PER>
__device__ __forceinline__ float gather_nb(const RevArgs& a, int i, int j, int k, bool ok) {
  constexpr int ax = EDer<C, T>::ax;
  const int n = axis_size(a.g, ax), lo = a.g.bc_lo[ax], hi = a.g.bc_hi[ax];
  const int m = pick(ax, i, j, k);
  const bool first = m > 0, inner = m + 1 < n - 1;
  const float qm = q_nb<C, T, CPML>(a, i, j, k, ok);
  const float qp = q_nb<C, T, CPML>(a, ax == 0 ? i + 1 : i, ax == 1 ? j + 1 : j, ax == 2 ? k + 1 : k,
                                    ok && m + 1 <= n - 1);
  // the samples across a periodic seam (used on the face lanes of periodic axes)
  float q_first = 0.0f, q_last = 0.0f;
  const bool use_last = !first && hi == MW_BC_PERIODIC, use_first = !inner && lo == MW_BC_PERIODIC;
  if (PER) {
    q_last = q_nb<C, T, CPML>(a, ax == 0 ? n - 1 : i, ax == 1 ? n - 1 : j, ax == 2 ? n - 1 : k, ok && use_last);
    q_first = q_nb<C, T, CPML>(a, ax == 0 ? 0 : i, ax == 1 ? 0 : j, ax == 2 ? 0 : k, ok && use_first);
  }
  // weights of q(m) and q(m+1): 1 inside; at a face the periodic wrap (1), the PMC
  // mirror (2) or nothing (electric.cu:23-41 transposed)
  const float wm = first ? 1.0f : (lo == MW_BC_PERIODIC ? 1.0f : (lo == MW_BC_PMC ? 2.0f : 0.0f));
  const float wp = inner ? 1.0f : (hi == MW_BC_PERIODIC ? 1.0f : (hi == MW_BC_PMC ? 2.0f : 0.0f));
  float s = 0.0f;
  s += wm * qm;
  s += q_last * keep_if(use_last);
  s -= wp * qp;
  s -= q_first * keep_if(use_first);
  return s;
}

// r of H sample (H; i,j,k) for term T (r_h), 0 when `valid` is false
template <int H, int T, bool CPML>
__device__ __forceinline__ float r_nb(const RevArgs& a, int i, int j, int k, bool valid) {
  constexpr int ax = HDer<H, T>::ax, q = HDer<H, T>::psi;
  const float s = h_sign(H, T);
  int sx, sy, sz;
  h_shape(a.g, H, sx, sy, sz);
  const uint32_t o = valid ? ((uint32_t)i * sy + j) * sz + k : 0u;
  const int x = valid ? pick(ax, i, j, k) : 0;
  const float nu = -coef_nb(a.cf.curl[3 + H], a.a.h[H], a.cf.curl_value[3 + H], o) * G(a.a.h[H])[o];
  float w = s * nu;
  if (CPML) {
    const int local = slab_local(x, a.cp.psi_low[q], a.cp.psi_high_start[q], a.cp.psi_high_len[q]);
    const bool in = valid && local >= 0;
    const uint32_t po = in ? slab_offset(ax, i, j, k, local, sy, sz, a.cp.psi_low[q] + a.cp.psi_high_len[q]) : 0u;
    const float lp = G(a.cp.psi[q] ? a.cp.psi[q] : a.a.h[H])[po] * keep_if(in);
    w = s * nu * G(a.cp.inv_kappa_h[ax])[x] + G(a.cp.c_h[ax])[x] * (lp + s * nu);
  }
  return w * G(a.g.inv_h[ax])[x] * keep_if(valid);
}

// fwd(H, T) of reverse_e_at for the E sample at (i,j,k) (which exists iff `ok`)
template <int H, int T, bool CPML>
__device__ __forceinline__ float fwd_nb(const RevArgs& a, int i, int j, int k, bool ok) {
  constexpr int ax = HDer<H, T>::ax;
  const int x = pick(ax, i, j, k), n = axis_size(a.g, ax);
  const float rm = r_nb<H, T, CPML>(a, ax == 0 ? i - 1 : i, ax == 1 ? j - 1 : j, ax == 2 ? k - 1 : k, ok && x >= 1);
  const float rc = r_nb<H, T, CPML>(a, i, j, k, ok && x <= n - 2);
  float s = 0.0f;
  s += rm;  // 0 when x = 0
  s -= rc;  // 0 when x = n-1
  return s;
}

// A psi sample to rewrite once every load of the node has issued.
struct PsiStore {
  float v;
  uint32_t o;
  bool in;
};

template <int H, int T>
__device__ __forceinline__ PsiStore psi_h_finish(const RevArgs& a, int i, int j, int k, bool ok, int sy, int sz,
                                                 float nu) {
  constexpr int ax = HDer<H, T>::ax, q = HDer<H, T>::psi;
  const int x = ok ? pick(ax, i, j, k) : 0;
  const int local = slab_local(x, a.cp.psi_low[q], a.cp.psi_high_start[q], a.cp.psi_high_len[q]);
  PsiStore st;
  st.in = ok && local >= 0;
  st.o = st.in ? slab_offset(ax, i, j, k, local, sy, sz, a.cp.psi_low[q] + a.cp.psi_high_len[q]) : 0u;
  const float pv = G(a.cp.psi[q] ? a.cp.psi[q] : a.a.h[H])[st.o];
  st.v = G(a.cp.b_h[ax])[x] * (pv + h_sign(H, T) * nu);
  return st;
}

template <int C, int T>
__device__ __forceinline__ PsiStore psi_e_finish(const RevArgs& a, int i, int j, int k, bool upd, int sy, int sz,
                                                 float mu) {
  constexpr int ax = EDer<C, T>::ax, q = EDer<C, T>::psi;
  const int x = upd ? pick(ax, i, j, k) : 0;
  const int local = slab_local(x, a.cp.psi_low[q], a.cp.psi_high_start[q], a.cp.psi_high_len[q]);
  PsiStore st;
  st.in = upd && local >= 0;
  st.o = st.in ? slab_offset(ax, i, j, k, local, sy, sz, a.cp.psi_low[q] + a.cp.psi_high_len[q]) : 0u;
  const float pv = G(a.cp.psi[q] ? a.cp.psi[q] : a.a.e[C])[st.o];
  st.v = G(a.cp.b_e[ax])[x] * (pv + e_sign(C, T) * mu);
  return st;
}

// Result of one component at one node: the new adjoint and its psi rewrites.
struct RevOut {
  float v;  // P2: the held/decayed part (the pole and curl parts follow)
  float f;  // P2: the curl-transpose part
  uint32_t o;
  bool ok;
  PsiStore p0, p1;
};

// P1 for component H at node (i,j,k), stores deferred
template <int H, bool CPML, bool PER>
__device__ __forceinline__ RevOut reverse_h_nb(const RevArgs& a, int i, int j, int k) {
  int sx, sy, sz;
  h_shape(a.g, H, sx, sy, sz);
  RevOut r;
  r.ok = i < sx && j < sy && k < sz;
  r.o = r.ok ? ((uint32_t)i * sy + j) * sz + k : 0u;
  const float lam = G(a.a.h[H])[r.o];
  const float dec = coef_nb(a.cf.decay[3 + H], a.a.h[H], a.cf.decay_value[3 + H], r.o);
  float g;
  if (H == 0) g = gather_nb<2, 1, CPML, PER>(a, i, j, k, r.ok) + gather_nb<1, 1, CPML, PER>(a, i, j, k, r.ok);  // Ez b, Ey b
  else if (H == 1) g = gather_nb<0, 1, CPML, PER>(a, i, j, k, r.ok) + gather_nb<2, 0, CPML, PER>(a, i, j, k, r.ok);
  else g = gather_nb<0, 0, CPML, PER>(a, i, j, k, r.ok) + gather_nb<1, 0, CPML, PER>(a, i, j, k, r.ok);
  r.p0.in = r.p1.in = false;
  if (CPML) {
    const float nu = -coef_nb(a.cf.curl[3 + H], a.a.h[H], a.cf.curl_value[3 + H], r.o) * lam;
    r.p0 = psi_h_finish<H, 0>(a, i, j, k, r.ok, sy, sz, nu);
    r.p1 = psi_h_finish<H, 1>(a, i, j, k, r.ok, sy, sz, nu);
  }
  float v = dec * lam;
  v += g;
  r.v = v;
  r.f = 0.0f;
  return r;
}

template <int H, bool CPML>
__device__ __forceinline__ void store_h(const RevArgs& a, const RevOut& r) {
  constexpr int q0 = HDer<H, 0>::psi, q1 = HDer<H, 1>::psi;
  if (CPML && r.p0.in) G(a.cp.psi[q0])[r.p0.o] = r.p0.v;
  if (CPML && r.p1.in) G(a.cp.psi[q1])[r.p1.o] = r.p1.v;
  if (r.ok) G(a.a.h[H])[r.o] = r.v;
}

// P2 for component C at node (i,j,k) without ADE, stores deferred
template <int C, bool CPML>
__device__ __forceinline__ RevOut reverse_e_nb(const RevArgs& a, int i, int j, int k, float& lam_out) {
  int sx, sy, sz;
  e_shape(a.g, C, sx, sy, sz);
  RevOut r;
  r.ok = i < sx && j < sy && k < sz;
  r.o = r.ok ? ((uint32_t)i * sy + j) * sz + k : 0u;
  const float lam = G(a.a.e[C])[r.o];
  const float dec = coef_nb(a.cf.decay[C], a.a.e[C], a.cf.decay_value[C], r.o);
  float f;
  if (C == 0) f = fwd_nb<1, 1, CPML>(a, i, j, k, r.ok) + fwd_nb<2, 1, CPML>(a, i, j, k, r.ok);       // Hy b, Hz b
  else if (C == 1) f = fwd_nb<0, 1, CPML>(a, i, j, k, r.ok) + fwd_nb<2, 0, CPML>(a, i, j, k, r.ok);  // Hx b, Hz a
  else f = fwd_nb<0, 0, CPML>(a, i, j, k, r.ok) + fwd_nb<1, 0, CPML>(a, i, j, k, r.ok);              // Hx a, Hy a
  const int mask = e_mask_nb<C>(a.g, i, j, k);
  // PEC: 0; none/PML face: lambda unchanged; active: decay * lambda
  float out = (dec * lam) * keep_if(mask == 0);
  out += lam * keep_if(mask == 2);
  r.p0.in = r.p1.in = false;
  if (CPML) {
    const float mu = coef_nb(a.cf.curl[C], a.a.e[C], a.cf.curl_value[C], r.o) * lam;
    r.p0 = psi_e_finish<C, 0>(a, i, j, k, r.ok && mask == 0, sy, sz, mu);
    r.p1 = psi_e_finish<C, 1>(a, i, j, k, r.ok && mask == 0, sy, sz, mu);
  }
  r.v = out;
  r.f = f;
  lam_out = lam;
  return r;
}

template <int C, bool CPML>
__device__ __forceinline__ void store_e(const RevArgs& a, const RevOut& r) {
  constexpr int q0 = EDer<C, 0>::psi, q1 = EDer<C, 1>::psi;
  if (CPML && r.p0.in) G(a.cp.psi[q0])[r.p0.o] = r.p0.v;
  if (CPML && r.p1.in) G(a.cp.psi[q1])[r.p1.o] = r.p1.v;
  if (r.ok) G(a.a.e[C])[r.o] = r.v;
}

// Interior forms of P1 and P2 for a wavefront whose samples all lie in the
// in_lo/in_hi box: no faces, no psi, every neighbour inside the grid, so each
// gather is a plain difference of mu = curl_E lambda_E (P1) or nu = -curl_H
// lambda_H (P2) samples -- the same sums reverse_h_at / reverse_e_at form, term by
// term.  The buffers are restrict-qualified parameters so the loads of a sample
// issue together.
__device__ __forceinline__ float cval(gptr<const float> __restrict__ arr, float value, uint32_t o) {
  return arr ? arr[o] : value;
}

__device__ __forceinline__ void reverse_h_fast(const RevArgs& a, int i, int j, int k, gptr<float> __restrict__ lhx,
                                               gptr<float> __restrict__ lhy, gptr<float> __restrict__ lhz,
                                               gptr<const float> __restrict__ lex, gptr<const float> __restrict__ ley,
                                               gptr<const float> __restrict__ lez, gptr<const float> __restrict__ cex,
                                               gptr<const float> __restrict__ cey, gptr<const float> __restrict__ cez,
                                               gptr<const float> __restrict__ dhx, gptr<const float> __restrict__ dhy,
                                               gptr<const float> __restrict__ dhz) {
  const int ny = a.g.ny, nz = a.g.nz;
  const uint32_t sEx = (uint32_t)ny * nz, sEy = (uint32_t)(ny - 1) * nz, sEz = (uint32_t)ny * (nz - 1);
  const uint32_t ox = i * sEx + (uint32_t)j * nz + k;
  const uint32_t oy = i * sEy + (uint32_t)j * nz + k;
  const uint32_t oz = i * sEz + (uint32_t)j * (nz - 1) + k;
  const float cx = a.cf.curl_value[0], cy = a.cf.curl_value[1], cz = a.cf.curl_value[2];
  auto mux = [&](uint32_t o) { return cval(cex, cx, o) * lex[o]; };
  auto muy = [&](uint32_t o) { return cval(cey, cy, o) * ley[o]; };
  auto muz = [&](uint32_t o) { return cval(cez, cz, o) * lez[o]; };
  const float mx = mux(ox), mx_k = mux(ox + 1), mx_j = mux(ox + nz);
  const float my = muy(oy), my_k = muy(oy + 1), my_i = muy(oy + sEy);
  const float mz = muz(oz), mz_j = muz(oz + (nz - 1)), mz_i = muz(oz + sEz);
  const float iex = G(a.g.inv_e[0])[i], iex1 = G(a.g.inv_e[0])[i + 1];
  const float iey = G(a.g.inv_e[1])[j], iey1 = G(a.g.inv_e[1])[j + 1];
  const float iez = G(a.g.inv_e[2])[k], iez1 = G(a.g.inv_e[2])[k + 1];
  const uint32_t ohx = ((uint32_t)i * (ny - 1) + j) * (nz - 1) + k;
  const uint32_t ohy = ((uint32_t)i * ny + j) * (nz - 1) + k;
  const uint32_t ohz = ((uint32_t)i * (ny - 1) + j) * nz + k;
  const float vx = cval(dhx, a.cf.decay_value[3], ohx) * lhx[ohx];
  const float vy = cval(dhy, a.cf.decay_value[4], ohy) * lhy[ohy];
  const float vz = cval(dhz, a.cf.decay_value[5], ohz) * lhz[ohz];
  lhx[ohx] = vx + (mz_j * iey1 - mz * iey) + (my * iez - my_k * iez1);   // Ez b, Ey b
  lhy[ohy] = vy + (mx_k * iez1 - mx * iez) + (mz * iex - mz_i * iex1);   // Ex b, Ez a
  lhz[ohz] = vz + (mx * iey - mx_j * iey1) + (my_i * iex1 - my * iex);   // Ex a, Ey a
}

template <bool ADE>
__device__ __forceinline__ void reverse_e_fast(const RevArgs& a, int i, int j, int k, gptr<float> __restrict__ lex,
                                               gptr<float> __restrict__ ley, gptr<float> __restrict__ lez,
                                               gptr<const float> __restrict__ lhx, gptr<const float> __restrict__ lhy,
                                               gptr<const float> __restrict__ lhz, gptr<const float> __restrict__ chx,
                                               gptr<const float> __restrict__ chy, gptr<const float> __restrict__ chz,
                                               gptr<const float> __restrict__ dex, gptr<const float> __restrict__ dey,
                                               gptr<const float> __restrict__ dez) {
  const int ny = a.g.ny, nz = a.g.nz;
  const uint32_t sHx = (uint32_t)(ny - 1) * (nz - 1), sHy = (uint32_t)ny * (nz - 1), sHz = (uint32_t)(ny - 1) * nz;
  const uint32_t ohx = i * sHx + (uint32_t)j * (nz - 1) + k;
  const uint32_t ohy = i * sHy + (uint32_t)j * (nz - 1) + k;
  const uint32_t ohz = i * sHz + (uint32_t)j * nz + k;
  const float cx = a.cf.curl_value[3], cy = a.cf.curl_value[4], cz = a.cf.curl_value[5];
  auto nux = [&](uint32_t o) { return -cval(chx, cx, o) * lhx[o]; };
  auto nuy = [&](uint32_t o) { return -cval(chy, cy, o) * lhy[o]; };
  auto nuz = [&](uint32_t o) { return -cval(chz, cz, o) * lhz[o]; };
  const float nx0 = nux(ohx), nx_k = nux(ohx - 1), nx_j = nux(ohx - (nz - 1));
  const float ny0 = nuy(ohy), ny_k = nuy(ohy - 1), ny_i = nuy(ohy - sHy);
  const float nz0 = nuz(ohz), nz_j = nuz(ohz - nz), nz_i = nuz(ohz - sHz);
  const float ihx = G(a.g.inv_h[0])[i], ihx0 = G(a.g.inv_h[0])[i - 1];
  const float ihy = G(a.g.inv_h[1])[j], ihy0 = G(a.g.inv_h[1])[j - 1];
  const float ihz = G(a.g.inv_h[2])[k], ihz0 = G(a.g.inv_h[2])[k - 1];
  const uint32_t ox = ((uint32_t)i * ny + j) * nz + k;
  const uint32_t oy = ((uint32_t)i * (ny - 1) + j) * nz + k;
  const uint32_t oz = ((uint32_t)i * ny + j) * (nz - 1) + k;
  const float lx = lex[ox], ly = ley[oy], lz = lez[oz];
  float outx = cval(dex, a.cf.decay_value[0], ox) * lx;
  float outy = cval(dey, a.cf.decay_value[1], oy) * ly;
  float outz = cval(dez, a.cf.decay_value[2], oz) * lz;
  if (ADE) {
    outx += reverse_ade(a, 0, ox, lx);
    outy += reverse_ade(a, 1, oy, ly);
    outz += reverse_ade(a, 2, oz, lz);
  }
  outx += (ny_k * ihz0 - ny0 * ihz) + (nz0 * ihy - nz_j * ihy0);   // Hy b, Hz b
  outy += (nx0 * ihz - nx_k * ihz0) + (nz_i * ihx0 - nz0 * ihx);   // Hx b, Hz a
  outz += (nx_j * ihy0 - nx0 * ihy) + (ny0 * ihx - ny_i * ihx0);   // Hx a, Hy a
  lex[ox] = outx;
  ley[oy] = outy;
  lez[oz] = outz;
}

// Interior launch: the in_lo/in_hi box, a wavefront per 64 z samples of a row,
// four rows per block, one plane per block (grid: z blocks, y blocks, planes).
__device__ __forceinline__ bool interior_node(const RevArgs& a, int& i, int& j, int& k) {
  i = a.in_lo[0] + (int)blockIdx.z;
  j = a.in_lo[1] + (int)blockIdx.y * 4 + wave_row();
  k = a.in_lo[2] + (int)blockIdx.x * 64 + (int)threadIdx.x;
  return i < a.in_hi[0] && j < a.in_hi[1] && k < a.in_hi[2];
}

template <bool CPML>
__global__ __launch_bounds__(256) void reverse_h_kernel(RevArgs a) {
  int i, j, k;
  if (!interior_node(a, i, j, k)) return;
  reverse_h_fast(a, i, j, k, G(a.a.h[0]), G(a.a.h[1]), G(a.a.h[2]), G(a.a.e[0]), G(a.a.e[1]), G(a.a.e[2]),
                 G(a.cf.curl[0]), G(a.cf.curl[1]), G(a.cf.curl[2]), G(a.cf.decay[3]), G(a.cf.decay[4]),
                 G(a.cf.decay[5]));
}

template <bool CPML, bool ADE>
__global__ __launch_bounds__(256) void reverse_e_kernel(RevArgs a) {
  int i, j, k;
  if (!interior_node(a, i, j, k)) return;
  reverse_e_fast<ADE>(a, i, j, k, G(a.a.e[0]), G(a.a.e[1]), G(a.a.e[2]), G(a.a.h[0]), G(a.a.h[1]), G(a.a.h[2]),
                      G(a.cf.curl[3]), G(a.cf.curl[4]), G(a.cf.curl[5]), G(a.cf.decay[0]), G(a.cf.decay[1]),
                      G(a.cf.decay[2]));
}

// General launch: block -> (box, plane, rows, columns); per-box lane width so a
// thin z slab does not leave most lanes idle.  Returns false for idle threads.
__device__ __forceinline__ bool box_node(const RevArgs& a, int& i, int& j, int& k) {
  const int logical = (int)blockIdx.x;
  if (logical >= a.total) return false;
  int b = 0;
  while (b + 1 < a.nboxes && logical >= a.box_first[b + 1]) ++b;
  const int t = logical - a.box_first[b];
  const int sh = a.box_shift[b], w = 1 << sh, rows = 256 >> sh;
  const int bz = t % a.box_nbz[b], by = (t / a.box_nbz[b]) % a.box_nby[b], bx = t / (a.box_nbz[b] * a.box_nby[b]);
  const int tid = (int)threadIdx.x + 64 * (int)threadIdx.y;
  k = a.box_lo[b][2] + bz * w + (tid & (w - 1));
  j = a.box_lo[b][1] + by * rows + (tid >> sh);
  i = a.box_lo[b][0] + bx;
  return k < a.box_hi[b][2] && j < a.box_hi[b][1];
}

template <bool CPML, bool PER>
__global__ __launch_bounds__(256) void reverse_h_box_kernel(RevArgs a) {
  int i, j, k;
  if (!box_node(a, i, j, k)) return;
  // P1 writes only lambda_H and the H-side psi of its own samples and reads
  // neither, so every load of the three components issues before any store
  const RevOut hx = reverse_h_nb<0, CPML, PER>(a, i, j, k);
  const RevOut hy = reverse_h_nb<1, CPML, PER>(a, i, j, k);
  const RevOut hz = reverse_h_nb<2, CPML, PER>(a, i, j, k);
  store_h<0, CPML>(a, hx);
  store_h<1, CPML>(a, hy);
  store_h<2, CPML>(a, hz);
}

template <bool CPML, bool ADE>
__global__ __launch_bounds__(256) void reverse_e_box_kernel(RevArgs a) {
  int i, j, k;
  if (!box_node(a, i, j, k)) return;
  // P2 writes only lambda_E, the E-side psi and the pole adjoints of its own
  // samples; the pole transposes (which read and write pole state) follow the
  // loads of all three components
  float lx, ly, lz;
  RevOut ex = reverse_e_nb<0, CPML>(a, i, j, k, lx);
  RevOut ey = reverse_e_nb<1, CPML>(a, i, j, k, ly);
  RevOut ez = reverse_e_nb<2, CPML>(a, i, j, k, lz);
  if (ADE) {
    if (ex.ok) ex.v += reverse_ade(a, 0, ex.o, lx);
    if (ey.ok) ey.v += reverse_ade(a, 1, ey.o, ly);
    if (ez.ok) ez.v += reverse_ade(a, 2, ez.o, lz);
  }
  ex.v += ex.f;
  ey.v += ey.f;
  ez.v += ez.f;
  store_e<0, CPML>(a, ex);
  store_e<1, CPML>(a, ey);
  store_e<2, CPML>(a, ez);
}

// Batched reverse phases: block b of the launch belongs to the scene s with
// prefix[s] <= b < prefix[s+1]; scene s's blocks tile (z: 64, y: 4, x: 1).
__device__ __forceinline__ int block_scene(const int* prefix, int count, int b) {
  int lo = 0, hi = count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (prefix[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <bool CPML>
__global__ __launch_bounds__(256) void reverse_h_batch_kernel(const RevArgs* __restrict__ table, const int* prefix,
                                                              int count, int total) {
  const int b = blockIdx.x;
  if (b >= total) return;
  const int s = block_scene(prefix, count, b);
  const RevArgs& a = table[s];
  const int t = b - prefix[s];
  const int bz = (a.g.nz + 63) / 64, by = (a.g.ny + 3) / 4;
  reverse_h_at<CPML>(a, t / (bz * by), (t / bz) % by * 4 + wave_row(), t % bz * 64 + threadIdx.x);
}

template <bool CPML, bool ADE>
__global__ __launch_bounds__(256) void reverse_e_batch_kernel(const RevArgs* __restrict__ table, const int* prefix,
                                                              int count, int total) {
  const int b = blockIdx.x;
  if (b >= total) return;
  const int s = block_scene(prefix, count, b);
  const RevArgs& a = table[s];
  const int t = b - prefix[s];
  const int bz = (a.g.nz + 63) / 64, by = (a.g.ny + 3) / 4;
  reverse_e_at<CPML, ADE>(a, t / (bz * by), (t / bz) % by * 4 + wave_row(), t % bz * 64 + threadIdx.x);
}

// --------------------------------------------------------------- seeds
// Transposes of the observer accumulations (fdtd/adjoint/seeds.py:922-990,
// kernels/adjoint.cu:1737-1858).  Several observers may name one sample, so the
// plane and point seeds add atomically (vector-memory float atomics).
struct SeedPlaneSet {
  int32_t count;
  int32_t pad;
  mw_plane_obs p[MW_MAX_PLANES];  // re/im = gradients of the sums
};

__global__ __launch_bounds__(256) void seed_planes_kernel(SeedPlaneSet set, FieldSet fs, int nf, const float* e_cos,
                                                          const float* e_sin, const float* h_cos, const float* h_sin,
                                                          int family) {
  const int qi = blockIdx.y;
  if (qi >= set.count) return;
  const mw_plane_obs& q = set.p[qi];
  const int c = q.component;
  if ((c < 3) != (family == 0)) return;
  const int sx = fs.shape[c][0], sy = fs.shape[c][1], sz = fs.shape[c][2];
  int na, nb;
  if (q.axis == 0) { na = sy; nb = sz; } else if (q.axis == 1) { na = sx; nb = sz; } else { na = sx; nb = sy; }
  const int64_t n = (int64_t)na * nb;
  const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= n) return;
  const int ia = (int)(l / nb), ib = (int)(l % nb);
  uint32_t o;
  if (q.axis == 0) o = ((uint32_t)q.index * sy + ia) * sz + ib;
  else if (q.axis == 1) o = ((uint32_t)ia * sy + q.index) * sz + ib;
  else o = ((uint32_t)ia * sy + ib) * sz + q.index;
  const float* wc = c < 3 ? e_cos : h_cos;
  const float* ws = c < 3 ? e_sin : h_sin;
  float v = 0.0f;
  for (int f = 0; f < nf; ++f) v += q.re[f * n + l] * wc[f] + q.im[f * n + l] * ws[f];
  atomicAdd(fs.ptr[c] + o, v);
}

__global__ __launch_bounds__(64) void seed_points_kernel(FieldSet fs, int count, const int32_t* comp,
                                                         const int32_t* lin, int nf, const float* e_cos,
                                                         const float* e_sin, const float* h_cos, const float* h_sin,
                                                         const float* gre, const float* gim, int family) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= count) return;
  const int c = comp[p];
  if ((c < 3) != (family == 0)) return;
  const float* wc = c < 3 ? e_cos : h_cos;
  const float* ws = c < 3 ? e_sin : h_sin;
  float v = 0.0f;
  for (int f = 0; f < nf; ++f) v += gre[f * count + p] * wc[f] + gim[f * count + p] * ws[f];
  atomicAdd(fs.ptr[c] + lin[p], v);
}

__global__ __launch_bounds__(256) void seed_dense_kernel(float* __restrict__ lam, int64_t n, int nf,
                                                         const float* __restrict__ wc, const float* __restrict__ ws,
                                                         const float* __restrict__ gre,
                                                         const float* __restrict__ gim) {
  const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= n) return;
  float v = 0.0f;
  for (int f = 0; f < nf; ++f) v += gre[f * n + l] * wc[f] + gim[f * n + l] * ws[f];
  lam[l] += v;
}

// ------------------------------------------------------- design window
// The window is the node box [lo, hi); E^n is captured on it and H^{n+1/2} on
// [lo-1, hi) (every H sample the E curl inside the window reads).
struct WindowArgs {
  int32_t lo[3], hi[3];
};

__global__ __launch_bounds__(256) void window_capture_kernel(WindowArgs w, mw_grid g, const float* e0, const float* e1,
                                                             const float* e2, float* dst, int halo) {
  const int wx = w.hi[0] - w.lo[0] + halo, wy = w.hi[1] - w.lo[1] + halo, wz = w.hi[2] - w.lo[2] + halo;
  const int64_t per = (int64_t)wx * wy * wz;
  const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (l >= per) return;
  const int wk = (int)(l % wz), wj = (int)((l / wz) % wy), wi = (int)(l / ((int64_t)wz * wy));
  const int i = w.lo[0] - halo + wi, j = w.lo[1] - halo + wj, k = w.lo[2] - halo + wk;
  int sx, sy, sz;
  const float* src;
  if (halo) { h_shape(g, c, sx, sy, sz); } else { e_shape(g, c, sx, sy, sz); }
  src = c == 0 ? e0 : (c == 1 ? e1 : e2);
  dst[c * per + l] = src[((uint32_t)i * sy + j) * sz + k];
}

// dL/dcurl_E += lambda_E * curl^n and dL/ddecay_E += lambda_E * E^n on the
// window (every window sample is an updated sample outside the CPML slabs).
__global__ __launch_bounds__(256) void window_gradient_kernel(WindowArgs w, mw_grid g, mw_fields lam,
                                                              const float* __restrict__ e_win,
                                                              const float* __restrict__ h_win, float* g_decay0,
                                                              float* g_decay1, float* g_decay2, float* g_curl0,
                                                              float* g_curl1, float* g_curl2) {
  const int wx = w.hi[0] - w.lo[0], wy = w.hi[1] - w.lo[1], wz = w.hi[2] - w.lo[2];
  const int64_t per = (int64_t)wx * wy * wz;
  const int hy = wy + 1, hz = wz + 1;
  const int64_t hper = (int64_t)(wx + 1) * hy * hz;
  const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (l >= per) return;
  const int wk = (int)(l % wz), wj = (int)((l / wz) % wy), wi = (int)(l / ((int64_t)wz * wy));
  const int i = w.lo[0] + wi, j = w.lo[1] + wj, k = w.lo[2] + wk;
  // H window index of global (i', j', k') is (i'-lo+1, ...)
  auto H = [&](int h, int di, int dj, int dk) {
    return h_win[h * hper + ((int64_t)(wi + 1 + di) * hy + (wj + 1 + dj)) * hz + (wk + 1 + dk)];
  };
  float curl;
  if (c == 0) curl = (H(2, 0, 0, 0) - H(2, 0, -1, 0)) * g.inv_e[1][j] - (H(1, 0, 0, 0) - H(1, 0, 0, -1)) * g.inv_e[2][k];
  else if (c == 1) curl = (H(0, 0, 0, 0) - H(0, 0, 0, -1)) * g.inv_e[2][k] - (H(2, 0, 0, 0) - H(2, -1, 0, 0)) * g.inv_e[0][i];
  else curl = (H(1, 0, 0, 0) - H(1, -1, 0, 0)) * g.inv_e[0][i] - (H(0, 0, 0, 0) - H(0, 0, -1, 0)) * g.inv_e[1][j];
  int sx, sy, sz;
  e_shape(g, c, sx, sy, sz);
  const float lv = lam.e[c][((uint32_t)i * sy + j) * sz + k];
  float* gd = c == 0 ? g_decay0 : (c == 1 ? g_decay1 : g_decay2);
  float* gc = c == 0 ? g_curl0 : (c == 1 ? g_curl1 : g_curl2);
  gd[l] += lv * e_win[c * per + l];
  gc[l] += lv * curl;
}

// ------------------------------------------------------ batched adjoint
// Seeds of one observer plane / point / full-field DFT component of one scene:
// lam += sum_f g_re[f] w_cos[n, f] + g_im[f] w_sin[n, f].
struct SeedPlane {
  int32_t axis, index, nf, family;  // family 0 = E, 1 = H
  int32_t shape[3];
  int32_t pad;
  float* lam;
  const float* gre;
  const float* gim;
  const float* w_cos;  // (steps, nf)
  const float* w_sin;
};

__global__ __launch_bounds__(256) void seed_planes_batch_kernel(const SeedPlane* __restrict__ planes, int count,
                                                                int step, int family) {
  const int qi = blockIdx.y;
  if (qi >= count) return;
  const SeedPlane& q = planes[qi];
  if (q.family != family) return;
  const int sx = q.shape[0], sy = q.shape[1], sz = q.shape[2];
  int na, nb;
  if (q.axis == 0) { na = sy; nb = sz; } else if (q.axis == 1) { na = sx; nb = sz; } else { na = sx; nb = sy; }
  const int64_t cells = (int64_t)na * nb;
  const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= cells) return;
  const int ia = (int)(l / nb), ib = (int)(l % nb);
  uint32_t o;
  if (q.axis == 0) o = ((uint32_t)q.index * sy + ia) * sz + ib;
  else if (q.axis == 1) o = ((uint32_t)ia * sy + q.index) * sz + ib;
  else o = ((uint32_t)ia * sy + ib) * sz + q.index;
  const float* wc = q.w_cos + (int64_t)step * q.nf;
  const float* ws = q.w_sin + (int64_t)step * q.nf;
  float v = 0.0f;
  for (int f = 0; f < q.nf; ++f) v += q.gre[f * cells + l] * wc[f] + q.gim[f * cells + l] * ws[f];
  atomicAdd(q.lam + o, v);
}

struct SeedPoint {
  int32_t linear, nf, stride, family;
  float* lam;
  const float* gre;  // offset to this point's column of (nf, stride)
  const float* gim;
  const float* w_cos;
  const float* w_sin;
};

__global__ __launch_bounds__(64) void seed_points_batch_kernel(const SeedPoint* __restrict__ points, int count,
                                                               int step, int family) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= count) return;
  const SeedPoint& q = points[p];
  if (q.family != family) return;
  const float* wc = q.w_cos + (int64_t)step * q.nf;
  const float* ws = q.w_sin + (int64_t)step * q.nf;
  float v = 0.0f;
  for (int f = 0; f < q.nf; ++f) v += q.gre[(int64_t)f * q.stride] * wc[f] + q.gim[(int64_t)f * q.stride] * ws[f];
  atomicAdd(q.lam + q.linear, v);
}

// Design window of one scene for the batched capture / gradient kernels.
struct BatchWindow {
  WindowArgs w;
  mw_grid g;
  const float* e_state[2][3];  // E buffers holding a state of even / odd index
  const float* h_state[2][3];
  float* lam_e[3];
  float* e_store;
  float* h_store;
  float* g_decay[3];
  float* g_curl[3];
};

// Capture E of state m (slot m - first) and, for m > first, H of the step that
// produced it (slot m - first - 1).  blockIdx.y = 3 scene + component.
__global__ __launch_bounds__(256) void window_capture_batch_kernel(const BatchWindow* __restrict__ wins, int count,
                                                                   int m, int first) {
  const int s = blockIdx.y / 3, c = blockIdx.y % 3;
  if (s >= count) return;
  const BatchWindow& b = wins[s];
  const int wx = b.w.hi[0] - b.w.lo[0], wy = b.w.hi[1] - b.w.lo[1], wz = b.w.hi[2] - b.w.lo[2];
  const int64_t per = (int64_t)wx * wy * wz, hper = (int64_t)(wx + 1) * (wy + 1) * (wz + 1);
  const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int par = m & 1;
  if (l < per) {
    const int wk = (int)(l % wz), wj = (int)((l / wz) % wy), wi = (int)(l / ((int64_t)wz * wy));
    int sx, sy, sz;
    e_shape(b.g, c, sx, sy, sz);
    b.e_store[(int64_t)(m - first) * 3 * per + c * per + l] =
        b.e_state[par][c][((uint32_t)(b.w.lo[0] + wi) * sy + (b.w.lo[1] + wj)) * sz + (b.w.lo[2] + wk)];
  }
  if (m > first && l < hper) {
    const int hy = wy + 1, hz = wz + 1;
    const int wk = (int)(l % hz), wj = (int)((l / hz) % hy), wi = (int)(l / ((int64_t)hz * hy));
    int sx, sy, sz;
    h_shape(b.g, c, sx, sy, sz);
    b.h_store[(int64_t)(m - first - 1) * 3 * hper + c * hper + l] =
        b.h_state[par][c][((uint32_t)(b.w.lo[0] - 1 + wi) * sy + (b.w.lo[1] - 1 + wj)) * sz + (b.w.lo[2] - 1 + wk)];
  }
}

__global__ __launch_bounds__(256) void window_gradient_batch_kernel(const BatchWindow* __restrict__ wins, int count,
                                                                    int slot) {
  const int s = blockIdx.y / 3, c = blockIdx.y % 3;
  if (s >= count) return;
  const BatchWindow& b = wins[s];
  const WindowArgs& w = b.w;
  const int wx = w.hi[0] - w.lo[0], wy = w.hi[1] - w.lo[1], wz = w.hi[2] - w.lo[2];
  const int64_t per = (int64_t)wx * wy * wz;
  const int hy = wy + 1, hz = wz + 1;
  const int64_t hper = (int64_t)(wx + 1) * hy * hz;
  const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= per) return;
  const int wk = (int)(l % wz), wj = (int)((l / wz) % wy), wi = (int)(l / ((int64_t)wz * wy));
  const int i = w.lo[0] + wi, j = w.lo[1] + wj, k = w.lo[2] + wk;
  const float* hw = b.h_store + (int64_t)slot * 3 * hper;
  auto H = [&](int h, int di, int dj, int dk) {
    return hw[h * hper + ((int64_t)(wi + 1 + di) * hy + (wj + 1 + dj)) * hz + (wk + 1 + dk)];
  };
  const mw_grid& g = b.g;
  float curl;
  if (c == 0) curl = (H(2, 0, 0, 0) - H(2, 0, -1, 0)) * g.inv_e[1][j] - (H(1, 0, 0, 0) - H(1, 0, 0, -1)) * g.inv_e[2][k];
  else if (c == 1) curl = (H(0, 0, 0, 0) - H(0, 0, 0, -1)) * g.inv_e[2][k] - (H(2, 0, 0, 0) - H(2, -1, 0, 0)) * g.inv_e[0][i];
  else curl = (H(1, 0, 0, 0) - H(1, -1, 0, 0)) * g.inv_e[0][i] - (H(0, 0, 0, 0) - H(0, 0, -1, 0)) * g.inv_e[1][j];
  int sx, sy, sz;
  e_shape(g, c, sx, sy, sz);
  const float lv = b.lam_e[c][((uint32_t)i * sy + j) * sz + k];
  b.g_decay[c][l] += lv * b.e_store[(int64_t)slot * 3 * per + c * per + l];
  b.g_curl[c][l] += lv * curl;
}

}  // namespace mw
