// Yee E/H update kernels for gfx950 (CDNA4).
//
// Layout: every component is its own z-fastest [x][y][z] array with the reference
// shapes (fdtd/runtime/stepping.py:1625-1630).  A 256-thread block covers a
// 64 (z, one wavefront) x 4 (y) column tile and marches along x over a chunk of
// planes, so the x-neighbour needed by the curl is carried in registers and each
// field value is read from HBM once per pass; y/z neighbours come from L1/L2.
// All three components of a family are updated in the same pass (one launch for
// H, one for E) instead of the reference's one launch per component
// (kernels/magnetic.cu:368-921, kernels/electric.cu:1542-2612).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "maxwell_fdtd.h"

namespace mw {

constexpr int kTileZ = 64;
constexpr int kTileY = 4;

// Row of the block handled by this wavefront.  Every step/reverse kernel runs
// blocks of 64 x rows threads, so threadIdx.y is constant across a wavefront;
// reading it through readfirstlane tells the compiler so, and everything derived
// from the row (spacings, CPML profiles, face tests) stays in scalar registers
// with scalar branches.
__device__ __forceinline__ int wave_row() { return __builtin_amdgcn_readfirstlane(threadIdx.y); }

// Global-memory views.  Pointers read out of argument tables in device memory
// (the scene batches) are generic to the compiler, which then emits flat loads
// and stores (64-bit VGPR addresses, counted against both the vector- and the
// scalar-memory counters).  Every buffer the kernels dereference is global
// memory, so each dereference goes through an address_space(1) view, which
// gives global_load/global_store with an SGPR base, as for pointers passed
// directly as kernel arguments.  Only device buffers are viewed this way, never
// the address of a kernel-argument struct member.
template <class T>
using gptr = __attribute__((address_space(1))) T*;
template <class T>
__device__ __forceinline__ gptr<T> G(T* p) {
  return (gptr<T>)p;
}

// Slab index of psi sample `coord` along its axis, or -1 outside the PML layers
// (reference kernels/common.cuh:65-78).
__device__ __forceinline__ int slab_local(int coord, int low, int high_start, int high_len) {
  if (coord < low) return coord;
  if (coord >= high_start && coord < high_start + high_len) return low + coord - high_start;
  return -1;
}

// Offset of psi sample (i,j,k) of a component with shape (sx, sy, sz) whose
// compressed axis is `axis` with compressed length `len`.
__device__ __forceinline__ uint32_t slab_offset(int axis, int i, int j, int k, int local, int sy, int sz, int len) {
  if (axis == 0) return ((uint32_t)local * sy + j) * sz + k;
  if (axis == 1) return ((uint32_t)i * len + local) * sz + k;
  return ((uint32_t)i * sy + j) * len + local;
}

struct PsiView {
  float* data;
  int low, high_start, high_len;
};

__device__ __forceinline__ PsiView psi_view(const mw_cpml& c, int q) {
  return PsiView{c.psi[q], c.psi_low[q], c.psi_high_start[q], c.psi_high_len[q]};
}

// psi <- b*psi + c*d at (i,j,k) when inside the slab; returns the new psi (0 outside).
__device__ __forceinline__ float psi_step(const PsiView& v, int axis, int coord, int i, int j, int k, int sy, int sz,
                                          float b, float c, float d) {
  const int local = slab_local(coord, v.low, v.high_start, v.high_len);
  if (local < 0) return 0.0f;
  const uint32_t o = slab_offset(axis, i, j, k, local, sy, sz, v.low + v.high_len);
  const float nv = b * G(v.data)[o] + c * d;
  G(v.data)[o] = nv;
  return nv;
}

__device__ __forceinline__ float coef(const float* arr, float value, uint32_t o) {
  return arr ? G(arr)[o] : value;
}

// ---------------------------------------------------------------- magnetic
// Appendix of the reference update (kernels/magnetic.cu:62-117):
//   Hx: a = dEz/dy, b = dEy/dz, curl = a - b
//   Hy: a = dEz/dx, b = dEx/dz, curl = b - a
//   Hz: a = dEy/dx, b = dEx/dy, curl = a - b
// with a -> a*kinv + psi_a (CPML) and H <- H*decay - curl_coef*curl.
template <bool CPML>
__global__ __launch_bounds__(256) void update_h_kernel(mw_grid g, mw_coeffs cf, mw_cpml cp, mw_fields f,
                                                       int comp_mask, int xchunk) {
  const int k = blockIdx.x * kTileZ + threadIdx.x;
  const int j = blockIdx.y * kTileY + wave_row();
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  if (k >= nz || j >= ny) return;
  const int i0 = blockIdx.z * xchunk;
  const int i1 = min(i0 + xchunk, nx);
  const float* __restrict__ ex = f.e[0];
  const float* __restrict__ ey = f.e[1];
  const float* __restrict__ ez = f.e[2];
  const bool jv = j < ny - 1, kv = k < nz - 1;
  const bool do_hx = (comp_mask & 1) && jv && kv;
  const bool do_hy = (comp_mask & 2) && kv;
  const bool do_hz = (comp_mask & 4) && jv;
  const float ihy = jv ? g.inv_h[1][j] : 0.0f;
  const float ihz = kv ? g.inv_h[2][k] : 0.0f;
  float bhy = 0, chy = 0, khy = 1, bhz = 0, chz = 0, khz = 1;
  if (CPML) {
    if (jv) { bhy = cp.b_h[1][j]; chy = cp.c_h[1][j]; khy = cp.inv_kappa_h[1][j]; }
    if (kv) { bhz = cp.b_h[2][k]; chz = cp.c_h[2][k]; khz = cp.inv_kappa_h[2][k]; }
  }
  // Carried x-neighbours: Ez(i,j,k) and Ey(i,j,k)
  float ez_c = (kv && i0 < nx) ? ez[((uint32_t)i0 * ny + j) * (nz - 1) + k] : 0.0f;
  float ey_c = (jv && i0 < nx) ? ey[((uint32_t)i0 * (ny - 1) + j) * nz + k] : 0.0f;
  for (int i = i0; i < i1; ++i) {
    const bool iv = i < nx - 1;
    float ez_n = 0.0f, ey_n = 0.0f;
    if (iv) {
      if (kv) ez_n = ez[((uint32_t)(i + 1) * ny + j) * (nz - 1) + k];
      if (jv) ey_n = ey[((uint32_t)(i + 1) * (ny - 1) + j) * nz + k];
    }
    if (do_hx) {
      const uint32_t o = ((uint32_t)i * (ny - 1) + j) * (nz - 1) + k;
      const float ez_jp = ez[((uint32_t)i * ny + j + 1) * (nz - 1) + k];
      const float ey_kp = ey[((uint32_t)i * (ny - 1) + j) * nz + k + 1];
      const float da = (ez_jp - ez_c) * ihy;
      const float db = (ey_kp - ey_c) * ihz;
      float curl;
      if (CPML) {
        const float pa = psi_step(psi_view(cp, 6), 1, j, i, j, k, ny - 1, nz - 1, bhy, chy, da);
        const float pb = psi_step(psi_view(cp, 7), 2, k, i, j, k, ny - 1, nz - 1, bhz, chz, db);
        curl = (da * khy + pa) - (db * khz + pb);
      } else {
        curl = da - db;
      }
      float* hx = f.h[0];
      hx[o] = hx[o] * coef(cf.decay[3], cf.decay_value[3], o) - coef(cf.curl[3], cf.curl_value[3], o) * curl;
    }
    if (iv && (do_hy || do_hz)) {
      const float ihx = g.inv_h[0][i];
      float bhx = 0, chx = 0, khx = 1;
      if (CPML) { bhx = cp.b_h[0][i]; chx = cp.c_h[0][i]; khx = cp.inv_kappa_h[0][i]; }
      const float ex_c = ex[((uint32_t)i * ny + j) * nz + k];
      if (do_hy) {
        const uint32_t o = ((uint32_t)i * ny + j) * (nz - 1) + k;
        const float ex_kp = ex[((uint32_t)i * ny + j) * nz + k + 1];
        const float da = (ez_n - ez_c) * ihx;
        const float db = (ex_kp - ex_c) * ihz;
        float curl;
        if (CPML) {
          const float pa = psi_step(psi_view(cp, 8), 0, i, i, j, k, ny, nz - 1, bhx, chx, da);
          const float pb = psi_step(psi_view(cp, 9), 2, k, i, j, k, ny, nz - 1, bhz, chz, db);
          curl = (db * khz + pb) - (da * khx + pa);
        } else {
          curl = db - da;
        }
        float* hy = f.h[1];
        hy[o] = hy[o] * coef(cf.decay[4], cf.decay_value[4], o) - coef(cf.curl[4], cf.curl_value[4], o) * curl;
      }
      if (do_hz) {
        const uint32_t o = ((uint32_t)i * (ny - 1) + j) * nz + k;
        const float ex_jp = ex[((uint32_t)i * ny + j + 1) * nz + k];
        const float da = (ey_n - ey_c) * ihx;
        const float db = (ex_jp - ex_c) * ihy;
        float curl;
        if (CPML) {
          const float pa = psi_step(psi_view(cp, 10), 0, i, i, j, k, ny - 1, nz, bhx, chx, da);
          const float pb = psi_step(psi_view(cp, 11), 1, j, i, j, k, ny - 1, nz, bhy, chy, db);
          curl = (da * khx + pa) - (db * khy + pb);
        } else {
          curl = da - db;
        }
        float* hz = f.h[2];
        hz[o] = hz[o] * coef(cf.decay[5], cf.decay_value[5], o) - coef(cf.curl[5], cf.curl_value[5], o) * curl;
      }
    }
    ez_c = ez_n;
    ey_c = ey_n;
  }
}

// ---------------------------------------------------------------- electric
// Boundary helpers (kernels/electric.cu:23-51).
__device__ __forceinline__ bool face_pec(int c, int size, int lo, int hi) {
  return (c == 0 && lo == MW_BC_PEC) || (c + 1 == size && hi == MW_BC_PEC);
}
__device__ __forceinline__ bool face_inactive(int c, int size, int lo, int hi) {
  return (c == 0 && (lo == MW_BC_NONE || lo == MW_BC_PML)) || (c + 1 == size && (hi == MW_BC_NONE || hi == MW_BC_PML));
}
// Derivative at a boundary sample (electric.cu:23-41): periodic wraps through the
// duplicated end node, PMC mirrors, everything else contributes zero.
__device__ __forceinline__ float edge_diff(bool low_end, int mode, float first, float last, float inv) {
  if (mode == MW_BC_PERIODIC) return (first - last) * inv;
  if (mode == MW_BC_PMC) return low_end ? 2.0f * first * inv : -2.0f * last * inv;
  return 0.0f;
}

// One pole's advance from E^n at sample o with its drive there; returns the
// pole's new polarisation current (kernels/dispersive.cu:18-70).
__device__ __forceinline__ float pole_step(const mw_ade& a, const mw_pole& pl, uint32_t o, float drive, float e_old) {
  if (pl.kind == 0) {
    const float prev = G(pl.polarization)[o];
    const float next = pl.decay * prev + drive * e_old;
    G(pl.polarization)[o] = next;
    const float cur = (next - prev) * (1.0f / a.dt);
    G(pl.current)[o] = cur;
    return cur;
  }
  if (pl.kind == 1) {
    const float cur = pl.decay * G(pl.current)[o] + drive * e_old;
    G(pl.current)[o] = cur;
    return cur;
  }
  const float pc = G(pl.current)[o];
  const float pp = G(pl.polarization)[o];
  const float nc = pl.decay * pc - pl.restoring * pp + drive * e_old;
  G(pl.current)[o] = nc;
  G(pl.polarization)[o] = pp + a.dt * nc;
  return nc;
}

// Fused ADE on one E sample: advance every pole from E^n, return the summed
// polarisation current (kernels/dispersive.cu:18-70).  Where a pole's drive is
// zero its state starts at zero and stays zero and it adds nothing, so it is
// skipped there (exact).  With the per-sample weight codes the drive is
// code/2 * drive_scale and no drive array is read; else the activity mask (when
// given) and the drive arrays are.
__device__ __forceinline__ float ade_advance(const mw_ade& a, int c, uint32_t o, float e_old) {
  float jsum = 0.0f;
  const int n = a.count[c];
  if (a.weight[c]) {
    const uint32_t code = G(a.weight[c])[o];
    if (code == 0u) return 0.0f;
    for (int p = 0; p < n; ++p) {
      const uint32_t w = (code >> (2 * p)) & 3u;
      if (w == 0u) continue;
      const mw_pole& pl = a.poles[c][p];
      jsum += pole_step(a, pl, o, w == 2u ? pl.drive_scale : 0.5f * pl.drive_scale, e_old);
    }
    return jsum;
  }
  // bit p: pole p acts on this sample (without the mask every pole is read)
  const uint32_t act = a.active[c] ? (uint32_t)G(a.active[c])[o] : 0xffu;
  if (act == 0u) return 0.0f;
  for (int p = 0; p < n; ++p) {
    if (((act >> p) & 1u) == 0u) continue;
    const mw_pole& pl = a.poles[c][p];
    const float drive = G(pl.drive)[o];
    if (drive == 0.0f) continue;
    jsum += pole_step(a, pl, o, drive, e_old);
  }
  return jsum;
}

// ade_advance with the weight code already read by the caller (`code`, used when
// the component has weight codes; otherwise the drive arrays are read as above).
__device__ __forceinline__ float ade_advance_coded(const mw_ade& a, int c, uint32_t o, float e_old, uint32_t code) {
  if (!a.weight[c]) return ade_advance(a, c, o, e_old);
  if (code == 0u) return 0.0f;
  float jsum = 0.0f;
  const int n = a.count[c];
  for (int p = 0; p < n; ++p) {
    const uint32_t w = (code >> (2 * p)) & 3u;
    if (w == 0u) continue;
    const mw_pole& pl = a.poles[c][p];
    jsum += pole_step(a, pl, o, w == 2u ? pl.drive_scale : 0.5f * pl.drive_scale, e_old);
  }
  return jsum;
}

// dt / eps_inf of sample o, the factor of the polarisation-current subtraction
// E -= dt J / eps_inf (kernels/dispersive.cu:72-83): on a sample with E decay 1
// (no conductivity) the E curl coefficient is exactly dt / eps, so it is used and
// 1/eps_inf is not read; otherwise the inv_eps array is.
__device__ __forceinline__ float ade_gain(const mw_ade& a, int c, uint32_t o, float decay, float curl) {
  return decay == 1.0f ? curl : a.dt * G(a.inv_eps[c])[o];
}

// E update for all three components at node (i,j,k) (kernels/electric.cu:463-553):
//   Ex: a = dHz/dy, b = dHy/dz, curl = a - b
//   Ey: a = dHz/dx, b = dHx/dz, curl = b - a
//   Ez: a = dHy/dx, b = dHx/dy, curl = a - b
// backward differences with boundary modes, E <- E*decay + curl_coef*curl, then
// E -= dt * inv_eps * sum(J) when ADE is fused.
template <bool CPML, bool ADE>
__global__ __launch_bounds__(256) void update_e_kernel(mw_grid g, mw_coeffs cf, mw_cpml cp, mw_ade ad, mw_fields f,
                                                       int comp_mask, int xchunk) {
  const int k = blockIdx.x * kTileZ + threadIdx.x;
  const int j = blockIdx.y * kTileY + wave_row();
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  if (k >= nz || j >= ny) return;
  const int i0 = blockIdx.z * xchunk;
  const int i1 = min(i0 + xchunk, nx);
  const float* __restrict__ hx = f.h[0];
  const float* __restrict__ hy = f.h[1];
  const float* __restrict__ hz = f.h[2];
  const int lx = g.bc_lo[0], hxm = g.bc_hi[0], ly = g.bc_lo[1], hym = g.bc_hi[1], lz = g.bc_lo[2], hzm = g.bc_hi[2];
  const bool jin = j > 0 && j < ny - 1;  // interior along y
  const bool kin = k > 0 && k < nz - 1;
  const float iey = g.inv_e[1][j];
  const float iez = g.inv_e[2][k];
  float bey = 0, cey = 0, key = 1, bez = 0, cez = 0, kez = 1;
  if (CPML) {
    bey = cp.b_e[1][j]; cey = cp.c_e[1][j]; key = cp.inv_kappa_e[1][j];
    bez = cp.b_e[2][k]; cez = cp.c_e[2][k]; kez = cp.inv_kappa_e[2][k];
  }
  const bool jH = j < ny - 1;  // Hz/Hx exist at this j
  const bool kH = k < nz - 1;  // Hy/Hx exist at this k
  // Carried x-neighbours at plane i-1: Hz(i-1,j,k), Hy(i-1,j,k)
  float hz_m = 0.0f, hy_m = 0.0f;
  if (i0 > 0 && i0 - 1 < nx - 1) {
    if (jH) hz_m = hz[((uint32_t)(i0 - 1) * (ny - 1) + j) * nz + k];
    if (kH) hy_m = hy[((uint32_t)(i0 - 1) * ny + j) * (nz - 1) + k];
  }
  for (int i = i0; i < i1; ++i) {
    const bool iH = i < nx - 1;  // Hy/Hz exist at this i
    const float hz_c = (iH && jH) ? hz[((uint32_t)i * (ny - 1) + j) * nz + k] : 0.0f;
    const float hy_c = (iH && kH) ? hy[((uint32_t)i * ny + j) * (nz - 1) + k] : 0.0f;
    const float hx_c = (jH && kH) ? hx[((uint32_t)i * (ny - 1) + j) * (nz - 1) + k] : 0.0f;
    const bool iin = i > 0 && i < nx - 1;
    // ---- Ex (nx-1, ny, nz)
    if ((comp_mask & 1) && iH) {
      const uint32_t o = ((uint32_t)i * ny + j) * nz + k;
      float* e = f.e[0];
      const float e_old = e[o];
      const float jsum = ADE ? ade_advance(ad, 0, o, e_old) : 0.0f;
      float e_new = e_old;
      if (face_pec(j, ny, ly, hym) || face_pec(k, nz, lz, hzm)) {
        e_new = 0.0f;
      } else if (!(face_inactive(j, ny, ly, hym) || face_inactive(k, nz, lz, hzm))) {
        float da, db;
        if (jin) da = (hz_c - hz[((uint32_t)i * (ny - 1) + j - 1) * nz + k]) * iey;
        else da = edge_diff(j == 0, j == 0 ? ly : hym, hz[((uint32_t)i * (ny - 1)) * nz + k],
                            hz[((uint32_t)i * (ny - 1) + ny - 2) * nz + k], iey);
        if (kin) db = (hy_c - hy[((uint32_t)i * ny + j) * (nz - 1) + k - 1]) * iez;
        else db = edge_diff(k == 0, k == 0 ? lz : hzm, hy[((uint32_t)i * ny + j) * (nz - 1)],
                            hy[((uint32_t)i * ny + j) * (nz - 1) + nz - 2], iez);
        float curl;
        if (CPML) {
          const float pa = psi_step(psi_view(cp, 0), 1, j, i, j, k, ny, nz, bey, cey, da);
          const float pb = psi_step(psi_view(cp, 1), 2, k, i, j, k, ny, nz, bez, cez, db);
          curl = (da * key + pa) - (db * kez + pb);
        } else {
          curl = da - db;
        }
        e_new = e_old * coef(cf.decay[0], cf.decay_value[0], o) + coef(cf.curl[0], cf.curl_value[0], o) * curl;
      }
      if (ADE) {
        e_new -= ad.dt * jsum * ad.inv_eps[0][o];
        if (face_pec(j, ny, ly, hym) || face_pec(k, nz, lz, hzm)) e_new = 0.0f;
      }
      e[o] = e_new;
    }
    // ---- Ey (nx, ny-1, nz)
    if ((comp_mask & 2) && jH) {
      const uint32_t o = ((uint32_t)i * (ny - 1) + j) * nz + k;
      float* e = f.e[1];
      const float e_old = e[o];
      const float jsum = ADE ? ade_advance(ad, 1, o, e_old) : 0.0f;
      float e_new = e_old;
      const bool pec = face_pec(i, nx, lx, hxm) || face_pec(k, nz, lz, hzm);
      if (pec) {
        e_new = 0.0f;
      } else if (!(face_inactive(i, nx, lx, hxm) || face_inactive(k, nz, lz, hzm))) {
        float da, db;
        if (iin) da = (hz_c - hz_m) * g.inv_e[0][i];
        else da = edge_diff(i == 0, i == 0 ? lx : hxm, hz[(uint32_t)j * nz + k],
                            hz[((uint32_t)(nx - 2) * (ny - 1) + j) * nz + k], g.inv_e[0][i]);
        if (kin) db = (hx_c - hx[((uint32_t)i * (ny - 1) + j) * (nz - 1) + k - 1]) * iez;
        else db = edge_diff(k == 0, k == 0 ? lz : hzm, hx[((uint32_t)i * (ny - 1) + j) * (nz - 1)],
                            hx[((uint32_t)i * (ny - 1) + j) * (nz - 1) + nz - 2], iez);
        float curl;
        if (CPML) {
          const float pa = psi_step(psi_view(cp, 2), 0, i, i, j, k, ny - 1, nz, cp.b_e[0][i], cp.c_e[0][i], da);
          const float pb = psi_step(psi_view(cp, 3), 2, k, i, j, k, ny - 1, nz, bez, cez, db);
          curl = (db * kez + pb) - (da * cp.inv_kappa_e[0][i] + pa);
        } else {
          curl = db - da;
        }
        e_new = e_old * coef(cf.decay[1], cf.decay_value[1], o) + coef(cf.curl[1], cf.curl_value[1], o) * curl;
      }
      if (ADE) {
        e_new -= ad.dt * jsum * ad.inv_eps[1][o];
        if (pec) e_new = 0.0f;
      }
      e[o] = e_new;
    }
    // ---- Ez (nx, ny, nz-1)
    if ((comp_mask & 4) && kH) {
      const uint32_t o = ((uint32_t)i * ny + j) * (nz - 1) + k;
      float* e = f.e[2];
      const float e_old = e[o];
      const float jsum = ADE ? ade_advance(ad, 2, o, e_old) : 0.0f;
      float e_new = e_old;
      const bool pec = face_pec(i, nx, lx, hxm) || face_pec(j, ny, ly, hym);
      if (pec) {
        e_new = 0.0f;
      } else if (!(face_inactive(i, nx, lx, hxm) || face_inactive(j, ny, ly, hym))) {
        float da, db;
        if (iin) da = (hy_c - hy_m) * g.inv_e[0][i];
        else da = edge_diff(i == 0, i == 0 ? lx : hxm, hy[(uint32_t)j * (nz - 1) + k],
                            hy[((uint32_t)(nx - 2) * ny + j) * (nz - 1) + k], g.inv_e[0][i]);
        if (jin) db = (hx_c - hx[((uint32_t)i * (ny - 1) + j - 1) * (nz - 1) + k]) * iey;
        else db = edge_diff(j == 0, j == 0 ? ly : hym, hx[((uint32_t)i * (ny - 1)) * (nz - 1) + k],
                            hx[((uint32_t)i * (ny - 1) + ny - 2) * (nz - 1) + k], iey);
        float curl;
        if (CPML) {
          const float pa = psi_step(psi_view(cp, 4), 0, i, i, j, k, ny, nz - 1, cp.b_e[0][i], cp.c_e[0][i], da);
          const float pb = psi_step(psi_view(cp, 5), 1, j, i, j, k, ny, nz - 1, bey, cey, db);
          curl = (da * cp.inv_kappa_e[0][i] + pa) - (db * key + pb);
        } else {
          curl = da - db;
        }
        e_new = e_old * coef(cf.decay[2], cf.decay_value[2], o) + coef(cf.curl[2], cf.curl_value[2], o) * curl;
      }
      if (ADE) {
        e_new -= ad.dt * jsum * ad.inv_eps[2][o];
        if (pec) e_new = 0.0f;
      }
      e[o] = e_new;
    }
    hz_m = hz_c;
    hy_m = hy_c;
  }
}

}  // namespace mw
