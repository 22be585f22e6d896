// The PML shell of the single-pass step: every cell outside the lean interior
// box, updated cell by cell in two launches (H, then E) over a list of boxes.
//
// The lean kernel (fused.h) advances the interior box, whose cells are free of
// PML, faces and boundary logic.  What is left is a thin shell: up to six boxes
// (x-low/high slabs over the whole y-z cross-section, y-low/high slabs over the
// lean x range, z-low/high slabs over the lean x-y range).  A tile-per-wavefront
// fused update would spend most of its lanes outside the grid on the thin z
// slabs, so the shell uses per-cell kernels instead, with the lane layout chosen
// per box: W consecutive z samples x 256/W rows per block (W = 64 for the x/y
// slabs, 16 or 32 for a z slab 12-30 samples wide), marching along x over a
// chunk of planes with the x-neighbour carried in registers.
//
//   shell_h: H^{n+1/2} of the shell cells from E^n / H^{n-1/2} (ping-pong psi_h)
//   lean   : interior H and E (recomputes its own H halo, reads only inputs)
//   shell_e: E^{n+1} of the shell cells from the complete H^{n+1/2}
//
// shell_h and lean read only the step's input buffers and write disjoint
// output cells; shell_e runs after both.  The update equations are those of the
// two-kernel path (yee.h: kernels/magnetic.cu:62-117, kernels/electric.cu:463-553)
// with the single-pass ordering of ADE, source terms, PEC and DFT.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fused.h"
#include "maxwell_fdtd.h"
#include "yee.h"

namespace mw {

constexpr int kShellMaxBoxes = 6;

struct ShellBox {
  int32_t lo[3], hi[3];  // cell box [lo, hi) (any component's samples inside it)
  int32_t w_shift;       // log2 of the z lanes per row (4, 5 or 6)
  int32_t chunk;         // planes per block along x
  int32_t nbz, nby, nbx; // blocks along z, y, x
  int32_t first;         // first block of this box in the launch
};

struct ShellArgs {
  FusedArgs f;  // grid, coefficients, CPML (psi[6..11] = H-side input), ADE, in/out, terms, DFT
  int32_t nboxes;
  int32_t total;  // blocks over all boxes
  ShellBox box[kShellMaxBoxes];
};

// Block -> (box, k, j, i range) for the calling thread.  Returns false for the
// padding blocks of the launch.
struct ShellCell {
  int b;            // box
  int k, j;         // this thread's column and row
  int i0, i1;       // planes of the block
  int krow0, krow1; // block's column range (for the term guards)
  int jrow0, jrow1; // block's row range
};

__device__ __forceinline__ bool shell_cell(const ShellArgs& a, int logical, ShellCell& c) {
  if (logical >= a.total) return false;
  int b = 0;
  while (b + 1 < a.nboxes && logical >= a.box[b + 1].first) ++b;
  const ShellBox& x = a.box[b];
  const int t = logical - x.first;
  const int bz = t % x.nbz, by = (t / x.nbz) % x.nby, bx = t / (x.nbz * x.nby);
  const int tid = threadIdx.x + 64 * threadIdx.y;
  const int w = 1 << x.w_shift;
  const int rows = 256 >> x.w_shift;
  c.b = b;
  c.krow0 = x.lo[2] + bz * w;
  c.krow1 = min(c.krow0 + w, x.hi[2]) - 1;
  c.jrow0 = x.lo[1] + by * rows;
  c.jrow1 = min(c.jrow0 + rows, x.hi[1]) - 1;
  c.k = c.krow0 + (tid & (w - 1));
  c.j = c.jrow0 + (tid >> x.w_shift);
  c.i0 = x.lo[0] + bx * x.chunk;
  c.i1 = min(c.i0 + x.chunk, x.hi[0]);
  return true;
}

// psi <- b*psi_in + c*d, written to psi_out (H-side ping-pong); 0 outside the slab.
__device__ __forceinline__ float psi_pp(const float* in, float* out, int low, int hs, int hl, int axis, int coord,
                                        int i, int j, int k, int sy, int sz, float b, float c, float d) {
  const int local = slab_local(coord, low, hs, hl);
  if (local < 0) return 0.0f;
  const uint32_t o = slab_offset(axis, i, j, k, local, sy, sz, low + hl);
  const float nv = b * in[o] + c * d;
  out[o] = nv;
  return nv;
}

template <bool CPML>
__global__ __launch_bounds__(256) void shell_h_kernel(ShellArgs sa) {
  ShellCell c;
  if (!shell_cell(sa, xcd_logical_block(), c)) return;
  const FusedArgs& a = sa.f;
  const ShellBox& bx = sa.box[c.b];
  const int nx = a.g.nx, ny = a.g.ny, nz = a.g.nz;
  const int j = c.j, k = c.k;
  const bool inside = j < bx.hi[1] && k < bx.hi[2];
  const bool jv = inside && j < ny - 1, kv = inside && k < nz - 1;
  const float* __restrict__ ex = a.in.e[0];
  const float* __restrict__ ey = a.in.e[1];
  const float* __restrict__ ez = a.in.e[2];
  const float ihy = jv ? a.g.inv_h[1][j] : 0.0f;
  const float ihz = kv ? a.g.inv_h[2][k] : 0.0f;
  float bhy = 0, chy = 0, khy = 1, bhz = 0, chz = 0, khz = 1;
  if (CPML) {
    if (jv) { bhy = a.cp.b_h[1][j]; chy = a.cp.c_h[1][j]; khy = a.cp.inv_kappa_h[1][j]; }
    if (kv) { bhz = a.cp.b_h[2][k]; chz = a.cp.c_h[2][k]; khz = a.cp.inv_kappa_h[2][k]; }
  }
  // carried x-neighbours Ez(i, j, k), Ey(i, j, k)
  float ez_c = (kv && inside) ? ez[((uint32_t)c.i0 * ny + j) * (nz - 1) + k] : 0.0f;
  float ey_c = (jv && inside) ? ey[((uint32_t)c.i0 * (ny - 1) + j) * nz + k] : 0.0f;
  for (int i = c.i0; i < c.i1; ++i) {
    const bool ht = terms_touch(a.h_terms, a.n_h_terms, i, c.jrow0, c.jrow1, c.krow0, c.krow1);
    const bool iv = i < nx - 1;
    float ez_n = 0.0f, ey_n = 0.0f;
    if (iv && inside) {
      if (kv) ez_n = ez[((uint32_t)(i + 1) * ny + j) * (nz - 1) + k];
      if (jv) ey_n = ey[((uint32_t)(i + 1) * (ny - 1) + j) * nz + k];
    }
    if (jv && kv) {  // Hx (nx, ny-1, nz-1): a = dEz/dy, b = dEy/dz, curl = a - b
      const uint32_t o = ((uint32_t)i * (ny - 1) + j) * (nz - 1) + k;
      const float da = (ez[((uint32_t)i * ny + j + 1) * (nz - 1) + k] - ez_c) * ihy;
      const float db = (ey[((uint32_t)i * (ny - 1) + j) * nz + k + 1] - ey_c) * ihz;
      float curl;
      if (CPML) {
        const float pa = psi_pp(a.cp.psi[6], a.psi_h_out[0], a.cp.psi_low[6], a.cp.psi_high_start[6],
                                a.cp.psi_high_len[6], 1, j, i, j, k, ny - 1, nz - 1, bhy, chy, da);
        const float pb = psi_pp(a.cp.psi[7], a.psi_h_out[1], a.cp.psi_low[7], a.cp.psi_high_start[7],
                                a.cp.psi_high_len[7], 2, k, i, j, k, ny - 1, nz - 1, bhz, chz, db);
        curl = (da * khy + pa) - (db * khz + pb);
      } else {
        curl = da - db;
      }
      float h = a.in.h[0][o] * coef(a.cf.decay[3], a.cf.decay_value[3], o) - coef(a.cf.curl[3], a.cf.curl_value[3], o) * curl;
      if (ht) h += term_value(a.h_terms, a.n_h_terms, 3, i, j, k, a.step, a.dt);
      a.out.h[0][o] = h;
    }
    if (iv && inside && (kv || jv)) {
      const float ihx = a.g.inv_h[0][i];
      float bhx = 0, chx = 0, khx = 1;
      if (CPML) { bhx = a.cp.b_h[0][i]; chx = a.cp.c_h[0][i]; khx = a.cp.inv_kappa_h[0][i]; }
      const float ex_c = ex[((uint32_t)i * ny + j) * nz + k];
      if (kv) {  // Hy (nx-1, ny, nz-1): a = dEz/dx, b = dEx/dz, curl = b - a
        const uint32_t o = ((uint32_t)i * ny + j) * (nz - 1) + k;
        const float da = (ez_n - ez_c) * ihx;
        const float db = (ex[((uint32_t)i * ny + j) * nz + k + 1] - ex_c) * ihz;
        float curl;
        if (CPML) {
          const float pa = psi_pp(a.cp.psi[8], a.psi_h_out[2], a.cp.psi_low[8], a.cp.psi_high_start[8],
                                  a.cp.psi_high_len[8], 0, i, i, j, k, ny, nz - 1, bhx, chx, da);
          const float pb = psi_pp(a.cp.psi[9], a.psi_h_out[3], a.cp.psi_low[9], a.cp.psi_high_start[9],
                                  a.cp.psi_high_len[9], 2, k, i, j, k, ny, nz - 1, bhz, chz, db);
          curl = (db * khz + pb) - (da * khx + pa);
        } else {
          curl = db - da;
        }
        float h = a.in.h[1][o] * coef(a.cf.decay[4], a.cf.decay_value[4], o) - coef(a.cf.curl[4], a.cf.curl_value[4], o) * curl;
        if (ht) h += term_value(a.h_terms, a.n_h_terms, 4, i, j, k, a.step, a.dt);
        a.out.h[1][o] = h;
      }
      if (jv) {  // Hz (nx-1, ny-1, nz): a = dEy/dx, b = dEx/dy, curl = a - b
        const uint32_t o = ((uint32_t)i * (ny - 1) + j) * nz + k;
        const float da = (ey_n - ey_c) * ihx;
        const float db = (ex[((uint32_t)i * ny + j + 1) * nz + k] - ex_c) * ihy;
        float curl;
        if (CPML) {
          const float pa = psi_pp(a.cp.psi[10], a.psi_h_out[4], a.cp.psi_low[10], a.cp.psi_high_start[10],
                                  a.cp.psi_high_len[10], 0, i, i, j, k, ny - 1, nz, bhx, chx, da);
          const float pb = psi_pp(a.cp.psi[11], a.psi_h_out[5], a.cp.psi_low[11], a.cp.psi_high_start[11],
                                  a.cp.psi_high_len[11], 1, j, i, j, k, ny - 1, nz, bhy, chy, db);
          curl = (da * khx + pa) - (db * khy + pb);
        } else {
          curl = da - db;
        }
        float h = a.in.h[2][o] * coef(a.cf.decay[5], a.cf.decay_value[5], o) - coef(a.cf.curl[5], a.cf.curl_value[5], o) * curl;
        if (ht) h += term_value(a.h_terms, a.n_h_terms, 5, i, j, k, a.step, a.dt);
        a.out.h[2][o] = h;
      }
    }
    ez_c = ez_n;
    ey_c = ey_n;
  }
}

// Derivative at a boundary sample of the single-pass step: PMC mirrors (the
// reference's kernels/electric.cu:23-41), every other non-periodic mode gives 0.
__device__ __forceinline__ float shell_edge(bool low_end, int mode, float first, float last, float inv) {
  if (mode != MW_BC_PMC) return 0.0f;
  return low_end ? 2.0f * first * inv : -2.0f * last * inv;
}

template <bool DFT>
__device__ __forceinline__ void shell_dft(const FusedArgs& a, int c, uint64_t n, uint32_t o, float e) {
  if (!DFT) return;
  for (int f = 0; f < a.nf_dft; ++f) {
    a.dft_re[c][f * n + o] += e * a.dft_cos[f];
    a.dft_im[c][f * n + o] += e * a.dft_sin[f];
  }
}

template <bool CPML, bool ADE, bool DFT>
__global__ __launch_bounds__(256) void shell_e_kernel(ShellArgs sa) {
  ShellCell c;
  if (!shell_cell(sa, xcd_logical_block(), c)) return;
  const FusedArgs& a = sa.f;
  const ShellBox& bx = sa.box[c.b];
  const int nx = a.g.nx, ny = a.g.ny, nz = a.g.nz;
  const int j = c.j, k = c.k;
  if (j >= bx.hi[1] || k >= bx.hi[2]) {
    // idle lanes still take part in nothing below (no shuffles in this kernel)
    return;
  }
  const float* __restrict__ hx = a.out.h[0];  // H^{n+1/2}, complete after shell_h and lean
  const float* __restrict__ hy = a.out.h[1];
  const float* __restrict__ hz = a.out.h[2];
  const int lx = a.g.bc_lo[0], hxm = a.g.bc_hi[0], ly = a.g.bc_lo[1], hym = a.g.bc_hi[1], lz = a.g.bc_lo[2],
            hzm = a.g.bc_hi[2];
  const bool jin = j > 0 && j < ny - 1;
  const bool kin = k > 0 && k < nz - 1;
  const float iey = a.g.inv_e[1][j];
  const float iez = a.g.inv_e[2][k];
  float bey = 0, cey = 0, key = 1, bez = 0, cez = 0, kez = 1;
  if (CPML) {
    bey = a.cp.b_e[1][j]; cey = a.cp.c_e[1][j]; key = a.cp.inv_kappa_e[1][j];
    bez = a.cp.b_e[2][k]; cez = a.cp.c_e[2][k]; kez = a.cp.inv_kappa_e[2][k];
  }
  const bool jH = j < ny - 1, kH = k < nz - 1;
  const uint32_t sEx = (uint32_t)ny * nz, sEy = (uint32_t)(ny - 1) * nz, sEz = (uint32_t)ny * (nz - 1);
  // carried Hz(i-1, j, k), Hy(i-1, j, k)
  float hz_m = 0.0f, hy_m = 0.0f;
  if (c.i0 > 0 && c.i0 - 1 < nx - 1) {
    if (jH) hz_m = hz[((uint32_t)(c.i0 - 1) * (ny - 1) + j) * nz + k];
    if (kH) hy_m = hy[((uint32_t)(c.i0 - 1) * ny + j) * (nz - 1) + k];
  }
  for (int i = c.i0; i < c.i1; ++i) {
    const bool et = terms_touch(a.e_terms, a.n_e_terms, i, c.jrow0, c.jrow1, c.krow0, c.krow1);
    const bool iH = i < nx - 1;
    const float hz_c = (iH && jH) ? hz[((uint32_t)i * (ny - 1) + j) * nz + k] : 0.0f;
    const float hy_c = (iH && kH) ? hy[((uint32_t)i * ny + j) * (nz - 1) + k] : 0.0f;
    const float hx_c = (jH && kH) ? hx[((uint32_t)i * (ny - 1) + j) * (nz - 1) + k] : 0.0f;
    const bool iin = i > 0 && i < nx - 1;
    if (iH) {  // Ex (nx-1, ny, nz): a = dHz/dy, b = dHy/dz, curl = a - b
      const uint32_t o = i * sEx + (uint32_t)j * nz + k;
      const float e_old = a.in.e[0][o];
      const float jsum = ADE ? ade_advance(a.ad, 0, o, e_old) : 0.0f;
      const bool pec = face_pec(j, ny, ly, hym) || face_pec(k, nz, lz, hzm);
      float e_new = e_old;
      if (!pec && !(face_inactive(j, ny, ly, hym) || face_inactive(k, nz, lz, hzm))) {
        const float hz_jm = j > 0 ? hz[((uint32_t)i * (ny - 1) + j - 1) * nz + k] : 0.0f;
        const float hy_km = k > 0 ? hy[((uint32_t)i * ny + j) * (nz - 1) + k - 1] : 0.0f;
        const float da = jin ? (hz_c - hz_jm) * iey : shell_edge(j == 0, j == 0 ? ly : hym, hz_c, hz_jm, iey);
        const float db = kin ? (hy_c - hy_km) * iez : shell_edge(k == 0, k == 0 ? lz : hzm, hy_c, hy_km, iez);
        float curl;
        if (CPML) {
          const float pa = psi_step(psi_view(a.cp, 0), 1, j, i, j, k, ny, nz, bey, cey, da);
          const float pb = psi_step(psi_view(a.cp, 1), 2, k, i, j, k, ny, nz, bez, cez, db);
          curl = (da * key + pa) - (db * kez + pb);
        } else {
          curl = da - db;
        }
        e_new = e_old * coef(a.cf.decay[0], a.cf.decay_value[0], o) + coef(a.cf.curl[0], a.cf.curl_value[0], o) * curl;
      }
      if (ADE) e_new -= a.ad.dt * jsum * a.ad.inv_eps[0][o];
      if (et) e_new += term_value(a.e_terms, a.n_e_terms, 0, i, j, k, a.step, a.dt);
      if (pec) e_new = 0.0f;
      a.out.e[0][o] = e_new;
      shell_dft<DFT>(a, 0, (uint64_t)sEx * (nx - 1), o, e_new);
    }
    if (jH) {  // Ey (nx, ny-1, nz): a = dHz/dx, b = dHx/dz, curl = b - a
      const uint32_t o = i * sEy + (uint32_t)j * nz + k;
      const float e_old = a.in.e[1][o];
      const float jsum = ADE ? ade_advance(a.ad, 1, o, e_old) : 0.0f;
      const bool pec = face_pec(i, nx, lx, hxm) || face_pec(k, nz, lz, hzm);
      float e_new = e_old;
      if (!pec && !(face_inactive(i, nx, lx, hxm) || face_inactive(k, nz, lz, hzm))) {
        const float iex = a.g.inv_e[0][i];
        const float hx_km = k > 0 ? hx[((uint32_t)i * (ny - 1) + j) * (nz - 1) + k - 1] : 0.0f;
        const float da = iin ? (hz_c - hz_m) * iex : shell_edge(i == 0, i == 0 ? lx : hxm, hz_c, hz_m, iex);
        const float db = kin ? (hx_c - hx_km) * iez : shell_edge(k == 0, k == 0 ? lz : hzm, hx_c, hx_km, iez);
        float curl;
        if (CPML) {
          const float pa = psi_step(psi_view(a.cp, 2), 0, i, i, j, k, ny - 1, nz, a.cp.b_e[0][i], a.cp.c_e[0][i], da);
          const float pb = psi_step(psi_view(a.cp, 3), 2, k, i, j, k, ny - 1, nz, bez, cez, db);
          curl = (db * kez + pb) - (da * a.cp.inv_kappa_e[0][i] + pa);
        } else {
          curl = db - da;
        }
        e_new = e_old * coef(a.cf.decay[1], a.cf.decay_value[1], o) + coef(a.cf.curl[1], a.cf.curl_value[1], o) * curl;
      }
      if (ADE) e_new -= a.ad.dt * jsum * a.ad.inv_eps[1][o];
      if (et) e_new += term_value(a.e_terms, a.n_e_terms, 1, i, j, k, a.step, a.dt);
      if (pec) e_new = 0.0f;
      a.out.e[1][o] = e_new;
      shell_dft<DFT>(a, 1, (uint64_t)sEy * nx, o, e_new);
    }
    if (kH) {  // Ez (nx, ny, nz-1): a = dHy/dx, b = dHx/dy, curl = a - b
      const uint32_t o = i * sEz + (uint32_t)j * (nz - 1) + k;
      const float e_old = a.in.e[2][o];
      const float jsum = ADE ? ade_advance(a.ad, 2, o, e_old) : 0.0f;
      const bool pec = face_pec(i, nx, lx, hxm) || face_pec(j, ny, ly, hym);
      float e_new = e_old;
      if (!pec && !(face_inactive(i, nx, lx, hxm) || face_inactive(j, ny, ly, hym))) {
        const float iex = a.g.inv_e[0][i];
        const float hx_jm = j > 0 ? hx[((uint32_t)i * (ny - 1) + j - 1) * (nz - 1) + k] : 0.0f;
        const float da = iin ? (hy_c - hy_m) * iex : shell_edge(i == 0, i == 0 ? lx : hxm, hy_c, hy_m, iex);
        const float db = jin ? (hx_c - hx_jm) * iey : shell_edge(j == 0, j == 0 ? ly : hym, hx_c, hx_jm, iey);
        float curl;
        if (CPML) {
          const float pa = psi_step(psi_view(a.cp, 4), 0, i, i, j, k, ny, nz - 1, a.cp.b_e[0][i], a.cp.c_e[0][i], da);
          const float pb = psi_step(psi_view(a.cp, 5), 1, j, i, j, k, ny, nz - 1, bey, cey, db);
          curl = (da * a.cp.inv_kappa_e[0][i] + pa) - (db * key + pb);
        } else {
          curl = da - db;
        }
        e_new = e_old * coef(a.cf.decay[2], a.cf.decay_value[2], o) + coef(a.cf.curl[2], a.cf.curl_value[2], o) * curl;
      }
      if (ADE) e_new -= a.ad.dt * jsum * a.ad.inv_eps[2][o];
      if (et) e_new += term_value(a.e_terms, a.n_e_terms, 2, i, j, k, a.step, a.dt);
      if (pec) e_new = 0.0f;
      a.out.e[2][o] = e_new;
      shell_dft<DFT>(a, 2, (uint64_t)sEz * nx, o, e_new);
    }
    hz_m = hz_c;
    hy_m = hy_c;
  }
}

}  // namespace mw
