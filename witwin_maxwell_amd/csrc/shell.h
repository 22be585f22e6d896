// The PML shell of the single-pass step: every cell outside the lean interior
// box, updated cell by cell in two launches (H, then E) over a list of boxes.
//
// The lean kernel (fused.h) advances the interior box, whose cells are free of
// PML, faces and boundary logic.  What is left is a thin shell: up to six boxes
// (x-low/high slabs over the whole y-z cross-section, y-low/high slabs over the
// lean x range and, unless the z ends are z-edge tiles (zedge.h), z-low/high
// slabs over the lean x-y range).  A tile-per-wavefront fused update would spend
// most of its lanes outside the grid on the thin z slabs, so the shell uses
// per-cell kernels instead, with the lane layout chosen per box: W consecutive z
// samples x 256/W rows per block (W = 16, 32 or 64, whichever leaves the fewest
// idle lanes over the box's z extent; 16 only with dispersive poles, whose E pass is
// heavy enough per cell that idle lanes cost more than the extra row segments per
// wavefront), marching along x over a chunk of planes with the x-neighbour carried
// in registers.
//
//   shell_h: H^{n+1/2} of the shell cells from E^n / H^{n-1/2} (ping-pong psi_h)
//   lean   : interior H and E (recomputes its own H halo, reads only inputs)
//   z edges: the z ends of the interior rows, when tiled that way (zedge.h)
//   shell_e: E^{n+1} of the shell cells from the complete H^{n+1/2}
//
// shell_h, lean and the z-edge tiles read only the step's input buffers and write
// disjoint output cells; shell_e runs after them, so it can read H across a
// periodic seam (periodic faces are shell cells).  The update equations are those of the
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
__device__ __forceinline__ float psi_pp(gptr<const float> __restrict__ in, gptr<float> __restrict__ out, int low,
                                        int hs, int hl, int axis, int coord, int i, int j, int k, int sy, int sz,
                                        float b, float c, float d) {
  const int local = slab_local(coord, low, hs, hl);
  if (local < 0) return 0.0f;
  const uint32_t o = slab_offset(axis, i, j, k, local, sy, sz, low + hl);
  const float nv = b * in[o] + c * d;
  out[o] = nv;
  return nv;
}

// psi <- b*psi + c*d in place (E side); 0 outside the slab.
__device__ __forceinline__ float psi_ip(gptr<float> __restrict__ psi, int low, int hs, int hl, int axis, int coord,
                                        int i, int j, int k, int sy, int sz, float b, float c, float d) {
  const int local = slab_local(coord, low, hs, hl);
  if (local < 0) return 0.0f;
  const uint32_t o = slab_offset(axis, i, j, k, local, sy, sz, low + hl);
  const float nv = b * psi[o] + c * d;
  psi[o] = nv;
  return nv;
}

__device__ __forceinline__ float coef_r(gptr<const float> __restrict__ arr, float value, uint32_t o) {
  return arr ? arr[o] : value;
}

// The shell kernels take every buffer as a restrict-qualified parameter of an
// inlined function, so the compiler may issue all loads of a plane before its
// stores instead of waiting for each store in case they alias.

template <bool CPML>
__device__ __forceinline__ void shell_h_cells(
    const ShellArgs& sa, const ShellCell& c, int step, gptr<const float> __restrict__ ex,
    gptr<const float> __restrict__ ey, gptr<const float> __restrict__ ez, gptr<const float> __restrict__ hxi,
    gptr<const float> __restrict__ hyi, gptr<const float> __restrict__ hzi, gptr<float> __restrict__ hxo,
    gptr<float> __restrict__ hyo, gptr<float> __restrict__ hzo, gptr<const float> __restrict__ pi0,
    gptr<const float> __restrict__ pi1, gptr<const float> __restrict__ pi2, gptr<const float> __restrict__ pi3,
    gptr<const float> __restrict__ pi4, gptr<const float> __restrict__ pi5, gptr<float> __restrict__ po0,
    gptr<float> __restrict__ po1, gptr<float> __restrict__ po2, gptr<float> __restrict__ po3,
    gptr<float> __restrict__ po4, gptr<float> __restrict__ po5) {
  const FusedArgs& a = sa.f;
  const mw_cpml& cp = a.cp;
  const ShellBox& bx = sa.box[c.b];
  const int nx = a.g.nx, ny = a.g.ny, nz = a.g.nz;
  const int j = c.j, k = c.k;
  const bool inside = j < bx.hi[1] && k < bx.hi[2];
  const bool jv = inside && j < ny - 1, kv = inside && k < nz - 1;
  // Every load of a plane is unconditional, at offset 0 when the sample does not
  // exist (the value is then discarded), so all of them issue before the plane's
  // stores; only the stores are predicated.
  const float ihy = G(a.g.inv_h[1])[jv ? j : 0];
  const float ihz = G(a.g.inv_h[2])[kv ? k : 0];
  float bhy = 0, chy = 0, khy = 1, bhz = 0, chz = 0, khz = 1;
  if (CPML) {
    bhy = G(cp.b_h[1])[jv ? j : 0]; chy = G(cp.c_h[1])[jv ? j : 0]; khy = G(cp.inv_kappa_h[1])[jv ? j : 0];
    bhz = G(cp.b_h[2])[kv ? k : 0]; chz = G(cp.c_h[2])[kv ? k : 0]; khz = G(cp.inv_kappa_h[2])[kv ? k : 0];
  }
  // psi slab positions that do not depend on the plane (-1 outside the slab)
  const int l6 = slab_local(j, cp.psi_low[6], cp.psi_high_start[6], cp.psi_high_len[6]);
  const int l7 = slab_local(k, cp.psi_low[7], cp.psi_high_start[7], cp.psi_high_len[7]);
  const int l9 = slab_local(k, cp.psi_low[9], cp.psi_high_start[9], cp.psi_high_len[9]);
  const int l11 = slab_local(j, cp.psi_low[11], cp.psi_high_start[11], cp.psi_high_len[11]);
  const uint32_t sEx = (uint32_t)ny * nz, sEy = (uint32_t)(ny - 1) * nz, sEz = (uint32_t)ny * (nz - 1);
  const uint32_t sHx = (uint32_t)(ny - 1) * (nz - 1), sHy = (uint32_t)ny * (nz - 1), sHz = (uint32_t)(ny - 1) * nz;
  float ez_c = ez[kv ? (uint32_t)c.i0 * sEz + (uint32_t)j * (nz - 1) + k : 0];
  float ey_c = ey[jv ? (uint32_t)c.i0 * sEy + (uint32_t)j * nz + k : 0];
  for (int i = c.i0; i < c.i1; ++i) {
    const bool ht = terms_touch(a.h_terms, a.n_h_terms, i, c.jrow0, c.jrow1, c.krow0, c.krow1);
    const bool iv = i < nx - 1;
    const bool dx = jv && kv, dy = iv && kv, dz = iv && jv;  // Hx, Hy, Hz exist here
    const uint32_t ox = dx ? i * sHx + (uint32_t)j * (nz - 1) + k : 0;
    const uint32_t oy = dy ? i * sHy + (uint32_t)j * (nz - 1) + k : 0;
    const uint32_t oz = dz ? i * sHz + (uint32_t)j * nz + k : 0;
    const uint32_t oex = (iv && inside) ? i * sEx + (uint32_t)j * nz + k : 0;
    // loads
    const float ez_n = ez[dy ? (i + 1) * sEz + (uint32_t)j * (nz - 1) + k : 0];
    const float ey_n = ey[dz ? (i + 1) * sEy + (uint32_t)j * nz + k : 0];
    const float ez_jp = ez[dx ? i * sEz + (uint32_t)(j + 1) * (nz - 1) + k : 0];
    const float ey_kp = ey[dx ? i * sEy + (uint32_t)j * nz + k + 1 : 0];
    const float ex_c = ex[oex];
    const float ex_kp = ex[dy ? oex + 1 : 0];
    const float ex_jp = ex[dz ? oex + nz : 0];
    const float hx_old = hxi[ox], hy_old = hyi[oy], hz_old = hzi[oz];
    const float ihx = G(a.g.inv_h[0])[iv ? i : 0];
    float bhx = 0, chx = 0, khx = 1;
    float p6 = 0, p7 = 0, p8 = 0, p9 = 0, p10 = 0, p11 = 0;
    uint32_t q6 = 0, q7 = 0, q8 = 0, q9 = 0, q10 = 0, q11 = 0;
    int l8 = -1, l10 = -1;
    if (CPML) {
      bhx = G(cp.b_h[0])[iv ? i : 0]; chx = G(cp.c_h[0])[iv ? i : 0]; khx = G(cp.inv_kappa_h[0])[iv ? i : 0];
      l8 = slab_local(i, cp.psi_low[8], cp.psi_high_start[8], cp.psi_high_len[8]);
      l10 = slab_local(i, cp.psi_low[10], cp.psi_high_start[10], cp.psi_high_len[10]);
      if (dx && l6 >= 0) q6 = slab_offset(1, i, j, k, l6, ny - 1, nz - 1, cp.psi_low[6] + cp.psi_high_len[6]);
      if (dx && l7 >= 0) q7 = slab_offset(2, i, j, k, l7, ny - 1, nz - 1, cp.psi_low[7] + cp.psi_high_len[7]);
      if (dy && l8 >= 0) q8 = slab_offset(0, i, j, k, l8, ny, nz - 1, cp.psi_low[8] + cp.psi_high_len[8]);
      if (dy && l9 >= 0) q9 = slab_offset(2, i, j, k, l9, ny, nz - 1, cp.psi_low[9] + cp.psi_high_len[9]);
      if (dz && l10 >= 0) q10 = slab_offset(0, i, j, k, l10, ny - 1, nz, cp.psi_low[10] + cp.psi_high_len[10]);
      if (dz && l11 >= 0) q11 = slab_offset(1, i, j, k, l11, ny - 1, nz, cp.psi_low[11] + cp.psi_high_len[11]);
      // an axis without PML layers has no psi buffer: the test is uniform
      if (pi0) p6 = pi0[q6];
      if (pi1) p7 = pi1[q7];
      if (pi2) p8 = pi2[q8];
      if (pi3) p9 = pi3[q9];
      if (pi4) p10 = pi4[q10];
      if (pi5) p11 = pi5[q11];
    }
    // Hx: a = dEz/dy, b = dEy/dz, curl = a - b
    {
      const float da = (ez_jp - ez_c) * ihy, db = (ey_kp - ey_c) * ihz;
      float curl = da - db;
      if (CPML) {
        const float pa = l6 >= 0 ? bhy * p6 + chy * da : 0.0f;
        const float pb = l7 >= 0 ? bhz * p7 + chz * db : 0.0f;
        if (dx && l6 >= 0) po0[q6] = pa;
        if (dx && l7 >= 0) po1[q7] = pb;
        curl = (da * khy + pa) - (db * khz + pb);
      }
      float h = hx_old * coef(a.cf.decay[3], a.cf.decay_value[3], ox) - coef(a.cf.curl[3], a.cf.curl_value[3], ox) * curl;
      if (ht && dx) h += term_value(a.h_terms, a.n_h_terms, 3, i, j, k, step, a.dt);
      if (dx) hxo[ox] = h;
    }
    // Hy: a = dEz/dx, b = dEx/dz, curl = b - a
    {
      const float da = (ez_n - ez_c) * ihx, db = (ex_kp - ex_c) * ihz;
      float curl = db - da;
      if (CPML) {
        const float pa = l8 >= 0 ? bhx * p8 + chx * da : 0.0f;
        const float pb = l9 >= 0 ? bhz * p9 + chz * db : 0.0f;
        if (dy && l8 >= 0) po2[q8] = pa;
        if (dy && l9 >= 0) po3[q9] = pb;
        curl = (db * khz + pb) - (da * khx + pa);
      }
      float h = hy_old * coef(a.cf.decay[4], a.cf.decay_value[4], oy) - coef(a.cf.curl[4], a.cf.curl_value[4], oy) * curl;
      if (ht && dy) h += term_value(a.h_terms, a.n_h_terms, 4, i, j, k, step, a.dt);
      if (dy) hyo[oy] = h;
    }
    // Hz: a = dEy/dx, b = dEx/dy, curl = a - b
    {
      const float da = (ey_n - ey_c) * ihx, db = (ex_jp - ex_c) * ihy;
      float curl = da - db;
      if (CPML) {
        const float pa = l10 >= 0 ? bhx * p10 + chx * da : 0.0f;
        const float pb = l11 >= 0 ? bhy * p11 + chy * db : 0.0f;
        if (dz && l10 >= 0) po4[q10] = pa;
        if (dz && l11 >= 0) po5[q11] = pb;
        curl = (da * khx + pa) - (db * khy + pb);
      }
      float h = hz_old * coef(a.cf.decay[5], a.cf.decay_value[5], oz) - coef(a.cf.curl[5], a.cf.curl_value[5], oz) * curl;
      if (ht && dz) h += term_value(a.h_terms, a.n_h_terms, 5, i, j, k, step, a.dt);
      if (dz) hzo[oz] = h;
    }
    ez_c = ez_n;
    ey_c = ey_n;
  }
}

// H of the shell cells of block `logical` for step `step`.
template <bool CPML>
__device__ __forceinline__ void shell_h_tile(const ShellArgs& sa, int logical, int step) {
  ShellCell c;
  if (!shell_cell(sa, logical, c)) return;
  const FusedArgs& a = sa.f;
  shell_h_cells<CPML>(sa, c, step, G(a.in.e[0]), G(a.in.e[1]), G(a.in.e[2]), G(a.in.h[0]), G(a.in.h[1]),
                      G(a.in.h[2]), G(a.out.h[0]), G(a.out.h[1]), G(a.out.h[2]), G(a.cp.psi[6]), G(a.cp.psi[7]),
                      G(a.cp.psi[8]), G(a.cp.psi[9]), G(a.cp.psi[10]), G(a.cp.psi[11]), G(a.psi_h_out[0]),
                      G(a.psi_h_out[1]), G(a.psi_h_out[2]), G(a.psi_h_out[3]), G(a.psi_h_out[4]),
                      G(a.psi_h_out[5]));
}

// Derivative at a boundary sample of the single-pass step (the reference's
// kernels/electric.cu:23-41): `cur` / `prev` are the H samples at the sample's own
// index and the one below, `wrap` the one across the periodic seam (index n-2 at
// the low end, 0 at the high end).  Periodic wraps through the duplicated end
// node, PMC mirrors, every other mode gives 0.  H^{n+1/2} is complete when the
// shell's E pass runs, so the seam sample is there to read.
__device__ __forceinline__ float shell_edge(bool low_end, int mode, float cur, float prev, float wrap, float inv) {
  if (mode == MW_BC_PMC) return low_end ? 2.0f * cur * inv : -2.0f * prev * inv;
  if (mode == MW_BC_PERIODIC) return low_end ? (cur - wrap) * inv : (wrap - prev) * inv;
  return 0.0f;
}

template <bool DFT>
__device__ __forceinline__ void shell_dft(const FusedArgs& a, gptr<float> __restrict__ re, gptr<float> __restrict__ im,
                                          uint64_t n, uint32_t o, float e, gptr<const float> __restrict__ wc,
                                          gptr<const float> __restrict__ ws) {
  if (!DFT) return;
  for (int f = 0; f < a.nf_dft; ++f) {
    re[f * n + o] += e * wc[f];
    im[f * n + o] += e * ws[f];
  }
}

template <bool CPML, bool ADE, bool DFT, bool PER>
__device__ __forceinline__ void shell_e_cells(
    const ShellArgs& sa, const ShellCell& c, int step, gptr<const float> __restrict__ wc,
    gptr<const float> __restrict__ ws, gptr<const float> __restrict__ hx, gptr<const float> __restrict__ hy,
    gptr<const float> __restrict__ hz, gptr<const float> __restrict__ exi, gptr<const float> __restrict__ eyi,
    gptr<const float> __restrict__ ezi, gptr<float> __restrict__ exo, gptr<float> __restrict__ eyo,
    gptr<float> __restrict__ ezo, gptr<float> __restrict__ ps0, gptr<float> __restrict__ ps1,
    gptr<float> __restrict__ ps2, gptr<float> __restrict__ ps3, gptr<float> __restrict__ ps4,
    gptr<float> __restrict__ ps5, gptr<const float> __restrict__ dx_, gptr<const float> __restrict__ dy_,
    gptr<const float> __restrict__ dz_, gptr<const float> __restrict__ cx_, gptr<const float> __restrict__ cy_,
    gptr<const float> __restrict__ cz_, gptr<float> __restrict__ dre0, gptr<float> __restrict__ dim0,
    gptr<float> __restrict__ dre1, gptr<float> __restrict__ dim1, gptr<float> __restrict__ dre2,
    gptr<float> __restrict__ dim2) {
  const FusedArgs& a = sa.f;
  const mw_cpml& cp = a.cp;
  const ShellBox& bx = sa.box[c.b];
  const int nx = a.g.nx, ny = a.g.ny, nz = a.g.nz;
  const int j = c.j, k = c.k;
  if (j >= bx.hi[1] || k >= bx.hi[2]) return;  // idle lane (no shuffles in this kernel)
  const int lx = a.g.bc_lo[0], hxm = a.g.bc_hi[0], ly = a.g.bc_lo[1], hym = a.g.bc_hi[1], lz = a.g.bc_lo[2],
            hzm = a.g.bc_hi[2];
  const bool jin = j > 0 && j < ny - 1;
  const bool kin = k > 0 && k < nz - 1;
  // a periodic axis has both faces periodic (the host checks it); PER is false
  // when no face of the grid is periodic, which keeps the six seam loads below
  // (and their registers) out of the kernel
  const bool per_x = PER && lx == MW_BC_PERIODIC, per_y = PER && ly == MW_BC_PERIODIC,
             per_z = PER && lz == MW_BC_PERIODIC;
  const float iey = G(a.g.inv_e[1])[j];
  const float iez = G(a.g.inv_e[2])[k];
  float bey = 0, cey = 0, key = 1, bez = 0, cez = 0, kez = 1;
  if (CPML) {
    bey = G(cp.b_e[1])[j]; cey = G(cp.c_e[1])[j]; key = G(cp.inv_kappa_e[1])[j];
    bez = G(cp.b_e[2])[k]; cez = G(cp.c_e[2])[k]; kez = G(cp.inv_kappa_e[2])[k];
  }
  const bool jH = j < ny - 1, kH = k < nz - 1;
  const uint32_t sEx = (uint32_t)ny * nz, sEy = (uint32_t)(ny - 1) * nz, sEz = (uint32_t)ny * (nz - 1);
  const uint32_t sHx = (uint32_t)(ny - 1) * (nz - 1), sHy = (uint32_t)ny * (nz - 1), sHz = (uint32_t)(ny - 1) * nz;
  // psi slab positions that do not depend on the plane (-1 outside the slab)
  const int l0 = slab_local(j, cp.psi_low[0], cp.psi_high_start[0], cp.psi_high_len[0]);
  const int l1 = slab_local(k, cp.psi_low[1], cp.psi_high_start[1], cp.psi_high_len[1]);
  const int l3 = slab_local(k, cp.psi_low[3], cp.psi_high_start[3], cp.psi_high_len[3]);
  const int l5 = slab_local(j, cp.psi_low[5], cp.psi_high_start[5], cp.psi_high_len[5]);
  // carried Hz(i-1, j, k), Hy(i-1, j, k); every load below is unconditional (offset
  // 0 for a sample that does not exist, whose value is discarded)
  const bool pv = c.i0 > 0 && c.i0 - 1 < nx - 1;
  float hz_m = hz[(pv && jH) ? (c.i0 - 1) * sHz + (uint32_t)j * nz + k : 0];
  float hy_m = hy[(pv && kH) ? (c.i0 - 1) * sHy + (uint32_t)j * (nz - 1) + k : 0];
  if (!(pv && jH)) hz_m = 0.0f;
  if (!(pv && kH)) hy_m = 0.0f;
  for (int i = c.i0; i < c.i1; ++i) {
    const bool et = terms_touch(a.e_terms, a.n_e_terms, i, c.jrow0, c.jrow1, c.krow0, c.krow1);
    const bool iH = i < nx - 1;
    const bool iin = i > 0 && i < nx - 1;
    const uint32_t ox = iH ? i * sEx + (uint32_t)j * nz + k : 0;
    const uint32_t oy = jH ? i * sEy + (uint32_t)j * nz + k : 0;
    const uint32_t oz = kH ? i * sEz + (uint32_t)j * (nz - 1) + k : 0;
    const uint32_t ohz = i * sHz + (uint32_t)j * nz + k;        // valid when iH && jH
    const uint32_t ohy = i * sHy + (uint32_t)j * (nz - 1) + k;  // valid when iH && kH
    const uint32_t ohx = i * sHx + (uint32_t)j * (nz - 1) + k;  // valid when jH && kH
    // loads
    float hz_c = hz[(iH && jH) ? ohz : 0];
    float hy_c = hy[(iH && kH) ? ohy : 0];
    float hx_c = hx[(jH && kH) ? ohx : 0];
    float hz_jm = hz[(iH && j > 0) ? ohz - nz : 0];
    float hy_km = hy[(iH && k > 0) ? ohy - 1 : 0];
    float hx_km = hx[(jH && k > 0) ? ohx - 1 : 0];
    float hx_jm = hx[(kH && j > 0) ? ohx - (nz - 1) : 0];
    const float exv = exi[ox], eyv = eyi[oy], ezv = ezi[oz];
    // pole weight codes with the plane's other loads (read before any store)
    uint32_t code_x = 0u, code_y = 0u, code_z = 0u;
    if (ADE) {
      if (a.ad.count[0] > 0 && a.ad.weight[0]) code_x = G(a.ad.weight[0])[ox];
      if (a.ad.count[1] > 0 && a.ad.weight[1]) code_y = G(a.ad.weight[1])[oy];
      if (a.ad.count[2] > 0 && a.ad.weight[2]) code_z = G(a.ad.weight[2])[oz];
    }
    const float cdx = dx_ ? dx_[ox] : a.cf.decay_value[0], ccx = cx_ ? cx_[ox] : a.cf.curl_value[0];
    const float cdy = dy_ ? dy_[oy] : a.cf.decay_value[1], ccy = cy_ ? cy_[oy] : a.cf.curl_value[1];
    const float cdz = dz_ ? dz_[oz] : a.cf.decay_value[2], ccz = cz_ ? cz_[oz] : a.cf.curl_value[2];
    const float iex = G(a.g.inv_e[0])[i];
    float bex = 0, cex = 0, kex = 1;
    float p0 = 0, p1 = 0, p2 = 0, p3 = 0, p4 = 0, p5 = 0;
    uint32_t q0 = 0, q1 = 0, q2 = 0, q3 = 0, q4 = 0, q5 = 0;
    int l2 = -1, l4 = -1;
    if (CPML) {
      bex = G(cp.b_e[0])[i]; cex = G(cp.c_e[0])[i]; kex = G(cp.inv_kappa_e[0])[i];
      l2 = slab_local(i, cp.psi_low[2], cp.psi_high_start[2], cp.psi_high_len[2]);
      l4 = slab_local(i, cp.psi_low[4], cp.psi_high_start[4], cp.psi_high_len[4]);
      if (iH && l0 >= 0) q0 = slab_offset(1, i, j, k, l0, ny, nz, cp.psi_low[0] + cp.psi_high_len[0]);
      if (iH && l1 >= 0) q1 = slab_offset(2, i, j, k, l1, ny, nz, cp.psi_low[1] + cp.psi_high_len[1]);
      if (jH && l2 >= 0) q2 = slab_offset(0, i, j, k, l2, ny - 1, nz, cp.psi_low[2] + cp.psi_high_len[2]);
      if (jH && l3 >= 0) q3 = slab_offset(2, i, j, k, l3, ny - 1, nz, cp.psi_low[3] + cp.psi_high_len[3]);
      if (kH && l4 >= 0) q4 = slab_offset(0, i, j, k, l4, ny, nz - 1, cp.psi_low[4] + cp.psi_high_len[4]);
      if (kH && l5 >= 0) q5 = slab_offset(1, i, j, k, l5, ny, nz - 1, cp.psi_low[5] + cp.psi_high_len[5]);
      // an axis without PML layers has no psi buffer: the test is uniform
      if (ps0) p0 = ps0[q0];
      if (ps1) p1 = ps1[q1];
      if (ps2) p2 = ps2[q2];
      if (ps3) p3 = ps3[q3];
      if (ps4) p4 = ps4[q4];
      if (ps5) p5 = ps5[q5];
    }
    if (!(iH && jH)) hz_c = 0.0f;
    if (!(iH && kH)) hy_c = 0.0f;
    if (!(jH && kH)) hx_c = 0.0f;
    if (!(iH && j > 0)) hz_jm = 0.0f;
    if (!(iH && k > 0)) hy_km = 0.0f;
    if (!(jH && k > 0)) hx_km = 0.0f;
    if (!(kH && j > 0)) hx_jm = 0.0f;
    // samples across a periodic seam, read only on the face samples of periodic
    // axes (offset 0 elsewhere, value discarded)
    const bool wi = per_x && (i == 0 || i == nx - 1), wj = per_y && (j == 0 || j == ny - 1),
               wk = per_z && (k == 0 || k == nz - 1);
    const int iw = i == 0 ? nx - 2 : 0, jw = j == 0 ? ny - 2 : 0, kw = k == 0 ? nz - 2 : 0;
    float wz_y = 0.0f, wy_z = 0.0f, wz_x = 0.0f, wx_z = 0.0f, wy_x = 0.0f, wx_y = 0.0f;
    if (PER) {
      wz_y = hz[(wj && iH) ? i * sHz + (uint32_t)jw * nz + k : 0];                 // Ex: Hz across y
      wy_z = hy[(wk && iH) ? i * sHy + (uint32_t)j * (nz - 1) + kw : 0];           // Ex: Hy across z
      wz_x = hz[(wi && jH) ? (uint32_t)iw * sHz + (uint32_t)j * nz + k : 0];       // Ey: Hz across x
      wx_z = hx[(wk && jH) ? i * sHx + (uint32_t)j * (nz - 1) + kw : 0];           // Ey: Hx across z
      wy_x = hy[(wi && kH) ? (uint32_t)iw * sHy + (uint32_t)j * (nz - 1) + k : 0];  // Ez: Hy across x
      wx_y = hx[(wj && kH) ? i * sHx + (uint32_t)jw * (nz - 1) + k : 0];           // Ez: Hx across y
    }
    if (iH) {  // Ex (nx-1, ny, nz): a = dHz/dy, b = dHy/dz, curl = a - b
      const float jsum = ADE ? ade_advance_coded(a.ad, 0, ox, exv, code_x) : 0.0f;
      const bool pec = face_pec(j, ny, ly, hym) || face_pec(k, nz, lz, hzm);
      float e_new = exv;
      if (!pec && !(face_inactive(j, ny, ly, hym) || face_inactive(k, nz, lz, hzm))) {
        const float da = jin ? (hz_c - hz_jm) * iey : shell_edge(j == 0, j == 0 ? ly : hym, hz_c, hz_jm, wz_y, iey);
        const float db = kin ? (hy_c - hy_km) * iez : shell_edge(k == 0, k == 0 ? lz : hzm, hy_c, hy_km, wy_z, iez);
        float curl = da - db;
        if (CPML) {
          const float pa = l0 >= 0 ? bey * p0 + cey * da : 0.0f;
          const float pb = l1 >= 0 ? bez * p1 + cez * db : 0.0f;
          if (l0 >= 0) ps0[q0] = pa;
          if (l1 >= 0) ps1[q1] = pb;
          curl = (da * key + pa) - (db * kez + pb);
        }
        e_new = exv * cdx + ccx * curl;
      }
      if (ADE && jsum != 0.0f) e_new -= ade_gain(a.ad, 0, ox, cdx, ccx) * jsum;
      if (et) e_new += term_value(a.e_terms, a.n_e_terms, 0, i, j, k, step, a.dt);
      if (pec) e_new = 0.0f;
      exo[ox] = e_new;
      shell_dft<DFT>(a, dre0, dim0, (uint64_t)sEx * (nx - 1), ox, e_new, wc, ws);
    }
    if (jH) {  // Ey (nx, ny-1, nz): a = dHz/dx, b = dHx/dz, curl = b - a
      const float jsum = ADE ? ade_advance_coded(a.ad, 1, oy, eyv, code_y) : 0.0f;
      const bool pec = face_pec(i, nx, lx, hxm) || face_pec(k, nz, lz, hzm);
      float e_new = eyv;
      if (!pec && !(face_inactive(i, nx, lx, hxm) || face_inactive(k, nz, lz, hzm))) {
        const float da = iin ? (hz_c - hz_m) * iex : shell_edge(i == 0, i == 0 ? lx : hxm, hz_c, hz_m, wz_x, iex);
        const float db = kin ? (hx_c - hx_km) * iez : shell_edge(k == 0, k == 0 ? lz : hzm, hx_c, hx_km, wx_z, iez);
        float curl = db - da;
        if (CPML) {
          const float pa = l2 >= 0 ? bex * p2 + cex * da : 0.0f;
          const float pb = l3 >= 0 ? bez * p3 + cez * db : 0.0f;
          if (l2 >= 0) ps2[q2] = pa;
          if (l3 >= 0) ps3[q3] = pb;
          curl = (db * kez + pb) - (da * kex + pa);
        }
        e_new = eyv * cdy + ccy * curl;
      }
      if (ADE && jsum != 0.0f) e_new -= ade_gain(a.ad, 1, oy, cdy, ccy) * jsum;
      if (et) e_new += term_value(a.e_terms, a.n_e_terms, 1, i, j, k, step, a.dt);
      if (pec) e_new = 0.0f;
      eyo[oy] = e_new;
      shell_dft<DFT>(a, dre1, dim1, (uint64_t)sEy * nx, oy, e_new, wc, ws);
    }
    if (kH) {  // Ez (nx, ny, nz-1): a = dHy/dx, b = dHx/dy, curl = a - b
      const float jsum = ADE ? ade_advance_coded(a.ad, 2, oz, ezv, code_z) : 0.0f;
      const bool pec = face_pec(i, nx, lx, hxm) || face_pec(j, ny, ly, hym);
      float e_new = ezv;
      if (!pec && !(face_inactive(i, nx, lx, hxm) || face_inactive(j, ny, ly, hym))) {
        const float da = iin ? (hy_c - hy_m) * iex : shell_edge(i == 0, i == 0 ? lx : hxm, hy_c, hy_m, wy_x, iex);
        const float db = jin ? (hx_c - hx_jm) * iey : shell_edge(j == 0, j == 0 ? ly : hym, hx_c, hx_jm, wx_y, iey);
        float curl = da - db;
        if (CPML) {
          const float pa = l4 >= 0 ? bex * p4 + cex * da : 0.0f;
          const float pb = l5 >= 0 ? bey * p5 + cey * db : 0.0f;
          if (l4 >= 0) ps4[q4] = pa;
          if (l5 >= 0) ps5[q5] = pb;
          curl = (da * kex + pa) - (db * key + pb);
        }
        e_new = ezv * cdz + ccz * curl;
      }
      if (ADE && jsum != 0.0f) e_new -= ade_gain(a.ad, 2, oz, cdz, ccz) * jsum;
      if (et) e_new += term_value(a.e_terms, a.n_e_terms, 2, i, j, k, step, a.dt);
      if (pec) e_new = 0.0f;
      ezo[oz] = e_new;
      shell_dft<DFT>(a, dre2, dim2, (uint64_t)sEz * nx, oz, e_new, wc, ws);
    }
    hz_m = hz_c;
    hy_m = hy_c;
  }
}

// E of the shell cells of block `logical` for step `step` with that step's DFT
// weights wc / ws.
template <bool CPML, bool ADE, bool DFT, bool PER>
__device__ __forceinline__ void shell_e_tile(const ShellArgs& sa, int logical, int step, const float* wc,
                                             const float* ws) {
  ShellCell c;
  if (!shell_cell(sa, logical, c)) return;
  const FusedArgs& a = sa.f;
  shell_e_cells<CPML, ADE, DFT, PER>(sa, c, step, G(wc), G(ws), G(a.out.h[0]), G(a.out.h[1]), G(a.out.h[2]), G(a.in.e[0]),
                                G(a.in.e[1]), G(a.in.e[2]), G(a.out.e[0]), G(a.out.e[1]), G(a.out.e[2]),
                                G(a.cp.psi[0]), G(a.cp.psi[1]), G(a.cp.psi[2]), G(a.cp.psi[3]), G(a.cp.psi[4]),
                                G(a.cp.psi[5]), G(a.cf.decay[0]), G(a.cf.decay[1]), G(a.cf.decay[2]),
                                G(a.cf.curl[0]), G(a.cf.curl[1]), G(a.cf.curl[2]), G(a.dft_re[0]), G(a.dft_im[0]),
                                G(a.dft_re[1]), G(a.dft_im[1]), G(a.dft_re[2]), G(a.dft_im[2]));
}

template <bool CPML>
__global__ __launch_bounds__(256) void shell_h_kernel(ShellArgs sa) {
  shell_h_tile<CPML>(sa, xcd_logical_block(), sa.f.step);
}

template <bool CPML, bool ADE, bool DFT, bool PER>
__global__ __launch_bounds__(256) void shell_e_kernel(ShellArgs sa) {
  shell_e_tile<CPML, ADE, DFT, PER>(sa, xcd_logical_block(), sa.f.step, sa.f.dft_cos, sa.f.dft_sin);
}

// Shell H and the interior tiles in one launch (both read only the step's inputs
// and write disjoint cells): blocks [0, lean_blocks) are interior tiles, the rest
// shell-H work, each range dealt to the XCDs in contiguous runs as the separate
// launches deal them (both counts are multiples of 8).  The shell blocks fill the
// interior tiles' tail instead of a launch of their own (dispatching them first
// measured slower: C2 544.4 vs 542.3 us, design region 216.9 vs 209.6 us).
//
// The launch can end with blocks that take the previous step's observers: first
// its z-plane gather (`zg.blocks` blocks; its stage row is not the one this
// step's interior tiles write), then its point observers (`pts.count` > 0).  The
// fields that step wrote are this step's inputs, which no block of the launch
// writes.
template <bool CPML, bool ADE, bool DFT, int COEF>
__global__ __launch_bounds__(256) void shell_h_lean_kernel(ShellArgs sa, int shell_blocks, int lean_blocks,
                                                           PointSet pts, ZGatherJob zg) {
  const int w0 = blockIdx.x;
  if (w0 >= lean_blocks + shell_blocks) {
    const int tid = (int)threadIdx.x + 64 * (int)threadIdx.y;
    const int e = w0 - lean_blocks - shell_blocks;
    if (e < zg.blocks) {
      gather_zplane_sample(zg.set, zg.fs, zg.samples, (int64_t)e * 256 + tid);
      return;
    }
    const int p = (e - zg.blocks) * 256 + tid;
    if (p < pts.count) accumulate_point(pts, p);
    return;
  }
  if (w0 >= lean_blocks) {
    const int w = w0 - lean_blocks;
    shell_h_tile<CPML>(sa, (w % 8) * (shell_blocks / 8) + w / 8, sa.f.step);
    return;
  }
  const int v = w0;
  const int logical = (v % 8) * (lean_blocks / 8) + v / 8;
  const FusedArgs& a = sa.f;
  const int bz = a.lz1 - a.lz0, by = a.ly1 - a.ly0, bx = a.lx1 - a.lx0;
  if (logical >= bz * by * bx) return;
  lean_tile<ADE, DFT, COEF>(a, a.step, a.dft_cos, a.dft_sin, a.lz0 + logical % bz, a.ly0 + (logical / bz) % by,
                            a.lx0 + logical / (bz * by));
}

// The merged launch for a batch of scenes: interior tiles of every scene first
// (lean_prefix), then shell-H blocks (shell_prefix), each range XCD-mapped as the
// separate batch launches map it; the step index comes from the device counter.
template <bool CPML, bool ADE, bool DFT, int COEF>
__global__ __launch_bounds__(256) void shell_h_lean_batch_kernel(
    const FusedArgs* __restrict__ args, const ShellArgs* __restrict__ shells, const int* __restrict__ lean_prefix,
    const int* __restrict__ shell_prefix, int nscenes, int lean_total, int shell_total, int lean_blocks,
    int shell_blocks, const int* __restrict__ counter) {
  const int w0 = blockIdx.x;
  const int n = *counter;
  if (w0 >= lean_blocks) {
    const int w = w0 - lean_blocks;
    const int logical = (w % 8) * (shell_blocks / 8) + w / 8;
    if (logical >= shell_total) return;
    const int s = table_scene(shell_prefix, nscenes, logical);
    shell_h_tile<CPML>(shells[2 * s + (n & 1)], logical - shell_prefix[s], n);
    return;
  }
  const int logical = (w0 % 8) * (lean_blocks / 8) + w0 / 8;
  if (logical >= lean_total) return;
  const int s = table_scene(lean_prefix, nscenes, logical);
  const FusedArgs& a = args[2 * s + (n & 1)];
  const int t = logical - lean_prefix[s];
  const int bz = a.lz1 - a.lz0, by = a.ly1 - a.ly0;
  const int64_t wofs = (int64_t)n * a.nf_dft;
  lean_tile<ADE, DFT, COEF>(a, n, a.dft_cos + wofs, a.dft_sin + wofs, a.lz0 + t % bz, a.ly0 + (t / bz) % by,
                            a.lx0 + t / (bz * by));
}

// Batches: args[2 s + p] is scene s's shell for steps of parity p (dft_cos /
// dft_sin the base of its weight tables), prefix[s] its first block; the step
// index comes from the device counter.
template <bool CPML>
__global__ __launch_bounds__(256) void shell_h_batch_kernel(const ShellArgs* __restrict__ args,
                                                            const int* __restrict__ prefix, int nscenes, int total,
                                                            const int* __restrict__ counter) {
  const int logical = xcd_logical_block();
  if (logical >= total) return;
  const int s = table_scene(prefix, nscenes, logical);
  const int n = *counter;
  shell_h_tile<CPML>(args[2 * s + (n & 1)], logical - prefix[s], n);
}

template <bool CPML, bool ADE, bool DFT, bool PER>
__global__ __launch_bounds__(256) void shell_e_batch_kernel(const ShellArgs* __restrict__ args,
                                                            const int* __restrict__ prefix, int nscenes, int total,
                                                            const int* __restrict__ counter) {
  const int logical = xcd_logical_block();
  if (logical >= total) return;
  const int s = table_scene(prefix, nscenes, logical);
  const int n = *counter;
  const ShellArgs& sa = args[2 * s + (n & 1)];
  const int64_t w = (int64_t)n * sa.f.nf_dft;
  shell_e_tile<CPML, ADE, DFT, PER>(sa, logical - prefix[s], n, sa.f.dft_cos + w, sa.f.dft_sin + w);
}

}  // namespace mw
