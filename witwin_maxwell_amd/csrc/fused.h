// Single-pass fused FDTD step for gfx950: one launch advances H and then E by a
// full time step; every field sample is read from HBM once and written once.
//
// Tiling.  Each wavefront is independent: it owns one y-row j of 63 z-columns
// (lanes 1..63; lane 0 is a z-halo) and marches along x over a chunk of planes.
// For E at row j it needs H^{n+1/2} at rows j and j-1, so it computes H on both
// rows itself (row j-1 is recomputed, not exchanged) -- no LDS, no barriers, and
// the register footprint stays small enough for high occupancy, which is what
// hides HBM latency on CDNA.  Neighbours:
//   * x: previous/next plane carried in registers;
//   * z (k+-1): lane shuffles;
//   * y (j-1, j+1): extra row loads that the 4 wavefronts of a block share in L1.
// Because halo H is recomputed from E^n, the step cannot update in place: E, H and
// the H-side psi are ping-ponged between two buffers per step.  E-side psi and ADE
// state are only touched by their owner and stay in place.
//
// Blocks are renumbered so that each XCD (8 per MI355X, each with its own L2)
// walks a contiguous run of neighbouring tiles, so rows re-read by adjacent tiles
// hit in that XCD's L2 instead of going back to HBM.
//
// Periodic faces need H from the far side of the grid, which a single pass does
// not have; scenes with periodic axes use the two-kernel path (yee.h).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "maxwell_fdtd.h"
#include "terms.h"
#include "yee.h"

namespace mw {

constexpr int kFRows = 4;    // wavefronts per block = owned rows per block
constexpr int kFLanes = 64;  // lane 0 = halo
constexpr int kFOwnCols = kFLanes - 1;

struct FusedArgs {
  mw_grid g;
  mw_coeffs cf;
  mw_cpml cp;          // psi[6..11] = H-side psi input buffers
  float* psi_h_out[6]; // H-side psi output buffers (same slab layout)
  mw_ade ad;
  mw_fields in;
  mw_fields out;
  const mw_term* h_terms;  // device copies of the descriptors
  const mw_term* e_terms;
  int n_h_terms, n_e_terms;
  int nf_dft;
  int step;
  const float* dft_cos;  // this step's weights (nf_dft)
  const float* dft_sin;
  float* dft_re[3];
  float* dft_im[3];
  double dt;
  int xchunk;
  int tiles_z, tiles_y, tiles_x;
  // tile origins: tile (tz, ty, tx) covers columns z_off + 63 tz - 1 .. +62 (lane 0 a
  // halo), rows y_off + 4 ty .. +3 and planes x_off + xchunk tx .. +xchunk-1; the
  // offsets align the lean box below with the PML interfaces
  int z_off, y_off, x_off;
  // tiles [lz0, lz1) x [ly0, ly1) x [lx0, lx1) run the lean kernel; the general
  // kernel skips them
  int lz0, lz1, ly0, ly1, lx0, lx1;
};

__device__ __forceinline__ float lane_up(float v) { return __shfl_up(v, 1, 64); }      // lane-1's value
__device__ __forceinline__ float lane_down(float v) { return __shfl_down(v, 1, 64); }  // lane+1's value

// Sum of all terms of one family that cover sample (i,j,k) of component c; the
// time-shifted ones were evaluated for this step by shifted_values_kernel.
__device__ __forceinline__ float term_value(const mw_term* __restrict__ terms, int n, int c, int i, int j, int k,
                                            int step, double dt) {
  float acc = 0.0f;
  for (int t = 0; t < n; ++t) {
    const mw_term& d = terms[t];
    if (d.component != c) continue;
    const int pi = i - d.offset[0], pj = j - d.offset[1], pk = k - d.offset[2];
    if (pi < 0 || pj < 0 || pk < 0 || pi >= d.shape[0] || pj >= d.shape[1] || pk >= d.shape[2]) continue;
    const int64_t p = ((int64_t)pi * d.shape[1] + pj) * d.shape[2] + pk;
    if (d.kind == MW_TERM_ADDITIVE) {
      acc += d.signal_table[step] * d.patch[p];
    } else if (d.kind == MW_TERM_LINE) {
      const int pa = d.line_axis == 0 ? pi : (d.line_axis == 1 ? pj : pk);
      acc += d.patch[p] * d.line[d.line_axis < 0 ? d.line_base : d.line_base + d.line_step * pa];
    } else {
      acc += d.values[p];
    }
  }
  return acc;
}

// psi update reading `in`, writing `out` (H-side ping-pong); write only if `store`.
__device__ __forceinline__ float psi_step_pp(const float* in, float* out, int low, int hs, int hl, int axis,
                                             int coord, int i, int j, int k, int sy, int sz, float b, float c,
                                             float d, bool store) {
  const int local = slab_local(coord, low, hs, hl);
  if (local < 0) return 0.0f;
  const uint32_t o = slab_offset(axis, i, j, k, local, sy, sz, low + hl);
  const float nv = b * in[o] + c * d;
  if (store) out[o] = nv;
  return nv;
}

// E^n samples of one row r at one plane that the H update of rows r-1/r needs.
struct ERow {
  float ex, ey, ez;  // at (r, k)
  float ex_k1, ey_k1;  // at (r, k+1), only meaningful in lane 63 (others shuffle)
};

// H^{n+1/2} at (i, r, k) for all three components; `store` writes H (and psi_h)
// to the output buffers.  e_r / e_r1 are E^n at rows r and r+1 of plane i, e_n
// the row-r samples at plane i+1.
template <bool CPML>
__device__ __forceinline__ void h_update(const FusedArgs& a, int i, int r, int k, bool lane63, const ERow& e_r,
                                         const ERow& e_r1, float ey_n, float ez_n, bool store, bool terms, float& hx,
                                         float& hy, float& hz, int step) {
  const int nx = a.g.nx, ny = a.g.ny, nz = a.g.nz;
  // shuffles first, while every lane of the wavefront is active
  float ey_k1 = lane_down(e_r.ey), ex_k1 = lane_down(e_r.ex);
  if (lane63) { ey_k1 = e_r.ey_k1; ex_k1 = e_r.ex_k1; }
  hx = hy = hz = 0.0f;
  if (r < 0 || r >= ny || k < 0 || k >= nz) return;
  const bool jv = r < ny - 1, kv = k < nz - 1, iv = i < nx - 1;
  const float ihy = jv ? a.g.inv_h[1][r] : 0.0f;
  const float ihz = kv ? a.g.inv_h[2][k] : 0.0f;
  if (jv && kv) {
    const uint32_t o = ((uint32_t)i * (ny - 1) + r) * (nz - 1) + k;
    const float da = (e_r1.ez - e_r.ez) * ihy;
    const float db = (ey_k1 - e_r.ey) * ihz;
    float curl;
    if (CPML) {
      const float pa = psi_step_pp(a.cp.psi[6], a.psi_h_out[0], a.cp.psi_low[6], a.cp.psi_high_start[6],
                                   a.cp.psi_high_len[6], 1, r, i, r, k, ny - 1, nz - 1, a.cp.b_h[1][r],
                                   a.cp.c_h[1][r], da, store);
      const float pb = psi_step_pp(a.cp.psi[7], a.psi_h_out[1], a.cp.psi_low[7], a.cp.psi_high_start[7],
                                   a.cp.psi_high_len[7], 2, k, i, r, k, ny - 1, nz - 1, a.cp.b_h[2][k],
                                   a.cp.c_h[2][k], db, store);
      curl = (da * a.cp.inv_kappa_h[1][r] + pa) - (db * a.cp.inv_kappa_h[2][k] + pb);
    } else {
      curl = da - db;
    }
    hx = a.in.h[0][o] * coef(a.cf.decay[3], a.cf.decay_value[3], o) - coef(a.cf.curl[3], a.cf.curl_value[3], o) * curl;
    if (terms) hx += term_value(a.h_terms, a.n_h_terms, 3, i, r, k, step, a.dt);
    if (store) a.out.h[0][o] = hx;
  }
  if (iv) {
    const float ihx = a.g.inv_h[0][i];
    if (kv) {
      const uint32_t o = ((uint32_t)i * ny + r) * (nz - 1) + k;
      const float da = (ez_n - e_r.ez) * ihx;
      const float db = (ex_k1 - e_r.ex) * ihz;
      float curl;
      if (CPML) {
        const float pa = psi_step_pp(a.cp.psi[8], a.psi_h_out[2], a.cp.psi_low[8], a.cp.psi_high_start[8],
                                     a.cp.psi_high_len[8], 0, i, i, r, k, ny, nz - 1, a.cp.b_h[0][i], a.cp.c_h[0][i],
                                     da, store);
        const float pb = psi_step_pp(a.cp.psi[9], a.psi_h_out[3], a.cp.psi_low[9], a.cp.psi_high_start[9],
                                     a.cp.psi_high_len[9], 2, k, i, r, k, ny, nz - 1, a.cp.b_h[2][k], a.cp.c_h[2][k],
                                     db, store);
        curl = (db * a.cp.inv_kappa_h[2][k] + pb) - (da * a.cp.inv_kappa_h[0][i] + pa);
      } else {
        curl = db - da;
      }
      hy = a.in.h[1][o] * coef(a.cf.decay[4], a.cf.decay_value[4], o) - coef(a.cf.curl[4], a.cf.curl_value[4], o) * curl;
      if (terms) hy += term_value(a.h_terms, a.n_h_terms, 4, i, r, k, step, a.dt);
      if (store) a.out.h[1][o] = hy;
    }
    if (jv) {
      const uint32_t o = ((uint32_t)i * (ny - 1) + r) * nz + k;
      const float da = (ey_n - e_r.ey) * ihx;
      const float db = (e_r1.ex - e_r.ex) * ihy;
      float curl;
      if (CPML) {
        const float pa = psi_step_pp(a.cp.psi[10], a.psi_h_out[4], a.cp.psi_low[10], a.cp.psi_high_start[10],
                                     a.cp.psi_high_len[10], 0, i, i, r, k, ny - 1, nz, a.cp.b_h[0][i],
                                     a.cp.c_h[0][i], da, store);
        const float pb = psi_step_pp(a.cp.psi[11], a.psi_h_out[5], a.cp.psi_low[11], a.cp.psi_high_start[11],
                                     a.cp.psi_high_len[11], 1, r, i, r, k, ny - 1, nz, a.cp.b_h[1][r],
                                     a.cp.c_h[1][r], db, store);
        curl = (da * a.cp.inv_kappa_h[0][i] + pa) - (db * a.cp.inv_kappa_h[1][r] + pb);
      } else {
        curl = da - db;
      }
      hz = a.in.h[2][o] * coef(a.cf.decay[5], a.cf.decay_value[5], o) - coef(a.cf.curl[5], a.cf.curl_value[5], o) * curl;
      if (terms) hz += term_value(a.h_terms, a.n_h_terms, 5, i, r, k, step, a.dt);
      if (store) a.out.h[2][o] = hz;
    }
  }
}

// Coefficient modes of the lean interior path (chosen per launch on the host):
//   0  general: arrays or scalars, decided per access
//   1  non-conductive, non-magnetic dielectric: E decay = H decay = 1, E curl an
//      array, H curl a scalar (the common inhomogeneous-dielectric scene)
//   2  homogeneous: every decay = 1 and every curl a scalar
template <int COEF>
__device__ __forceinline__ float lean_e(const FusedArgs& a, int c, uint32_t o, float e_old, float curl) {
  if (COEF == 1) return e_old + a.cf.curl[c][o] * curl;
  if (COEF == 2) return e_old + a.cf.curl_value[c] * curl;
  return e_old * coef(a.cf.decay[c], a.cf.decay_value[c], o) + coef(a.cf.curl[c], a.cf.curl_value[c], o) * curl;
}
template <int COEF>
__device__ __forceinline__ float lean_h(const FusedArgs& a, int c, uint32_t o, float h_old, float curl) {
  if (COEF == 1 || COEF == 2) return h_old - a.cf.curl_value[c] * curl;
  return h_old * coef(a.cf.decay[c], a.cf.decay_value[c], o) - coef(a.cf.curl[c], a.cf.curl_value[c], o) * curl;
}

// H on rows j-1, j and E on row j of plane i for a wavefront whose samples are
// all strictly inside the grid, outside every PML slab and untouched by source
// terms: plain Yee curls, no psi, no boundary modes, no terms.
template <int COEF, bool ADE, bool DFT>
__device__ __forceinline__ void lean_plane(const FusedArgs& a, int i, int j, int k, bool store, const ERow& em,
                                           const ERow& ec, const ERow& ep, const ERow& nm, const ERow& nc,
                                           float& hy_prev, float& hz_prev, bool h_terms, bool e_terms, int step,
                                           const float* __restrict__ wc_t, const float* __restrict__ ws_t) {
  const int nx = a.g.nx, ny = a.g.ny, nz = a.g.nz;
  const uint32_t sHx = (uint32_t)(ny - 1) * (nz - 1), sHy = (uint32_t)ny * (nz - 1), sHz = (uint32_t)(ny - 1) * nz;
  const uint32_t sEx = (uint32_t)ny * nz, sEy = (uint32_t)(ny - 1) * nz, sEz = (uint32_t)ny * (nz - 1);
  const float ihx = a.g.inv_h[0][i];
  const float ihz = a.g.inv_h[2][k];
  const float ihy_m = a.g.inv_h[1][j - 1], ihy_c = a.g.inv_h[1][j];
  // k+1 neighbours (lane 63 gets its own extra samples)
  const bool l63 = threadIdx.x == kFLanes - 1;
  float eym_k1 = lane_down(em.ey), exm_k1 = lane_down(em.ex);
  float eyc_k1 = lane_down(ec.ey), exc_k1 = lane_down(ec.ex);
  if (l63) { eym_k1 = em.ey_k1; exm_k1 = em.ex_k1; eyc_k1 = ec.ey_k1; exc_k1 = ec.ex_k1; }
  // row j-1 (recomputed halo) -- only Hx and Hz feed row j's E
  const uint32_t oxm = i * sHx + (uint32_t)(j - 1) * (nz - 1) + k;
  const uint32_t ozm = i * sHz + (uint32_t)(j - 1) * nz + k;
  float hx_m = lean_h<COEF>(a, 3, oxm, a.in.h[0][oxm], (ec.ez - em.ez) * ihy_m - (eym_k1 - em.ey) * ihz);
  float hz_m = lean_h<COEF>(a, 5, ozm, a.in.h[2][ozm], (nm.ey - em.ey) * ihx - (ec.ex - em.ex) * ihy_m);
  // row j
  const uint32_t oxc = i * sHx + (uint32_t)j * (nz - 1) + k;
  const uint32_t oyc = i * sHy + (uint32_t)j * (nz - 1) + k;
  const uint32_t ozc = i * sHz + (uint32_t)j * nz + k;
  float hx_c = lean_h<COEF>(a, 3, oxc, a.in.h[0][oxc], (ep.ez - ec.ez) * ihy_c - (eyc_k1 - ec.ey) * ihz);
  float hy_c = lean_h<COEF>(a, 4, oyc, a.in.h[1][oyc], (exc_k1 - ec.ex) * ihz - (nc.ez - ec.ez) * ihx);
  float hz_c = lean_h<COEF>(a, 5, ozc, a.in.h[2][ozc], (nc.ey - ec.ey) * ihx - (ep.ex - ec.ex) * ihy_c);
  if (h_terms) {  // wave-uniform: this plane/row range carries H source terms
    hx_m += term_value(a.h_terms, a.n_h_terms, 3, i, j - 1, k, step, a.dt);
    hz_m += term_value(a.h_terms, a.n_h_terms, 5, i, j - 1, k, step, a.dt);
    hx_c += term_value(a.h_terms, a.n_h_terms, 3, i, j, k, step, a.dt);
    hy_c += term_value(a.h_terms, a.n_h_terms, 4, i, j, k, step, a.dt);
    hz_c += term_value(a.h_terms, a.n_h_terms, 5, i, j, k, step, a.dt);
  }
  if (store) {
    a.out.h[0][oxc] = hx_c;
    a.out.h[1][oyc] = hy_c;
    a.out.h[2][ozc] = hz_c;
  }
  const float hy_km = lane_up(hy_c);
  const float hx_km = lane_up(hx_c);
  if (store) {
    const float iex = a.g.inv_e[0][i], iey = a.g.inv_e[1][j], iez = a.g.inv_e[2][k];
    const uint32_t ox = i * sEx + (uint32_t)j * nz + k;
    const uint32_t oy = i * sEy + (uint32_t)j * nz + k;
    const uint32_t oz = i * sEz + (uint32_t)j * (nz - 1) + k;
    float ex = lean_e<COEF>(a, 0, ox, ec.ex, (hz_c - hz_m) * iey - (hy_c - hy_km) * iez);
    float ey = lean_e<COEF>(a, 1, oy, ec.ey, (hx_c - hx_km) * iez - (hz_c - hz_prev) * iex);
    float ez = lean_e<COEF>(a, 2, oz, ec.ez, (hy_c - hy_prev) * iex - (hx_c - hx_m) * iey);
    if (ADE) {
      ex -= a.ad.dt * ade_advance(a.ad, 0, ox, ec.ex) * (a.ad.count[0] ? a.ad.inv_eps[0][ox] : 0.0f);
      ey -= a.ad.dt * ade_advance(a.ad, 1, oy, ec.ey) * (a.ad.count[1] ? a.ad.inv_eps[1][oy] : 0.0f);
      ez -= a.ad.dt * ade_advance(a.ad, 2, oz, ec.ez) * (a.ad.count[2] ? a.ad.inv_eps[2][oz] : 0.0f);
    }
    if (e_terms) {
      ex += term_value(a.e_terms, a.n_e_terms, 0, i, j, k, step, a.dt);
      ey += term_value(a.e_terms, a.n_e_terms, 1, i, j, k, step, a.dt);
      ez += term_value(a.e_terms, a.n_e_terms, 2, i, j, k, step, a.dt);
    }
    a.out.e[0][ox] = ex;
    a.out.e[1][oy] = ey;
    a.out.e[2][oz] = ez;
    if (DFT) {
      const uint64_t nxs = (uint64_t)sEx * (nx - 1), nys = (uint64_t)sEy * nx, nzs = (uint64_t)sEz * nx;
      for (int f = 0; f < a.nf_dft; ++f) {
        const float wc = wc_t[f], ws = ws_t[f];
        a.dft_re[0][f * nxs + ox] += ex * wc;
        a.dft_im[0][f * nxs + ox] += ex * ws;
        a.dft_re[1][f * nys + oy] += ey * wc;
        a.dft_im[1][f * nys + oy] += ey * ws;
        a.dft_re[2][f * nzs + oz] += ez * wc;
        a.dft_im[2][f * nzs + oz] += ez * ws;
      }
    }
  }
  hy_prev = hy_c;
  hz_prev = hz_c;
}

// Does any term of a family touch plane i, rows [r0, r1] and columns [k0, k1]?
__device__ __forceinline__ bool terms_touch(const mw_term* __restrict__ t, int n, int i, int r0, int r1, int k0,
                                            int k1) {
  for (int q = 0; q < n; ++q) {
    const mw_term& d = t[q];
    if (i >= d.offset[0] && i < d.offset[0] + d.shape[0] && r1 >= d.offset[1] && r0 < d.offset[1] + d.shape[1] &&
        k1 >= d.offset[2] && k0 < d.offset[2] + d.shape[2])
      return true;
  }
  return false;
}

__device__ __forceinline__ float pmc_edge(bool low_end, int mode, float own, float prev, float inv) {
  // PMC mirror at a boundary sample (kernels/electric.cu:23-41); other modes are
  // inactive or PEC there and never reach this
  if (mode != MW_BC_PMC) return 0.0f;
  return low_end ? 2.0f * own * inv : -2.0f * prev * inv;
}

// XCD-aware tile order: the hardware hands block w to XCD w%8; (w%8, w/8) ->
// logical is a bijection on the padded launch (a multiple of 8 blocks), so each
// XCD walks a contiguous run of tiles (z fastest, then y, then x-chunk).
__device__ __forceinline__ int xcd_logical_block() {
  const int w = blockIdx.x;
  const int per_xcd = (int)(gridDim.x / 8);
  return (w % 8) * per_xcd + (w / 8);
}

// One tile (tz, ty, tx) of the general kernel for step `step` with that step's
// DFT weights wc_t / ws_t.
template <bool CPML, bool ADE, bool DFT>
__device__ __forceinline__ void fused_tile(const FusedArgs& a, int step, const float* __restrict__ wc_t,
                                           const float* __restrict__ ws_t, int tz, int ty, int tx) {
  const int nx = a.g.nx, ny = a.g.ny, nz = a.g.nz;
  const int lane = threadIdx.x;  // 0..63
  const int wv = wave_row();  // 0..3

  const int kc0 = a.z_off + tz * kFOwnCols - 1;  // column of lane 0
  const int k = kc0 + lane;
  const int j = a.y_off + ty * kFRows + wv;
  const int i0 = max(a.x_off + tx * a.xchunk, 0);
  const int i1 = min(a.x_off + (tx + 1) * a.xchunk, nx);
  if (i1 <= i0) return;
  const bool kin = k >= 0 && k < nz;
  const bool jin = j >= 0 && j < ny;
  const bool owned = kin && jin && lane > 0;
  const bool lane63 = lane == kFLanes - 1;

  const float* __restrict__ exi = a.in.e[0];
  const float* __restrict__ eyi = a.in.e[1];
  const float* __restrict__ ezi = a.in.e[2];
  const uint32_t sEx = (uint32_t)ny * nz, sEy = (uint32_t)(ny - 1) * nz, sEz = (uint32_t)ny * (nz - 1);

  // E^n of row r at plane i (zeros outside the grid)
  auto load_row = [&](int i, int r) {
    ERow e{0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    if (i < 0 || i >= nx || r < 0 || r >= ny || !kin) return e;
    const bool ixv = i < nx - 1, jv = r < ny - 1, kv = k < nz - 1;
    if (ixv) e.ex = exi[i * sEx + (uint32_t)r * nz + k];
    if (jv) e.ey = eyi[i * sEy + (uint32_t)r * nz + k];
    if (kv) e.ez = ezi[i * sEz + (uint32_t)r * (nz - 1) + k];
    if (lane63 && k + 1 < nz) {
      if (ixv) e.ex_k1 = exi[i * sEx + (uint32_t)r * nz + k + 1];
      if (jv) e.ey_k1 = eyi[i * sEy + (uint32_t)r * nz + k + 1];
    }
    return e;
  };

  const int istart = i0 > 0 ? i0 - 1 : 0;  // one halo plane of H before the chunk
  // rows j-1 (m), j (c), j+1 (p) of E^n at the current plane
  ERow em = load_row(istart, j - 1), ec = load_row(istart, j), ep = load_row(istart, j + 1);
  float hy_prev = 0.0f, hz_prev = 0.0f;  // H^{n+1/2} at (i-1, j)

  const int ly = a.g.bc_lo[1], hym = a.g.bc_hi[1], lz = a.g.bc_lo[2], hzm = a.g.bc_hi[2];
  const int lx = a.g.bc_lo[0], hxm = a.g.bc_hi[0];

  for (int i = istart; i < i1; ++i) {
    // next plane of the three rows (only Ey/Ez of rows j-1, j are used from it)
    const ERow nm = load_row(i + 1, j - 1), nc = load_row(i + 1, j), np = load_row(i + 1, j + 1);
    // H on rows j-1 (halo, never stored) and j (stored when owned, inside the chunk)
    // wave-uniform: does this plane/row range carry source terms at all?
    const bool ht = terms_touch(a.h_terms, a.n_h_terms, i, j - 1, j, kc0, kc0 + kFLanes - 1);
    const bool et = terms_touch(a.e_terms, a.n_e_terms, i, j, j, kc0, kc0 + kFLanes - 1);
    float hx_m, hy_m, hz_m, hx_c, hy_c, hz_c;
    h_update<CPML>(a, i, j - 1, k, lane63, em, ec, nm.ey, nm.ez, false, ht, hx_m, hy_m, hz_m, step);
    h_update<CPML>(a, i, j, k, lane63, ec, ep, nc.ey, nc.ez, owned && i >= i0, ht, hx_c, hy_c, hz_c, step);
    (void)hy_m;
    const float hy_km = lane_up(hy_c);
    const float hx_km = lane_up(hx_c);

    if (owned && i >= i0) {
      const float iey = a.g.inv_e[1][j];
      const float iez = a.g.inv_e[2][k];
      const float iex = a.g.inv_e[0][i];
      // Ex (nx-1, ny, nz): a = dHz/dy, b = dHy/dz
      if (i < nx - 1) {
        const uint32_t o = i * sEx + (uint32_t)j * nz + k;
        const float e_old = ec.ex;
        const float jsum = ADE ? ade_advance(a.ad, 0, o, e_old) : 0.0f;
        const bool pec = face_pec(j, ny, ly, hym) || face_pec(k, nz, lz, hzm);
        float e_new = e_old;
        if (!pec && !(face_inactive(j, ny, ly, hym) || face_inactive(k, nz, lz, hzm))) {
          const float da = (j > 0 && j < ny - 1) ? (hz_c - hz_m) * iey : pmc_edge(j == 0, j == 0 ? ly : hym, hz_c, hz_m, iey);
          const float db = (k > 0 && k < nz - 1) ? (hy_c - hy_km) * iez : pmc_edge(k == 0, k == 0 ? lz : hzm, hy_c, hy_km, iez);
          float curl;
          if (CPML) {
            const float pa = psi_step(psi_view(a.cp, 0), 1, j, i, j, k, ny, nz, a.cp.b_e[1][j], a.cp.c_e[1][j], da);
            const float pb = psi_step(psi_view(a.cp, 1), 2, k, i, j, k, ny, nz, a.cp.b_e[2][k], a.cp.c_e[2][k], db);
            curl = (da * a.cp.inv_kappa_e[1][j] + pa) - (db * a.cp.inv_kappa_e[2][k] + pb);
          } else {
            curl = da - db;
          }
          e_new = e_old * coef(a.cf.decay[0], a.cf.decay_value[0], o) + coef(a.cf.curl[0], a.cf.curl_value[0], o) * curl;
        }
        if (ADE) e_new -= a.ad.dt * jsum * a.ad.inv_eps[0][o];
        if (et) e_new += term_value(a.e_terms, a.n_e_terms, 0, i, j, k, step, a.dt);
        if (pec) e_new = 0.0f;
        a.out.e[0][o] = e_new;
        if (DFT) {
          const uint64_t n = (uint64_t)sEx * (nx - 1);
          for (int f = 0; f < a.nf_dft; ++f) {
            a.dft_re[0][f * n + o] += e_new * wc_t[f];
            a.dft_im[0][f * n + o] += e_new * ws_t[f];
          }
        }
      }
      // Ey (nx, ny-1, nz): a = dHz/dx, b = dHx/dz, curl = b - a
      if (j < ny - 1) {
        const uint32_t o = i * sEy + (uint32_t)j * nz + k;
        const float e_old = ec.ey;
        const float jsum = ADE ? ade_advance(a.ad, 1, o, e_old) : 0.0f;
        const bool pec = face_pec(i, nx, lx, hxm) || face_pec(k, nz, lz, hzm);
        float e_new = e_old;
        if (!pec && !(face_inactive(i, nx, lx, hxm) || face_inactive(k, nz, lz, hzm))) {
          const float da = (i > 0 && i < nx - 1) ? (hz_c - hz_prev) * iex : pmc_edge(i == 0, i == 0 ? lx : hxm, hz_c, hz_prev, iex);
          const float db = (k > 0 && k < nz - 1) ? (hx_c - hx_km) * iez : pmc_edge(k == 0, k == 0 ? lz : hzm, hx_c, hx_km, iez);
          float curl;
          if (CPML) {
            const float pa = psi_step(psi_view(a.cp, 2), 0, i, i, j, k, ny - 1, nz, a.cp.b_e[0][i], a.cp.c_e[0][i], da);
            const float pb = psi_step(psi_view(a.cp, 3), 2, k, i, j, k, ny - 1, nz, a.cp.b_e[2][k], a.cp.c_e[2][k], db);
            curl = (db * a.cp.inv_kappa_e[2][k] + pb) - (da * a.cp.inv_kappa_e[0][i] + pa);
          } else {
            curl = db - da;
          }
          e_new = e_old * coef(a.cf.decay[1], a.cf.decay_value[1], o) + coef(a.cf.curl[1], a.cf.curl_value[1], o) * curl;
        }
        if (ADE) e_new -= a.ad.dt * jsum * a.ad.inv_eps[1][o];
        if (et) e_new += term_value(a.e_terms, a.n_e_terms, 1, i, j, k, step, a.dt);
        if (pec) e_new = 0.0f;
        a.out.e[1][o] = e_new;
        if (DFT) {
          const uint64_t n = (uint64_t)sEy * nx;
          for (int f = 0; f < a.nf_dft; ++f) {
            a.dft_re[1][f * n + o] += e_new * wc_t[f];
            a.dft_im[1][f * n + o] += e_new * ws_t[f];
          }
        }
      }
      // Ez (nx, ny, nz-1): a = dHy/dx, b = dHx/dy
      if (k < nz - 1) {
        const uint32_t o = i * sEz + (uint32_t)j * (nz - 1) + k;
        const float e_old = ec.ez;
        const float jsum = ADE ? ade_advance(a.ad, 2, o, e_old) : 0.0f;
        const bool pec = face_pec(i, nx, lx, hxm) || face_pec(j, ny, ly, hym);
        float e_new = e_old;
        if (!pec && !(face_inactive(i, nx, lx, hxm) || face_inactive(j, ny, ly, hym))) {
          const float da = (i > 0 && i < nx - 1) ? (hy_c - hy_prev) * iex : pmc_edge(i == 0, i == 0 ? lx : hxm, hy_c, hy_prev, iex);
          const float db = (j > 0 && j < ny - 1) ? (hx_c - hx_m) * iey : pmc_edge(j == 0, j == 0 ? ly : hym, hx_c, hx_m, iey);
          float curl;
          if (CPML) {
            const float pa = psi_step(psi_view(a.cp, 4), 0, i, i, j, k, ny, nz - 1, a.cp.b_e[0][i], a.cp.c_e[0][i], da);
            const float pb = psi_step(psi_view(a.cp, 5), 1, j, i, j, k, ny, nz - 1, a.cp.b_e[1][j], a.cp.c_e[1][j], db);
            curl = (da * a.cp.inv_kappa_e[0][i] + pa) - (db * a.cp.inv_kappa_e[1][j] + pb);
          } else {
            curl = da - db;
          }
          e_new = e_old * coef(a.cf.decay[2], a.cf.decay_value[2], o) + coef(a.cf.curl[2], a.cf.curl_value[2], o) * curl;
        }
        if (ADE) e_new -= a.ad.dt * jsum * a.ad.inv_eps[2][o];
        if (et) e_new += term_value(a.e_terms, a.n_e_terms, 2, i, j, k, step, a.dt);
        if (pec) e_new = 0.0f;
        a.out.e[2][o] = e_new;
        if (DFT) {
          const uint64_t n = (uint64_t)sEz * nx;
          for (int f = 0; f < a.nf_dft; ++f) {
            a.dft_re[2][f * n + o] += e_new * wc_t[f];
            a.dft_im[2][f * n + o] += e_new * ws_t[f];
          }
        }
      }
    }
    hy_prev = hy_c;
    hz_prev = hz_c;
    em = nm;
    ec = nc;
    ep = np;
  }
}

// Lean interior kernel: tiles of the box [lz0,lz1) x [ly0,ly1) x [lx0,lx1), whose
// samples (plus the halo row/column and the x halo plane) are all strictly inside
// the grid and outside every PML slab, so no bounds, psi or boundary logic is
// needed and the register footprint stays small (high occupancy hides HBM
// latency).  Source terms are applied under wave-uniform guards.
template <bool ADE, bool DFT, int COEF>
__device__ __forceinline__ void lean_tile(const FusedArgs& a, int step, const float* __restrict__ wc_t,
                                          const float* __restrict__ ws_t, int tz, int ty, int tx) {
  const int ny = a.g.ny, nz = a.g.nz;
  const int lane = threadIdx.x;
  const int wv = wave_row();
  const int kc0 = a.z_off + tz * kFOwnCols - 1;  // column of lane 0
  const int k = kc0 + lane;
  const int j = a.y_off + ty * kFRows + wv;
  const int i0 = a.x_off + tx * a.xchunk;
  const int i1 = i0 + a.xchunk;
  const bool lane63 = lane == kFLanes - 1;
  const float* __restrict__ exi = a.in.e[0];
  const float* __restrict__ eyi = a.in.e[1];
  const float* __restrict__ ezi = a.in.e[2];
  const uint32_t sEx = (uint32_t)ny * nz, sEy = (uint32_t)(ny - 1) * nz, sEz = (uint32_t)ny * (nz - 1);
  const uint32_t rEx = (uint32_t)j * nz + k, rEy = (uint32_t)j * nz + k, rEz = (uint32_t)j * (nz - 1) + k;

  auto load3 = [&](int i, int dr, bool k1) {
    ERow e;
    const uint32_t pe = i * sEx + rEx + dr * nz, py = i * sEy + rEy + dr * nz, pz = i * sEz + rEz + dr * (nz - 1);
    e.ex = exi[pe];
    e.ey = eyi[py];
    e.ez = ezi[pz];
    e.ex_k1 = 0.0f;
    e.ey_k1 = 0.0f;
    if (k1 && lane63) {
      e.ex_k1 = exi[pe + 1];
      e.ey_k1 = eyi[py + 1];
    }
    return e;
  };
  ERow em = load3(i0 - 1, -1, true), ec = load3(i0 - 1, 0, true), ep = load3(i0 - 1, 1, false);
  float hy_prev = 0.0f, hz_prev = 0.0f;
  for (int i = i0 - 1; i < i1; ++i) {
    const ERow nm = load3(i + 1, -1, true), nc = load3(i + 1, 0, true), np = load3(i + 1, 1, false);
    const bool ht = terms_touch(a.h_terms, a.n_h_terms, i, j - 1, j, kc0, kc0 + kFLanes - 1);
    const bool et = terms_touch(a.e_terms, a.n_e_terms, i, j, j, kc0, kc0 + kFLanes - 1);
    lean_plane<COEF, ADE, DFT>(a, i, j, k, lane > 0 && i >= i0, em, ec, ep, nm, nc, hy_prev, hz_prev, ht, et, step,
                               wc_t, ws_t);
    em = nm;
    ec = nc;
    ep = np;
  }
}

template <bool CPML, bool ADE, bool DFT>
__global__ __launch_bounds__(256) void fused_step_kernel(FusedArgs a) {
  const int logical = xcd_logical_block();
  if (logical >= a.tiles_z * a.tiles_y * a.tiles_x) return;
  const int tz = logical % a.tiles_z;
  const int ty = (logical / a.tiles_z) % a.tiles_y;
  const int tx = logical / (a.tiles_z * a.tiles_y);
  if (tz >= a.lz0 && tz < a.lz1 && ty >= a.ly0 && ty < a.ly1 && tx >= a.lx0 && tx < a.lx1) return;
  fused_tile<CPML, ADE, DFT>(a, a.step, a.dft_cos, a.dft_sin, tz, ty, tx);
}

template <bool ADE, bool DFT, int COEF>
__global__ __launch_bounds__(256) void lean_step_kernel(FusedArgs a) {
  const int bz = a.lz1 - a.lz0, by = a.ly1 - a.ly0, bx = a.lx1 - a.lx0;
  const int logical = xcd_logical_block();
  if (logical >= bz * by * bx) return;
  lean_tile<ADE, DFT, COEF>(a, a.step, a.dft_cos, a.dft_sin, a.lz0 + logical % bz, a.ly0 + (logical / bz) % by,
                            a.lx0 + logical / (bz * by));
}

// ---------------------------------------------------------------- batches
// Many independent scenes advance in one launch: args[2 s + p] holds scene s's
// arguments for steps of parity p (buffers swapped), with dft_cos/dft_sin the
// base of its (steps, nf) weight tables; prefix[s] is the first logical tile of
// scene s (general tiles in the general launch, lean-box tiles in the lean one).
// The step index is read from a device counter so a captured graph of steps
// replays without host arguments.
struct BatchView {
  const FusedArgs* args;
  const int* prefix;  // nscenes + 1
  int nscenes;
  int total;
  const int* counter;
};

__device__ __forceinline__ int batch_scene(const BatchView& b, int logical) {
  int lo = 0, hi = b.nscenes - 1;  // last s with prefix[s] <= logical
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (b.prefix[mid] <= logical) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <bool CPML, bool ADE, bool DFT>
__global__ __launch_bounds__(256) void fused_batch_kernel(BatchView b) {
  const int logical = xcd_logical_block();
  if (logical >= b.total) return;
  const int s = batch_scene(b, logical);
  const int n = *b.counter;
  const FusedArgs& a = b.args[2 * s + (n & 1)];
  const int t = logical - b.prefix[s];
  const int tz = t % a.tiles_z;
  const int ty = (t / a.tiles_z) % a.tiles_y;
  const int tx = t / (a.tiles_z * a.tiles_y);
  if (tz >= a.lz0 && tz < a.lz1 && ty >= a.ly0 && ty < a.ly1 && tx >= a.lx0 && tx < a.lx1) return;
  const int64_t w = (int64_t)n * a.nf_dft;
  fused_tile<CPML, ADE, DFT>(a, n, a.dft_cos + w, a.dft_sin + w, tz, ty, tx);
}

template <bool ADE, bool DFT, int COEF>
__global__ __launch_bounds__(256) void lean_batch_kernel(BatchView b) {
  const int logical = xcd_logical_block();
  if (logical >= b.total) return;
  const int s = batch_scene(b, logical);
  const int n = *b.counter;
  const FusedArgs& a = b.args[2 * s + (n & 1)];
  const int t = logical - b.prefix[s];
  const int bz = a.lz1 - a.lz0, by = a.ly1 - a.ly0;
  const int64_t w = (int64_t)n * a.nf_dft;
  lean_tile<ADE, DFT, COEF>(a, n, a.dft_cos + w, a.dft_sin + w, a.lz0 + t % bz, a.ly0 + (t / bz) % by,
                            a.lx0 + t / (bz * by));
}

}  // namespace mw
