// Single-pass FDTD step for gfx950, interior part: one launch advances H and then
// E by a full time step over the lean box (every cell strictly inside the grid
// and outside every PML slab); every field sample there is read from HBM once and
// written once.  The cells around the box (PML slabs, faces) are advanced by the
// per-cell shell kernels of shell.h in the same step.
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
// Periodic faces need H from the far side of the grid: they are shell cells, and
// the shell's E pass runs after every H of the step is stored (shell.h).
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
  // tiles [lz0, lz1) x [ly0, ly1) x [lx0, lx1) run the lean kernel
  int lz0, lz1, ly0, ly1, lx0, lx1;
  // 1: the z tiles outside [lz0, lz1) over the lean y-x range are z-edge tiles
  // (zedge.h); 0: those columns belong to the per-cell shell (shell.h)
  int zedge;
  // 1: the z ends [0, zend_lo) and [zend_hi, nz) over the lean y-x range are
  // z-end tiles (zedge.h: both ends of two rows in one wavefront); 0: see zedge
  int zends;
  int zend_lo, zend_hi;
  // z-normal plane samples that the interior tiles store into this step's stage
  // row (zrow == nullptr: none): zrow[zoff[q] + i * rows(zcomp[q]) + j] = component
  // zcomp[q] at (i, j, zk[q]) for the interior cells of the launch; the gather
  // takes the other samples
  float* zrow;
  int zcount;
  int zcomp[MW_MAX_PLANES];
  int zk[MW_MAX_PLANES];
  int64_t zoff[MW_MAX_PLANES];
};

__device__ __forceinline__ float lane_up(float v) { return __shfl_up(v, 1, 64); }      // lane-1's value
__device__ __forceinline__ float lane_down(float v) { return __shfl_down(v, 1, 64); }  // lane+1's value

// Sum of all terms of one family that cover sample (i,j,k) of component c; the
// time-shifted ones were evaluated for this step by shifted_values_kernel.
__device__ __forceinline__ float term_value(const mw_term* __restrict__ terms, int n, int c, int i, int j, int k,
                                            int step, double dt) {
  float acc = 0.0f;
  for (int t = 0; t < n; ++t) {
    const auto& d = G(terms)[t];
    if (d.component != c) continue;
    const int pi = i - d.offset[0], pj = j - d.offset[1], pk = k - d.offset[2];
    if (pi < 0 || pj < 0 || pk < 0 || pi >= d.shape[0] || pj >= d.shape[1] || pk >= d.shape[2]) continue;
    const int64_t p = ((int64_t)pi * d.shape[1] + pj) * d.shape[2] + pk;
    if (d.kind == MW_TERM_ADDITIVE) {
      acc += G(d.signal_table)[step] * G(d.patch)[p];
    } else if (d.kind == MW_TERM_LINE) {
      const int pa = d.line_axis == 0 ? pi : (d.line_axis == 1 ? pj : pk);
      acc += G(d.patch)[p] * G(d.line)[d.line_axis < 0 ? d.line_base : d.line_base + d.line_step * pa];
    } else {
      acc += G(d.values)[p];
    }
  }
  return acc;
}

// E^n samples of one row r at one plane that the H update of rows r-1/r needs.
struct ERow {
  float ex, ey, ez;  // at (r, k)
  float ex_k1, ey_k1;  // at (r, k+1), only meaningful in lane 63 (others shuffle)
};

// Coefficient modes of the lean interior path (chosen per launch on the host):
//   0  general: arrays or scalars, decided per access
//   1  non-conductive, non-magnetic dielectric: E decay = H decay = 1, E curl an
//      array, H curl a scalar (the common inhomogeneous-dielectric scene)
//   2  homogeneous: every decay = 1 and every curl a scalar
template <int COEF>
__device__ __forceinline__ float lean_e(const FusedArgs& a, int c, uint32_t o, float e_old, float curl,
                                        gptr<const float> __restrict__ curl_coef) {
  if (COEF == 1) return e_old + curl_coef[o] * curl;
  if (COEF == 2) return e_old + a.cf.curl_value[c] * curl;
  return e_old * coef(a.cf.decay[c], a.cf.decay_value[c], o) + coef(a.cf.curl[c], a.cf.curl_value[c], o) * curl;
}
// dt / eps_inf of the polarisation-current subtraction: modes 1 and 2 have E
// decay 1, so it is the E curl coefficient; mode 0 decides per sample (ade_gain).
template <int COEF>
__device__ __forceinline__ float lean_ade_gain(const FusedArgs& a, int c, uint32_t o,
                                               gptr<const float> __restrict__ curl_coef) {
  if (COEF == 1) return curl_coef[o];
  if (COEF == 2) return a.cf.curl_value[c];
  return ade_gain(a.ad, c, o, coef(a.cf.decay[c], a.cf.decay_value[c], o), coef(a.cf.curl[c], a.cf.curl_value[c], o));
}

template <int COEF>
__device__ __forceinline__ float lean_h(const FusedArgs& a, int c, uint32_t o, float h_old, float curl) {
  if (COEF == 1 || COEF == 2) return h_old - a.cf.curl_value[c] * curl;
  return h_old * coef(a.cf.decay[c], a.cf.decay_value[c], o) - coef(a.cf.curl[c], a.cf.curl_value[c], o) * curl;
}

// H on rows j-1, j and E on row j of plane i for a wavefront whose samples are
// all strictly inside the grid, outside every PML slab and untouched by source
// terms: plain Yee curls, no psi, no boundary modes, no terms.
// The buffers are restrict-qualified parameters of these inlined functions, so the
// compiler knows no store of a plane can change a later load: every load of a
// plane (and the next plane's E loads) issues before the plane's stores instead
// of waiting for each store in case the two alias.
template <int COEF, bool ADE, bool DFT>
__device__ __forceinline__ void lean_plane(const FusedArgs& a, int i, int j, int k, bool store, const ERow& em,
                                           const ERow& ec, const ERow& ep, const ERow& nm, const ERow& nc,
                                           float& hy_prev, float& hz_prev, bool h_terms, bool e_terms, int step,
                                           uint32_t zmask,
                                           gptr<const float> __restrict__ wc_t, gptr<const float> __restrict__ ws_t,
                                           gptr<const float> __restrict__ hxi, gptr<const float> __restrict__ hyi,
                                           gptr<const float> __restrict__ hzi, gptr<float> __restrict__ exo,
                                           gptr<float> __restrict__ eyo, gptr<float> __restrict__ ezo,
                                           gptr<float> __restrict__ hxo, gptr<float> __restrict__ hyo,
                                           gptr<float> __restrict__ hzo, gptr<const float> __restrict__ cex,
                                           gptr<const float> __restrict__ cey, gptr<const float> __restrict__ cez,
                                           gptr<float> __restrict__ dre0, gptr<float> __restrict__ dim0,
                                           gptr<float> __restrict__ dre1, gptr<float> __restrict__ dim1,
                                           gptr<float> __restrict__ dre2, gptr<float> __restrict__ dim2) {
  const int nx = a.g.nx, ny = a.g.ny, nz = a.g.nz;
  const uint32_t sHx = (uint32_t)(ny - 1) * (nz - 1), sHy = (uint32_t)ny * (nz - 1), sHz = (uint32_t)(ny - 1) * nz;
  const uint32_t sEx = (uint32_t)ny * nz, sEy = (uint32_t)(ny - 1) * nz, sEz = (uint32_t)ny * (nz - 1);
  const float ihx = G(a.g.inv_h[0])[i];
  const float ihz = G(a.g.inv_h[2])[k];
  const float ihy_m = G(a.g.inv_h[1])[j - 1], ihy_c = G(a.g.inv_h[1])[j];
  // k+1 neighbours (lane 63 gets its own extra samples)
  const bool l63 = threadIdx.x == kFLanes - 1;
  float eym_k1 = lane_down(em.ey), exm_k1 = lane_down(em.ex);
  float eyc_k1 = lane_down(ec.ey), exc_k1 = lane_down(ec.ex);
  if (l63) { eym_k1 = em.ey_k1; exm_k1 = em.ex_k1; eyc_k1 = ec.ey_k1; exc_k1 = ec.ex_k1; }
  // the pole weight codes of row j's E samples, read with the plane's other loads
  // (a code read after the H stores would wait for them: three more memory round
  // trips per plane in every wavefront, pole-free ones included)
  uint32_t code_x = 0u, code_y = 0u, code_z = 0u;
  if (ADE) {
    const uint32_t cx = i * sEx + (uint32_t)j * nz + k, cy = i * sEy + (uint32_t)j * nz + k;
    const uint32_t cz = i * sEz + (uint32_t)j * (nz - 1) + k;
    if (a.ad.count[0] > 0 && a.ad.weight[0]) code_x = G(a.ad.weight[0])[cx];
    if (a.ad.count[1] > 0 && a.ad.weight[1]) code_y = G(a.ad.weight[1])[cy];
    if (a.ad.count[2] > 0 && a.ad.weight[2]) code_z = G(a.ad.weight[2])[cz];
  }
  // row j-1 (recomputed halo) -- only Hx and Hz feed row j's E
  const uint32_t oxm = i * sHx + (uint32_t)(j - 1) * (nz - 1) + k;
  const uint32_t ozm = i * sHz + (uint32_t)(j - 1) * nz + k;
  float hx_m = lean_h<COEF>(a, 3, oxm, hxi[oxm], (ec.ez - em.ez) * ihy_m - (eym_k1 - em.ey) * ihz);
  float hz_m = lean_h<COEF>(a, 5, ozm, hzi[ozm], (nm.ey - em.ey) * ihx - (ec.ex - em.ex) * ihy_m);
  // row j
  const uint32_t oxc = i * sHx + (uint32_t)j * (nz - 1) + k;
  const uint32_t oyc = i * sHy + (uint32_t)j * (nz - 1) + k;
  const uint32_t ozc = i * sHz + (uint32_t)j * nz + k;
  float hx_c = lean_h<COEF>(a, 3, oxc, hxi[oxc], (ep.ez - ec.ez) * ihy_c - (eyc_k1 - ec.ey) * ihz);
  float hy_c = lean_h<COEF>(a, 4, oyc, hyi[oyc], (exc_k1 - ec.ex) * ihz - (nc.ez - ec.ez) * ihx);
  float hz_c = lean_h<COEF>(a, 5, ozc, hzi[ozc], (nc.ey - ec.ey) * ihx - (ep.ex - ec.ex) * ihy_c);
  if (h_terms) {  // wave-uniform: this plane/row range carries H source terms
    hx_m += term_value(a.h_terms, a.n_h_terms, 3, i, j - 1, k, step, a.dt);
    hz_m += term_value(a.h_terms, a.n_h_terms, 5, i, j - 1, k, step, a.dt);
    hx_c += term_value(a.h_terms, a.n_h_terms, 3, i, j, k, step, a.dt);
    hy_c += term_value(a.h_terms, a.n_h_terms, 4, i, j, k, step, a.dt);
    hz_c += term_value(a.h_terms, a.n_h_terms, 5, i, j, k, step, a.dt);
  }
  if (store) {
    hxo[oxc] = hx_c;
    hyo[oyc] = hy_c;
    hzo[ozc] = hz_c;
  }
  const float hy_km = lane_up(hy_c);
  const float hx_km = lane_up(hx_c);
  if (store) {
    const float iex = G(a.g.inv_e[0])[i], iey = G(a.g.inv_e[1])[j], iez = G(a.g.inv_e[2])[k];
    const uint32_t ox = i * sEx + (uint32_t)j * nz + k;
    const uint32_t oy = i * sEy + (uint32_t)j * nz + k;
    const uint32_t oz = i * sEz + (uint32_t)j * (nz - 1) + k;
    float ex = lean_e<COEF>(a, 0, ox, ec.ex, (hz_c - hz_m) * iey - (hy_c - hy_km) * iez, cex);
    float ey = lean_e<COEF>(a, 1, oy, ec.ey, (hx_c - hx_km) * iez - (hz_c - hz_prev) * iex, cey);
    float ez = lean_e<COEF>(a, 2, oz, ec.ez, (hy_c - hy_prev) * iex - (hx_c - hx_m) * iey, cez);
    if (ADE) {  // dt/eps_inf: the curl coefficient on lossless samples
      const float jx = ade_advance_coded(a.ad, 0, ox, ec.ex, code_x), jy = ade_advance_coded(a.ad, 1, oy, ec.ey, code_y);
      const float jz = ade_advance_coded(a.ad, 2, oz, ec.ez, code_z);
      if (jx != 0.0f) ex -= lean_ade_gain<COEF>(a, 0, ox, cex) * jx;
      if (jy != 0.0f) ey -= lean_ade_gain<COEF>(a, 1, oy, cey) * jy;
      if (jz != 0.0f) ez -= lean_ade_gain<COEF>(a, 2, oz, cez) * jz;
    }
    if (e_terms) {
      ex += term_value(a.e_terms, a.n_e_terms, 0, i, j, k, step, a.dt);
      ey += term_value(a.e_terms, a.n_e_terms, 1, i, j, k, step, a.dt);
      ez += term_value(a.e_terms, a.n_e_terms, 2, i, j, k, step, a.dt);
    }
    exo[ox] = ex;
    eyo[oy] = ey;
    ezo[oz] = ez;
    // staged z-normal plane samples of this tile's columns (zmask: wave-uniform)
    for (uint32_t m = zmask; m != 0u; m &= m - 1u) {
      const int q = __builtin_ctz(m);
      if (k != a.zk[q]) continue;
      const int c = a.zcomp[q];
      const float v = c == 0 ? ex : c == 1 ? ey : c == 2 ? ez : c == 3 ? hx_c : c == 4 ? hy_c : hz_c;
      const int rows = (c == 1 || c == 3 || c == 5) ? ny - 1 : ny;
      G(a.zrow)[a.zoff[q] + (int64_t)i * rows + j] = v;
    }
    if (DFT) {
      const uint64_t nxs = (uint64_t)sEx * (nx - 1), nys = (uint64_t)sEy * nx, nzs = (uint64_t)sEz * nx;
      for (int f = 0; f < a.nf_dft; ++f) {
        const float wc = wc_t[f], ws = ws_t[f];
        dre0[f * nxs + ox] += ex * wc;
        dim0[f * nxs + ox] += ex * ws;
        dre1[f * nys + oy] += ey * wc;
        dim1[f * nys + oy] += ey * ws;
        dre2[f * nzs + oz] += ez * wc;
        dim2[f * nzs + oz] += ez * ws;
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
    const auto& d = G(t)[q];
    if (i >= d.offset[0] && i < d.offset[0] + d.shape[0] && r1 >= d.offset[1] && r0 < d.offset[1] + d.shape[1] &&
        k1 >= d.offset[2] && k0 < d.offset[2] + d.shape[2])
      return true;
  }
  return false;
}

// XCD-aware tile order: the hardware hands block w to XCD w%8; (w%8, w/8) ->
// logical is a bijection on the padded launch (a multiple of 8 blocks), so each
// XCD walks a contiguous run of tiles (z fastest, then y, then x-chunk).
__device__ __forceinline__ int xcd_logical_block() {
  const int w = blockIdx.x;
  const int per_xcd = (int)(gridDim.x / 8);
  return (w % 8) * per_xcd + (w / 8);
}

// Lean interior kernel: tiles of the box [lz0,lz1) x [ly0,ly1) x [lx0,lx1), whose
// samples (plus the halo row/column and the x halo plane) are all strictly inside
// the grid and outside every PML slab, so no bounds, psi or boundary logic is
// needed and the register footprint stays small (high occupancy hides HBM
// latency).  Source terms are applied under wave-uniform guards.
template <bool ADE, bool DFT, int COEF>
__device__ __forceinline__ void lean_march(
    const FusedArgs& a, int step, int tz, int ty, int tx, gptr<const float> __restrict__ wc_t,
    gptr<const float> __restrict__ ws_t, gptr<const float> __restrict__ exi, gptr<const float> __restrict__ eyi,
    gptr<const float> __restrict__ ezi, gptr<const float> __restrict__ hxi, gptr<const float> __restrict__ hyi,
    gptr<const float> __restrict__ hzi, gptr<float> __restrict__ exo, gptr<float> __restrict__ eyo,
    gptr<float> __restrict__ ezo, gptr<float> __restrict__ hxo, gptr<float> __restrict__ hyo,
    gptr<float> __restrict__ hzo, gptr<const float> __restrict__ cex, gptr<const float> __restrict__ cey,
    gptr<const float> __restrict__ cez, gptr<float> __restrict__ dre0, gptr<float> __restrict__ dim0,
    gptr<float> __restrict__ dre1, gptr<float> __restrict__ dim1, gptr<float> __restrict__ dre2,
    gptr<float> __restrict__ dim2) {
  const int ny = a.g.ny, nz = a.g.nz;
  const int lane = threadIdx.x;
  const int wv = wave_row();
  const int kc0 = a.z_off + tz * kFOwnCols - 1;  // column of lane 0
  const int k = kc0 + lane;
  const int j = a.y_off + ty * kFRows + wv;
  const int i0 = a.x_off + tx * a.xchunk;
  const int i1 = i0 + a.xchunk;
  const bool lane63 = lane == kFLanes - 1;
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
  // staged planes whose column this tile owns (lanes 1..63)
  uint32_t zmask = 0u;
  if (a.zrow)
    for (int q = 0; q < a.zcount; ++q)
      if (a.zk[q] > kc0 && a.zk[q] < kc0 + kFLanes) zmask |= 1u << q;
  for (int i = i0 - 1; i < i1; ++i) {
    const ERow nm = load3(i + 1, -1, true), nc = load3(i + 1, 0, true), np = load3(i + 1, 1, false);
    const bool ht = terms_touch(a.h_terms, a.n_h_terms, i, j - 1, j, kc0, kc0 + kFLanes - 1);
    const bool et = terms_touch(a.e_terms, a.n_e_terms, i, j, j, kc0, kc0 + kFLanes - 1);
    lean_plane<COEF, ADE, DFT>(a, i, j, k, lane > 0 && i >= i0, em, ec, ep, nm, nc, hy_prev, hz_prev, ht, et, step,
                               zmask,
                               wc_t, ws_t, hxi, hyi, hzi, exo, eyo, ezo, hxo, hyo, hzo, cex, cey, cez, dre0, dim0,
                               dre1, dim1, dre2, dim2);
    em = nm;
    ec = nc;
    ep = np;
  }
}

template <bool ADE, bool DFT, int COEF>
__device__ __forceinline__ void lean_tile(const FusedArgs& a, int step, const float* wc_t, const float* ws_t, int tz,
                                          int ty, int tx) {
  lean_march<ADE, DFT, COEF>(a, step, tz, ty, tx, G(wc_t), G(ws_t), G(a.in.e[0]), G(a.in.e[1]), G(a.in.e[2]),
                             G(a.in.h[0]), G(a.in.h[1]), G(a.in.h[2]), G(a.out.e[0]), G(a.out.e[1]), G(a.out.e[2]),
                             G(a.out.h[0]), G(a.out.h[1]), G(a.out.h[2]), G(a.cf.curl[0]), G(a.cf.curl[1]),
                             G(a.cf.curl[2]), G(a.dft_re[0]), G(a.dft_im[0]), G(a.dft_re[1]), G(a.dft_im[1]),
                             G(a.dft_re[2]), G(a.dft_im[2]));
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
// scene s among the launch's tiles.  The step index is read from a device counter
// so a captured graph of steps replays without host arguments.  The tables are
// passed as restrict-qualified kernel arguments: nothing the step writes aliases
// them, so their fields stay in scalar registers across the x march.
__device__ __forceinline__ int table_scene(const int* __restrict__ prefix, int nscenes, int logical) {
  int lo = 0, hi = nscenes - 1;  // last s with prefix[s] <= logical
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (prefix[mid] <= logical) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <bool ADE, bool DFT, int COEF>
__global__ __launch_bounds__(256) void lean_batch_kernel(const FusedArgs* __restrict__ args,
                                                         const int* __restrict__ prefix, int nscenes, int total,
                                                         const int* __restrict__ counter) {
  const int logical = xcd_logical_block();
  if (logical >= total) return;
  const int s = table_scene(prefix, nscenes, logical);
  const int n = *counter;
  const FusedArgs& a = args[2 * s + (n & 1)];
  const int t = logical - prefix[s];
  const int bz = a.lz1 - a.lz0, by = a.ly1 - a.ly0;
  const int64_t w = (int64_t)n * a.nf_dft;
  lean_tile<ADE, DFT, COEF>(a, n, a.dft_cos + w, a.dft_sin + w, a.lz0 + t % bz, a.ly0 + (t / bz) % by,
                            a.lx0 + t / (bz * by));
}

}  // namespace mw
