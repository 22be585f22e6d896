// Interior tiles at the two ends of the z axis.
//
// The lean box (fused.h) holds whole 63-column tiles strictly inside every PML
// slab, so along z it stops short of the z-low / z-high CPML slabs and faces.
// Those columns share cache lines with the first and last lean columns; updating
// them in a separate per-cell pass would fetch every such line twice.  The z-edge
// tiles are the partial tiles at both ends of each lean row: the same wavefront
// layout and x march as the lean kernel (so the shared lines are read once, by
// neighbouring wavefronts on one XCD), plus what the lean kernel leaves out along
// z only: lanes outside the grid, Hx/Hy/Ez absent at k = nz-1, the z-derivative
// CPML terms (psi_hx_z, psi_hy_z ping-ponged; psi_ex_z, psi_ey_z in place) and the
// Ex/Ey boundary modes of the z faces (never periodic: the host keeps periodic z
// faces in the shell).  Along x and y these tiles are inside the
// lean box's ranges, so no x/y psi or x/y faces occur.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fused.h"
#include "maxwell_fdtd.h"
#include "shell.h"
#include "yee.h"

namespace mw {

// H coefficient of the z-edge tiles in the lean kernel's coefficient modes
// (fused.h): modes 1 and 2 need no per-sample load at all.
template <int COEF>
__device__ __forceinline__ float zedge_h(const FusedArgs& a, int c, uint32_t o, float h_old, float curl) {
  if (COEF == 1 || COEF == 2) return h_old - a.cf.curl_value[c] * curl;
  return h_old * coef(a.cf.decay[c], a.cf.decay_value[c], o) - coef(a.cf.curl[c], a.cf.curl_value[c], o) * curl;
}

// Every load of a plane is issued together at the top of the plane, at offset 0
// for a sample that does not exist (the value is then unused) and through a
// stand-in buffer for a psi array the scene does not have, so the plane costs one
// memory round trip: a load behind a run-time branch (a null test, a lane
// condition) cannot be hoisted above the previous component's stores.
template <bool ADE, bool DFT, int COEF>
__device__ __forceinline__ void zedge_march(
    const FusedArgs& a, int step, int tz, int ty, int tx, gptr<const float> __restrict__ wc_t,
    gptr<const float> __restrict__ ws_t, gptr<const float> __restrict__ exi, gptr<const float> __restrict__ eyi,
    gptr<const float> __restrict__ ezi, gptr<const float> __restrict__ hxi, gptr<const float> __restrict__ hyi,
    gptr<const float> __restrict__ hzi, gptr<float> __restrict__ exo, gptr<float> __restrict__ eyo,
    gptr<float> __restrict__ ezo, gptr<float> __restrict__ hxo, gptr<float> __restrict__ hyo,
    gptr<float> __restrict__ hzo, gptr<const float> __restrict__ p7i, gptr<const float> __restrict__ p9i,
    gptr<float> __restrict__ p7o, gptr<float> __restrict__ p9o, gptr<float> __restrict__ p1,
    gptr<float> __restrict__ p3, gptr<const float> __restrict__ kcx, gptr<const float> __restrict__ kcy,
    gptr<const float> __restrict__ kcz, gptr<float> __restrict__ dre0, gptr<float> __restrict__ dim0,
    gptr<float> __restrict__ dre1, gptr<float> __restrict__ dim1, gptr<float> __restrict__ dre2,
    gptr<float> __restrict__ dim2) {
  const int nx = a.g.nx, ny = a.g.ny, nz = a.g.nz;
  const mw_cpml& cp = a.cp;
  const int lane = threadIdx.x;
  const int wv = wave_row();
  const int kc0 = a.z_off + tz * kFOwnCols - 1;  // column of lane 0
  const int k = kc0 + lane;
  const int j = a.y_off + ty * kFRows + wv;
  const int i0 = a.x_off + tx * a.xchunk;
  const int i1 = i0 + a.xchunk;
  const bool lane63 = lane == kFLanes - 1;
  const bool kin = k >= 0 && k < nz;      // Ex, Ey, Hz columns
  const bool kv = k >= 0 && k < nz - 1;   // Ez, Hx, Hy columns
  const bool kint = k > 0 && k < nz - 1;  // z-interior E sample
  const int lz = a.g.bc_lo[2], hzm = a.g.bc_hi[2];
  const bool cpml = cp.enabled != 0;
  const uint32_t sEx = (uint32_t)ny * nz, sEy = (uint32_t)(ny - 1) * nz, sEz = (uint32_t)ny * (nz - 1);
  const uint32_t sHx = (uint32_t)(ny - 1) * (nz - 1), sHy = (uint32_t)ny * (nz - 1), sHz = (uint32_t)(ny - 1) * nz;
  const uint32_t rEx = (uint32_t)j * nz + k, rEz = (uint32_t)j * (nz - 1) + k;
  // per-lane z quantities (offset 0 for a column outside the grid; the value is unused)
  const float ihz = G(a.g.inv_h[2])[kv ? k : 0];
  const float iez = G(a.g.inv_e[2])[kin ? k : 0];
  float bhz = 0.0f, chz = 0.0f, khz = 1.0f, bez = 0.0f, cez = 0.0f, kez = 1.0f;
  int l7 = -1, l9 = -1, l1 = -1, l3 = -1;
  if (cpml) {
    bhz = G(cp.b_h[2])[kv ? k : 0]; chz = G(cp.c_h[2])[kv ? k : 0]; khz = G(cp.inv_kappa_h[2])[kv ? k : 0];
    bez = G(cp.b_e[2])[kin ? k : 0]; cez = G(cp.c_e[2])[kin ? k : 0]; kez = G(cp.inv_kappa_e[2])[kin ? k : 0];
    if (kv) {
      l7 = slab_local(k, cp.psi_low[7], cp.psi_high_start[7], cp.psi_high_len[7]);
      l9 = slab_local(k, cp.psi_low[9], cp.psi_high_start[9], cp.psi_high_len[9]);
    }
    if (kin) {
      l1 = slab_local(k, cp.psi_low[1], cp.psi_high_start[1], cp.psi_high_len[1]);
      l3 = slab_local(k, cp.psi_low[3], cp.psi_high_start[3], cp.psi_high_len[3]);
    }
  }
  // psi arrays the scene does not have read the E^n field instead (offset 0, never
  // used: their slab positions are all -1); the stores test the real pointer
  gptr<const float> p7r = p7i ? p7i : exi;
  gptr<const float> p9r = p9i ? p9i : exi;
  gptr<const float> p1r = p1 ? (gptr<const float>)p1 : exi;
  gptr<const float> p3r = p3 ? (gptr<const float>)p3 : exi;
  const int len7 = cp.psi_low[7] + cp.psi_high_len[7], len9 = cp.psi_low[9] + cp.psi_high_len[9];
  const int len1 = cp.psi_low[1] + cp.psi_high_len[1], len3 = cp.psi_low[3] + cp.psi_high_len[3];

  auto load3 = [&](int i, int dr, bool k1) {
    ERow e;
    const uint32_t pe = i * sEx + rEx + dr * nz, py = i * sEy + rEx + dr * nz, pz = i * sEz + rEz + dr * (nz - 1);
    e.ex = exi[kin ? pe : 0];
    e.ey = eyi[kin ? py : 0];
    e.ez = ezi[kv ? pz : 0];
    if (!kin) { e.ex = 0.0f; e.ey = 0.0f; }
    if (!kv) e.ez = 0.0f;
    e.ex_k1 = 0.0f;
    e.ey_k1 = 0.0f;
    if (k1 && lane63 && k + 1 >= 0 && k + 1 < nz) {
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
    const bool store = lane > 0 && i >= i0;
    const float ihx = G(a.g.inv_h[0])[i];
    const float ihy_m = G(a.g.inv_h[1])[j - 1], ihy_c = G(a.g.inv_h[1])[j];
    const float iex = G(a.g.inv_e[0])[i], iey = G(a.g.inv_e[1])[j];
    float eym_k1 = lane_down(em.ey);
    float eyc_k1 = lane_down(ec.ey), exc_k1 = lane_down(ec.ex);
    if (lane63) { eym_k1 = em.ey_k1; eyc_k1 = ec.ey_k1; exc_k1 = ec.ex_k1; }
    // every load of this plane
    const uint32_t oxm = kv ? i * sHx + (uint32_t)(j - 1) * (nz - 1) + k : 0;
    const uint32_t ozm = kin ? i * sHz + (uint32_t)(j - 1) * nz + k : 0;
    const uint32_t oxc = kv ? i * sHx + (uint32_t)j * (nz - 1) + k : 0;
    const uint32_t oyc = kv ? i * sHy + (uint32_t)j * (nz - 1) + k : 0;
    const uint32_t ozc = kin ? i * sHz + (uint32_t)j * nz + k : 0;
    const float hxm_old = hxi[oxm], hzm_old = hzi[ozm], hxc_old = hxi[oxc], hyc_old = hyi[oyc], hzc_old = hzi[ozc];
    const uint32_t q7m = l7 >= 0 ? slab_offset(2, i, j - 1, k, l7, ny - 1, nz - 1, len7) : 0;
    const uint32_t q7c = l7 >= 0 ? slab_offset(2, i, j, k, l7, ny - 1, nz - 1, len7) : 0;
    const uint32_t q9c = l9 >= 0 ? slab_offset(2, i, j, k, l9, ny, nz - 1, len9) : 0;
    const float p7m = p7r[q7m], p7c = p7r[q7c], p9c = p9r[q9c];
    const uint32_t ox = kin ? i * sEx + (uint32_t)j * nz + k : 0;
    const uint32_t oy = kin ? i * sEy + (uint32_t)j * nz + k : 0;
    const uint32_t oz = kv ? i * sEz + (uint32_t)j * (nz - 1) + k : 0;
    const uint32_t q1 = l1 >= 0 ? slab_offset(2, i, j, k, l1, ny, nz, len1) : 0;
    const uint32_t q3 = l3 >= 0 ? slab_offset(2, i, j, k, l3, ny - 1, nz, len3) : 0;
    const float p1v = p1r[q1], p3v = p3r[q3];
    // pole weight codes with the plane's other loads (offset 0 where the sample
    // does not exist; read before any store of the plane)
    uint32_t code_x = 0u, code_y = 0u, code_z = 0u;
    if (ADE) {
      if (a.ad.count[0] > 0 && a.ad.weight[0]) code_x = G(a.ad.weight[0])[ox];
      if (a.ad.count[1] > 0 && a.ad.weight[1]) code_y = G(a.ad.weight[1])[oy];
      if (a.ad.count[2] > 0 && a.ad.weight[2]) code_z = G(a.ad.weight[2])[oz];
    }
    float ccx = 0.0f, ccy = 0.0f, ccz = 0.0f;  // E curl coefficients of mode 1
    if (COEF == 1) { ccx = kcx[ox]; ccy = kcy[oy]; ccz = kcz[oz]; }
    // H, row j-1 (Hx, Hz: what row j's E needs)
    float hx_m, hz_m, hx_c, hy_c, hz_c;
    {
      const float da = (ec.ez - em.ez) * ihy_m, db = (eym_k1 - em.ey) * ihz;
      const float pz = l7 >= 0 ? bhz * p7m + chz * db : 0.0f;
      hx_m = kv ? zedge_h<COEF>(a, 3, oxm, hxm_old, da - (db * khz + pz)) : 0.0f;
    }
    hz_m = kin ? zedge_h<COEF>(a, 5, ozm, hzm_old, (nm.ey - em.ey) * ihx - (ec.ex - em.ex) * ihy_m) : 0.0f;
    // H, row j
    float p7n, p9n;
    {
      const float da = (ep.ez - ec.ez) * ihy_c, db = (eyc_k1 - ec.ey) * ihz;
      p7n = l7 >= 0 ? bhz * p7c + chz * db : 0.0f;
      hx_c = kv ? zedge_h<COEF>(a, 3, oxc, hxc_old, da - (db * khz + p7n)) : 0.0f;
    }
    {
      const float da = (nc.ez - ec.ez) * ihx, db = (exc_k1 - ec.ex) * ihz;
      p9n = l9 >= 0 ? bhz * p9c + chz * db : 0.0f;
      hy_c = kv ? zedge_h<COEF>(a, 4, oyc, hyc_old, (db * khz + p9n) - da) : 0.0f;
    }
    hz_c = kin ? zedge_h<COEF>(a, 5, ozc, hzc_old, (nc.ey - ec.ey) * ihx - (ep.ex - ec.ex) * ihy_c) : 0.0f;
    if (ht) {
      if (kv) {
        hx_m += term_value(a.h_terms, a.n_h_terms, 3, i, j - 1, k, step, a.dt);
        hx_c += term_value(a.h_terms, a.n_h_terms, 3, i, j, k, step, a.dt);
        hy_c += term_value(a.h_terms, a.n_h_terms, 4, i, j, k, step, a.dt);
      }
      if (kin) {
        hz_m += term_value(a.h_terms, a.n_h_terms, 5, i, j - 1, k, step, a.dt);
        hz_c += term_value(a.h_terms, a.n_h_terms, 5, i, j, k, step, a.dt);
      }
    }
    if (store) {
      if (kv) {
        hxo[oxc] = hx_c;
        hyo[oyc] = hy_c;
        if (l7 >= 0 && p7o) p7o[q7c] = p7n;
        if (l9 >= 0 && p9o) p9o[q9c] = p9n;
      }
      if (kin) hzo[ozc] = hz_c;
    }
    const float hy_km = lane_up(hy_c);
    const float hx_km = lane_up(hx_c);
    // E of row j: the face modes of the z faces, z CPML on Ex / Ey
    const bool zpec = face_pec(k, nz, lz, hzm);
    const bool zupd = !zpec && !face_inactive(k, nz, lz, hzm);
    float ex, ey, ez;
    {  // Ex: a = dHz/dy, b = dHy/dz, curl = a - b
      const float da = (hz_c - hz_m) * iey;
      const float db = kint ? (hy_c - hy_km) * iez : shell_edge(k == 0, k == 0 ? lz : hzm, hy_c, hy_km, 0.0f, iez);
      const float pb = l1 >= 0 ? bez * p1v + cez * db : 0.0f;
      if (store && zupd && l1 >= 0 && p1) p1[q1] = pb;
      const float curl = da - (db * kez + pb);
      if (COEF == 1) ex = ec.ex + ccx * curl;
      else if (COEF == 2) ex = ec.ex + a.cf.curl_value[0] * curl;
      else ex = ec.ex * coef(a.cf.decay[0], a.cf.decay_value[0], ox) + coef(a.cf.curl[0], a.cf.curl_value[0], ox) * curl;
      if (!zupd) ex = ec.ex;
    }
    {  // Ey: a = dHz/dx, b = dHx/dz, curl = b - a
      const float da = (hz_c - hz_prev) * iex;
      const float db = kint ? (hx_c - hx_km) * iez : shell_edge(k == 0, k == 0 ? lz : hzm, hx_c, hx_km, 0.0f, iez);
      const float pb = l3 >= 0 ? bez * p3v + cez * db : 0.0f;
      if (store && zupd && l3 >= 0 && p3) p3[q3] = pb;
      const float curl = (db * kez + pb) - da;
      if (COEF == 1) ey = ec.ey + ccy * curl;
      else if (COEF == 2) ey = ec.ey + a.cf.curl_value[1] * curl;
      else ey = ec.ey * coef(a.cf.decay[1], a.cf.decay_value[1], oy) + coef(a.cf.curl[1], a.cf.curl_value[1], oy) * curl;
      if (!zupd) ey = ec.ey;
    }
    {  // Ez: a = dHy/dx, b = dHx/dy, curl = a - b (no z derivative)
      const float curl = (hy_c - hy_prev) * iex - (hx_c - hx_m) * iey;
      if (COEF == 1) ez = ec.ez + ccz * curl;
      else if (COEF == 2) ez = ec.ez + a.cf.curl_value[2] * curl;
      else ez = ec.ez * coef(a.cf.decay[2], a.cf.decay_value[2], oz) + coef(a.cf.curl[2], a.cf.curl_value[2], oz) * curl;
    }
    if (store) {
      if (ADE) {
        if (kin) {
          const float jx = ade_advance_coded(a.ad, 0, ox, ec.ex, code_x);
          const float jy = ade_advance_coded(a.ad, 1, oy, ec.ey, code_y);
          if (jx != 0.0f) ex -= lean_ade_gain<COEF>(a, 0, ox, kcx) * jx;
          if (jy != 0.0f) ey -= lean_ade_gain<COEF>(a, 1, oy, kcy) * jy;
        }
        if (kv) {
          const float jz = ade_advance_coded(a.ad, 2, oz, ec.ez, code_z);
          if (jz != 0.0f) ez -= lean_ade_gain<COEF>(a, 2, oz, kcz) * jz;
        }
      }
      if (et) {
        if (kin) {
          ex += term_value(a.e_terms, a.n_e_terms, 0, i, j, k, step, a.dt);
          ey += term_value(a.e_terms, a.n_e_terms, 1, i, j, k, step, a.dt);
        }
        if (kv) ez += term_value(a.e_terms, a.n_e_terms, 2, i, j, k, step, a.dt);
      }
      if (zpec) { ex = 0.0f; ey = 0.0f; }
      if (kin) {
        exo[ox] = ex;
        eyo[oy] = ey;
      }
      if (kv) ezo[oz] = ez;
      if (DFT) {
        const uint64_t nxs = (uint64_t)sEx * (nx - 1), nys = (uint64_t)sEy * nx, nzs = (uint64_t)sEz * nx;
        for (int f = 0; f < a.nf_dft; ++f) {
          const float wc = wc_t[f], ws = ws_t[f];
          if (kin) {
            dre0[f * nxs + ox] += ex * wc;
            dim0[f * nxs + ox] += ex * ws;
            dre1[f * nys + oy] += ey * wc;
            dim1[f * nys + oy] += ey * ws;
          }
          if (kv) {
            dre2[f * nzs + oz] += ez * wc;
            dim2[f * nzs + oz] += ez * ws;
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

// Z-end tiles.  With the lean tiles aligned to the z PML (the z ends then hold
// just the CPML layers and the faces, 13-16 columns at each end on the dense
// workloads), a wavefront covers BOTH z ends of TWO rows: lanes 0-31 row j, lanes
// 32-63 row j+1; in each half, segment A = a halo lane (k = -1, outside the grid)
// and the columns [0, zend_lo), then segment B = a halo lane (k = zend_hi - 1,
// whose H the first column of B needs) and the columns [zend_hi, nz).  A block
// (four wavefronts) covers eight rows and marches along x over one chunk of the
// lean x range.  The last lane of each segment loads its k+1 neighbours itself
// (the next lane belongs to another segment or row).  Same update as the z-edge
// tiles otherwise; the host uses this layout only when both segments fit in 32
// lanes and the z faces are not periodic.
template <bool ADE, bool DFT, int COEF>
__device__ __forceinline__ void zends_march(
    const FusedArgs& a, int step, int t, gptr<const float> __restrict__ wc_t, gptr<const float> __restrict__ ws_t,
    gptr<const float> __restrict__ exi, gptr<const float> __restrict__ eyi, gptr<const float> __restrict__ ezi,
    gptr<const float> __restrict__ hxi, gptr<const float> __restrict__ hyi, gptr<const float> __restrict__ hzi,
    gptr<float> __restrict__ exo, gptr<float> __restrict__ eyo, gptr<float> __restrict__ ezo,
    gptr<float> __restrict__ hxo, gptr<float> __restrict__ hyo, gptr<float> __restrict__ hzo,
    gptr<const float> __restrict__ p7i, gptr<const float> __restrict__ p9i, gptr<float> __restrict__ p7o,
    gptr<float> __restrict__ p9o, gptr<float> __restrict__ p1, gptr<float> __restrict__ p3,
    gptr<const float> __restrict__ kcx, gptr<const float> __restrict__ kcy, gptr<const float> __restrict__ kcz,
    gptr<float> __restrict__ dre0, gptr<float> __restrict__ dim0, gptr<float> __restrict__ dre1,
    gptr<float> __restrict__ dim1, gptr<float> __restrict__ dre2, gptr<float> __restrict__ dim2) {
  const int nx = a.g.nx, ny = a.g.ny, nz = a.g.nz;
  const mw_cpml& cp = a.cp;
  const int Y0 = a.y_off + kFRows * a.ly0, Y1 = a.y_off + kFRows * a.ly1;
  const int nby = (Y1 - Y0 + 2 * kFRows - 1) / (2 * kFRows);
  const int i0 = a.x_off + a.xchunk * (a.lx0 + t / nby);
  const int i1 = i0 + a.xchunk;
  const int lane = threadIdx.x;
  const int jw = Y0 + 2 * kFRows * (t % nby) + 2 * wave_row();  // first row of this wavefront
  const int j = jw + (lane >> 5);
  const int sub = lane & 31;
  const int na = a.zend_lo + 1, nb = nz - a.zend_hi + 1;
  const bool in_a = sub < na, in_b = !in_a && sub < na + nb;
  const int k = in_a ? sub - 1 : (in_b ? a.zend_hi - 1 + (sub - na) : -1);
  const bool halo = sub == 0 || sub == na;
  const bool seg_last = sub == na - 1 || sub == na + nb - 1;
  const bool row_ok = j < Y1;
  const bool own = (in_a || in_b) && !halo && row_ok;
  const bool kin = k >= 0 && k < nz && row_ok;      // Ex, Ey, Hz columns
  const bool kv = k >= 0 && k < nz - 1 && row_ok;   // Ez, Hx, Hy columns
  const bool kint = k > 0 && k < nz - 1;             // z-interior E sample
  const int jl = row_ok ? j : Y0;                    // a row that exists, for idle lanes' addresses
  const int lz = a.g.bc_lo[2], hzm = a.g.bc_hi[2];
  const bool cpml = cp.enabled != 0;
  const uint32_t sEx = (uint32_t)ny * nz, sEy = (uint32_t)(ny - 1) * nz, sEz = (uint32_t)ny * (nz - 1);
  const uint32_t sHx = (uint32_t)(ny - 1) * (nz - 1), sHy = (uint32_t)ny * (nz - 1), sHz = (uint32_t)(ny - 1) * nz;
  const uint32_t rEx = (uint32_t)jl * nz + (kin ? k : 0), rEz = (uint32_t)jl * (nz - 1) + (kv ? k : 0);
  const float ihz = G(a.g.inv_h[2])[kv ? k : 0];
  const float iez = G(a.g.inv_e[2])[kin ? k : 0];
  float bhz = 0.0f, chz = 0.0f, khz = 1.0f, bez = 0.0f, cez = 0.0f, kez = 1.0f;
  int l7 = -1, l9 = -1, l1 = -1, l3 = -1;
  if (cpml) {
    bhz = G(cp.b_h[2])[kv ? k : 0]; chz = G(cp.c_h[2])[kv ? k : 0]; khz = G(cp.inv_kappa_h[2])[kv ? k : 0];
    bez = G(cp.b_e[2])[kin ? k : 0]; cez = G(cp.c_e[2])[kin ? k : 0]; kez = G(cp.inv_kappa_e[2])[kin ? k : 0];
    if (kv) {
      l7 = slab_local(k, cp.psi_low[7], cp.psi_high_start[7], cp.psi_high_len[7]);
      l9 = slab_local(k, cp.psi_low[9], cp.psi_high_start[9], cp.psi_high_len[9]);
    }
    if (kin) {
      l1 = slab_local(k, cp.psi_low[1], cp.psi_high_start[1], cp.psi_high_len[1]);
      l3 = slab_local(k, cp.psi_low[3], cp.psi_high_start[3], cp.psi_high_len[3]);
    }
  }
  gptr<const float> p7r = p7i ? p7i : exi;
  gptr<const float> p9r = p9i ? p9i : exi;
  gptr<const float> p1r = p1 ? (gptr<const float>)p1 : exi;
  gptr<const float> p3r = p3 ? (gptr<const float>)p3 : exi;
  const int len7 = cp.psi_low[7] + cp.psi_high_len[7], len9 = cp.psi_low[9] + cp.psi_high_len[9];
  const int len1 = cp.psi_low[1] + cp.psi_high_len[1], len3 = cp.psi_low[3] + cp.psi_high_len[3];
  const bool k1_ok = seg_last && k + 1 < nz && row_ok && k >= -1;

  auto load3 = [&](int i, int dr, bool k1) {
    ERow e;
    const uint32_t pe = i * sEx + rEx + dr * nz, py = i * sEy + rEx + dr * nz, pz = i * sEz + rEz + dr * (nz - 1);
    e.ex = exi[kin ? pe : 0];
    e.ey = eyi[kin ? py : 0];
    e.ez = ezi[kv ? pz : 0];
    if (!kin) { e.ex = 0.0f; e.ey = 0.0f; }
    if (!kv) e.ez = 0.0f;
    e.ex_k1 = 0.0f;
    e.ey_k1 = 0.0f;
    if (k1 && k1_ok) {  // k+1 < nz: the column exists (k = -1 loads column 0)
      const uint32_t q1 = i * sEx + (uint32_t)(jl + dr) * nz + (k + 1), q2 = i * sEy + (uint32_t)(jl + dr) * nz + (k + 1);
      e.ex_k1 = exi[q1];
      e.ey_k1 = eyi[q2];
    }
    return e;
  };
  ERow em = load3(i0 - 1, -1, true), ec = load3(i0 - 1, 0, true), ep = load3(i0 - 1, 1, false);
  float hy_prev = 0.0f, hz_prev = 0.0f;
  for (int i = i0 - 1; i < i1; ++i) {
    const ERow nm = load3(i + 1, -1, true), nc = load3(i + 1, 0, true), np = load3(i + 1, 1, false);
    // term guards over the wavefront's rows (jw-1 .. jw+1) and every column
    const bool ht = terms_touch(a.h_terms, a.n_h_terms, i, jw - 1, jw + 1, 0, nz - 1);
    const bool et = terms_touch(a.e_terms, a.n_e_terms, i, jw, jw + 1, 0, nz - 1);
    const bool store = own && i >= i0;
    const float ihx = G(a.g.inv_h[0])[i];
    const float ihy_m = G(a.g.inv_h[1])[jl - 1], ihy_c = G(a.g.inv_h[1])[jl];
    const float iex = G(a.g.inv_e[0])[i], iey = G(a.g.inv_e[1])[jl];
    float eym_k1 = lane_down(em.ey);
    float eyc_k1 = lane_down(ec.ey), exc_k1 = lane_down(ec.ex);
    if (seg_last) { eym_k1 = em.ey_k1; eyc_k1 = ec.ey_k1; exc_k1 = ec.ex_k1; }
    // every load of this plane
    const uint32_t oxm = kv ? i * sHx + (uint32_t)(jl - 1) * (nz - 1) + k : 0;
    const uint32_t ozm = kin ? i * sHz + (uint32_t)(jl - 1) * nz + k : 0;
    const uint32_t oxc = kv ? i * sHx + (uint32_t)jl * (nz - 1) + k : 0;
    const uint32_t oyc = kv ? i * sHy + (uint32_t)jl * (nz - 1) + k : 0;
    const uint32_t ozc = kin ? i * sHz + (uint32_t)jl * nz + k : 0;
    const float hxm_old = hxi[oxm], hzm_old = hzi[ozm], hxc_old = hxi[oxc], hyc_old = hyi[oyc], hzc_old = hzi[ozc];
    const uint32_t q7m = l7 >= 0 ? slab_offset(2, i, jl - 1, k, l7, ny - 1, nz - 1, len7) : 0;
    const uint32_t q7c = l7 >= 0 ? slab_offset(2, i, jl, k, l7, ny - 1, nz - 1, len7) : 0;
    const uint32_t q9c = l9 >= 0 ? slab_offset(2, i, jl, k, l9, ny, nz - 1, len9) : 0;
    const float p7m = p7r[q7m], p7c = p7r[q7c], p9c = p9r[q9c];
    const uint32_t ox = kin ? i * sEx + (uint32_t)jl * nz + k : 0;
    const uint32_t oy = kin ? i * sEy + (uint32_t)jl * nz + k : 0;
    const uint32_t oz = kv ? i * sEz + (uint32_t)jl * (nz - 1) + k : 0;
    const uint32_t q1 = l1 >= 0 ? slab_offset(2, i, jl, k, l1, ny, nz, len1) : 0;
    const uint32_t q3 = l3 >= 0 ? slab_offset(2, i, jl, k, l3, ny - 1, nz, len3) : 0;
    const float p1v = p1r[q1], p3v = p3r[q3];
    // pole weight codes with the plane's other loads (offset 0 where the sample
    // does not exist; read before any store of the plane)
    uint32_t code_x = 0u, code_y = 0u, code_z = 0u;
    if (ADE) {
      if (a.ad.count[0] > 0 && a.ad.weight[0]) code_x = G(a.ad.weight[0])[ox];
      if (a.ad.count[1] > 0 && a.ad.weight[1]) code_y = G(a.ad.weight[1])[oy];
      if (a.ad.count[2] > 0 && a.ad.weight[2]) code_z = G(a.ad.weight[2])[oz];
    }
    float ccx = 0.0f, ccy = 0.0f, ccz = 0.0f;  // E curl coefficients of mode 1
    if (COEF == 1) { ccx = kcx[ox]; ccy = kcy[oy]; ccz = kcz[oz]; }
    float hx_m, hz_m, hx_c, hy_c, hz_c;
    {
      const float da = (ec.ez - em.ez) * ihy_m, db = (eym_k1 - em.ey) * ihz;
      const float pz = l7 >= 0 ? bhz * p7m + chz * db : 0.0f;
      hx_m = kv ? zedge_h<COEF>(a, 3, oxm, hxm_old, da - (db * khz + pz)) : 0.0f;
    }
    hz_m = kin ? zedge_h<COEF>(a, 5, ozm, hzm_old, (nm.ey - em.ey) * ihx - (ec.ex - em.ex) * ihy_m) : 0.0f;
    float p7n, p9n;
    {
      const float da = (ep.ez - ec.ez) * ihy_c, db = (eyc_k1 - ec.ey) * ihz;
      p7n = l7 >= 0 ? bhz * p7c + chz * db : 0.0f;
      hx_c = kv ? zedge_h<COEF>(a, 3, oxc, hxc_old, da - (db * khz + p7n)) : 0.0f;
    }
    {
      const float da = (nc.ez - ec.ez) * ihx, db = (exc_k1 - ec.ex) * ihz;
      p9n = l9 >= 0 ? bhz * p9c + chz * db : 0.0f;
      hy_c = kv ? zedge_h<COEF>(a, 4, oyc, hyc_old, (db * khz + p9n) - da) : 0.0f;
    }
    hz_c = kin ? zedge_h<COEF>(a, 5, ozc, hzc_old, (nc.ey - ec.ey) * ihx - (ep.ex - ec.ex) * ihy_c) : 0.0f;
    if (ht) {
      if (kv) {
        hx_m += term_value(a.h_terms, a.n_h_terms, 3, i, jl - 1, k, step, a.dt);
        hx_c += term_value(a.h_terms, a.n_h_terms, 3, i, jl, k, step, a.dt);
        hy_c += term_value(a.h_terms, a.n_h_terms, 4, i, jl, k, step, a.dt);
      }
      if (kin) {
        hz_m += term_value(a.h_terms, a.n_h_terms, 5, i, jl - 1, k, step, a.dt);
        hz_c += term_value(a.h_terms, a.n_h_terms, 5, i, jl, k, step, a.dt);
      }
    }
    if (store) {
      if (kv) {
        hxo[oxc] = hx_c;
        hyo[oyc] = hy_c;
        if (l7 >= 0 && p7o) p7o[q7c] = p7n;
        if (l9 >= 0 && p9o) p9o[q9c] = p9n;
      }
      if (kin) hzo[ozc] = hz_c;
    }
    const float hy_km = lane_up(hy_c);
    const float hx_km = lane_up(hx_c);
    const bool zpec = face_pec(k, nz, lz, hzm);
    const bool zupd = !zpec && !face_inactive(k, nz, lz, hzm);
    float ex, ey, ez;
    {  // Ex: a = dHz/dy, b = dHy/dz, curl = a - b
      const float da = (hz_c - hz_m) * iey;
      const float db = kint ? (hy_c - hy_km) * iez : shell_edge(k == 0, k == 0 ? lz : hzm, hy_c, hy_km, 0.0f, iez);
      const float pb = l1 >= 0 ? bez * p1v + cez * db : 0.0f;
      if (store && zupd && l1 >= 0 && p1) p1[q1] = pb;
      const float curl = da - (db * kez + pb);
      if (COEF == 1) ex = ec.ex + ccx * curl;
      else if (COEF == 2) ex = ec.ex + a.cf.curl_value[0] * curl;
      else ex = ec.ex * coef(a.cf.decay[0], a.cf.decay_value[0], ox) + coef(a.cf.curl[0], a.cf.curl_value[0], ox) * curl;
      if (!zupd) ex = ec.ex;
    }
    {  // Ey: a = dHz/dx, b = dHx/dz, curl = b - a
      const float da = (hz_c - hz_prev) * iex;
      const float db = kint ? (hx_c - hx_km) * iez : shell_edge(k == 0, k == 0 ? lz : hzm, hx_c, hx_km, 0.0f, iez);
      const float pb = l3 >= 0 ? bez * p3v + cez * db : 0.0f;
      if (store && zupd && l3 >= 0 && p3) p3[q3] = pb;
      const float curl = (db * kez + pb) - da;
      if (COEF == 1) ey = ec.ey + ccy * curl;
      else if (COEF == 2) ey = ec.ey + a.cf.curl_value[1] * curl;
      else ey = ec.ey * coef(a.cf.decay[1], a.cf.decay_value[1], oy) + coef(a.cf.curl[1], a.cf.curl_value[1], oy) * curl;
      if (!zupd) ey = ec.ey;
    }
    {  // Ez: a = dHy/dx, b = dHx/dy, curl = a - b (no z derivative)
      const float curl = (hy_c - hy_prev) * iex - (hx_c - hx_m) * iey;
      if (COEF == 1) ez = ec.ez + ccz * curl;
      else if (COEF == 2) ez = ec.ez + a.cf.curl_value[2] * curl;
      else ez = ec.ez * coef(a.cf.decay[2], a.cf.decay_value[2], oz) + coef(a.cf.curl[2], a.cf.curl_value[2], oz) * curl;
    }
    if (store) {
      if (ADE) {
        if (kin) {
          const float jx = ade_advance_coded(a.ad, 0, ox, ec.ex, code_x);
          const float jy = ade_advance_coded(a.ad, 1, oy, ec.ey, code_y);
          if (jx != 0.0f) ex -= lean_ade_gain<COEF>(a, 0, ox, kcx) * jx;
          if (jy != 0.0f) ey -= lean_ade_gain<COEF>(a, 1, oy, kcy) * jy;
        }
        if (kv) {
          const float jz = ade_advance_coded(a.ad, 2, oz, ec.ez, code_z);
          if (jz != 0.0f) ez -= lean_ade_gain<COEF>(a, 2, oz, kcz) * jz;
        }
      }
      if (et) {
        if (kin) {
          ex += term_value(a.e_terms, a.n_e_terms, 0, i, jl, k, step, a.dt);
          ey += term_value(a.e_terms, a.n_e_terms, 1, i, jl, k, step, a.dt);
        }
        if (kv) ez += term_value(a.e_terms, a.n_e_terms, 2, i, jl, k, step, a.dt);
      }
      if (zpec) { ex = 0.0f; ey = 0.0f; }
      if (kin) {
        exo[ox] = ex;
        eyo[oy] = ey;
      }
      if (kv) ezo[oz] = ez;
      if (DFT) {
        const uint64_t nxs = (uint64_t)sEx * (nx - 1), nys = (uint64_t)sEy * nx, nzs = (uint64_t)sEz * nx;
        for (int f = 0; f < a.nf_dft; ++f) {
          const float wc = wc_t[f], ws = ws_t[f];
          if (kin) {
            dre0[f * nxs + ox] += ex * wc;
            dim0[f * nxs + ox] += ex * ws;
            dre1[f * nys + oy] += ey * wc;
            dim1[f * nys + oy] += ey * ws;
          }
          if (kv) {
            dre2[f * nzs + oz] += ez * wc;
            dim2[f * nzs + oz] += ez * ws;
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

// Tile t of the z-edge launch: the z index runs over the low edge tiles
// [0, lz0) then the high ones [lz1, tiles_z), over the lean box's y and x tiles.
template <bool ADE, bool DFT, int COEF>
__device__ __forceinline__ void zedge_tile(const FusedArgs& a, int step, const float* wc_t, const float* ws_t, int t) {
  const int nze = a.lz0 + (a.tiles_z - a.lz1);
  const int by = a.ly1 - a.ly0;
  const int zi = t % nze;
  const int tz = zi < a.lz0 ? zi : a.lz1 + (zi - a.lz0);
  const int ty = a.ly0 + (t / nze) % by;
  const int tx = a.lx0 + t / (nze * by);
  zedge_march<ADE, DFT, COEF>(a, step, tz, ty, tx, G(wc_t), G(ws_t), G(a.in.e[0]), G(a.in.e[1]), G(a.in.e[2]),
                              G(a.in.h[0]), G(a.in.h[1]), G(a.in.h[2]), G(a.out.e[0]), G(a.out.e[1]),
                              G(a.out.e[2]), G(a.out.h[0]), G(a.out.h[1]), G(a.out.h[2]), G(a.cp.psi[7]),
                              G(a.cp.psi[9]), G(a.psi_h_out[1]), G(a.psi_h_out[3]), G(a.cp.psi[1]), G(a.cp.psi[3]),
                              G(a.cf.curl[0]), G(a.cf.curl[1]), G(a.cf.curl[2]), G(a.dft_re[0]), G(a.dft_im[0]),
                              G(a.dft_re[1]), G(a.dft_im[1]), G(a.dft_re[2]), G(a.dft_im[2]));
}

__host__ __device__ inline int zedge_tiles(const FusedArgs& a) {
  if (a.zends) {
    const int rows = kFRows * (a.ly1 - a.ly0);
    return ((rows + 2 * kFRows - 1) / (2 * kFRows)) * (a.lx1 - a.lx0);
  }
  return a.zedge ? (a.lz0 + (a.tiles_z - a.lz1)) * (a.ly1 - a.ly0) * (a.lx1 - a.lx0) : 0;
}

template <bool ADE, bool DFT, int COEF>
__device__ __forceinline__ void zends_tile(const FusedArgs& a, int step, const float* wc_t, const float* ws_t, int t) {
  zends_march<ADE, DFT, COEF>(a, step, t, G(wc_t), G(ws_t), G(a.in.e[0]), G(a.in.e[1]), G(a.in.e[2]), G(a.in.h[0]),
                              G(a.in.h[1]), G(a.in.h[2]), G(a.out.e[0]), G(a.out.e[1]), G(a.out.e[2]), G(a.out.h[0]),
                              G(a.out.h[1]), G(a.out.h[2]), G(a.cp.psi[7]), G(a.cp.psi[9]), G(a.psi_h_out[1]),
                              G(a.psi_h_out[3]), G(a.cp.psi[1]), G(a.cp.psi[3]), G(a.cf.curl[0]), G(a.cf.curl[1]),
                              G(a.cf.curl[2]), G(a.dft_re[0]), G(a.dft_im[0]), G(a.dft_re[1]), G(a.dft_im[1]),
                              G(a.dft_re[2]), G(a.dft_im[2]));
}

template <bool ADE, bool DFT, int COEF>
__global__ __launch_bounds__(256) void zends_step_kernel(FusedArgs a) {
  const int logical = xcd_logical_block();
  if (logical >= zedge_tiles(a)) return;
  zends_tile<ADE, DFT, COEF>(a, a.step, a.dft_cos, a.dft_sin, logical);
}

template <bool ADE, bool DFT, int COEF>
__global__ __launch_bounds__(256) void zedge_step_kernel(FusedArgs a) {
  const int logical = xcd_logical_block();
  if (logical >= zedge_tiles(a)) return;
  zedge_tile<ADE, DFT, COEF>(a, a.step, a.dft_cos, a.dft_sin, logical);
}

template <bool ADE, bool DFT, int COEF>
__global__ __launch_bounds__(256) void zedge_batch_kernel(const FusedArgs* __restrict__ args,
                                                          const int* __restrict__ prefix, int nscenes, int total,
                                                          const int* __restrict__ counter) {
  const int logical = xcd_logical_block();
  if (logical >= total) return;
  const int s = table_scene(prefix, nscenes, logical);
  const int n = *counter;
  const FusedArgs& a = args[2 * s + (n & 1)];
  const int64_t w = (int64_t)n * a.nf_dft;
  zedge_tile<ADE, DFT, COEF>(a, n, a.dft_cos + w, a.dft_sin + w, logical - prefix[s]);
}

}  // namespace mw
