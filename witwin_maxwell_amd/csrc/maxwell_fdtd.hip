// C ABI entry points of the FDTD runtime (declarations: include/maxwell_fdtd.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "maxwell_fdtd.h"
#include "adjoint.h"
#include "batch.h"
#include "fused.h"
#include "shell.h"
#include "terms.h"
#include "zedge.h"
#include "yee.h"

namespace {

thread_local std::string g_error;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_error = buf;
  return code;
}

int check_launch(const char* what) {
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return fail(MW_ERR_LAUNCH, "%s: launch failed: %s", what, hipGetErrorString(err));
  return MW_OK;
}

constexpr int64_t kMaxElems = (int64_t(1) << 31) - 1;

int check_grid(const mw_grid* g) {
  if (!g) return fail(MW_ERR_ARGUMENT, "grid is NULL");
  if (g->nx < 2 || g->ny < 2 || g->nz < 2) return fail(MW_ERR_ARGUMENT, "grid needs >= 2 nodes per axis, got (%d, %d, %d)", g->nx, g->ny, g->nz);
  if ((int64_t)g->nx * g->ny * g->nz > kMaxElems) return fail(MW_ERR_ARGUMENT, "grid has more than 2^31 nodes");
  for (int a = 0; a < 3; ++a) {
    if (!g->inv_e[a] || !g->inv_h[a]) return fail(MW_ERR_ARGUMENT, "grid spacing array %d is NULL", a);
    for (int s = 0; s < 2; ++s) {
      const int m = s ? g->bc_hi[a] : g->bc_lo[a];
      if (m != MW_BC_NONE && m != MW_BC_PML && m != MW_BC_PERIODIC && m != MW_BC_PEC && m != MW_BC_PMC)
        return fail(MW_ERR_ARGUMENT, "unsupported boundary code %d on axis %d", m, a);
    }
  }
  return MW_OK;
}

void component_shape(const mw_grid* g, int c, int out[3]) {
  out[0] = g->nx; out[1] = g->ny; out[2] = g->nz;
  switch (c) {
    case 0: out[0] -= 1; break;
    case 1: out[1] -= 1; break;
    case 2: out[2] -= 1; break;
    case 3: out[1] -= 1; out[2] -= 1; break;
    case 4: out[0] -= 1; out[2] -= 1; break;
    default: out[0] -= 1; out[1] -= 1; break;
  }
}

// psi q -> (component, axis)
constexpr int kPsiComp[12] = {0, 0, 1, 1, 2, 2, 3, 3, 4, 4, 5, 5};
constexpr int kPsiAxis[12] = {1, 2, 0, 2, 0, 1, 1, 2, 0, 2, 0, 1};

int check_cpml(const mw_grid* g, const mw_cpml* c, bool electric) {
  if (!c || !c->enabled) return MW_OK;
  for (int a = 0; a < 3; ++a) {
    if (electric && (!c->inv_kappa_e[a] || !c->b_e[a] || !c->c_e[a])) return fail(MW_ERR_ARGUMENT, "CPML E profile %d is NULL", a);
    if (!electric && (!c->inv_kappa_h[a] || !c->b_h[a] || !c->c_h[a])) return fail(MW_ERR_ARGUMENT, "CPML H profile %d is NULL", a);
  }
  for (int q = electric ? 0 : 6; q < (electric ? 6 : 12); ++q) {
    int shp[3];
    component_shape(g, kPsiComp[q], shp);
    const int size = shp[kPsiAxis[q]];
    const int low = c->psi_low[q], hs = c->psi_high_start[q], hl = c->psi_high_len[q];
    if (low < 0 || hl < 0 || low > size || hs < low || hs + hl > size)
      return fail(MW_ERR_ARGUMENT, "psi %d slab layout (%d, %d, %d) does not fit axis size %d", q, low, hs, hl, size);
    if ((low + hl) > 0 && !c->psi[q]) return fail(MW_ERR_ARGUMENT, "psi %d is NULL", q);
  }
  return MW_OK;
}

int pick_xchunk(const mw_grid* g) {
  // enough blocks to fill 256 CUs several times over, long enough x-runs to
  // amortise the carried-neighbour preload
  const int64_t cols = (int64_t)((g->nz + mw::kTileZ - 1) / mw::kTileZ) * ((g->ny + mw::kTileY - 1) / mw::kTileY);
  int chunk = 16;
  while (chunk > 2 && cols * ((g->nx + chunk - 1) / chunk) < 4096) chunk /= 2;
  return chunk;
}

int launch_h(const mw_grid* g, const mw_coeffs* cf, const mw_cpml* cp, const mw_fields* f, int mask, hipStream_t s) {
  const int chunk = pick_xchunk(g);
  dim3 block(mw::kTileZ, mw::kTileY, 1);
  dim3 grid((g->nz + mw::kTileZ - 1) / mw::kTileZ, (g->ny + mw::kTileY - 1) / mw::kTileY, (g->nx + chunk - 1) / chunk);
  mw_cpml cpv{};
  if (cp) cpv = *cp;
  if (cpv.enabled) hipLaunchKernelGGL((mw::update_h_kernel<true>), grid, block, 0, s, *g, *cf, cpv, *f, mask, chunk);
  else hipLaunchKernelGGL((mw::update_h_kernel<false>), grid, block, 0, s, *g, *cf, cpv, *f, mask, chunk);
  return check_launch("update_magnetic");
}

int launch_e(const mw_grid* g, const mw_coeffs* cf, const mw_cpml* cp, const mw_ade* ad, const mw_fields* f, int mask,
             hipStream_t s) {
  const int chunk = pick_xchunk(g);
  dim3 block(mw::kTileZ, mw::kTileY, 1);
  dim3 grid((g->nz + mw::kTileZ - 1) / mw::kTileZ, (g->ny + mw::kTileY - 1) / mw::kTileY, (g->nx + chunk - 1) / chunk);
  mw_cpml cpv{};
  if (cp) cpv = *cp;
  mw_ade adv{};
  const bool ade = ad && (ad->count[0] + ad->count[1] + ad->count[2]) > 0;
  if (ade) adv = *ad;
  if (cpv.enabled) {
    if (ade) hipLaunchKernelGGL((mw::update_e_kernel<true, true>), grid, block, 0, s, *g, *cf, cpv, adv, *f, mask, chunk);
    else hipLaunchKernelGGL((mw::update_e_kernel<true, false>), grid, block, 0, s, *g, *cf, cpv, adv, *f, mask, chunk);
  } else {
    if (ade) hipLaunchKernelGGL((mw::update_e_kernel<false, true>), grid, block, 0, s, *g, *cf, cpv, adv, *f, mask, chunk);
    else hipLaunchKernelGGL((mw::update_e_kernel<false, false>), grid, block, 0, s, *g, *cf, cpv, adv, *f, mask, chunk);
  }
  return check_launch("update_electric");
}

int check_fields(const mw_fields* f, bool e, bool h) {
  if (!f) return fail(MW_ERR_ARGUMENT, "fields is NULL");
  for (int c = 0; c < 3; ++c) {
    if (!f->e[c] || !f->h[c]) return fail(MW_ERR_ARGUMENT, "field pointer %d is NULL", c);
  }
  (void)e; (void)h;
  return MW_OK;
}

int check_ade(const mw_grid* g, const mw_ade* a) {
  if (!a) return MW_OK;
  for (int c = 0; c < 3; ++c) {
    if (a->count[c] < 0 || a->count[c] > MW_MAX_POLES) return fail(MW_ERR_ARGUMENT, "ADE pole count %d out of range", a->count[c]);
    if (a->count[c] > 0 && !a->inv_eps[c]) return fail(MW_ERR_ARGUMENT, "ADE inv_eps %d is NULL", c);
    for (int p = 0; p < a->count[c]; ++p) {
      const mw_pole& pl = a->poles[c][p];
      if (pl.kind < 0 || pl.kind > 2) return fail(MW_ERR_ARGUMENT, "ADE pole kind %d invalid", pl.kind);
      if (!pl.drive || !pl.current || (pl.kind != 1 && !pl.polarization))
        return fail(MW_ERR_ARGUMENT, "ADE pole (%d, %d) has a NULL state array", c, p);
    }
  }
  (void)g;
  return MW_OK;
}

int grid_blocks(int64_t n) { return (int)((n + 255) / 256); }

mw::FieldSet field_set(const mw_grid* g, const mw_fields* f) {
  mw::FieldSet fs;
  for (int c = 0; c < 6; ++c) {
    fs.ptr[c] = c < 3 ? f->e[c] : f->h[c - 3];
    component_shape(g, c, fs.shape[c]);
  }
  return fs;
}

int check_terms(const mw_grid* g, const mw_term* t, int n, const char* which, int total_steps) {
  if (n < 0 || n > MW_MAX_TERMS) return fail(MW_ERR_ARGUMENT, "%s: %d terms (max %d)", which, n, MW_MAX_TERMS);
  if (n > 0 && !t) return fail(MW_ERR_ARGUMENT, "%s: term array is NULL", which);
  for (int i = 0; i < n; ++i) {
    const mw_term& x = t[i];
    if (x.component < 0 || x.component > 5) return fail(MW_ERR_ARGUMENT, "%s[%d]: bad component", which, i);
    int shp[3];
    component_shape(g, x.component, shp);
    for (int a = 0; a < 3; ++a) {
      if (x.shape[a] <= 0 || x.offset[a] < 0 || x.offset[a] + x.shape[a] > shp[a])
        return fail(MW_ERR_ARGUMENT, "%s[%d]: patch [%d, %d) exceeds axis %d of size %d", which, i, x.offset[a],
                    x.offset[a] + x.shape[a], a, shp[a]);
    }
    if (!x.patch) return fail(MW_ERR_ARGUMENT, "%s[%d]: patch is NULL", which, i);
    if (x.kind == MW_TERM_ADDITIVE && !x.signal_table) return fail(MW_ERR_ARGUMENT, "%s[%d]: signal table is NULL", which, i);
    if (x.kind == MW_TERM_SHIFTED && !x.delay) return fail(MW_ERR_ARGUMENT, "%s[%d]: delay is NULL", which, i);
    if (x.kind == MW_TERM_LINE) {
      if (!x.line || x.line_len <= 0) return fail(MW_ERR_ARGUMENT, "%s[%d]: line is NULL", which, i);
      if (x.line_axis > 2) return fail(MW_ERR_ARGUMENT, "%s[%d]: bad line axis %d", which, i, x.line_axis);
      const int last = x.line_axis < 0 ? x.line_base : x.line_base + x.line_step * (x.shape[x.line_axis] - 1);
      if (x.line_base < 0 || x.line_base >= x.line_len || last < 0 || last >= x.line_len)
        return fail(MW_ERR_ARGUMENT, "%s[%d]: line samples [%d, %d] outside the line of %d", which, i, x.line_base,
                    last, x.line_len);
    }
    if (x.kind != MW_TERM_ADDITIVE && x.kind != MW_TERM_SHIFTED && x.kind != MW_TERM_LINE)
      return fail(MW_ERR_ARGUMENT, "%s[%d]: bad kind", which, i);
  }
  (void)total_steps;
  return MW_OK;
}

int launch_terms(const mw_grid* g, const mw_fields* f, const mw_term* t, int n, int step, double dt, hipStream_t s) {
  if (n <= 0) return MW_OK;
  mw::TermSet set{};
  set.count = n;
  int64_t cells = 0;
  for (int i = 0; i < n; ++i) {
    set.t[i] = t[i];
    cells = std::max<int64_t>(cells, (int64_t)t[i].shape[0] * t[i].shape[1] * t[i].shape[2]);
  }
  dim3 grid(grid_blocks(cells), n, 1);
  hipLaunchKernelGGL(mw::apply_terms_kernel, grid, dim3(256), 0, s, set, field_set(g, f), step, dt);
  return check_launch("apply_terms");
}

int coef_mode(const mw_coeffs& c) {
  // coefficient mode of the lean interior path (see fused.h)
  bool decay_one = true, h_curl_scalar = true, e_curl_scalar = true, e_curl_array = true;
  for (int q = 0; q < 6; ++q)
    if (c.decay[q] || c.decay_value[q] != 1.0f) decay_one = false;
  for (int q = 3; q < 6; ++q)
    if (c.curl[q]) h_curl_scalar = false;
  for (int q = 0; q < 3; ++q) {
    if (c.curl[q]) e_curl_scalar = false;
    else e_curl_array = false;
  }
  if (decay_one && h_curl_scalar && e_curl_scalar) return 2;
  if (decay_one && h_curl_scalar && e_curl_array) return 1;
  return 0;
}


template <bool ADE, bool DFT>
void dispatch_lean(dim3 grid, dim3 block, hipStream_t s, const mw::FusedArgs& a, int mode) {
  if (mode == 2) hipLaunchKernelGGL((mw::lean_step_kernel<ADE, DFT, 2>), grid, block, 0, s, a);
  else if (mode == 1) hipLaunchKernelGGL((mw::lean_step_kernel<ADE, DFT, 1>), grid, block, 0, s, a);
  else hipLaunchKernelGGL((mw::lean_step_kernel<ADE, DFT, 0>), grid, block, 0, s, a);
}

#ifndef MW_MERGE_SHELL_H
#define MW_MERGE_SHELL_H 1
#endif

template <bool CPML, bool ADE, bool DFT>
void dispatch_shell_h_lean(int sb, int lb, hipStream_t s, const mw::ShellArgs& sa, int mode, const mw::PointSet& pts,
                           const mw::ZGatherJob& zg) {
  const int pb = pts.count > 0 ? (pts.count + 255) / 256 : 0;
  const dim3 grid(sb + lb + zg.blocks + pb), block(mw::kFLanes, mw::kFRows, 1);
  if (mode == 2) hipLaunchKernelGGL((mw::shell_h_lean_kernel<CPML, ADE, DFT, 2>), grid, block, 0, s, sa, sb, lb, pts, zg);
  else if (mode == 1) hipLaunchKernelGGL((mw::shell_h_lean_kernel<CPML, ADE, DFT, 1>), grid, block, 0, s, sa, sb, lb, pts, zg);
  else hipLaunchKernelGGL((mw::shell_h_lean_kernel<CPML, ADE, DFT, 0>), grid, block, 0, s, sa, sb, lb, pts, zg);
}

template <bool ADE, bool DFT>
void dispatch_zedge(dim3 grid, dim3 block, hipStream_t s, const mw::FusedArgs& a, int mode) {
  if (a.zends) {
    if (mode == 2) hipLaunchKernelGGL((mw::zends_step_kernel<ADE, DFT, 2>), grid, block, 0, s, a);
    else if (mode == 1) hipLaunchKernelGGL((mw::zends_step_kernel<ADE, DFT, 1>), grid, block, 0, s, a);
    else hipLaunchKernelGGL((mw::zends_step_kernel<ADE, DFT, 0>), grid, block, 0, s, a);
    return;
  }
  if (mode == 2) hipLaunchKernelGGL((mw::zedge_step_kernel<ADE, DFT, 2>), grid, block, 0, s, a);
  else if (mode == 1) hipLaunchKernelGGL((mw::zedge_step_kernel<ADE, DFT, 1>), grid, block, 0, s, a);
  else hipLaunchKernelGGL((mw::zedge_step_kernel<ADE, DFT, 0>), grid, block, 0, s, a);
}

// Tile origin and lean-box extent along one axis.  Tile t covers `width`
// samples starting at off + width*t; the lean box holds the tiles whose samples,
// plus `lead` samples before them (the halo), all lie in [lo, hi).
struct AxisTiling {
  int off, count, lean0, lean1;
};
AxisTiling tile_axis(int size, int width, int lead, int lo, int hi) {
  AxisTiling t;
  const int first = lo + lead;  // first owned sample of the first lean tile
  const int r = first % width;
  t.off = r ? r - width : 0;
  t.count = (size - t.off + width - 1) / width;
  t.lean0 = (first - t.off) / width;
  int n = 0;
  while (first + width * (n + 1) <= hi) ++n;
  t.lean1 = std::max(t.lean0, std::min(t.lean0 + n, t.count));
  return t;
}

// z tiling of the single-pass step: every tile is either a lean tile (all of
// its samples, the halo column before them included, inside [lo, hi)) or a
// z-edge tile (zedge.h), which costs more per tile.  Pick the origin among the
// `width` possible ones with the fewest tiles, counting a z-edge tile as 1.3 and
// preferring any origin that leaves at least one lean tile.
AxisTiling tile_axis_z(int size, int width, int lo, int hi) {
  AxisTiling best{0, 0, 0, 0};
  double best_cost = 1e30;
  for (int off = -(width - 1); off <= 0; ++off) {
    const int count = (size - off + width - 1) / width;
    int t0 = -1, t1 = -1;
    for (int t = 0; t < count; ++t) {
      const int first = off + width * t, last = first + width - 1;
      if (first - 1 >= lo && last <= hi - 1) {
        if (t0 < 0) t0 = t;
        t1 = t + 1;
      }
    }
    if (t0 < 0) t0 = t1 = 0;
    // without a lean tile there are no z-edge tiles either: every column then goes
    // to the per-cell shell, the dearest path, so such an origin is a last resort
    const double cost = t1 > t0 ? (t1 - t0) + 1.3 * (count - (t1 - t0)) : 1e9 + count;
    if (cost < best_cost - 1e-9) {
      best_cost = cost;
      best = AxisTiling{off, count, t0, t1};
    }
  }
  return best;
}

// This step's values of the time-shifted terms of both families (H then E), in
// as few launches as the term set's capacity allows (one on every workload here).
int launch_shifted_values(const mw_term* h, int nh, const mw_term* e, int ne, int step, double dt, hipStream_t s) {
  mw::TermSet set{};
  int64_t cells = 0;
  auto flush = [&]() {
    if (set.count == 0) return MW_OK;
    hipLaunchKernelGGL(mw::shifted_values_kernel, dim3(grid_blocks(cells), set.count), dim3(256), 0, s, set, step, dt);
    set.count = 0;
    cells = 0;
    return check_launch("shifted_values");
  };
  for (int f = 0; f < 2; ++f) {
    const mw_term* t = f ? e : h;
    const int n = f ? ne : nh;
    for (int i = 0; i < n; ++i) {
      if (t[i].kind != MW_TERM_SHIFTED) continue;
      if (set.count == MW_MAX_TERMS)
        if (int rc = flush()) return rc;
      set.t[set.count++] = t[i];
      cells = std::max<int64_t>(cells, (int64_t)t[i].shape[0] * t[i].shape[1] * t[i].shape[2]);
    }
  }
  return flush();
}

// Arguments of the single-pass step for step n (buffers by parity, this step's
// DFT weights) and the host tiling: lean box aligned with the PML interfaces,
// x-chunk chosen near `target` planes (chunk_override > 0 forces it).
int build_fused_args(const mw_program* p, int n, int chunk_override, mw::FusedArgs* out, int* nlean_out) {
  mw::FusedArgs& a = *out;
  memset(&a, 0, sizeof(a));
  a.g = p->grid;
  a.cf = p->coeffs;
  a.cp = p->cpml;
  const bool odd = (n & 1) != 0;
  a.in = odd ? p->fields_alt : p->fields;
  a.out = odd ? p->fields : p->fields_alt;
  if (p->cpml.enabled) {
    for (int q = 0; q < 6; ++q) {
      a.cp.psi[6 + q] = odd ? p->psi_h_alt[q] : p->cpml.psi[6 + q];
      a.psi_h_out[q] = odd ? p->cpml.psi[6 + q] : p->psi_h_alt[q];
    }
  }
  const bool ade = p->has_ade && (p->ade.count[0] + p->ade.count[1] + p->ade.count[2]) > 0;
  if (ade) a.ad = p->ade;
  a.h_terms = p->h_terms_dev;
  a.e_terms = p->e_terms_dev;
  a.n_h_terms = p->n_h_terms;
  a.n_e_terms = p->n_e_terms;
  a.nf_dft = p->nf_dft;
  a.step = n;
  if (p->nf_dft > 0) {
    a.dft_cos = p->dft_cos + (int64_t)n * p->nf_dft;
    a.dft_sin = p->dft_sin + (int64_t)n * p->nf_dft;
    for (int c = 0; c < 3; ++c) { a.dft_re[c] = p->dft_re[c]; a.dft_im[c] = p->dft_im[c]; }
  }
  a.dt = p->dt;
  const mw_grid* g = &p->grid;
  const mw_cpml& cp = p->cpml;
  // lean ranges [lo, hi) per axis: strictly interior and outside every psi slab
  auto low = [&](int q) { return cp.enabled ? cp.psi_low[q] : 0; };
  auto high = [&](int q, int size) { return (cp.enabled && cp.psi_high_len[q] > 0) ? cp.psi_high_start[q] : size; };
  // x and y stop two samples short of the high face: the march loads E of plane
  // i+1 (Ex has nx-1 planes) and of row j+1 (Ey has ny-1 rows) for every owned
  // plane and row, so the last owned ones are nx-3 and ny-3 (without PML layers
  // on an axis nothing else bounds it there)
  const int ilo = std::max({1, low(8), low(2)}), ihi = std::min({g->nx - 2, high(8, g->nx - 1), high(2, g->nx)});
  const int jlo = std::max({2, low(6) + 1, low(0)}), jhi = std::min({g->ny - 2, high(6, g->ny - 1), high(0, g->ny)});
  const int klo = std::max({1, low(7), low(1)}), khi = std::min({g->nz - 1, high(7, g->nz - 1), high(1, g->nz)});
  // z: owned columns start one after the halo lane; rows: row j-1 is recomputed
  // (jlo already accounts for it); planes: one halo plane before each chunk
  // z-edge tiles handle PML/PEC/PMC/none z faces; periodic z faces stay in the shell
  const bool zedge = p->z_tiling == 1 && g->bc_lo[2] != MW_BC_PERIODIC;
  const AxisTiling tz = zedge ? tile_axis_z(g->nz, mw::kFOwnCols, klo, khi) : tile_axis(g->nz, mw::kFOwnCols, 1, klo, khi);
  const AxisTiling ty = tile_axis(g->ny, mw::kFRows, 0, jlo, jhi);
  // x chunk: short marches keep the planes that the blocks resident on one XCD
  // touch at a time few and contiguous (measured best at 3-6 planes on the dense
  // workloads); among 3..8 take the one leaving the fewest interior planes to the
  // shell, nearest to 4 on ties
  const int target = 4;
  int chunk = target;
  if (chunk_override > 0) {
    chunk = chunk_override;
  } else {
    const int span = ihi - (ilo + 1);
    int best = span >= 0 ? span % target : 0;
    for (int c = 3; c <= 8 && span >= 3; ++c) {
      const int left = span % c;
      if (left < best || (left == best && std::abs(c - target) < std::abs(chunk - target))) { best = left; chunk = c; }
    }
  }
  const AxisTiling tx = tile_axis(g->nx, chunk, 1, ilo, ihi);
  a.z_off = tz.off;
  a.y_off = ty.off;
  a.x_off = tx.off;
  a.tiles_z = tz.count;
  a.tiles_y = ty.count;
  a.tiles_x = tx.count;
  a.xchunk = chunk;
  a.lz0 = tz.lean0; a.lz1 = tz.lean1;
  a.ly0 = ty.lean0; a.ly1 = ty.lean1;
  a.lx0 = tx.lean0; a.lx1 = tx.lean1;
  a.zedge = zedge ? 1 : 0;
  int nlean = (a.lz1 - a.lz0) * (a.ly1 - a.ly0) * (a.lx1 - a.lx0);
  if (nlean <= 0) { nlean = 0; a.lz0 = a.lz1 = a.ly0 = a.ly1 = a.lx0 = a.lx1 = 0; }
  // z-end tiles (z_tiling 3 and 4): both z ends of two rows per wavefront, when each
  // end plus its halo lane fits half a wavefront; otherwise the z ends stay shell cells
  a.zends = 0;
  a.zend_lo = a.zend_hi = 0;
  if ((p->z_tiling == 3 || p->z_tiling == 4) && nlean > 0 && g->bc_lo[2] != MW_BC_PERIODIC) {
    const int z0 = a.z_off + mw::kFOwnCols * a.lz0, z1 = a.z_off + mw::kFOwnCols * a.lz1;
    if (z0 >= 1 && z1 < g->nz && (z0 + 1) + (g->nz - z1 + 1) <= 32) {
      a.zends = 1;
      a.zend_lo = z0;
      a.zend_hi = z1;
    }
  }
  *nlean_out = nlean;
  return MW_OK;
}

// Default x planes per shell block: short marches keep the planes that the
// blocks resident on one XCD touch at the same time few and contiguous.
constexpr int kShellChunk = 4;

// The cells outside the lean box (and its z-edge tiles, when the z ends are
// tiled that way) as at most six boxes: x slabs over the whole cross-section, y
// slabs over the lean x range and, without z-edge tiles, z slabs over the lean
// x-y range; each with its lane layout and x chunk.  Without a lean box, one box
// holds the whole grid.
void build_shell_args(const mw::FusedArgs& a, int nlean, int chunk, mw::ShellArgs* out) {
  mw::ShellArgs& sa = *out;
  sa.f = a;
  sa.nboxes = 0;
  sa.total = 0;
  const int nx = a.g.nx, ny = a.g.ny, nz = a.g.nz;
  int X0 = nx, X1 = nx, Y0 = 0, Y1 = 0, Z0 = 0, Z1 = 0;
  if (nlean > 0) {
    X0 = a.x_off + a.xchunk * a.lx0; X1 = a.x_off + a.xchunk * a.lx1;
    Y0 = a.y_off + mw::kFRows * a.ly0; Y1 = a.y_off + mw::kFRows * a.ly1;
    Z0 = a.z_off + mw::kFOwnCols * a.lz0; Z1 = a.z_off + mw::kFOwnCols * a.lz1;
  }
  // with a lean box, its z ends [0, Z0) and [Z1, nz) over the lean x-y range are
  // the z-edge or z-end tiles (zedge.h) when those are on, not shell boxes
  const int boxes[6][6] = {{0, X0, 0, ny, 0, nz},  {X1, nx, 0, ny, 0, nz},  {X0, X1, 0, Y0, 0, nz},
                           {X0, X1, Y1, ny, 0, nz}, {X0, X1, Y0, Y1, 0, Z0}, {X0, X1, Y0, Y1, Z1, nz}};
  const int min_shift = (a.ad.count[0] + a.ad.count[1] + a.ad.count[2]) > 0 ? 4 : 5;
  for (int bi = 0; bi < ((a.zedge || a.zends) ? 4 : 6); ++bi) {
    const int* bb = boxes[bi];
    const int xe = bb[1] - bb[0], ye = bb[3] - bb[2], ze = bb[5] - bb[4];
    if (xe <= 0 || ye <= 0 || ze <= 0) continue;
    mw::ShellBox& b = sa.box[sa.nboxes++];
    b.lo[0] = bb[0]; b.hi[0] = bb[1];
    b.lo[1] = bb[2]; b.hi[1] = bb[3];
    b.lo[2] = bb[4]; b.hi[2] = bb[5];
    // lanes per row: the width with the fewest idle lanes over the box's z extent
    // (ties within 5% go to the wider rows), from 16, 32 or 64 with dispersive poles
    // and from 32 or 64 without.  16-lane rows idle fewer lanes on the thin z slabs
    // but put four row segments in every wavefront instead of two; per step with
    // 16- vs 32-lane rows on those slabs (13-16 planes): dielectric boxes 281^3 nodes
    // 644 vs 624 us, vacuum 85^3 58 vs 53 us, design region 225x225x145 252 vs 245
    // us; with pole state in the E pass (dispersive 409x409x281) 2679 vs 2758 us.
    b.w_shift = 6;
    double best_use = 0.0;
    for (int sh = 6; sh >= min_shift; --sh) {
      const int w = 1 << sh;
      const double use = (double)ze / (double)(((ze + w - 1) / w) * w);
      if (use > best_use * 1.05) { best_use = use; b.w_shift = sh; }
    }
    const int w = 1 << b.w_shift, rows = 256 >> b.w_shift;
    b.chunk = std::min(xe, chunk > 0 ? chunk : kShellChunk);
    b.nbz = (ze + w - 1) / w;
    b.nby = (ye + rows - 1) / rows;
    b.nbx = (xe + b.chunk - 1) / b.chunk;
    b.first = sa.total;
    sa.total += b.nbz * b.nby * b.nbx;
  }
}

// Any periodic face: the shell E pass then reads H across the seams (the
// kernel variant without them leaves those loads out).
bool grid_periodic(const mw_grid& g) {
  for (int q = 0; q < 3; ++q)
    if (g.bc_lo[q] == MW_BC_PERIODIC || g.bc_hi[q] == MW_BC_PERIODIC) return true;
  return false;
}

template <bool CPML, bool ADE, bool DFT>
int launch_shell_e(dim3 grid, hipStream_t s, const mw::ShellArgs& sa) {
  if (grid_periodic(sa.f.g)) hipLaunchKernelGGL((mw::shell_e_kernel<CPML, ADE, DFT, true>), grid, dim3(64, 4, 1), 0, s, sa);
  else hipLaunchKernelGGL((mw::shell_e_kernel<CPML, ADE, DFT, false>), grid, dim3(64, 4, 1), 0, s, sa);
  return check_launch("fused_step (shell E)");
}

// A second stream for the z-end launch (z_tiling 4).  The z-end tiles, like the
// interior + shell-H launch, read only the step's inputs and write their own
// cells, so the two launches run side by side; shell E reads the z ends' H and
// waits for them.  A call takes a stream and its fork/join events from a
// per-device pool for its whole duration and gives them back at its end, so
// callers on different host threads never share one and the pool grows only to
// the number of concurrent callers.  The streams are non-blocking and of default
// priority; the library keeps them until the process ends.
struct SideLane {
  int device;
  hipStream_t stream;
  hipEvent_t fork, join;
};

std::mutex g_side_mu;
std::vector<SideLane*> g_side_free;

// *out stays NULL (the z ends run on `s`) while `s` is being captured or when
// `s` belongs to another device than the current one.
int side_acquire(hipStream_t s, SideLane** out) {
  *out = nullptr;
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &st) != hipSuccess) {
    (void)hipGetLastError();
    return MW_OK;
  }
  if (st != hipStreamCaptureStatusNone) return MW_OK;
  int dev = 0;
  hipDevice_t sdev = 0;
  if (hipGetDevice(&dev) != hipSuccess || hipStreamGetDevice(s, &sdev) != hipSuccess) {
    (void)hipGetLastError();
    return MW_OK;
  }
  if ((int)sdev != dev) return MW_OK;
  {
    std::lock_guard<std::mutex> lock(g_side_mu);
    for (size_t i = 0; i < g_side_free.size(); ++i)
      if (g_side_free[i]->device == dev) {
        *out = g_side_free[i];
        g_side_free.erase(g_side_free.begin() + (std::ptrdiff_t)i);
        return MW_OK;
      }
  }
  SideLane* lane = new SideLane{dev, nullptr, nullptr, nullptr};
  if (hipStreamCreateWithFlags(&lane->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&lane->fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&lane->join, hipEventDisableTiming) != hipSuccess) {
    const hipError_t err = hipGetLastError();
    if (lane->join) (void)hipEventDestroy(lane->join);
    if (lane->fork) (void)hipEventDestroy(lane->fork);
    if (lane->stream) (void)hipStreamDestroy(lane->stream);
    delete lane;
    return fail(MW_ERR_RUNTIME, "z-end stream: %s", hipGetErrorString(err));
  }
  *out = lane;
  return MW_OK;
}

struct SideGuard {
  SideLane* lane = nullptr;
  ~SideGuard() {
    if (!lane) return;
    std::lock_guard<std::mutex> lock(g_side_mu);
    g_side_free.push_back(lane);
  }
};

// One single-pass step.  With `zs` (its row set to this step's stage row), the
// interior tiles store the staged z-plane samples of their cells, and `zs`'s
// own box is set to those cells (empty when there are no interior tiles) so the
// gather skips them.
// With `side`, the z-end launch runs on its stream between a fork after the term
// values and a join before shell E.  With `pts` (the previous step's point
// observers) and `zg` (the previous step's z-plane gather), the merged interior +
// shell-H launch takes them and `*obs_taken` is set; otherwise the caller
// launches them.
int launch_fused(const mw_program* p, int n, hipStream_t s, mw_fields* out_fields, mw::ZGatherSet* zs = nullptr,
                 SideLane* side = nullptr, const mw::PointSet* pts = nullptr, const mw::ZGatherJob* zg = nullptr,
                 bool* obs_taken = nullptr) {
  if (obs_taken) *obs_taken = false;
  int rc;
  if ((rc = launch_shifted_values(p->h_terms, p->n_h_terms, p->e_terms, p->n_e_terms, n, p->dt, s))) return rc;
  mw::FusedArgs a;
  int nlean = 0;
  build_fused_args(p, n, p->xchunk, &a, &nlean);
  if (zs) {
    for (int q = 0; q < 3; ++q) zs->own_lo[q] = zs->own_hi[q] = 0;
    if (nlean > 0 && zs->count > 0) {
      a.zrow = zs->row;
      a.zcount = zs->count;
      for (int q = 0; q < zs->count; ++q) {
        a.zcomp[q] = zs->comp[q];
        a.zk[q] = zs->k[q];
        a.zoff[q] = zs->off[q];
      }
      zs->own_lo[0] = a.x_off + a.xchunk * a.lx0; zs->own_hi[0] = a.x_off + a.xchunk * a.lx1;
      zs->own_lo[1] = a.y_off + mw::kFRows * a.ly0; zs->own_hi[1] = a.y_off + mw::kFRows * a.ly1;
      zs->own_lo[2] = a.z_off + mw::kFOwnCols * a.lz0; zs->own_hi[2] = a.z_off + mw::kFOwnCols * a.lz1;
    }
  }
  mw::ShellArgs sa;
  build_shell_args(a, nlean, p->shell_chunk, &sa);
  const bool ade = p->has_ade && (p->ade.count[0] + p->ade.count[1] + p->ade.count[2]) > 0;
  const bool cpml = p->cpml.enabled != 0, dft = p->nf_dft > 0;
  const int mode = coef_mode(p->coeffs);
  const dim3 block(mw::kFLanes, mw::kFRows, 1);
  const dim3 sgrid(((sa.total + 7) / 8) * 8);
  const int nze = mw::zedge_tiles(a);
  auto launch_zends = [&](hipStream_t zs_stream) {
    const dim3 zgrid(((nze + 7) / 8) * 8);
    if (ade) { if (dft) dispatch_zedge<true, true>(zgrid, block, zs_stream, a, mode); else dispatch_zedge<true, false>(zgrid, block, zs_stream, a, mode); }
    else { if (dft) dispatch_zedge<false, true>(zgrid, block, zs_stream, a, mode); else dispatch_zedge<false, false>(zgrid, block, zs_stream, a, mode); }
    return check_launch("fused_step (z edges)");
  };
  // the z ends on the side stream: they start with the launch below, and the join
  // is recorded right after them so that every later return path can wait for it
  const bool forked = side && a.zends && nze > 0;
  if (forked) {
    if (hipEventRecord(side->fork, s) != hipSuccess || hipStreamWaitEvent(side->stream, side->fork, 0) != hipSuccess)
      return fail(MW_ERR_RUNTIME, "z-end stream: fork failed");
    const int zrc = launch_zends(side->stream);
    if (hipEventRecord(side->join, side->stream) != hipSuccess) return fail(MW_ERR_RUNTIME, "z-end stream: join failed");
    if (zrc) {
      (void)hipStreamWaitEvent(s, side->join, 0);
      return zrc;
    }
  }
  // shell H and lean read only this step's inputs and write disjoint cells: one
  // launch when both have work
  const bool merged = MW_MERGE_SHELL_H && sa.total > 0 && nlean > 0;
  if (merged) {
    const int sb = ((sa.total + 7) / 8) * 8, lb = ((nlean + 7) / 8) * 8;
    mw::PointSet none{};
    const mw::PointSet& ps = pts ? *pts : none;
    mw::ZGatherJob no_gather{};
    const mw::ZGatherJob& zj = zg ? *zg : no_gather;
    if (cpml) {
      if (ade) { if (dft) dispatch_shell_h_lean<true, true, true>(sb, lb, s, sa, mode, ps, zj); else dispatch_shell_h_lean<true, true, false>(sb, lb, s, sa, mode, ps, zj); }
      else { if (dft) dispatch_shell_h_lean<true, false, true>(sb, lb, s, sa, mode, ps, zj); else dispatch_shell_h_lean<true, false, false>(sb, lb, s, sa, mode, ps, zj); }
    } else {
      if (ade) { if (dft) dispatch_shell_h_lean<false, true, true>(sb, lb, s, sa, mode, ps, zj); else dispatch_shell_h_lean<false, true, false>(sb, lb, s, sa, mode, ps, zj); }
      else { if (dft) dispatch_shell_h_lean<false, false, true>(sb, lb, s, sa, mode, ps, zj); else dispatch_shell_h_lean<false, false, false>(sb, lb, s, sa, mode, ps, zj); }
    }
    if (obs_taken) *obs_taken = true;
    if ((rc = check_launch("fused_step (shell H + lean)"))) {
      if (forked) (void)hipStreamWaitEvent(s, side->join, 0);
      return rc;
    }
  }
  if (sa.total > 0 && !merged) {
    if (cpml) hipLaunchKernelGGL(mw::shell_h_kernel<true>, sgrid, dim3(64, 4, 1), 0, s, sa);
    else hipLaunchKernelGGL(mw::shell_h_kernel<false>, sgrid, dim3(64, 4, 1), 0, s, sa);
    if ((rc = check_launch("fused_step (shell H)"))) {
      if (forked) (void)hipStreamWaitEvent(s, side->join, 0);
      return rc;
    }
  }
  if (nlean > 0 && !merged) {
    dim3 lgrid(((nlean + 7) / 8) * 8);
    if (ade) { if (dft) dispatch_lean<true, true>(lgrid, block, s, a, mode); else dispatch_lean<true, false>(lgrid, block, s, a, mode); }
    else { if (dft) dispatch_lean<false, true>(lgrid, block, s, a, mode); else dispatch_lean<false, false>(lgrid, block, s, a, mode); }
    if ((rc = check_launch("fused_step (lean)"))) {
      if (forked) (void)hipStreamWaitEvent(s, side->join, 0);
      return rc;
    }
  }
  if (forked) {
    // shell E reads the z ends' H^{n+1/2}
    if (hipStreamWaitEvent(s, side->join, 0) != hipSuccess) return fail(MW_ERR_RUNTIME, "z-end stream: join failed");
  } else if (nze > 0) {
    if ((rc = launch_zends(s))) return rc;
  }
  // shell E reads the complete H^{n+1/2}
  if (sa.total > 0) {
    if (cpml) {
      if (ade) rc = dft ? launch_shell_e<true, true, true>(sgrid, s, sa) : launch_shell_e<true, true, false>(sgrid, s, sa);
      else rc = dft ? launch_shell_e<true, false, true>(sgrid, s, sa) : launch_shell_e<true, false, false>(sgrid, s, sa);
    } else {
      if (ade) rc = dft ? launch_shell_e<false, true, true>(sgrid, s, sa) : launch_shell_e<false, true, false>(sgrid, s, sa);
      else rc = dft ? launch_shell_e<false, false, true>(sgrid, s, sa) : launch_shell_e<false, false, false>(sgrid, s, sa);
    }
    if (rc) return rc;
  }
  *out_fields = a.out;
  return MW_OK;
}

// The single-pass step reads H across a periodic seam from the face on the other
// side, so a periodic axis must have both faces periodic.
int check_periodic_pairs(const mw_grid* g) {
  for (int a = 0; a < 3; ++a)
    if ((g->bc_lo[a] == MW_BC_PERIODIC) != (g->bc_hi[a] == MW_BC_PERIODIC))
      return fail(MW_ERR_ARGUMENT, "axis %d is periodic on one face only", a);
  return MW_OK;
}

// state of step n (fused mode double-buffers E/H by step parity)
const mw_fields& state_fields(const mw_program* p, int n) { return (p->fused && (n & 1)) ? p->fields_alt : p->fields; }

mw::WindowArgs window_args(const mw_window* w) {
  mw::WindowArgs a;
  for (int x = 0; x < 3; ++x) { a.lo[x] = w->lo[x]; a.hi[x] = w->hi[x]; }
  return a;
}

int64_t window_cells(const mw_window* w, int halo) {
  return (int64_t)(w->hi[0] - w->lo[0] + halo) * (w->hi[1] - w->lo[1] + halo) * (w->hi[2] - w->lo[2] + halo);
}

int check_window(const mw_program* p, const mw_window* w) {
  const mw_grid* g = &p->grid;
  const int n[3] = {g->nx, g->ny, g->nz};
  for (int x = 0; x < 3; ++x) {
    if (w->lo[x] < 1 || w->hi[x] > n[x] - 1 || w->hi[x] <= w->lo[x])
      return fail(MW_ERR_ARGUMENT, "window axis %d: [%d, %d) must lie in [1, %d)", x, w->lo[x], w->hi[x], n[x] - 1);
  }
  if (p->cpml.enabled) {
    // E-side psi q=0..5 differentiate along axes y z x z x y
    static const int axis_of[6] = {1, 2, 0, 2, 0, 1};
    for (int q = 0; q < 6; ++q) {
      const int ax = axis_of[q];
      const int low = p->cpml.psi_low[q], hs = p->cpml.psi_high_start[q], hl = p->cpml.psi_high_len[q];
      if (w->lo[ax] < low || (hl > 0 && w->hi[ax] > hs))
        return fail(MW_ERR_ARGUMENT, "window axis %d [%d, %d) overlaps the CPML layers", ax, w->lo[ax], w->hi[ax]);
    }
  }
  if (w->capacity <= 0 || !w->e_store || !w->h_store) return fail(MW_ERR_ARGUMENT, "window stores missing");
  return MW_OK;
}

template <bool ADE, bool DFT>
void launch_batch_zedge(dim3 grid, hipStream_t s, const mw::FusedArgs* args, const int* prefix, int count, int total,
                        const int* counter, int mode) {
  const dim3 block(mw::kFLanes, mw::kFRows, 1);
  if (mode == 2) hipLaunchKernelGGL((mw::zedge_batch_kernel<ADE, DFT, 2>), grid, block, 0, s, args, prefix, count, total, counter);
  else if (mode == 1) hipLaunchKernelGGL((mw::zedge_batch_kernel<ADE, DFT, 1>), grid, block, 0, s, args, prefix, count, total, counter);
  else hipLaunchKernelGGL((mw::zedge_batch_kernel<ADE, DFT, 0>), grid, block, 0, s, args, prefix, count, total, counter);
}

template <bool ADE, bool DFT>
void launch_batch_lean(dim3 grid, hipStream_t s, const mw::FusedArgs* args, const int* prefix, int count, int total,
                       const int* counter, int mode) {
  const dim3 block(mw::kFLanes, mw::kFRows, 1);
  if (mode == 2) hipLaunchKernelGGL((mw::lean_batch_kernel<ADE, DFT, 2>), grid, block, 0, s, args, prefix, count, total, counter);
  else if (mode == 1) hipLaunchKernelGGL((mw::lean_batch_kernel<ADE, DFT, 1>), grid, block, 0, s, args, prefix, count, total, counter);
  else hipLaunchKernelGGL((mw::lean_batch_kernel<ADE, DFT, 0>), grid, block, 0, s, args, prefix, count, total, counter);
}

template <bool CPML, bool ADE, bool DFT>
void launch_batch_shell_e(dim3 grid, hipStream_t s, const mw::ShellArgs* args, const int* prefix, int count, int total,
                          const int* counter, bool periodic) {
  if (periodic)
    hipLaunchKernelGGL((mw::shell_e_batch_kernel<CPML, ADE, DFT, true>), grid, dim3(64, 4, 1), 0, s, args, prefix, count,
                       total, counter);
  else
    hipLaunchKernelGGL((mw::shell_e_batch_kernel<CPML, ADE, DFT, false>), grid, dim3(64, 4, 1), 0, s, args, prefix,
                       count, total, counter);
}

}  // namespace

struct mw_batch {
  int count = 0;
  int total_steps = 0;
  bool cpml = false, ade = false, dft = false;
  bool periodic = false;  // some scene has a periodic face
  int mode = 0;
  mw::FusedArgs* d_args = nullptr;   // 2 per scene (step parity): lean interior
  mw::ShellArgs* d_shell = nullptr;  // 2 per scene: PML shell
  int* d_shell_prefix = nullptr;
  int* d_lean_prefix = nullptr;
  int* d_zedge_prefix = nullptr;
  int total_shell = 0, total_lean = 0, total_zedge = 0;
  int* d_counter = nullptr;
  mw::BatchTerm* d_terms = nullptr;
  int n_terms = 0;
  int64_t max_term_cells = 0;
  mw_aux_line* d_aux = nullptr;
  int n_aux = 0, max_aux = 0;
  mw::BatchPlane* d_planes = nullptr;
  int n_planes = 0;
  int64_t max_plane_cells = 0;
  mw::BatchPoint* d_points = nullptr;
  int n_points = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int graph_steps = 0;
};

namespace {

void batch_free(mw_batch* b) {
  if (!b) return;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  void* bufs[] = {b->d_args, b->d_shell, b->d_shell_prefix, b->d_lean_prefix, b->d_zedge_prefix, b->d_counter,
                  b->d_terms, b->d_aux, b->d_planes, b->d_points};
  for (void* q : bufs)
    if (q) (void)hipFree(q);
  if (b->ev_in) (void)hipEventDestroy(b->ev_in);
  if (b->ev_out) (void)hipEventDestroy(b->ev_out);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

template <typename T>
int upload(const std::vector<T>& host, T** dst, const char* what) {
  if (host.empty()) { *dst = nullptr; return MW_OK; }
  if (hipMalloc((void**)dst, host.size() * sizeof(T)) != hipSuccess)
    return fail(MW_ERR_RUNTIME, "batch: cannot allocate %s", what);
  if (hipMemcpy(*dst, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
    return fail(MW_ERR_RUNTIME, "batch: cannot upload %s", what);
  return MW_OK;
}

// One step of every scene, reading the step index from the device counter
// (observers off for the adjoint's replay: the forward sums are already taken).
template <bool CPML, bool ADE, bool DFT>
void launch_batch_shell_h_lean(const mw_batch* b, hipStream_t s) {
  const int lb = ((b->total_lean + 7) / 8) * 8, sb = ((b->total_shell + 7) / 8) * 8;
  const dim3 grid(lb + sb), block(mw::kFLanes, mw::kFRows, 1);
  if (b->mode == 2)
    hipLaunchKernelGGL((mw::shell_h_lean_batch_kernel<CPML, ADE, DFT, 2>), grid, block, 0, s, b->d_args, b->d_shell,
                       b->d_lean_prefix, b->d_shell_prefix, b->count, b->total_lean, b->total_shell, lb, sb, b->d_counter);
  else if (b->mode == 1)
    hipLaunchKernelGGL((mw::shell_h_lean_batch_kernel<CPML, ADE, DFT, 1>), grid, block, 0, s, b->d_args, b->d_shell,
                       b->d_lean_prefix, b->d_shell_prefix, b->count, b->total_lean, b->total_shell, lb, sb, b->d_counter);
  else
    hipLaunchKernelGGL((mw::shell_h_lean_batch_kernel<CPML, ADE, DFT, 0>), grid, block, 0, s, b->d_args, b->d_shell,
                       b->d_lean_prefix, b->d_shell_prefix, b->count, b->total_lean, b->total_shell, lb, sb, b->d_counter);
}

int batch_step(const mw_batch* b, hipStream_t s, bool observe = true) {
  int rc;
  if (b->n_aux > 0) {
    hipLaunchKernelGGL(mw::aux_line_h_batch_kernel, dim3((b->max_aux + 255) / 256, b->n_aux), dim3(256), 0, s,
                       b->d_aux, b->n_aux);
    if ((rc = check_launch("batch aux_line_h"))) return rc;
  }
  if (b->n_terms > 0) {
    hipLaunchKernelGGL(mw::shifted_values_batch_kernel, dim3(grid_blocks(b->max_term_cells), b->n_terms), dim3(256),
                       0, s, b->d_terms, b->n_terms, b->d_counter);
    if ((rc = check_launch("batch shifted_values"))) return rc;
  }
  // shell H and lean read only the step's inputs; shell E reads the complete H
  const dim3 sgrid(((b->total_shell + 7) / 8) * 8);
  const bool merged = MW_MERGE_SHELL_H && b->total_shell > 0 && b->total_lean > 0;
  if (merged) {
    if (b->cpml) {
      if (b->ade) { if (b->dft) launch_batch_shell_h_lean<true, true, true>(b, s); else launch_batch_shell_h_lean<true, true, false>(b, s); }
      else { if (b->dft) launch_batch_shell_h_lean<true, false, true>(b, s); else launch_batch_shell_h_lean<true, false, false>(b, s); }
    } else {
      if (b->ade) { if (b->dft) launch_batch_shell_h_lean<false, true, true>(b, s); else launch_batch_shell_h_lean<false, true, false>(b, s); }
      else { if (b->dft) launch_batch_shell_h_lean<false, false, true>(b, s); else launch_batch_shell_h_lean<false, false, false>(b, s); }
    }
    if ((rc = check_launch("batch fused_step (shell H + lean)"))) return rc;
  }
  if (b->total_shell > 0 && !merged) {
    if (b->cpml) hipLaunchKernelGGL(mw::shell_h_batch_kernel<true>, sgrid, dim3(64, 4, 1), 0, s, b->d_shell,
                                    b->d_shell_prefix, b->count, b->total_shell, b->d_counter);
    else hipLaunchKernelGGL(mw::shell_h_batch_kernel<false>, sgrid, dim3(64, 4, 1), 0, s, b->d_shell,
                            b->d_shell_prefix, b->count, b->total_shell, b->d_counter);
    if ((rc = check_launch("batch fused_step (shell H)"))) return rc;
  }
  if (b->total_lean > 0 && !merged) {
    const dim3 lgrid(((b->total_lean + 7) / 8) * 8);
    const int* lp = b->d_lean_prefix;
    if (b->ade) {
      if (b->dft) launch_batch_lean<true, true>(lgrid, s, b->d_args, lp, b->count, b->total_lean, b->d_counter, b->mode);
      else launch_batch_lean<true, false>(lgrid, s, b->d_args, lp, b->count, b->total_lean, b->d_counter, b->mode);
    } else {
      if (b->dft) launch_batch_lean<false, true>(lgrid, s, b->d_args, lp, b->count, b->total_lean, b->d_counter, b->mode);
      else launch_batch_lean<false, false>(lgrid, s, b->d_args, lp, b->count, b->total_lean, b->d_counter, b->mode);
    }
    if ((rc = check_launch("batch fused_step (lean)"))) return rc;
  }
  if (b->total_zedge > 0) {
    const dim3 zgrid(((b->total_zedge + 7) / 8) * 8);
    const int* zp = b->d_zedge_prefix;
    if (b->ade) {
      if (b->dft) launch_batch_zedge<true, true>(zgrid, s, b->d_args, zp, b->count, b->total_zedge, b->d_counter, b->mode);
      else launch_batch_zedge<true, false>(zgrid, s, b->d_args, zp, b->count, b->total_zedge, b->d_counter, b->mode);
    } else {
      if (b->dft) launch_batch_zedge<false, true>(zgrid, s, b->d_args, zp, b->count, b->total_zedge, b->d_counter, b->mode);
      else launch_batch_zedge<false, false>(zgrid, s, b->d_args, zp, b->count, b->total_zedge, b->d_counter, b->mode);
    }
    if ((rc = check_launch("batch fused_step (z edges)"))) return rc;
  }
  if (b->total_shell > 0) {
    const mw::ShellArgs* sa = b->d_shell;
    const int* sp = b->d_shell_prefix;
    const int n = b->count, t = b->total_shell;
    const int* ctr = b->d_counter;
    if (b->cpml) {
      if (b->ade) { if (b->dft) launch_batch_shell_e<true, true, true>(sgrid, s, sa, sp, n, t, ctr, b->periodic); else launch_batch_shell_e<true, true, false>(sgrid, s, sa, sp, n, t, ctr, b->periodic); }
      else { if (b->dft) launch_batch_shell_e<true, false, true>(sgrid, s, sa, sp, n, t, ctr, b->periodic); else launch_batch_shell_e<true, false, false>(sgrid, s, sa, sp, n, t, ctr, b->periodic); }
    } else {
      if (b->ade) { if (b->dft) launch_batch_shell_e<false, true, true>(sgrid, s, sa, sp, n, t, ctr, b->periodic); else launch_batch_shell_e<false, true, false>(sgrid, s, sa, sp, n, t, ctr, b->periodic); }
      else { if (b->dft) launch_batch_shell_e<false, false, true>(sgrid, s, sa, sp, n, t, ctr, b->periodic); else launch_batch_shell_e<false, false, false>(sgrid, s, sa, sp, n, t, ctr, b->periodic); }
    }
    if ((rc = check_launch("batch fused_step (shell E)"))) return rc;
  }
  if (b->n_aux > 0) {
    hipLaunchKernelGGL(mw::aux_line_e_batch_kernel, dim3((b->max_aux + 255) / 256, b->n_aux), dim3(256), 0, s,
                       b->d_aux, b->n_aux, b->d_counter);
    if ((rc = check_launch("batch aux_line_e"))) return rc;
  }
  if (observe && b->n_planes > 0) {
    hipLaunchKernelGGL(mw::accumulate_planes_batch_kernel, dim3(grid_blocks(b->max_plane_cells), b->n_planes),
                       dim3(256), 0, s, b->d_planes, b->n_planes, b->d_counter);
    if ((rc = check_launch("batch accumulate_planes"))) return rc;
  }
  if (observe && b->n_points > 0) {
    hipLaunchKernelGGL(mw::accumulate_points_batch_kernel, dim3((b->n_points + 63) / 64), dim3(64), 0, s,
                       b->d_points, b->n_points, b->d_counter);
    if ((rc = check_launch("batch accumulate_points"))) return rc;
  }
  hipLaunchKernelGGL(mw::step_counter_kernel, dim3(1), dim3(64), 0, s, b->d_counter, 0, 0);
  return check_launch("batch step_counter");
}

// Per-scene descriptor tables of a batch, gathered into a sink: VectorSink
// (host vectors, uploaded to buffers the batch handle owns) or the fixed-capacity
// sinks of the no-allocation entry point.  Fills the handle's counts and flags.
struct VectorSink {
  std::vector<mw::FusedArgs> args;
  std::vector<mw::ShellArgs> shells;
  std::vector<int> shell_prefix{0}, lean_prefix{0}, zedge_prefix{0};
  std::vector<mw::BatchTerm> terms;
  std::vector<mw_aux_line> aux;
  std::vector<mw::BatchPlane> planes;
  std::vector<mw::BatchPoint> points;
  bool arg(const mw::FusedArgs& a) { args.push_back(a); return true; }
  bool shell(const mw::ShellArgs& a) { shells.push_back(a); return true; }
  bool prefixes(int shell_blocks, int lean_tiles, int zedge_tiles) {
    shell_prefix.push_back(shell_prefix.back() + shell_blocks);
    lean_prefix.push_back(lean_prefix.back() + lean_tiles);
    zedge_prefix.push_back(zedge_prefix.back() + zedge_tiles);
    return true;
  }
  bool term(const mw::BatchTerm& t) { terms.push_back(t); return true; }
  bool line(const mw_aux_line& l) { aux.push_back(l); return true; }
  bool plane(const mw::BatchPlane& q) { planes.push_back(q); return true; }
  bool point(const mw::BatchPoint& q) { points.push_back(q); return true; }
};

template <class Sink>
int collect_batch(const mw_program* const* programs, int count, int xchunk, Sink& sink, mw_batch* b) {
  if (xchunk < 0 || xchunk > 256) return fail(MW_ERR_ARGUMENT, "batch: xchunk %d out of range", xchunk);
  int rc;
  const char* full = "batch: the scenes' tables exceed the no-allocation scratch (MW_NOALLOC_SCRATCH_BYTES)";
  int shell_total = 0, lean_total = 0, zedge_total = 0;
  b->count = count;
  for (int si = 0; si < count; ++si) {
    const mw_program* p = programs[si];
    if (!p) return fail(MW_ERR_ARGUMENT, "batch: program %d is NULL", si);
    const mw_grid* g = &p->grid;
    if ((rc = check_grid(g)) || (rc = check_fields(&p->fields, true, true)) ||
        (rc = check_fields(&p->fields_alt, true, true)) || (rc = check_cpml(g, &p->cpml, true)) ||
        (rc = check_cpml(g, &p->cpml, false)) || (rc = check_ade(g, p->has_ade ? &p->ade : nullptr)) ||
        (rc = check_terms(g, p->h_terms, p->n_h_terms, "h_terms", p->total_steps)) ||
        (rc = check_terms(g, p->e_terms, p->n_e_terms, "e_terms", p->total_steps)))
      return rc;
    if (!p->fused) return fail(MW_ERR_ARGUMENT, "batch: scene %d is not a single-pass program", si);
    if ((rc = check_periodic_pairs(g))) return rc;
    if ((p->n_h_terms > 0 && !p->h_terms_dev) || (p->n_e_terms > 0 && !p->e_terms_dev))
      return fail(MW_ERR_ARGUMENT, "batch: scene %d has no device term descriptors", si);
    const bool cpml = p->cpml.enabled != 0;
    const bool ade = p->has_ade && (p->ade.count[0] + p->ade.count[1] + p->ade.count[2]) > 0;
    const bool dft = p->nf_dft > 0;
    const int mode = coef_mode(p->coeffs);
    if (si == 0) {
      b->cpml = cpml; b->ade = ade; b->dft = dft; b->mode = mode; b->total_steps = p->total_steps;
      b->periodic = false;
    } else if (cpml != b->cpml || ade != b->ade || dft != b->dft) {
      return fail(MW_ERR_ARGUMENT, "batch: scene %d differs in CPML/ADE/full-field-DFT use", si);
    }
    if (mode != b->mode) b->mode = 0;
    if (grid_periodic(*g)) b->periodic = true;
    b->total_steps = std::min(b->total_steps, p->total_steps);
    for (int par = 0; par < 2; ++par) {
      mw::FusedArgs a;
      int nlean = 0;
      // the batch launches have no z-end tiles: z_tiling 3 and 4 run as 2 there
      mw_program q = *p;
      if (q.z_tiling >= 3) q.z_tiling = 2;
      build_fused_args(&q, par, xchunk > 0 ? xchunk : p->xchunk, &a, &nlean);
      a.dft_cos = p->dft_cos;  // the batch kernels add step * nf
      a.dft_sin = p->dft_sin;
      mw::ShellArgs sh;
      build_shell_args(a, nlean, p->shell_chunk, &sh);
      if (!sink.arg(a) || !sink.shell(sh)) return fail(MW_ERR_ARGUMENT, "%s", full);
      if (par == 0) {
        const int zt = mw::zedge_tiles(a);
        if (!sink.prefixes(sh.total, nlean, zt)) return fail(MW_ERR_ARGUMENT, "%s", full);
        shell_total += sh.total;
        lean_total += nlean;
        zedge_total += zt;
      }
    }
    for (int f = 0; f < 2; ++f) {
      const mw_term* t = f ? p->e_terms : p->h_terms;
      for (int i = 0; i < (f ? p->n_e_terms : p->n_h_terms); ++i) {
        if (t[i].kind != MW_TERM_SHIFTED) continue;
        if (!t[i].values) return fail(MW_ERR_ARGUMENT, "batch: scene %d term %d has no value buffer", si, i);
        if (!sink.term(mw::BatchTerm{t[i], p->dt})) return fail(MW_ERR_ARGUMENT, "%s", full);
        b->n_terms++;
        b->max_term_cells = std::max<int64_t>(b->max_term_cells, (int64_t)t[i].shape[0] * t[i].shape[1] * t[i].shape[2]);
      }
    }
    if (p->aux.n_e > 0) {
      const mw_aux_line& l = p->aux;
      if (!l.e || !l.h || !l.decay_e || !l.curl_e || !l.decay_h || !l.curl_h || !l.source)
        return fail(MW_ERR_ARGUMENT, "batch: scene %d auxiliary line arrays missing", si);
      if (!sink.line(l)) return fail(MW_ERR_ARGUMENT, "%s", full);
      b->n_aux++;
      b->max_aux = std::max(b->max_aux, l.n_e);
    }
    if (p->n_planes < 0 || p->n_planes > MW_MAX_PLANES) return fail(MW_ERR_ARGUMENT, "batch: scene %d plane count", si);
    if ((p->n_planes > 0 || p->n_points > 0) &&
        (p->nf_obs <= 0 || !p->obs_e_cos || !p->obs_e_sin || !p->obs_h_cos || !p->obs_h_sin))
      return fail(MW_ERR_ARGUMENT, "batch: scene %d observer weight tables missing", si);
    auto comp_ptr = [](const mw_fields& f, int c) -> const float* { return c < 3 ? f.e[c] : f.h[c - 3]; };
    for (int q = 0; q < p->n_planes; ++q) {
      const mw_plane_obs& o = p->planes[q];
      if (o.component < 0 || o.component > 5 || o.axis < 0 || o.axis > 2 || !o.re || !o.im)
        return fail(MW_ERR_ARGUMENT, "batch: scene %d plane %d descriptor invalid", si, q);
      mw::BatchPlane bp{};
      bp.component = o.component; bp.axis = o.axis; bp.index = o.index; bp.nf = p->nf_obs;
      component_shape(g, o.component, bp.shape);
      if (o.index < 0 || o.index >= bp.shape[o.axis])
        return fail(MW_ERR_ARGUMENT, "batch: scene %d plane %d index outside axis", si, q);
      bp.field[0] = comp_ptr(p->fields_alt, o.component);  // state after an even step
      bp.field[1] = comp_ptr(p->fields, o.component);
      bp.re = o.re; bp.im = o.im;
      bp.w_cos = o.component < 3 ? p->obs_e_cos : p->obs_h_cos;
      bp.w_sin = o.component < 3 ? p->obs_e_sin : p->obs_h_sin;
      const int* sh = bp.shape;
      const int64_t cells = o.axis == 0 ? (int64_t)sh[1] * sh[2] : (o.axis == 1 ? (int64_t)sh[0] * sh[2] : (int64_t)sh[0] * sh[1]);
      b->max_plane_cells = std::max(b->max_plane_cells, cells);
      if (!sink.plane(bp)) return fail(MW_ERR_ARGUMENT, "%s", full);
      b->n_planes++;
    }
    if (p->n_points > 0) {
      if (!p->point_component || !p->point_linear || !p->point_re || !p->point_im)
        return fail(MW_ERR_ARGUMENT, "batch: scene %d point observer arrays missing", si);
      for (int q = 0; q < p->n_points; ++q) {
        int32_t comp = -1, lin = -1;
        if (hipMemcpy(&comp, p->point_component + q, sizeof(int32_t), hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(&lin, p->point_linear + q, sizeof(int32_t), hipMemcpyDeviceToHost) != hipSuccess)
          return fail(MW_ERR_RUNTIME, "batch: cannot read scene %d point tables", si);
        if (comp < 0 || comp > 5) return fail(MW_ERR_ARGUMENT, "batch: scene %d point %d component", si, q);
        int sh[3];
        component_shape(g, comp, sh);
        if (lin < 0 || (int64_t)lin >= (int64_t)sh[0] * sh[1] * sh[2])
          return fail(MW_ERR_ARGUMENT, "batch: scene %d point %d index outside the field", si, q);
        mw::BatchPoint bq{};
        bq.linear = lin; bq.nf = p->nf_obs; bq.stride = p->n_points;
        bq.field[0] = comp_ptr(p->fields_alt, comp);
        bq.field[1] = comp_ptr(p->fields, comp);
        bq.re = p->point_re + q; bq.im = p->point_im + q;
        bq.w_cos = comp < 3 ? p->obs_e_cos : p->obs_h_cos;
        bq.w_sin = comp < 3 ? p->obs_e_sin : p->obs_h_sin;
        if (!sink.point(bq)) return fail(MW_ERR_ARGUMENT, "%s", full);
        b->n_points++;
      }
    }
  }
  b->total_shell = shell_total;
  b->total_lean = lean_total;
  b->total_zedge = zedge_total;
  if (b->n_terms > 65535 || b->n_aux > 65535 || b->n_planes > 65535)
    return fail(MW_ERR_ARGUMENT, "batch: too many terms, lines or planes for one launch");
  return MW_OK;
}

}  // namespace

extern "C" {

const char* mw_last_error(void) { return g_error.c_str(); }

int32_t mw_abi_version(void) { return 4; }

int32_t mw_abi_struct_sizes(int64_t* out, int32_t capacity) {
  const int64_t sizes[] = {(int64_t)sizeof(mw_grid),     (int64_t)sizeof(mw_coeffs),   (int64_t)sizeof(mw_cpml),
                           (int64_t)sizeof(mw_fields),   (int64_t)sizeof(mw_pole),     (int64_t)sizeof(mw_ade),
                           (int64_t)sizeof(mw_term),     (int64_t)sizeof(mw_plane_obs), (int64_t)sizeof(mw_program),
                           (int64_t)sizeof(mw_window),   (int64_t)sizeof(mw_adjoint)};
  const int32_t n = (int32_t)(sizeof(sizes) / sizeof(sizes[0]));
  for (int32_t i = 0; i < n && i < capacity; ++i) out[i] = sizes[i];
  return n;
}

int mw_update_magnetic(const mw_grid* grid, const mw_coeffs* coeffs, const mw_cpml* cpml, const mw_fields* fields,
                       int32_t comp_mask, void* stream) {
  int rc;
  if ((rc = check_grid(grid)) || (rc = check_fields(fields, false, true)) || (rc = check_cpml(grid, cpml, false))) return rc;
  if (!coeffs) return fail(MW_ERR_ARGUMENT, "coeffs is NULL");
  return launch_h(grid, coeffs, cpml, fields, comp_mask & 7, (hipStream_t)stream);
}

int mw_update_electric(const mw_grid* grid, const mw_coeffs* coeffs, const mw_cpml* cpml, const mw_ade* ade,
                       const mw_fields* fields, int32_t comp_mask, void* stream) {
  int rc;
  if ((rc = check_grid(grid)) || (rc = check_fields(fields, true, false)) || (rc = check_cpml(grid, cpml, true)) ||
      (rc = check_ade(grid, ade)))
    return rc;
  if (!coeffs) return fail(MW_ERR_ARGUMENT, "coeffs is NULL");
  return launch_e(grid, coeffs, cpml, ade, fields, comp_mask & 7, (hipStream_t)stream);
}

int mw_add_source_patch(float* field, int32_t fnx, int32_t fny, int32_t fnz, const float* patch, int32_t pnx,
                        int32_t pny, int32_t pnz, int32_t oi, int32_t oj, int32_t ok, float signal, void* stream) {
  if (!field || !patch) return fail(MW_ERR_ARGUMENT, "add_source_patch: NULL pointer");
  if (pnx <= 0 || pny <= 0 || pnz <= 0 || oi < 0 || oj < 0 || ok < 0 || oi + pnx > fnx || oj + pny > fny ||
      ok + pnz > fnz)
    return fail(MW_ERR_ARGUMENT, "add_source_patch: patch (%d,%d,%d)+(%d,%d,%d) outside field (%d,%d,%d)", oi, oj, ok,
                pnx, pny, pnz, fnx, fny, fnz);
  const int64_t cells = (int64_t)pnx * pny * pnz;
  hipLaunchKernelGGL(mw::add_patch_kernel, dim3(grid_blocks(cells)), dim3(256), 0, (hipStream_t)stream, field, fny, fnz,
                     patch, pnx, pny, pnz, oi, oj, ok, signal);
  return check_launch("add_source_patch");
}

int mw_add_time_shifted_patch(float* field, int32_t fnx, int32_t fny, int32_t fnz, const float* patch,
                              const double* delay, int32_t pnx, int32_t pny, int32_t pnz, int32_t oi, int32_t oj,
                              int32_t ok, double time, int32_t kind, double frequency, double fwidth,
                              double amplitude, double phase, double delay0, void* stream) {
  if (!field || !patch || !delay) return fail(MW_ERR_ARGUMENT, "add_time_shifted_patch: NULL pointer");
  if (pnx <= 0 || pny <= 0 || pnz <= 0 || oi < 0 || oj < 0 || ok < 0 || oi + pnx > fnx || oj + pny > fny ||
      ok + pnz > fnz)
    return fail(MW_ERR_ARGUMENT, "add_time_shifted_patch: patch outside field");
  if (kind < 0 || kind > 2) return fail(MW_ERR_ARGUMENT, "add_time_shifted_patch: waveform kind %d", kind);
  const int64_t cells = (int64_t)pnx * pny * pnz;
  hipLaunchKernelGGL(mw::add_shifted_kernel, dim3(grid_blocks(cells)), dim3(256), 0, (hipStream_t)stream, field, fny,
                     fnz, patch, delay, pnx, pny, pnz, oi, oj, ok, time, kind, frequency, fwidth, amplitude, phase,
                     delay0);
  return check_launch("add_time_shifted_patch");
}

int mw_accumulate_dft(const float* field, int64_t count, int32_t nf, const float* wcos, const float* wsin, float* re,
                      float* im, void* stream) {
  if (!field || !wcos || !wsin || !re || !im) return fail(MW_ERR_ARGUMENT, "accumulate_dft: NULL pointer");
  if (count <= 0 || count > kMaxElems || nf <= 0) return fail(MW_ERR_ARGUMENT, "accumulate_dft: bad sizes");
  hipLaunchKernelGGL(mw::accumulate_kernel, dim3(grid_blocks(count)), dim3(256), 0, (hipStream_t)stream, field, count,
                     nf, wcos, wsin, re, im);
  return check_launch("accumulate_dft");
}

int mw_accumulate_plane(const float* field, int32_t nx, int32_t ny, int32_t nz, int32_t axis, int32_t index,
                        int32_t nf, const float* wcos, const float* wsin, float* re, float* im, void* stream) {
  if (!field || !wcos || !wsin || !re || !im) return fail(MW_ERR_ARGUMENT, "accumulate_plane: NULL pointer");
  if (axis < 0 || axis > 2 || nf <= 0) return fail(MW_ERR_ARGUMENT, "accumulate_plane: bad axis/nf");
  const int size = axis == 0 ? nx : (axis == 1 ? ny : nz);
  if (index < 0 || index >= size) return fail(MW_ERR_ARGUMENT, "accumulate_plane: index %d outside [0,%d)", index, size);
  const int64_t n = axis == 0 ? (int64_t)ny * nz : (axis == 1 ? (int64_t)nx * nz : (int64_t)nx * ny);
  hipLaunchKernelGGL(mw::accumulate_plane_single_kernel, dim3(grid_blocks(n)), dim3(256), 0, (hipStream_t)stream, field,
                     nx, ny, nz, axis, index, nf, wcos, wsin, re, im);
  return check_launch("accumulate_plane");
}

int mw_accumulate_points(const float* field, const int32_t* linear_index, int32_t count, int32_t nf, const float* wcos,
                         const float* wsin, float* re, float* im, void* stream) {
  if (!field || !linear_index || !wcos || !wsin || !re || !im) return fail(MW_ERR_ARGUMENT, "accumulate_points: NULL pointer");
  if (count <= 0 || nf <= 0) return fail(MW_ERR_ARGUMENT, "accumulate_points: bad sizes");
  hipLaunchKernelGGL(mw::accumulate_points_single_kernel, dim3((count + 63) / 64), dim3(64), 0, (hipStream_t)stream,
                     field, linear_index, count, nf, wcos, wsin, re, im);
  return check_launch("accumulate_points");
}

int mw_update_pole_current(int32_t kind, const float* e, const float* drive, float* polarization, float* current,
                           int64_t count, float decay, float restoring, float dt, void* stream) {
  if (!e || !drive || !current || (kind != 1 && !polarization))
    return fail(MW_ERR_ARGUMENT, "update_pole_current: NULL pointer");
  if (kind < 0 || kind > 2) return fail(MW_ERR_ARGUMENT, "update_pole_current: pole kind %d", kind);
  if (count <= 0 || count > kMaxElems) return fail(MW_ERR_ARGUMENT, "update_pole_current: bad sample count");
  if (kind != 1 && !(dt > 0.0f)) return fail(MW_ERR_ARGUMENT, "update_pole_current: dt must be > 0");
  hipLaunchKernelGGL(mw::pole_current_kernel, dim3(grid_blocks(count)), dim3(256), 0, (hipStream_t)stream, kind, e,
                     drive, polarization, current, count, decay, restoring, dt);
  return check_launch("update_pole_current");
}

int mw_apply_polarization_current(float* e, const float* current, const float* inv_eps, int64_t count, float dt,
                                  void* stream) {
  if (!e || !current || !inv_eps) return fail(MW_ERR_ARGUMENT, "apply_polarization_current: NULL pointer");
  if (count <= 0 || count > kMaxElems) return fail(MW_ERR_ARGUMENT, "apply_polarization_current: bad sample count");
  hipLaunchKernelGGL(mw::apply_polarization_kernel, dim3(grid_blocks(count)), dim3(256), 0, (hipStream_t)stream, e,
                     current, inv_eps, count, dt);
  return check_launch("apply_polarization_current");
}

int mw_update_aux_magnetic(float* h, const float* e, const float* decay, const float* curl, int32_t n_h,
                           void* stream) {
  if (!h || !e || !decay || !curl) return fail(MW_ERR_ARGUMENT, "update_aux_magnetic: NULL pointer");
  if (n_h <= 0) return fail(MW_ERR_ARGUMENT, "update_aux_magnetic: %d samples", n_h);
  hipLaunchKernelGGL(mw::aux_magnetic_kernel, dim3(grid_blocks(n_h)), dim3(256), 0, (hipStream_t)stream, h, e, decay,
                     curl, n_h);
  return check_launch("update_aux_magnetic");
}

int mw_update_aux_electric(float* e, const float* h, const float* decay, const float* curl, int32_t n_e,
                           int32_t source_index, float source_value, void* stream) {
  if (!e || !h || !decay || !curl) return fail(MW_ERR_ARGUMENT, "update_aux_electric: NULL pointer");
  if (n_e < 2) return fail(MW_ERR_ARGUMENT, "update_aux_electric: %d samples", n_e);
  hipLaunchKernelGGL(mw::aux_electric_kernel, dim3(grid_blocks(n_e)), dim3(256), 0, (hipStream_t)stream, e, h, decay,
                     curl, n_e, source_index, source_value);
  return check_launch("update_aux_electric");
}

int mw_run_steps(const mw_program* p, int32_t first, int32_t count, void* stream) {
  if (!p) return fail(MW_ERR_ARGUMENT, "program is NULL");
  int rc;
  const mw_grid* g = &p->grid;
  if ((rc = check_grid(g)) || (rc = check_fields(&p->fields, true, true)) || (rc = check_cpml(g, &p->cpml, true)) ||
      (rc = check_cpml(g, &p->cpml, false)) || (rc = check_ade(g, p->has_ade ? &p->ade : nullptr)))
    return rc;
  if (first < 0 || count < 0 || first + count > p->total_steps)
    return fail(MW_ERR_ARGUMENT, "steps [%d, %d) outside the program's %d steps", first, first + count, p->total_steps);
  if (p->xchunk < 0 || p->xchunk > 256 || p->shell_chunk < 0 || p->shell_chunk > 256)
    return fail(MW_ERR_ARGUMENT, "xchunk %d / shell_chunk %d outside [0, 256]", p->xchunk, p->shell_chunk);
  if (p->z_tiling < 0 || p->z_tiling > 4) return fail(MW_ERR_ARGUMENT, "z_tiling %d outside [0, 4]", p->z_tiling);
  if ((rc = check_terms(g, p->h_terms, p->n_h_terms, "h_terms", p->total_steps)) ||
      (rc = check_terms(g, p->e_terms, p->n_e_terms, "e_terms", p->total_steps)))
    return rc;
  if (p->n_planes < 0 || p->n_planes > MW_MAX_PLANES) return fail(MW_ERR_ARGUMENT, "%d planes (max %d)", p->n_planes, MW_MAX_PLANES);
  // planes with axis 2 are staged when the program gives a stage; the others are
  // added to their sums after each step
  const bool staging = p->plane_stage && p->n_planes > 0;
  if (staging && (p->plane_stage_depth < 1 || p->plane_stage_depth > mw::kMaxStageDepth))
    return fail(MW_ERR_ARGUMENT, "plane_stage_depth %d outside [1, %d]", p->plane_stage_depth, mw::kMaxStageDepth);
  mw::PlaneSet planes{};
  mw::ZGatherSet zstage{};
  mw::ZFlushSet zflush{};
  int64_t plane_cells = 0, zstage_cells = 0, row_floats = 0;
  for (int q = 0; q < p->n_planes; ++q) {
    const mw_plane_obs& o = p->planes[q];
    if (o.component < 0 || o.component > 5 || o.axis < 0 || o.axis > 2 || !o.re || !o.im)
      return fail(MW_ERR_ARGUMENT, "plane %d descriptor invalid", q);
    int shp[3];
    component_shape(g, o.component, shp);
    if (o.index < 0 || o.index >= shp[o.axis]) return fail(MW_ERR_ARGUMENT, "plane %d index %d outside axis", q, o.index);
    const int64_t n = o.axis == 0 ? (int64_t)shp[1] * shp[2] : (o.axis == 1 ? (int64_t)shp[0] * shp[2] : (int64_t)shp[0] * shp[1]);
    if (staging && o.axis == 2) {
      // one stage plane per (component, column); planes that repeat one share it
      int z = 0;
      while (z < zstage.count && !(zstage.comp[z] == o.component && zstage.k[z] == o.index)) ++z;
      if (z == zstage.count) {
        zstage.comp[z] = o.component;
        zstage.k[z] = o.index;
        zstage.off[z] = row_floats;
        ++zstage.count;
        row_floats += n;
      }
      zflush.p[zflush.count++] = mw::ZFlushPlane{o.component, 0, zstage.off[z], n, o.re, o.im};
      zstage_cells = std::max(zstage_cells, n);
    } else {
      planes.p[planes.count++] = o;
      plane_cells = std::max(plane_cells, n);
    }
  }
  if (zstage.count > 0) {
    if ((int64_t)p->plane_stage_depth * row_floats > p->plane_stage_floats)
      return fail(MW_ERR_ARGUMENT, "plane_stage holds %lld floats; %d rows of the %d z planes need %lld",
                  (long long)p->plane_stage_floats, p->plane_stage_depth, zstage.count,
                  (long long)((int64_t)p->plane_stage_depth * row_floats));
    zflush.row_floats = row_floats;
    zflush.stage = p->plane_stage;
  }
  if ((p->n_planes > 0 || p->n_points > 0) && (p->nf_obs <= 0 || !p->obs_e_cos || !p->obs_e_sin || !p->obs_h_cos || !p->obs_h_sin))
    return fail(MW_ERR_ARGUMENT, "observer weight tables missing");
  if (p->n_points > 0 && (!p->point_component || !p->point_linear || !p->point_re || !p->point_im))
    return fail(MW_ERR_ARGUMENT, "point observer arrays missing");
  if (p->nf_dft > 0) {
    if (!p->dft_cos || !p->dft_sin) return fail(MW_ERR_ARGUMENT, "DFT weight tables missing");
    for (int c = 0; c < 3; ++c)
      if (!p->dft_re[c] || !p->dft_im[c]) return fail(MW_ERR_ARGUMENT, "DFT accumulator %d missing", c);
  }
  if (p->fused) {
    if ((rc = check_periodic_pairs(g))) return rc;
    for (int c = 0; c < 3; ++c)
      if (!p->fields_alt.e[c] || !p->fields_alt.h[c]) return fail(MW_ERR_ARGUMENT, "fields_alt %d is NULL", c);
    if (p->cpml.enabled)
      for (int q = 0; q < 6; ++q)
        if ((p->cpml.psi_low[6 + q] + p->cpml.psi_high_len[6 + q]) > 0 && !p->psi_h_alt[q])
          return fail(MW_ERR_ARGUMENT, "psi_h_alt %d is NULL", q);
    if ((p->n_h_terms > 0 && !p->h_terms_dev) || (p->n_e_terms > 0 && !p->e_terms_dev))
      return fail(MW_ERR_ARGUMENT, "device term descriptors missing");
    for (int f = 0; f < 2; ++f) {
      const mw_term* t = f ? p->e_terms : p->h_terms;
      for (int i = 0; i < (f ? p->n_e_terms : p->n_h_terms); ++i)
        if (t[i].kind == MW_TERM_SHIFTED && !t[i].values) return fail(MW_ERR_ARGUMENT, "term %d has no value buffer", i);
    }
  }
  const mw_aux_line& aux = p->aux;
  if (aux.n_e < 0 || aux.n_e == 1 ||
      (aux.n_e > 0 && (!aux.e || !aux.h || !aux.decay_e || !aux.curl_e || !aux.decay_h || !aux.curl_h || !aux.source)))
    return fail(MW_ERR_ARGUMENT, "auxiliary line arrays missing");
  hipStream_t s = (hipStream_t)stream;
  const mw_ade* ade = p->has_ade ? &p->ade : nullptr;
  SideGuard side;
  if (p->fused && p->z_tiling == 4 && count > 0 && (rc = side_acquire(s, &side.lane))) return rc;
  // the point observers of step n ride in step n+1's merged interior + shell-H
  // launch when there is one (that step reads the fields step n wrote and writes
  // the other buffers); the rest, and the call's last step, get a launch of their own
  mw::PointSet pts{};
  bool pts_pending = false;
  auto point_set = [&](const mw::FieldSet& fs, int n) {
    mw::PointSet ps{};
    for (int c = 0; c < 6; ++c) ps.ptr[c] = fs.ptr[c];
    const int64_t off = (int64_t)n * p->nf_obs;
    ps.comp = p->point_component;
    ps.lin = p->point_linear;
    ps.e_cos = p->obs_e_cos + off;
    ps.e_sin = p->obs_e_sin + off;
    ps.h_cos = p->obs_h_cos + off;
    ps.h_sin = p->obs_h_sin + off;
    ps.re = p->point_re;
    ps.im = p->point_im;
    ps.count = p->n_points;
    ps.nf = p->nf_obs;
    return ps;
  };
  auto launch_points = [&](const mw::PointSet& ps) {
    hipLaunchKernelGGL(mw::accumulate_points_kernel, dim3((ps.count + 63) / 64), dim3(64), 0, s, ps);
    return check_launch("accumulate_points");
  };
  // likewise the z-plane gather of a step inside a block of staged steps (its
  // stage row is not the next step's); a block's last step and the call's last
  // step gather in a launch of their own, ahead of the fold
  mw::ZGatherJob zjob{};
  bool zjob_pending = false;
  auto launch_gather = [&](const mw::ZGatherJob& j) {
    hipLaunchKernelGGL(mw::gather_zplanes_kernel, dim3(j.blocks), dim3(256), 0, s, j.set, j.fs, j.samples);
    return check_launch("gather_zplanes");
  };
  for (int n = first; n < first + count; ++n) {
    if (aux.n_e > 0) {
      hipLaunchKernelGGL(mw::aux_line_h_kernel, dim3((aux.n_e + 254) / 256), dim3(256), 0, s, aux);
      if ((rc = check_launch("aux_line_h"))) return rc;
    }
    mw_fields cur = p->fields;
    if (zstage.count > 0) {
      zstage.row = p->plane_stage + (int64_t)(n % p->plane_stage_depth) * row_floats;
      for (int q = 0; q < 3; ++q) zstage.own_lo[q] = zstage.own_hi[q] = 0;
    }
    if (p->fused) {
      bool taken = false;
      if ((rc = launch_fused(p, n, s, &cur, zstage.count > 0 ? &zstage : nullptr, side.lane,
                             pts_pending ? &pts : nullptr, zjob_pending ? &zjob : nullptr, &taken)))
        return rc;
      if (!taken) {
        if (zjob_pending && (rc = launch_gather(zjob))) return rc;
        if (pts_pending && (rc = launch_points(pts))) return rc;
      }
      pts_pending = zjob_pending = false;
    } else {
      if ((rc = launch_h(g, &p->coeffs, &p->cpml, &p->fields, 7, s))) return rc;
      if ((rc = launch_terms(g, &p->fields, p->h_terms, p->n_h_terms, n, p->dt, s))) return rc;
      if ((rc = launch_e(g, &p->coeffs, &p->cpml, ade, &p->fields, 7, s))) return rc;
      if ((rc = launch_terms(g, &p->fields, p->e_terms, p->n_e_terms, n, p->dt, s))) return rc;
    }
    if (aux.n_e > 0) {
      hipLaunchKernelGGL(mw::aux_line_e_kernel, dim3((aux.n_e + 255) / 256), dim3(256), 0, s, aux, n);
      if ((rc = check_launch("aux_line_e"))) return rc;
    }
    const mw::FieldSet fs = field_set(g, &cur);
    if (p->nf_dft > 0 && !p->fused) {
      for (int c = 0; c < 3; ++c) {
        int shp[3];
        component_shape(g, c, shp);
        const int64_t cnt = (int64_t)shp[0] * shp[1] * shp[2];
        hipLaunchKernelGGL(mw::accumulate_kernel, dim3(grid_blocks(cnt)), dim3(256), 0, s, fs.ptr[c], cnt, p->nf_dft,
                           p->dft_cos + (int64_t)n * p->nf_dft, p->dft_sin + (int64_t)n * p->nf_dft, p->dft_re[c],
                           p->dft_im[c]);
      }
      if ((rc = check_launch("accumulate_dft"))) return rc;
    }
    if (planes.count > 0) {
      const int64_t off = (int64_t)n * p->nf_obs;
      hipLaunchKernelGGL(mw::accumulate_planes_kernel, dim3(grid_blocks(plane_cells), planes.count), dim3(256), 0, s,
                         planes, fs, p->nf_obs, p->obs_e_cos + off, p->obs_e_sin + off, p->obs_h_cos + off,
                         p->obs_h_sin + off);
      if ((rc = check_launch("accumulate_planes"))) return rc;
    }
    if (zstage.count > 0) {
      const int row = n % p->plane_stage_depth;
      const bool block_end = row == p->plane_stage_depth - 1, call_end = n == first + count - 1;
      const mw::ZGatherJob job{zstage, fs, zstage_cells, grid_blocks(zstage_cells * zstage.count), 0};
      if (p->fused && !block_end && !call_end) {
        zjob = job;
        zjob_pending = true;
      } else if ((rc = launch_gather(job))) {
        return rc;
      }
      // fold the block's rows when it completes or the call ends inside it (then
      // clear them: rows of steps already folded must add nothing at the block's end)
      if (block_end || call_end) {
        const int64_t off = (int64_t)(n - row) * p->nf_obs;
        hipLaunchKernelGGL(mw::flush_zstage_kernel, dim3(grid_blocks(zstage_cells), zflush.count), dim3(256), 0, s,
                           zflush, row + 1, p->nf_obs, p->obs_e_cos + off, p->obs_e_sin + off, p->obs_h_cos + off,
                           p->obs_h_sin + off);
        if ((rc = check_launch("flush_zstage"))) return rc;
        if (!block_end) {
          const int64_t folded = (int64_t)(row + 1) * row_floats;
          hipLaunchKernelGGL(mw::clear_floats_kernel, dim3(grid_blocks(folded)), dim3(256), 0, s, p->plane_stage,
                             folded);
          if ((rc = check_launch("clear_zstage"))) return rc;
        }
      }
    }
    if (p->n_points > 0) {
      pts = point_set(fs, n);
      if (p->fused) pts_pending = true;
      else if ((rc = launch_points(pts))) return rc;
    }
  }
  if (pts_pending && (rc = launch_points(pts))) return rc;
  return MW_OK;
}

int mw_capture_steps(const mw_program* p, int32_t first, int32_t count, const mw_window* w, void* stream) {
  if (!p || !w) return fail(MW_ERR_ARGUMENT, "program or window is NULL");
  int rc;
  const mw_grid* g = &p->grid;
  if ((rc = check_window(p, w))) return rc;
  if (count < 0 || count > w->capacity)
    return fail(MW_ERR_ARGUMENT, "capture of %d steps exceeds the window capacity %d", count, w->capacity);
  // replay without observers: the forward sums were taken by the first pass
  mw_program q = *p;
  q.n_planes = 0;
  q.n_points = 0;
  q.nf_dft = 0;
  hipStream_t s = (hipStream_t)stream;
  const mw::WindowArgs wa = window_args(w);
  const int64_t we = window_cells(w, 0), wh = window_cells(w, 1);
  auto capture = [&](const mw_fields& f, bool h, float* dst) {
    const float* const* src = h ? (const float* const*)f.h : (const float* const*)f.e;
    hipLaunchKernelGGL(mw::window_capture_kernel, dim3(grid_blocks(h ? wh : we), 3), dim3(256), 0, s, wa, *g, src[0],
                       src[1], src[2], dst, h ? 1 : 0);
    return check_launch("window_capture");
  };
  if ((rc = capture(state_fields(p, first), false, w->e_store))) return rc;
  for (int n = first; n < first + count; ++n) {
    if ((rc = mw_run_steps(&q, n, 1, stream))) return rc;
    const mw_fields& cur = state_fields(p, n + 1);
    if ((rc = capture(cur, false, w->e_store + (int64_t)(n - first + 1) * 3 * we))) return rc;
    if ((rc = capture(cur, true, w->h_store + (int64_t)(n - first) * 3 * wh))) return rc;
  }
  return MW_OK;
}

int mw_reverse_steps(const mw_program* p, const mw_adjoint* adj, const mw_window* w, int32_t first, int32_t count,
                     void* stream) {
  if (!p || !adj) return fail(MW_ERR_ARGUMENT, "program or adjoint is NULL");
  int rc;
  const mw_grid* g = &p->grid;
  if ((rc = check_grid(g)) || (rc = check_fields(&adj->fields, true, true))) return rc;
  if (first < 0 || count < 0 || first + count > p->total_steps)
    return fail(MW_ERR_ARGUMENT, "steps [%d, %d) outside the program's %d steps", first, first + count, p->total_steps);
  if (w) {
    if ((rc = check_window(p, w))) return rc;
    if (count > w->capacity) return fail(MW_ERR_ARGUMENT, "%d steps exceed the window capacity %d", count, w->capacity);
    for (int c = 0; c < 3; ++c)
      if (!w->g_decay[c] || !w->g_curl[c]) return fail(MW_ERR_ARGUMENT, "window gradient %d is NULL", c);
  }
  mw::RevArgs a{};
  a.g = *g;
  a.cf = p->coeffs;
  a.cp = p->cpml;
  a.a = adj->fields;
  a.cpml = p->cpml.enabled ? 1 : 0;
  if (a.cpml) {
    for (int q = 0; q < 12; ++q) {
      a.cp.psi[q] = adj->psi[q];
      if (p->cpml.psi_low[q] + p->cpml.psi_high_len[q] > 0 && !adj->psi[q])
        return fail(MW_ERR_ARGUMENT, "adjoint psi %d is NULL", q);
    }
  }
  a.ade = p->has_ade ? 1 : 0;
  if (a.ade) {
    if ((rc = check_ade(g, &p->ade))) return rc;
    a.ad = p->ade;
    for (int c = 0; c < 3; ++c)
      for (int q = 0; q < p->ade.count[c]; ++q) {
        mw_pole& pl = a.ad.poles[c][q];
        pl.polarization = adj->polarization[c][q];
        pl.current = adj->current[c][q];
        if (!pl.current || (pl.kind != 1 && !pl.polarization))
          return fail(MW_ERR_ARGUMENT, "adjoint pole (%d, %d) has a NULL state array", c, q);
      }
  }
  // observer seeds
  if (p->n_planes < 0 || p->n_planes > MW_MAX_PLANES) return fail(MW_ERR_ARGUMENT, "%d planes (max %d)", p->n_planes, MW_MAX_PLANES);
  mw::SeedPlaneSet planes{};
  int64_t plane_cells = 0;
  for (int q = 0; q < p->n_planes; ++q) {
    if (!adj->g_plane_re[q] && !adj->g_plane_im[q]) continue;
    if (!adj->g_plane_re[q] || !adj->g_plane_im[q])
      return fail(MW_ERR_ARGUMENT, "plane %d: give both seed gradients or neither", q);
    mw_plane_obs o = p->planes[q];
    if (o.component < 0 || o.component > 5 || o.axis < 0 || o.axis > 2)
      return fail(MW_ERR_ARGUMENT, "plane %d descriptor invalid", q);
    {
      int shp[3];
      component_shape(g, o.component, shp);
      if (o.index < 0 || o.index >= shp[o.axis]) return fail(MW_ERR_ARGUMENT, "plane %d index %d outside axis", q, o.index);
    }
    if (p->nf_obs <= 0 || !p->obs_e_cos || !p->obs_e_sin || !p->obs_h_cos || !p->obs_h_sin)
      return fail(MW_ERR_ARGUMENT, "observer weight tables missing");
    o.re = (float*)adj->g_plane_re[q];
    o.im = (float*)adj->g_plane_im[q];
    int shp[3];
    component_shape(g, o.component, shp);
    const int64_t n = o.axis == 0 ? (int64_t)shp[1] * shp[2] : (o.axis == 1 ? (int64_t)shp[0] * shp[2] : (int64_t)shp[0] * shp[1]);
    plane_cells = std::max(plane_cells, n);
    planes.p[planes.count++] = o;
  }
  const bool point_seeds = p->n_points > 0 && adj->g_point_re && adj->g_point_im;
  if (point_seeds && (!p->point_component || !p->point_linear || p->nf_obs <= 0 || !p->obs_e_cos))
    return fail(MW_ERR_ARGUMENT, "point observer arrays missing");
  if (p->n_points > 0 && (!adj->g_point_re != !adj->g_point_im))
    return fail(MW_ERR_ARGUMENT, "points: give both seed gradients or neither");
  bool dft_seeds = false;
  for (int c = 0; c < 3 && p->nf_dft > 0; ++c) {
    if (!adj->g_dft_re[c] != !adj->g_dft_im[c]) return fail(MW_ERR_ARGUMENT, "DFT %d: give both seed gradients or neither", c);
    dft_seeds = dft_seeds || adj->g_dft_re[c];
  }
  hipStream_t s = (hipStream_t)stream;
  const mw::FieldSet fs = field_set(g, &adj->fields);
  auto seeds = [&](int n, int family) {
    const int64_t off = (int64_t)n * p->nf_obs;
    if (planes.count > 0) {
      hipLaunchKernelGGL(mw::seed_planes_kernel, dim3(grid_blocks(plane_cells), planes.count), dim3(256), 0, s, planes,
                         fs, p->nf_obs, p->obs_e_cos + off, p->obs_e_sin + off, p->obs_h_cos + off, p->obs_h_sin + off,
                         family);
      if (int r = check_launch("seed_planes")) return r;
    }
    if (point_seeds) {
      hipLaunchKernelGGL(mw::seed_points_kernel, dim3((p->n_points + 63) / 64), dim3(64), 0, s, fs, p->n_points,
                         p->point_component, p->point_linear, p->nf_obs, p->obs_e_cos + off, p->obs_e_sin + off,
                         p->obs_h_cos + off, p->obs_h_sin + off, adj->g_point_re, adj->g_point_im, family);
      if (int r = check_launch("seed_points")) return r;
    }
    if (family == 0 && dft_seeds) {
      for (int c = 0; c < 3; ++c) {
        if (!adj->g_dft_re[c]) continue;
        int shp[3];
        component_shape(g, c, shp);
        const int64_t cnt = (int64_t)shp[0] * shp[1] * shp[2];
        hipLaunchKernelGGL(mw::seed_dense_kernel, dim3(grid_blocks(cnt)), dim3(256), 0, s, adj->fields.e[c], cnt,
                           p->nf_dft, p->dft_cos + (int64_t)n * p->nf_dft, p->dft_sin + (int64_t)n * p->nf_dft,
                           adj->g_dft_re[c], adj->g_dft_im[c]);
      }
      if (int r = check_launch("seed_dft")) return r;
    }
    return MW_OK;
  };
  const dim3 block(64, 4, 1);
  // interior box of the reverse phases: per axis, coordinates whose update reads
  // no face, no psi and no sample outside the grid
  {
    static const int axis_of[6] = {1, 2, 0, 2, 0, 1};
    const int n[3] = {g->nx, g->ny, g->nz};
    for (int ax = 0; ax < 3; ++ax) {
      int lo = 1, hi = n[ax] - 2;
      if (a.cpml) {
        for (int q = 0; q < 12; ++q) {
          if (axis_of[q % 6] != ax) continue;
          const int low = p->cpml.psi_low[q], hs = p->cpml.psi_high_start[q], hl = p->cpml.psi_high_len[q];
          if (low + hl <= 0) continue;
          lo = std::max(lo, q < 6 ? low : low + 1);
          if (hl > 0) hi = std::min(hi, q < 6 ? hs - 1 : hs);
        }
      }
      a.in_lo[ax] = lo;
      a.in_hi[ax] = std::max(lo, hi);
    }
    bool empty = false;
    for (int ax = 0; ax < 3; ++ax) empty = empty || a.in_hi[ax] <= a.in_lo[ax];
    if (empty)
      for (int ax = 0; ax < 3; ++ax) a.in_lo[ax] = a.in_hi[ax] = 0;
    // the nodes outside the interior box (all of them when it is empty)
    const int X0 = empty ? n[0] : a.in_lo[0], X1 = empty ? n[0] : a.in_hi[0];
    const int Y0 = a.in_lo[1], Y1 = a.in_hi[1], Z0 = a.in_lo[2], Z1 = a.in_hi[2];
    const int boxes[6][6] = {{0, X0, 0, n[1], 0, n[2]},  {X1, n[0], 0, n[1], 0, n[2]}, {X0, X1, 0, Y0, 0, n[2]},
                             {X0, X1, Y1, n[1], 0, n[2]}, {X0, X1, Y0, Y1, 0, Z0},     {X0, X1, Y0, Y1, Z1, n[2]}};
    a.nboxes = 0;
    a.total = 0;
    for (const auto& bb : boxes) {
      const int xe = bb[1] - bb[0], ye = bb[3] - bb[2], ze = bb[5] - bb[4];
      if (xe <= 0 || ye <= 0 || ze <= 0) continue;
      const int q = a.nboxes++;
      for (int ax = 0; ax < 3; ++ax) { a.box_lo[q][ax] = bb[2 * ax]; a.box_hi[q][ax] = bb[2 * ax + 1]; }
      int shift = 6;
      double best_use = 0.0;
      for (int sh = 6; sh >= 4; --sh) {
        const int w = 1 << sh;
        const double use = (double)ze / (double)(((ze + w - 1) / w) * w);
        if (use > best_use * 1.05) { best_use = use; shift = sh; }
      }
      a.box_shift[q] = shift;
      a.box_nbz[q] = (ze + (1 << shift) - 1) >> shift;
      a.box_nby[q] = (ye + (256 >> shift) - 1) / (256 >> shift);
      a.box_first[q] = a.total;
      a.total += a.box_nbz[q] * a.box_nby[q] * xe;
    }
  }
  const dim3 grid((a.in_hi[2] - a.in_lo[2] + 63) / 64, (a.in_hi[1] - a.in_lo[1] + 3) / 4, a.in_hi[0] - a.in_lo[0]);
  const bool interior = a.in_hi[0] > a.in_lo[0];
  const dim3 bgrid(a.total);
  const bool per = grid_periodic(*g);
  const mw::WindowArgs wa = w ? window_args(w) : mw::WindowArgs{};
  const int64_t we = w ? window_cells(w, 0) : 0, wh = w ? window_cells(w, 1) : 0;
  for (int n = first + count - 1; n >= first; --n) {
    if ((rc = seeds(n, 0))) return rc;
    if (interior) {
      if (a.cpml) hipLaunchKernelGGL(mw::reverse_h_kernel<true>, grid, block, 0, s, a);
      else hipLaunchKernelGGL(mw::reverse_h_kernel<false>, grid, block, 0, s, a);
      if ((rc = check_launch("reverse_h"))) return rc;
    }
    if (a.total > 0) {
      if (per) {
        if (a.cpml) hipLaunchKernelGGL((mw::reverse_h_box_kernel<true, true>), bgrid, block, 0, s, a);
        else hipLaunchKernelGGL((mw::reverse_h_box_kernel<false, true>), bgrid, block, 0, s, a);
      } else {
        if (a.cpml) hipLaunchKernelGGL((mw::reverse_h_box_kernel<true, false>), bgrid, block, 0, s, a);
        else hipLaunchKernelGGL((mw::reverse_h_box_kernel<false, false>), bgrid, block, 0, s, a);
      }
      if ((rc = check_launch("reverse_h (boundary boxes)"))) return rc;
    }
    if ((rc = seeds(n, 1))) return rc;
    if (w) {
      const int64_t si = n - first;
      hipLaunchKernelGGL(mw::window_gradient_kernel, dim3(grid_blocks(we), 3), dim3(256), 0, s, wa, *g, adj->fields,
                         w->e_store + si * 3 * we, w->h_store + si * 3 * wh, w->g_decay[0], w->g_decay[1],
                         w->g_decay[2], w->g_curl[0], w->g_curl[1], w->g_curl[2]);
      if ((rc = check_launch("window_gradient"))) return rc;
    }
    if (interior) {
      if (a.cpml) {
        if (a.ade) hipLaunchKernelGGL((mw::reverse_e_kernel<true, true>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((mw::reverse_e_kernel<true, false>), grid, block, 0, s, a);
      } else {
        if (a.ade) hipLaunchKernelGGL((mw::reverse_e_kernel<false, true>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((mw::reverse_e_kernel<false, false>), grid, block, 0, s, a);
      }
      if ((rc = check_launch("reverse_e"))) return rc;
    }
    if (a.total > 0) {
      if (a.cpml) {
        if (a.ade) hipLaunchKernelGGL((mw::reverse_e_box_kernel<true, true>), bgrid, block, 0, s, a);
        else hipLaunchKernelGGL((mw::reverse_e_box_kernel<true, false>), bgrid, block, 0, s, a);
      } else {
        if (a.ade) hipLaunchKernelGGL((mw::reverse_e_box_kernel<false, true>), bgrid, block, 0, s, a);
        else hipLaunchKernelGGL((mw::reverse_e_box_kernel<false, false>), bgrid, block, 0, s, a);
      }
      if ((rc = check_launch("reverse_e (boundary boxes)"))) return rc;
    }
  }
  return MW_OK;
}

int mw_batch_create(const mw_program* const* programs, int32_t count, int32_t xchunk, mw_batch** out) {
  if (!programs || !out || count <= 0) return fail(MW_ERR_ARGUMENT, "batch: need programs, a count > 0 and an output");
  *out = nullptr;
  mw_batch* b = new mw_batch();
  auto bail = [&](int rc) { batch_free(b); return rc; };
  int rc;
  VectorSink v;
  if ((rc = collect_batch(programs, count, xchunk, v, b))) return bail(rc);
  if ((rc = upload(v.args, &b->d_args, "arguments")) || (rc = upload(v.shells, &b->d_shell, "shell arguments")) ||
      (rc = upload(v.shell_prefix, &b->d_shell_prefix, "shell block table")) ||
      (rc = upload(v.zedge_prefix, &b->d_zedge_prefix, "z-edge tile table")) ||
      (rc = upload(v.lean_prefix, &b->d_lean_prefix, "lean tile table")) || (rc = upload(v.terms, &b->d_terms, "terms")) ||
      (rc = upload(v.aux, &b->d_aux, "auxiliary lines")) || (rc = upload(v.planes, &b->d_planes, "planes")) ||
      (rc = upload(v.points, &b->d_points, "points")))
    return bail(rc);
  if (hipMalloc((void**)&b->d_counter, sizeof(int)) != hipSuccess) return bail(fail(MW_ERR_RUNTIME, "batch: counter"));
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev_in, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev_out, hipEventDisableTiming) != hipSuccess)
    return bail(fail(MW_ERR_RUNTIME, "batch: cannot create its stream/events"));
  *out = b;
  return MW_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// No-allocation batch run.  The descriptor tables and the step counter of a
// batch live in `scratch`, a static device array (each GPU holds its own copy
// until the process ends); its address on the current device comes from
// hipGetSymbolAddress once per device and is used wherever host code needs the
// array.  The tables are assembled in a static host staging buffer of the same
// layout and copied in before the steps.

static __device__ __align__(256) unsigned char scratch[MW_NOALLOC_SCRATCH_BYTES];

namespace {

alignas(256) unsigned char g_stage[MW_NOALLOC_SCRATCH_BYTES];
std::mutex g_noalloc_mutex;
constexpr int kMaxDevices = 64;
void* g_scratch_ptr[kMaxDevices];

// First pass: table sizes only.
struct CountSink {
  int64_t args = 0, shells = 0, prefixes_n = 1, terms = 0, lines = 0, planes = 0, points = 0;
  bool arg(const mw::FusedArgs&) { ++args; return true; }
  bool shell(const mw::ShellArgs&) { ++shells; return true; }
  bool prefixes(int, int, int) { ++prefixes_n; return true; }
  bool term(const mw::BatchTerm&) { ++terms; return true; }
  bool line(const mw_aux_line&) { ++lines; return true; }
  bool plane(const mw::BatchPlane&) { ++planes; return true; }
  bool point(const mw::BatchPoint&) { ++points; return true; }
};

// Byte offsets of each table in the scratch (256-byte aligned).
struct ScratchLayout {
  size_t counter, args, shells, shell_prefix, lean_prefix, zedge_prefix, terms, aux, planes, points, total;
};

ScratchLayout scratch_layout(const CountSink& c) {
  ScratchLayout L{};
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = off;
    off += (bytes + 255) & ~size_t(255);
    return o;
  };
  L.counter = take(sizeof(int));
  L.args = take(c.args * sizeof(mw::FusedArgs));
  L.shells = take(c.shells * sizeof(mw::ShellArgs));
  L.shell_prefix = take(c.prefixes_n * sizeof(int));
  L.lean_prefix = take(c.prefixes_n * sizeof(int));
  L.zedge_prefix = take(c.prefixes_n * sizeof(int));
  L.terms = take(c.terms * sizeof(mw::BatchTerm));
  L.aux = take(c.lines * sizeof(mw_aux_line));
  L.planes = take(c.planes * sizeof(mw::BatchPlane));
  L.points = take(c.points * sizeof(mw::BatchPoint));
  L.total = off;
  return L;
}

// Second pass: writes each table into the host staging buffer at its offset,
// refusing anything beyond the counts of the first pass.
struct StageSink {
  unsigned char* base;
  ScratchLayout L;
  CountSink cap;
  CountSink n;
  template <typename T>
  bool put(size_t off, int64_t& i, int64_t limit, const T& v) {
    if (i >= limit) return false;
    memcpy(base + off + (size_t)i * sizeof(T), &v, sizeof(T));
    ++i;
    return true;
  }
  bool arg(const mw::FusedArgs& a) { return put(L.args, n.args, cap.args, a); }
  bool shell(const mw::ShellArgs& a) { return put(L.shells, n.shells, cap.shells, a); }
  bool prefixes(int shell_blocks, int lean_tiles, int zedge_tiles) {
    if (n.prefixes_n >= cap.prefixes_n) return false;
    int* sp = (int*)(base + L.shell_prefix);
    int* lp = (int*)(base + L.lean_prefix);
    int* zp = (int*)(base + L.zedge_prefix);
    const int64_t i = n.prefixes_n++;
    sp[i] = sp[i - 1] + shell_blocks;
    lp[i] = lp[i - 1] + lean_tiles;
    zp[i] = zp[i - 1] + zedge_tiles;
    return true;
  }
  bool term(const mw::BatchTerm& t) { return put(L.terms, n.terms, cap.terms, t); }
  bool line(const mw_aux_line& l) { return put(L.aux, n.lines, cap.lines, l); }
  bool plane(const mw::BatchPlane& q) { return put(L.planes, n.planes, cap.planes, q); }
  bool point(const mw::BatchPoint& q) { return put(L.points, n.points, cap.points, q); }
};

}  // namespace

extern "C" {

int mw_batch_run_noalloc(const mw_program* const* programs, int32_t count, int32_t xchunk, int32_t first,
                         int32_t steps, void* stream) {
  if (!programs || count <= 0) return fail(MW_ERR_ARGUMENT, "no-allocation batch: need programs and a count > 0");
  if (first < 0 || steps < 0) return fail(MW_ERR_ARGUMENT, "no-allocation batch: negative first step or step count");
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
    return fail(MW_ERR_RUNTIME, "no-allocation batch: no usable current device (%d)", dev);
  std::lock_guard<std::mutex> lock(g_noalloc_mutex);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(MW_ERR_RUNTIME, "no-allocation batch: cannot resolve the scratch array on device %d: %s", dev,
                  hipGetErrorString(e));
    g_scratch_ptr[dev] = ptr;
  }
  unsigned char* dptr = (unsigned char*)g_scratch_ptr[dev];
  int rc;
  mw_batch sized;
  CountSink counts;
  if ((rc = collect_batch(programs, count, xchunk, counts, &sized))) return rc;
  const ScratchLayout L = scratch_layout(counts);
  if (L.total > (size_t)MW_NOALLOC_SCRATCH_BYTES)
    return fail(MW_ERR_ARGUMENT, "no-allocation batch: the %d scenes' tables need %zu bytes, the scratch holds %d",
                count, L.total, MW_NOALLOC_SCRATCH_BYTES);
  mw_batch b;
  StageSink stage{g_stage, L, counts, CountSink{}};
  memset(g_stage + L.shell_prefix, 0, sizeof(int));
  memset(g_stage + L.lean_prefix, 0, sizeof(int));
  memset(g_stage + L.zedge_prefix, 0, sizeof(int));
  if ((rc = collect_batch(programs, count, xchunk, stage, &b))) return rc;
  if (first + (int64_t)steps > b.total_steps)
    return fail(MW_ERR_ARGUMENT, "no-allocation batch: steps [%d, %d) outside the %d steps every scene holds", first,
                first + steps, b.total_steps);
  hipStream_t s = (hipStream_t)stream;
  if (hipMemcpyAsync(dptr, g_stage, L.total, hipMemcpyHostToDevice, s) != hipSuccess)
    return fail(MW_ERR_RUNTIME, "no-allocation batch: cannot copy the tables into the scratch");
  // the staging buffer is the next call's: wait until this copy has read it
  if (hipStreamSynchronize(s) != hipSuccess)
    return fail(MW_ERR_RUNTIME, "no-allocation batch: the table copy did not complete");
  b.d_counter = (int*)(dptr + L.counter);
  b.d_args = (mw::FusedArgs*)(dptr + L.args);
  b.d_shell = (mw::ShellArgs*)(dptr + L.shells);
  b.d_shell_prefix = (int*)(dptr + L.shell_prefix);
  b.d_lean_prefix = (int*)(dptr + L.lean_prefix);
  b.d_zedge_prefix = (int*)(dptr + L.zedge_prefix);
  b.d_terms = (mw::BatchTerm*)(dptr + L.terms);
  b.d_aux = (mw_aux_line*)(dptr + L.aux);
  b.d_planes = (mw::BatchPlane*)(dptr + L.planes);
  b.d_points = (mw::BatchPoint*)(dptr + L.points);
  hipLaunchKernelGGL(mw::step_counter_kernel, dim3(1), dim3(64), 0, s, b.d_counter, 1, first);
  if ((rc = check_launch("no-allocation batch counter"))) return rc;
  for (int i = 0; i < steps; ++i)
    if ((rc = batch_step(&b, s))) return rc;
  return MW_OK;
}

int mw_batch_run(mw_batch* b, int32_t first, int32_t steps, int32_t graph_steps, void* stream) {
  if (!b) return fail(MW_ERR_ARGUMENT, "batch is NULL");
  if (first < 0 || steps < 0 || first + steps > b->total_steps)
    return fail(MW_ERR_ARGUMENT, "batch: steps [%d, %d) outside the %d steps every scene holds", first, first + steps,
                b->total_steps);
  if (graph_steps < 0 || graph_steps > 4096) return fail(MW_ERR_ARGUMENT, "batch: graph_steps %d out of range", graph_steps);
  int rc;
  hipStream_t caller = (hipStream_t)stream;
  if (hipEventRecord(b->ev_in, caller) != hipSuccess || hipStreamWaitEvent(b->stream, b->ev_in, 0) != hipSuccess)
    return fail(MW_ERR_RUNTIME, "batch: cannot order after the caller's stream");
  hipLaunchKernelGGL(mw::step_counter_kernel, dim3(1), dim3(64), 0, b->stream, b->d_counter, 1, first);
  if ((rc = check_launch("batch counter"))) return rc;
  int done = 0;
  if (graph_steps > 1 && steps >= graph_steps) {
    if (!b->exec || b->graph_steps != graph_steps) {
      if (b->exec) { (void)hipGraphExecDestroy(b->exec); b->exec = nullptr; }
      if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
      if (hipStreamBeginCapture(b->stream, hipStreamCaptureModeRelaxed) != hipSuccess)
        return fail(MW_ERR_RUNTIME, "batch: cannot begin graph capture");
      int crc = MW_OK;
      for (int i = 0; i < graph_steps && crc == MW_OK; ++i) crc = batch_step(b, b->stream);
      hipGraph_t graph = nullptr;
      const hipError_t e = hipStreamEndCapture(b->stream, &graph);
      if (crc != MW_OK) { if (graph) (void)hipGraphDestroy(graph); return crc; }
      if (e != hipSuccess || !graph) return fail(MW_ERR_RUNTIME, "batch: graph capture failed: %s", hipGetErrorString(e));
      if (hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGraphDestroy(graph);
        return fail(MW_ERR_RUNTIME, "batch: cannot instantiate the step graph");
      }
      b->graph = graph;
      b->graph_steps = graph_steps;
    }
    for (; done + graph_steps <= steps; done += graph_steps)
      if (hipGraphLaunch(b->exec, b->stream) != hipSuccess) return fail(MW_ERR_LAUNCH, "batch: graph launch failed");
  }
  for (; done < steps; ++done)
    if ((rc = batch_step(b, b->stream))) return rc;
  if (hipEventRecord(b->ev_out, b->stream) != hipSuccess || hipStreamWaitEvent(caller, b->ev_out, 0) != hipSuccess)
    return fail(MW_ERR_RUNTIME, "batch: cannot order the caller's stream after the batch");
  return MW_OK;
}

}  // extern "C"

struct mw_adjoint_batch {
  mw_batch* fwd = nullptr;
  int count = 0;
  bool cpml = false, ade = false;
  mw::RevArgs* d_rev = nullptr;
  int* d_prefix = nullptr;
  int total = 0;
  mw::SeedPlane* d_planes = nullptr;
  int n_planes = 0;
  int64_t max_plane_cells = 0;
  mw::SeedPoint* d_points = nullptr;
  int n_points = 0;
  mw::BatchWindow* d_windows = nullptr;
  int64_t max_window = 0, max_hwindow = 0;
  int capacity = 0;
};

namespace {
void adjoint_batch_free(mw_adjoint_batch* b) {
  if (!b) return;
  void* bufs[] = {b->d_rev, b->d_prefix, b->d_planes, b->d_points, b->d_windows};
  for (void* q : bufs)
    if (q) (void)hipFree(q);
  delete b;
}
}  // namespace

extern "C" {

int mw_adjoint_batch_create(mw_batch* fwd, const mw_program* const* programs, const mw_adjoint* const* adjoints,
                            const mw_window* const* windows, int32_t count, mw_adjoint_batch** out) {
  if (!fwd || !programs || !adjoints || !windows || !out || count != fwd->count)
    return fail(MW_ERR_ARGUMENT, "adjoint batch: need the forward batch and one program/adjoint/window per scene");
  *out = nullptr;
  mw_adjoint_batch* b = new mw_adjoint_batch();
  auto bail = [&](int rc) { adjoint_batch_free(b); return rc; };
  b->fwd = fwd;
  b->count = count;
  std::vector<mw::RevArgs> rev;
  std::vector<int> prefix(1, 0);
  std::vector<mw::SeedPlane> planes;
  std::vector<mw::SeedPoint> points;
  std::vector<mw::BatchWindow> wins;
  int rc;
  for (int si = 0; si < count; ++si) {
    const mw_program* p = programs[si];
    const mw_adjoint* adj = adjoints[si];
    const mw_window* w = windows[si];
    if (!p || !adj || !w) return bail(fail(MW_ERR_ARGUMENT, "adjoint batch: scene %d has a NULL input", si));
    const mw_grid* g = &p->grid;
    if ((rc = check_grid(g)) || (rc = check_fields(&adj->fields, true, true)) || (rc = check_window(p, w))) return bail(rc);
    if (p->nf_dft > 0) return bail(fail(MW_ERR_ARGUMENT, "adjoint batch: full-field DFT seeds are not batched (scene %d)", si));
    for (int c = 0; c < 3; ++c)
      if (!w->g_decay[c] || !w->g_curl[c]) return bail(fail(MW_ERR_ARGUMENT, "adjoint batch: scene %d window gradients", si));
    mw::RevArgs a{};
    a.g = *g;
    a.cf = p->coeffs;
    a.cp = p->cpml;
    a.a = adj->fields;
    a.cpml = p->cpml.enabled ? 1 : 0;
    if (a.cpml)
      for (int q = 0; q < 12; ++q) {
        a.cp.psi[q] = adj->psi[q];
        if (p->cpml.psi_low[q] + p->cpml.psi_high_len[q] > 0 && !adj->psi[q])
          return bail(fail(MW_ERR_ARGUMENT, "adjoint batch: scene %d adjoint psi %d is NULL", si, q));
      }
    a.ade = p->has_ade ? 1 : 0;
    if (a.ade) {
      if ((rc = check_ade(g, &p->ade))) return bail(rc);
      a.ad = p->ade;
      for (int c = 0; c < 3; ++c)
        for (int q = 0; q < p->ade.count[c]; ++q) {
          mw_pole& pl = a.ad.poles[c][q];
          pl.polarization = adj->polarization[c][q];
          pl.current = adj->current[c][q];
          if (!pl.current || (pl.kind != 1 && !pl.polarization))
            return bail(fail(MW_ERR_ARGUMENT, "adjoint batch: scene %d pole (%d, %d) state is NULL", si, c, q));
        }
    }
    if (si == 0) { b->cpml = a.cpml; b->ade = a.ade; }
    else if ((bool)a.cpml != b->cpml || (bool)a.ade != b->ade)
      return bail(fail(MW_ERR_ARGUMENT, "adjoint batch: scene %d differs in CPML/ADE use", si));
    rev.push_back(a);
    prefix.push_back(prefix.back() + ((g->nz + 63) / 64) * ((g->ny + 3) / 4) * g->nx);
    // seeds
    if (p->n_planes < 0 || p->n_planes > MW_MAX_PLANES) return bail(fail(MW_ERR_ARGUMENT, "adjoint batch: plane count"));
    for (int q = 0; q < p->n_planes; ++q) {
      if (!adj->g_plane_re[q] && !adj->g_plane_im[q]) continue;
      if (!adj->g_plane_re[q] || !adj->g_plane_im[q])
        return bail(fail(MW_ERR_ARGUMENT, "adjoint batch: scene %d plane %d: give both seed gradients or neither", si, q));
      const mw_plane_obs& o = p->planes[q];
      if (o.component < 0 || o.component > 5 || o.axis < 0 || o.axis > 2)
        return bail(fail(MW_ERR_ARGUMENT, "adjoint batch: scene %d plane %d descriptor invalid", si, q));
      mw::SeedPlane sp{};
      sp.axis = o.axis; sp.index = o.index; sp.nf = p->nf_obs; sp.family = o.component < 3 ? 0 : 1;
      component_shape(g, o.component, sp.shape);
      if (o.index < 0 || o.index >= sp.shape[o.axis]) return bail(fail(MW_ERR_ARGUMENT, "adjoint batch: plane index"));
      sp.lam = o.component < 3 ? adj->fields.e[o.component] : adj->fields.h[o.component - 3];
      sp.gre = adj->g_plane_re[q]; sp.gim = adj->g_plane_im[q];
      sp.w_cos = o.component < 3 ? p->obs_e_cos : p->obs_h_cos;
      sp.w_sin = o.component < 3 ? p->obs_e_sin : p->obs_h_sin;
      if (!sp.w_cos || !sp.w_sin) return bail(fail(MW_ERR_ARGUMENT, "adjoint batch: observer weight tables missing"));
      const int* sh = sp.shape;
      b->max_plane_cells = std::max(b->max_plane_cells, o.axis == 0 ? (int64_t)sh[1] * sh[2]
                                                        : (o.axis == 1 ? (int64_t)sh[0] * sh[2] : (int64_t)sh[0] * sh[1]));
      planes.push_back(sp);
    }
    if (p->n_points > 0 && adj->g_point_re && adj->g_point_im) {
      std::vector<int32_t> comp(p->n_points), lin(p->n_points);
      if (hipMemcpy(comp.data(), p->point_component, p->n_points * sizeof(int32_t), hipMemcpyDeviceToHost) != hipSuccess ||
          hipMemcpy(lin.data(), p->point_linear, p->n_points * sizeof(int32_t), hipMemcpyDeviceToHost) != hipSuccess)
        return bail(fail(MW_ERR_RUNTIME, "adjoint batch: cannot read scene %d point tables", si));
      for (int q = 0; q < p->n_points; ++q) {
        if (comp[q] < 0 || comp[q] > 5) return bail(fail(MW_ERR_ARGUMENT, "adjoint batch: point component"));
        int sh[3];
        component_shape(g, comp[q], sh);
        if (lin[q] < 0 || (int64_t)lin[q] >= (int64_t)sh[0] * sh[1] * sh[2])
          return bail(fail(MW_ERR_ARGUMENT, "adjoint batch: point index outside the field"));
        mw::SeedPoint sp{};
        sp.linear = lin[q]; sp.nf = p->nf_obs; sp.stride = p->n_points; sp.family = comp[q] < 3 ? 0 : 1;
        sp.lam = comp[q] < 3 ? adj->fields.e[comp[q]] : adj->fields.h[comp[q] - 3];
        sp.gre = adj->g_point_re + q; sp.gim = adj->g_point_im + q;
        sp.w_cos = comp[q] < 3 ? p->obs_e_cos : p->obs_h_cos;
        sp.w_sin = comp[q] < 3 ? p->obs_e_sin : p->obs_h_sin;
        points.push_back(sp);
      }
    }
    mw::BatchWindow bw{};
    bw.w = window_args(w);
    bw.g = *g;
    for (int c = 0; c < 3; ++c) {
      bw.e_state[0][c] = p->fields.e[c];
      bw.e_state[1][c] = p->fused ? p->fields_alt.e[c] : p->fields.e[c];
      bw.h_state[0][c] = p->fields.h[c];
      bw.h_state[1][c] = p->fused ? p->fields_alt.h[c] : p->fields.h[c];
      bw.lam_e[c] = adj->fields.e[c];
      bw.g_decay[c] = w->g_decay[c];
      bw.g_curl[c] = w->g_curl[c];
    }
    bw.e_store = w->e_store;
    bw.h_store = w->h_store;
    b->capacity = si == 0 ? w->capacity : std::min(b->capacity, w->capacity);
    b->max_window = std::max(b->max_window, window_cells(w, 0));
    b->max_hwindow = std::max(b->max_hwindow, window_cells(w, 1));
    wins.push_back(bw);
  }
  b->total = prefix.back();
  b->n_planes = (int)planes.size();
  b->n_points = (int)points.size();
  if ((rc = upload(rev, &b->d_rev, "reverse arguments")) || (rc = upload(prefix, &b->d_prefix, "block table")) ||
      (rc = upload(planes, &b->d_planes, "seed planes")) || (rc = upload(points, &b->d_points, "seed points")) ||
      (rc = upload(wins, &b->d_windows, "windows")))
    return bail(rc);
  *out = b;
  return MW_OK;
}

int mw_adjoint_batch_capture(mw_adjoint_batch* b, int32_t first, int32_t count, void* stream) {
  if (!b) return fail(MW_ERR_ARGUMENT, "adjoint batch is NULL");
  mw_batch* f = b->fwd;
  if (first < 0 || count < 0 || first + count > f->total_steps || count > b->capacity)
    return fail(MW_ERR_ARGUMENT, "adjoint batch: capture [%d, %d) outside the steps or the window capacity %d", first,
                first + count, b->capacity);
  hipStream_t s = (hipStream_t)stream;
  int rc;
  hipLaunchKernelGGL(mw::step_counter_kernel, dim3(1), dim3(64), 0, s, f->d_counter, 1, first);
  if ((rc = check_launch("adjoint batch counter"))) return rc;
  const int64_t cells = std::max(b->max_window, b->max_hwindow);
  const dim3 cgrid(grid_blocks(cells), 3 * b->count);
  hipLaunchKernelGGL(mw::window_capture_batch_kernel, cgrid, dim3(256), 0, s, b->d_windows, b->count, first, first);
  if ((rc = check_launch("batch window capture"))) return rc;
  for (int m = first + 1; m <= first + count; ++m) {
    if ((rc = batch_step(f, s, false))) return rc;
    hipLaunchKernelGGL(mw::window_capture_batch_kernel, cgrid, dim3(256), 0, s, b->d_windows, b->count, m, first);
    if ((rc = check_launch("batch window capture"))) return rc;
  }
  return MW_OK;
}

int mw_adjoint_batch_reverse(mw_adjoint_batch* b, int32_t first, int32_t count, void* stream) {
  if (!b) return fail(MW_ERR_ARGUMENT, "adjoint batch is NULL");
  if (first < 0 || count < 0 || first + count > b->fwd->total_steps || count > b->capacity)
    return fail(MW_ERR_ARGUMENT, "adjoint batch: reverse [%d, %d) outside the steps or the window capacity", first,
                first + count);
  hipStream_t s = (hipStream_t)stream;
  int rc;
  const dim3 block(64, 4, 1);
  const dim3 grid(b->total);
  auto seeds = [&](int n, int family) {
    if (b->n_planes > 0) {
      hipLaunchKernelGGL(mw::seed_planes_batch_kernel, dim3(grid_blocks(b->max_plane_cells), b->n_planes), dim3(256),
                         0, s, b->d_planes, b->n_planes, n, family);
      if (int r = check_launch("batch seed planes")) return r;
    }
    if (b->n_points > 0) {
      hipLaunchKernelGGL(mw::seed_points_batch_kernel, dim3((b->n_points + 63) / 64), dim3(64), 0, s, b->d_points,
                         b->n_points, n, family);
      if (int r = check_launch("batch seed points")) return r;
    }
    return MW_OK;
  };
  for (int n = first + count - 1; n >= first; --n) {
    if ((rc = seeds(n, 0))) return rc;
    if (b->cpml) hipLaunchKernelGGL(mw::reverse_h_batch_kernel<true>, grid, block, 0, s, b->d_rev, b->d_prefix, b->count, b->total);
    else hipLaunchKernelGGL(mw::reverse_h_batch_kernel<false>, grid, block, 0, s, b->d_rev, b->d_prefix, b->count, b->total);
    if ((rc = check_launch("batch reverse_h"))) return rc;
    if ((rc = seeds(n, 1))) return rc;
    hipLaunchKernelGGL(mw::window_gradient_batch_kernel, dim3(grid_blocks(b->max_window), 3 * b->count), dim3(256), 0, s,
                       b->d_windows, b->count, n - first);
    if ((rc = check_launch("batch window gradient"))) return rc;
    if (b->cpml) {
      if (b->ade) hipLaunchKernelGGL((mw::reverse_e_batch_kernel<true, true>), grid, block, 0, s, b->d_rev, b->d_prefix, b->count, b->total);
      else hipLaunchKernelGGL((mw::reverse_e_batch_kernel<true, false>), grid, block, 0, s, b->d_rev, b->d_prefix, b->count, b->total);
    } else {
      if (b->ade) hipLaunchKernelGGL((mw::reverse_e_batch_kernel<false, true>), grid, block, 0, s, b->d_rev, b->d_prefix, b->count, b->total);
      else hipLaunchKernelGGL((mw::reverse_e_batch_kernel<false, false>), grid, block, 0, s, b->d_rev, b->d_prefix, b->count, b->total);
    }
    if ((rc = check_launch("batch reverse_e"))) return rc;
  }
  return MW_OK;
}

int mw_adjoint_batch_destroy(mw_adjoint_batch* b) {
  // the tables may still be read by queued launches: wait for the device first
  if (b) (void)hipDeviceSynchronize();
  adjoint_batch_free(b);
  return MW_OK;
}

int mw_batch_destroy(mw_batch* b) {
  if (!b) return MW_OK;
  // the batch's tables are read by launches on its own stream and, through
  // mw_adjoint_batch_capture's replay, on a caller's stream: wait for the device
  (void)hipDeviceSynchronize();
  batch_free(b);
  return MW_OK;
}

}  // extern "C"
