// hops_select.hip — the adaptive basis selection of a ragged batch on the
// device: the per-auxiliary error bound, the stable set, the boundary
// candidates with the f_discard pre-trim, their binning by target auxiliary
// and the per-state sums of the state half.  One launch of each kernel covers
// every listed slot (tiles of 256 auxiliaries, or one workgroup per slot).
//
// Restated from the reference (MesoHOPS v1.6.0, src/mesohops/):
//   basis/hops_basis.py:484-558      _define_hierarchy_basis
//   basis/hops_basis.py:560-665      hier_stable_error
//   basis/hops_basis.py:667-793      _determine_boundary_hier (pre-trim,
//                                    binning by target auxiliary)
//   basis/hops_basis.py:795-912      state_stable_error
//   basis/basis_functions.py:8-41    determine_error_thresh
//   basis/basis_functions_adaptive.py:76-390   flux up / down / state flux
//
// Every value is computed with the operations, in the order, of the host
// selection in hops_runtime.cpp (select_hier, select_states), and without
// fused multiply-adds, so each error, threshold and set is bitwise the
// host's.  Sums over auxiliaries are blocked: a sequential sum over each
// 256-entry block, then a sequential sum of the blocks.  The thresholds scan
// the ascending-sorted values in order, as the reference's cumsum does.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_segmented_radix_sort.hpp>

#include <cmath>
#include <cstdio>

#include "hops_select.h"

extern "C" int hops_internal_set_error(const char *msg);

#pragma clang fp contract(off)

namespace {

constexpr int kT = 256;  // entries per tile

int sel_fail(const char *what, hipError_t e) {
  char buf[256];
  snprintf(buf, sizeof(buf), "%s: %s", what, hipGetErrorString(e));
  return hops_internal_set_error(buf);
}

#define SEL_LAUNCH(what)                             \
  do {                                               \
    hipError_t e_ = hipGetLastError();               \
    if (e_ != hipSuccess) return sel_fail(what, e_); \
  } while (0)

__device__ inline double n2(double re, double im) { return re * re + im * im; }

__device__ inline double readlane_f64(double v, int l) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

__device__ inline int aux_value(const hops_op_t &op, int k, int m) {
  for (int q = op.aux_ptr_d[k]; q < op.aux_ptr_d[k + 1]; ++q)
    if (op.aux_mode_d[q] == m) return op.aux_val_d[q];
  return 0;
}

// The flux terms of auxiliary k, relative mode by relative mode (ascending):
// f(lm, m, kv, sum_up, eu, sum_dn, ed, cand_up, cand_dn).  sum_*: the term
// enters the error of k; cand_*: k +- e_m is a boundary candidate (outside
// the basis, nonzero flux).  Same terms and arithmetic as select_hier.
template <class F>
__device__ void flux_terms(const hops_sel_seg_t &g, const hops_op_t &op, const hops_sel_mode_t *md,
                           const double2 *pk, int k, double colsum, double h2, F &&f) {
  const int a0 = op.aux_ptr_d[k], a1 = op.aux_ptr_d[k + 1];
  int depth = 0;
  for (int q = a0; q < a1; ++q) depth += op.aux_val_d[q];
  const bool top = depth == g.maxhier, low = depth < g.maxhier;
  // Markovian filter (hops_fluxfilters.py:354-397): no flux up from any
  // auxiliary but the zeroth along a filtered mode, none at all from e_m of
  // a filtered mode m
  bool mk_first = false;
  if (g.markov && a1 - a0 == 1 && op.aux_val_d[a0] == 1) {
    const int m = op.aux_mode_d[a0];
    for (int lm = 0; lm < g.nm; ++lm)
      if (op.mode_abs_d[lm] == m) {
        mk_first = md[lm].markov != 0;
        break;
      }
  }
  int qa = a0, qu = op.up_ptr_d[k], qd = op.dn_ptr_d[k];
  const int qu1 = op.up_ptr_d[k + 1], qd1 = op.dn_ptr_d[k + 1];
  for (int lm = 0; lm < g.nm; ++lm) {
    const int m = op.mode_abs_d[lm];
    while (qa < a1 && op.aux_mode_d[qa] < m) ++qa;
    const int kv = (qa < a1 && op.aux_mode_d[qa] == m) ? op.aux_val_d[qa] : 0;
    while (qu < qu1 && op.up_mode_d[qu] < lm) ++qu;
    const bool in_up = qu < qu1 && op.up_mode_d[qu] == lm;
    while (qd < qd1 && op.dn_mode_d[qd] < lm) ++qd;
    const bool in_dn = qd < qd1 && op.dn_mode_d[qd] == lm;
    const hops_sel_mode_t c = md[lm];
    const double cs = c.m2s >= 0 ? n2(pk[c.m2s].x, pk[c.m2s].y) : 0.0;
    const bool fu = !(g.markov && ((c.markov && k >= 1) || mk_first));
    double eu = 0.0, ed = 0.0;
    if (fu && (!top || low)) {
      const double p_up = c.m2s >= 0 ? c.m2u * cs : 0.0;
      const double kw = c.aw * (1.0 + kv);
      eu = (kw * kw * p_up) / h2;
    }
    const bool has_dn = kv > 0;
    if (has_dn) {
      double p_dn = c.m2s >= 0 ? c.m2d * cs : 0.0;
      if (c.nX != 0) p_dn += c.nX * (colsum - (c.m2s >= 0 ? cs : 0.0));
      ed = (c.gw2 * p_dn) / h2;
    }
    f(lm, m, kv, !top && fu, eu, has_dn, ed, low && fu && eu != 0 && !in_up, has_dn && ed != 0 && !in_dn);
  }
}

struct RowCtx {
  hops_sel_seg_t g;
  const hops_op_t *op;
  const double2 *pk;
  double colsum;
};

__device__ inline RowCtx row_ctx(const hops_sel_work_t &w, const hops_sel_seg_t &g, int k) {
  RowCtx r;
  r.g = g;
  r.op = &w.ops[g.slot];
  r.pk = reinterpret_cast<const double2 *>(w.pn) + g.pn_off + (int64_t)k * g.n;
  double cs = 0.0;
  for (int s = 0; s < g.n; ++s) cs += n2(r.pk[s].x, r.pk[s].y);
  r.colsum = cs;
  return r;
}

// exclusive scan in place of v over [0, n) by one workgroup of kT threads;
// returns the total (every thread)
template <class Get, class Put>
__device__ int64_t block_exscan(int n, Get get, Put put, int64_t *sh) {
  int64_t carry = 0;
  for (int base = 0; base < n; base += kT) {
    const int j = base + (int)threadIdx.x;
    const int64_t v = j < n ? get(j) : 0;
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < kT; off <<= 1) {
      const int64_t t = (int)threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
      __syncthreads();
      sh[threadIdx.x] += t;
      __syncthreads();
    }
    if (j < n) put(j, carry + sh[threadIdx.x] - v);
    carry += sh[kT - 1];
    __syncthreads();
  }
  return carry;
}

// ---- hierarchy half ---------------------------------------------------------

__global__ __launch_bounds__(kT) void k_sel_err(hops_sel_work_t w, const int2 *tiles) {
  const int2 tl = tiles[blockIdx.x];
  const hops_sel_seg_t g = w.segs[tl.x];
  const int k = tl.y * kT + (int)threadIdx.x;
  if (k >= g.nh) return;
  const RowCtx r = row_ctx(w, g, k);
  const double h2 = w.hbar * w.hbar;
  double err = w.e_aux[g.e_off + k];
  // flux out of the state basis through H (basis_functions_adaptive.py:76-163)
  if (g.n_sc > 0) {
    const int32_t *sp = w.sc_ptr + g.sc_off;
    const hops_sel_sc_t *se = w.sc_ent + g.sce_off;
    double acc = 0.0;
    for (int i = 0; i < g.n_sc; ++i) {
      double zr = 0.0, zi = 0.0;
      for (int q = sp[i]; q < sp[i + 1]; ++q) {
        const hops_sel_sc_t e = se[q];
        const double2 p = r.pk[e.s];
        zr += e.re * p.x - e.im * p.y;
        zi += e.re * p.y + e.im * p.x;
      }
      zr = zr / w.hbar;
      zi = zi / w.hbar;
      acc += n2(zr, zi);
    }
    err += acc;
  }
  double acc_up = 0.0, acc_dn = 0.0;
  flux_terms(g, *r.op, w.modes + g.m_off, r.pk, k, r.colsum, h2,
             [&](int, int, int, bool su, double eu, bool sd, double ed, bool, bool) {
               if (su) acc_up += eu;
               if (sd) acc_dn += ed;
             });
  err += acc_up;
  err += acc_dn;
  w.err[g.h_off + k] = err;
}

__global__ __launch_bounds__(kT) void k_sel_stable(hops_sel_work_t w, const int2 *tiles, int want_c1) {
  __shared__ double lu[kT];
  __shared__ uint8_t lf[kT];
  const int2 tl = tiles[blockIdx.x];
  const hops_sel_seg_t g = w.segs[tl.x];
  const int k = tl.y * kT + (int)threadIdx.x;
  const double thr = w.scal[8 * tl.x + SC_THR];
  double u = 0.0;
  uint8_t st = 0;
  if (k < g.nh) {
    const double e = w.err[g.h_off + k];
    st = e > thr ? 1 : 0;
    int32_t c = 0;
    if (st) {
      const RowCtx r = row_ctx(w, g, k);
      flux_terms(g, *r.op, w.modes + g.m_off, r.pk, k, r.colsum, w.hbar * w.hbar,
                 [&](int, int, int, bool, double, bool, double, bool cu, bool cd) { c += (int)cu + (int)cd; });
    } else {
      u = e;
    }
    w.flag[g.h_off + k] = st;
    w.cnt[g.h_off + k] = c;
  }
  lu[threadIdx.x] = u;
  lf[threadIdx.x] = st;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int q = 0; q < kT; ++q) s += lu[q];
    w.part_u[g.t_off + tl.y] = s;
  }
  if (want_c1) {
    const double2 *pn = reinterpret_cast<const double2 *>(w.pn) + g.pn_off;
    const int rows = min(kT, g.nh - tl.y * kT);
    for (int s = (int)threadIdx.x; s < g.n; s += kT) {
      double a = 0.0;
      for (int q = 0; q < rows; ++q)
        if (lf[q]) {
          const double2 p = pn[(int64_t)(tl.y * kT + q) * g.n + s];
          a += n2(p.x, p.y);
        }
      w.part_c[g.tc_off + (int64_t)tl.y * g.n + s] = a;
    }
  }
}

__global__ __launch_bounds__(kT) void k_sel_fin1(hops_sel_work_t w, int want_c1) {
  __shared__ int64_t sh[kT];
  __shared__ int nst;
  const int i = blockIdx.x;
  const hops_sel_seg_t g = w.segs[i];
  if (threadIdx.x == 0) {
    double u = 0.0;
    for (int t = 0; t < g.n_tiles; ++t) u += w.part_u[g.t_off + t];
    w.scal[8 * i + SC_UNSTABLE] = u;
    w.scal[8 * i + SC_DBOUND] = g.dsq - u;
    nst = 0;
  }
  if (want_c1)
    for (int s = (int)threadIdx.x; s < g.n; s += kT) {
      double a = 0.0;
      for (int t = 0; t < g.n_tiles; ++t) a += w.part_c[g.tc_off + (int64_t)t * g.n + s];
      w.c1[g.s_off + s] = a;
    }
  __syncthreads();
  int mine = 0;
  for (int k = (int)threadIdx.x; k < g.nh; k += kT) mine += w.flag[g.h_off + k];
  atomicAdd(&nst, mine);
  int32_t *cnt = w.cnt + g.h_off;
  const int64_t total = block_exscan(
      g.nh, [&](int j) { return (int64_t)cnt[j]; }, [&](int j, int64_t v) { cnt[j] = (int32_t)v; }, sh);
  if (threadIdx.x == 0) {
    w.tot[8 * i + TT_STABLE] = nst;
    w.tot[8 * i + TT_NZ] = total;
  }
}

__global__ __launch_bounds__(kT) void k_sel_cand(hops_sel_work_t w, const int2 *tiles) {
  const int2 tl = tiles[blockIdx.x];
  const hops_sel_seg_t g = w.segs[tl.x];
  const int k = tl.y * kT + (int)threadIdx.x;
  if (k >= g.nh || !w.flag[g.h_off + k]) return;
  const RowCtx r = row_ctx(w, g, k);
  double *out = w.nz + g.nz_off + w.cnt[g.h_off + k];
  int j = 0;
  flux_terms(g, *r.op, w.modes + g.m_off, r.pk, k, r.colsum, w.hbar * w.hbar,
             [&](int, int, int, bool, double eu, bool, double ed, bool cu, bool cd) {
               if (cu) out[j++] = eu;
               if (cd) out[j++] = ed;
             });
}

// partial sums of the candidates at or below the pre-trim threshold, per
// 256-entry block of the candidate list
__global__ __launch_bounds__(kT) void k_sel_disc(hops_sel_work_t w, const int2 *tiles) {
  __shared__ double lv[kT];
  const int2 tl = tiles[blockIdx.x];
  const hops_sel_seg_t g = w.segs[tl.x];
  const int64_t n = w.tot[8 * tl.x + TT_NZ];
  const double thr1 = w.scal[8 * tl.x + SC_THR1];
  const int64_t j = (int64_t)tl.y * kT + threadIdx.x;
  double v = 0.0;
  if (j < n) {
    const double x = w.nz[g.nz_off + j];
    if (x <= thr1) v = x;
  }
  lv[threadIdx.x] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int q = 0; q < kT; ++q) s += lv[q];
    w.part_d[g.nzt_off + tl.y] = s;
  }
}

// survivors (candidates above the pre-trim threshold) per auxiliary
__global__ __launch_bounds__(kT) void k_sel_scount(hops_sel_work_t w, const int2 *tiles) {
  const int2 tl = tiles[blockIdx.x];
  const hops_sel_seg_t g = w.segs[tl.x];
  const int k = tl.y * kT + (int)threadIdx.x;
  if (k >= g.nh) return;
  int32_t c = 0;
  if (w.flag[g.h_off + k]) {
    const int64_t b = w.cnt[g.h_off + k];
    const int64_t e = k + 1 < g.nh ? (int64_t)w.cnt[g.h_off + k + 1] : w.tot[8 * tl.x + TT_NZ];
    const double thr1 = w.scal[8 * tl.x + SC_THR1];
    for (int64_t q = b; q < e; ++q) c += w.nz[g.nz_off + q] > thr1 ? 1 : 0;
  }
  w.scnt[g.h_off + k] = c;
}

__global__ __launch_bounds__(kT) void k_sel_fin2(hops_sel_work_t w) {
  __shared__ int64_t sh[kT];
  const int i = blockIdx.x;
  const hops_sel_seg_t g = w.segs[i];
  if (threadIdx.x == 0) {
    const int64_t nt = (w.tot[8 * i + TT_NZ] + kT - 1) / kT;
    double s = 0.0;
    for (int64_t t = 0; t < nt; ++t) s += w.part_d[g.nzt_off + t];
    w.scal[8 * i + SC_DISC] = s;
  }
  int32_t *sc = w.scnt + g.h_off;
  const int64_t total = block_exscan(
      g.nh, [&](int j) { return (int64_t)sc[j]; }, [&](int j, int64_t v) { sc[j] = (int32_t)v; }, sh);
  if (threadIdx.x == 0) w.tot[8 * i + TT_SURV] = total;
}

__device__ inline void key_of(const hops_sel_seg_t &g, const hops_op_t &op, int k, uint64_t &a, uint64_t &b) {
  a = 0;
  b = 0;
  for (int q = op.aux_ptr_d[k]; q < op.aux_ptr_d[k + 1]; ++q) {
    const int z = op.aux_mode_d[q] * 256 + op.aux_val_d[q];
    a ^= g.zob_a[z];
    b ^= g.zob_b[z];
  }
}

// the survivors of each stable auxiliary in the reference's insertion order
// (flux up by mode, then flux down by mode), with the Zobrist key of their
// target auxiliary
__global__ __launch_bounds__(kT) void k_sel_surv(hops_sel_work_t w, const int2 *tiles) {
  const int2 tl = tiles[blockIdx.x];
  const hops_sel_seg_t g = w.segs[tl.x];
  const int k = tl.y * kT + (int)threadIdx.x;
  if (k >= g.nh || !w.flag[g.h_off + k]) return;
  const double thr1 = w.scal[8 * tl.x + SC_THR1];
  const RowCtx r = row_ctx(w, g, k);
  uint64_t ka, kb;
  key_of(g, *r.op, k, ka, kb);
  int64_t p = w.scnt[g.h_off + k];
  const int64_t o = g.sv_off;
  for (int d = 0; d < 2; ++d)
    flux_terms(g, *r.op, w.modes + g.m_off, r.pk, k, r.colsum, w.hbar * w.hbar,
               [&](int lm, int m, int kv, bool, double eu, bool, double ed, bool cu, bool cd) {
                 const bool c = d == 0 ? cu : cd;
                 const double v = d == 0 ? eu : ed;
                 if (!c || !(v > thr1)) return;
                 const int v1 = d == 0 ? kv + 1 : kv - 1;
                 const int z0 = m * 256 + kv, z1 = m * 256 + v1;
                 w.ka[o + p] = ka ^ g.zob_a[z0] ^ g.zob_a[z1];
                 w.kb[o + p] = kb ^ g.zob_b[z0] ^ g.zob_b[z1];
                 w.pos[o + p] = (uint32_t)p;
                 w.sval[o + p] = v;
                 w.desc[o + p] = (uint64_t)(uint32_t)k | ((uint64_t)lm << 32) | ((uint64_t)(d == 0) << 48);
                 ++p;
               });
}

// one target auxiliary per key: its first survivor (in insertion order)
// sums the fluxes of every survivor with that key, in insertion order
__global__ __launch_bounds__(kT) void k_sel_group(hops_sel_work_t w, const int2 *tiles) {
  const int2 tl = tiles[blockIdx.x];
  const hops_sel_seg_t g = w.segs[tl.x];
  const int64_t n = w.tot[8 * tl.x + TT_SURV];
  const int64_t r = (int64_t)tl.y * kT + threadIdx.x;
  if (r >= n) return;
  const int64_t o = g.sv_off;
  const uint64_t a = w.ka_sorted[o + r];
  const uint64_t b = w.kb[o + w.pos_sorted[o + r]];
  bool lead = true;
  for (int64_t j = r - 1; j >= 0 && w.ka_sorted[o + j] == a; --j)
    if (w.kb[o + w.pos_sorted[o + j]] == b) {
      lead = false;
      break;
    }
  double s = 0.0;
  if (lead)
    for (int64_t j = r; j < n && w.ka_sorted[o + j] == a; ++j) {
      const uint32_t pj = w.pos_sorted[o + j];
      if (w.kb[o + pj] == b) s += w.sval[o + pj];
    }
  w.lead[o + r] = lead ? 1 : 0;
  w.gv[o + r] = s;
}

__global__ __launch_bounds__(kT) void k_sel_fin3(hops_sel_work_t w) {
  __shared__ int64_t sh[kT];
  const int i = blockIdx.x;
  const hops_sel_seg_t g = w.segs[i];
  const int64_t n = w.tot[8 * i + TT_SURV];
  const int64_t o = g.sv_off;
  const int64_t total = block_exscan(
      (int)n, [&](int j) { return (int64_t)w.lead[o + j]; },
      [&](int j, int64_t q) {
        if (w.lead[o + j]) {
          w.gvals[o + q] = w.gv[o + j];
          w.gdesc[o + q] = w.desc[o + w.pos_sorted[o + j]];
        }
      },
      sh);
  if (threadIdx.x == 0) {
    w.tot[8 * i + TT_LEAD] = total;
    w.seg_beg[i] = o;
    w.seg_end[i] = o + total;
  }
}

__global__ __launch_bounds__(kT) void k_sel_bcount(hops_sel_work_t w) {
  __shared__ int cnt;
  const int i = blockIdx.x;
  const hops_sel_seg_t g = w.segs[i];
  const int64_t n = w.tot[8 * i + TT_LEAD];
  const double thr2 = w.scal[8 * i + SC_THR2];
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  int mine = 0;
  for (int64_t j = threadIdx.x; j < n; j += kT) mine += w.gvals[g.sv_off + j] > thr2 ? 1 : 0;
  atomicAdd(&cnt, mine);
  __syncthreads();
  if (threadIdx.x == 0) w.tot[8 * i + TT_BOUND] = cnt;
}

__global__ void k_sel_boff(hops_sel_work_t w) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  int64_t s = 0;
  w.bound_off[0] = 0;
  for (int i = 0; i < w.n_list; ++i) {
    s += w.tot[8 * i + TT_BOUND];
    w.bound_off[i + 1] = s;
  }
}

__global__ __launch_bounds__(kT) void k_sel_bwrite(hops_sel_work_t w) {
  __shared__ int64_t sh[kT];
  const int i = blockIdx.x;
  const hops_sel_seg_t g = w.segs[i];
  const int64_t n = w.tot[8 * i + TT_LEAD];
  const double thr2 = w.scal[8 * i + SC_THR2];
  const int64_t o = g.sv_off, bo = w.bound_off[i];
  block_exscan(
      (int)n, [&](int j) { return (int64_t)(w.gvals[o + j] > thr2 ? 1 : 0); },
      [&](int j, int64_t q) {
        if (w.gvals[o + j] > thr2) w.bound[bo + q] = w.gdesc[o + j];
      },
      sh);
}

// error_threshold (basis_functions.py:8-41) of each listed slot's sorted
// segment: the running sum in ascending order, as np.cumsum adds it, up to
// its first crossing of max_err - offset
__global__ __launch_bounds__(64) void k_sel_thr(hops_sel_work_t w, const double *sorted, int kind) {
  const int i = blockIdx.x;
  const hops_sel_seg_t g = w.segs[i];
  int64_t b, n;
  double max_err, offset = 0.0;
  int out;
  if (kind == 0) {
    b = g.h_off;
    n = g.nh;
    max_err = g.dsq / 2.0;
    out = SC_THR;
  } else if (kind == 1) {
    b = g.nz_off;
    n = w.tot[8 * i + TT_NZ];
    max_err = w.scal[8 * i + SC_DBOUND] * g.f_discard * g.f_discard;
    out = SC_THR1;
  } else {
    b = g.sv_off;
    n = w.tot[8 * i + TT_LEAD];
    max_err = w.scal[8 * i + SC_DBOUND];
    offset = w.scal[8 * i + SC_DISC];
    out = SC_THR2;
  }
  double thr = 0.0;
  if (n > 0) {
    double cum = 0.0;
    int64_t found = -1;
    for (int64_t base = 0; base < n; base += 64) {
      const int64_t j = base + threadIdx.x;
      const double v = j < n ? sorted[b + j] : 0.0;
      const int m = (int)(n - base < 64 ? n - base : 64);
      int hit = -1;
      for (int l = 0; l < m; ++l) {
        cum += readlane_f64(v, l);
        if (cum + offset >= max_err) {
          hit = l;
          break;
        }
      }
      if (hit >= 0) {
        found = base + hit;
        break;
      }
    }
    if (found >= 0)
      thr = found > 0 ? sorted[b + found - 1] : 0.0;
    else
      thr = sorted[b] + offset < max_err ? INFINITY : 0.0;
  }
  if (threadIdx.x == 0) w.scal[8 * i + out] = thr;
}

// ---- state half ---------------------------------------------------------------

__global__ __launch_bounds__(kT) void k_state_err(hops_sel_work_t w) {
  const int i = blockIdx.x;
  const hops_sel_seg_t g = w.segs[i];
  const hops_op_t &op = w.ops[g.slot];
  const hops_sel_mode_t *md = w.modes + g.m_off;
  const double2 *pn = reinterpret_cast<const double2 *>(w.pn) + g.pn_off;
  const uint32_t *fl = w.fl + 2 * g.fl_off;
  const double h2 = w.hbar * w.hbar;
  for (int s = (int)threadIdx.x; s < g.n; s += kT) {
    double err = w.e_state[g.s_off + s];
    err += w.v1[g.s_off + s] * w.c1[g.s_off + s] / h2;
    for (int e = 0; e < g.n_fl; ++e) {
      const uint32_t code = fl[2 * e];
      const int k = (int)fl[2 * e + 1];
      const int lm = (int)(code >> 2);
      const hops_sel_mode_t c = md[lm];
      const double2 p = pn[(int64_t)k * g.n + s];
      const double c2 = n2(p.x, p.y);
      if (code & 2u) {
        const double a_dn = s == c.m2s ? c.m2d : c.nX;
        err += (a_dn * c.gw2) * c2 / h2;
      }
      if (code & 1u) {
        const double a_up = s == c.m2s ? c.m2u : 0.0;
        const double kw = c.aw * (1.0 + aux_value(op, k, op.mode_abs_d[lm]));
        err += (a_up * kw * kw) * c2 / h2;
      }
    }
    w.err_s[g.s_off + s] = err;
  }
}

// flux into each outside coupled state from the stable states, per tile of
// stable auxiliaries (hops_basis.py:400-482)
__global__ __launch_bounds__(kT) void k_state_flux(hops_sel_work_t w, const int2 *tiles) {
  const int2 tl = tiles[blockIdx.x];
  const hops_sel_seg_t g = w.segs[tl.x];
  const double2 *pn = reinterpret_cast<const double2 *>(w.pn) + g.pn_off;
  const int rows = min(kT, g.nh - tl.y * kT);
  for (int ii = (int)threadIdx.x; ii < g.n_sc; ii += kT) {
    const hops_sel_sc_t *q = w.q_ent + g.q_off + (int64_t)ii * g.n_q;
    double a = 0.0;
    for (int r = 0; r < rows; ++r) {
      const int k = tl.y * kT + r;
      if (!w.flag[g.h_off + k]) continue;
      double zr = 0.0, zi = 0.0;
      for (int e = 0; e < g.n_q && q[e].s >= 0; ++e) {
        const double2 p = pn[(int64_t)k * g.n + q[e].s];
        const double br = p.x / w.hbar, bi = p.y / w.hbar;
        zr += q[e].re * br - q[e].im * bi;
        zi += q[e].re * bi + q[e].im * br;
      }
      a += n2(zr, zi);
    }
    w.part_e[g.te_off + (int64_t)tl.y * g.n_sc + ii] = a;
  }
}

__global__ __launch_bounds__(kT) void k_state_fin(hops_sel_work_t w) {
  const int i = blockIdx.x;
  const hops_sel_seg_t g = w.segs[i];
  for (int ii = (int)threadIdx.x; ii < g.n_sc; ii += kT) {
    double a = 0.0;
    for (int t = 0; t < g.n_tiles; ++t) a += w.part_e[g.te_off + (int64_t)t * g.n_sc + ii];
    w.e_sc[g.esc_off + ii] = a;
  }
}

__global__ __launch_bounds__(kT) void k_gather_psi(hops_sel_work_t w, double *out, const int64_t *off) {
  const int i = blockIdx.x;
  const hops_sel_seg_t g = w.segs[i];
  const double2 *pn = reinterpret_cast<const double2 *>(w.pn) + g.pn_off;
  double2 *o = reinterpret_cast<double2 *>(out) + off[i];
  for (int s = (int)threadIdx.x; s < g.n; s += kT) o[s] = pn[s];
}

}  // namespace

int hops_sel_err(const hops_sel_work_t &w, const int2 *tiles, int n_tiles, hipStream_t st) {
  if (n_tiles <= 0) return 0;
  hipLaunchKernelGGL(k_sel_err, dim3(n_tiles), dim3(kT), 0, st, w, tiles);
  SEL_LAUNCH("k_sel_err");
  return 0;
}

int hops_sel_stable(const hops_sel_work_t &w, const int2 *tiles, int n_tiles, int want_c1, hipStream_t st) {
  if (n_tiles <= 0) return 0;
  hipLaunchKernelGGL(k_sel_stable, dim3(n_tiles), dim3(kT), 0, st, w, tiles, want_c1);
  SEL_LAUNCH("k_sel_stable");
  return 0;
}

int hops_sel_fin1(const hops_sel_work_t &w, int want_c1, hipStream_t st) {
  hipLaunchKernelGGL(k_sel_fin1, dim3(w.n_list), dim3(kT), 0, st, w, want_c1);
  SEL_LAUNCH("k_sel_fin1");
  return 0;
}

int hops_sel_cand(const hops_sel_work_t &w, const int2 *tiles, int n_tiles, hipStream_t st) {
  if (n_tiles <= 0) return 0;
  hipLaunchKernelGGL(k_sel_cand, dim3(n_tiles), dim3(kT), 0, st, w, tiles);
  SEL_LAUNCH("k_sel_cand");
  return 0;
}

int hops_sel_disc(const hops_sel_work_t &w, const int2 *nz_tiles, int n_nz_tiles, const int2 *tiles, int n_tiles,
                  hipStream_t st) {
  if (n_nz_tiles > 0) {
    hipLaunchKernelGGL(k_sel_disc, dim3(n_nz_tiles), dim3(kT), 0, st, w, nz_tiles);
    SEL_LAUNCH("k_sel_disc");
  }
  if (n_tiles > 0) {
    hipLaunchKernelGGL(k_sel_scount, dim3(n_tiles), dim3(kT), 0, st, w, tiles);
    SEL_LAUNCH("k_sel_scount");
  }
  return 0;
}

int hops_sel_fin2(const hops_sel_work_t &w, hipStream_t st) {
  hipLaunchKernelGGL(k_sel_fin2, dim3(w.n_list), dim3(kT), 0, st, w);
  SEL_LAUNCH("k_sel_fin2");
  return 0;
}

int hops_sel_surv(const hops_sel_work_t &w, const int2 *tiles, int n_tiles, hipStream_t st) {
  if (n_tiles <= 0) return 0;
  hipLaunchKernelGGL(k_sel_surv, dim3(n_tiles), dim3(kT), 0, st, w, tiles);
  SEL_LAUNCH("k_sel_surv");
  return 0;
}

int hops_sel_group(const hops_sel_work_t &w, const int2 *sv_tiles, int n_sv_tiles, hipStream_t st) {
  if (n_sv_tiles <= 0) return 0;
  hipLaunchKernelGGL(k_sel_group, dim3(n_sv_tiles), dim3(kT), 0, st, w, sv_tiles);
  SEL_LAUNCH("k_sel_group");
  return 0;
}

int hops_sel_fin3(const hops_sel_work_t &w, hipStream_t st) {
  hipLaunchKernelGGL(k_sel_fin3, dim3(w.n_list), dim3(kT), 0, st, w);
  SEL_LAUNCH("k_sel_fin3");
  return 0;
}

int hops_sel_bound(const hops_sel_work_t &w, hipStream_t st) {
  hipLaunchKernelGGL(k_sel_bcount, dim3(w.n_list), dim3(kT), 0, st, w);
  SEL_LAUNCH("k_sel_bcount");
  hipLaunchKernelGGL(k_sel_boff, dim3(1), dim3(64), 0, st, w);
  SEL_LAUNCH("k_sel_boff");
  hipLaunchKernelGGL(k_sel_bwrite, dim3(w.n_list), dim3(kT), 0, st, w);
  SEL_LAUNCH("k_sel_bwrite");
  return 0;
}

int hops_sel_threshold(const hops_sel_work_t &w, const double *sorted, int kind, hipStream_t st) {
  hipLaunchKernelGGL(k_sel_thr, dim3(w.n_list), dim3(64), 0, st, w, sorted, kind);
  SEL_LAUNCH("k_sel_thr");
  return 0;
}

int hops_sel_sort_keys(void *tmp, size_t *tmp_bytes, const double *in, double *out, size_t size, int segs,
                       const int64_t *beg, const int64_t *end, hipStream_t st) {
  const hipError_t e = rocprim::segmented_radix_sort_keys(tmp, *tmp_bytes, in, out, (unsigned int)size,
                                                          (unsigned int)segs, beg, end, 0, 64, st);
  if (e != hipSuccess) return sel_fail("segmented_radix_sort_keys", e);
  return 0;
}

int hops_sel_sort_pairs(void *tmp, size_t *tmp_bytes, const uint64_t *kin, uint64_t *kout, const uint32_t *vin,
                        uint32_t *vout, size_t size, int segs, const int64_t *beg, const int64_t *end,
                        hipStream_t st) {
  const hipError_t e = rocprim::segmented_radix_sort_pairs(tmp, *tmp_bytes, kin, kout, vin, vout,
                                                           (unsigned int)size, (unsigned int)segs, beg, end, 0,
                                                           64, st);
  if (e != hipSuccess) return sel_fail("segmented_radix_sort_pairs", e);
  return 0;
}

int hops_sel_state_err(const hops_sel_work_t &w, hipStream_t st) {
  hipLaunchKernelGGL(k_state_err, dim3(w.n_list), dim3(kT), 0, st, w);
  SEL_LAUNCH("k_state_err");
  return 0;
}

int hops_sel_state_flux(const hops_sel_work_t &w, const int2 *tiles, int n_tiles, hipStream_t st) {
  if (n_tiles > 0) {
    hipLaunchKernelGGL(k_state_flux, dim3(n_tiles), dim3(kT), 0, st, w, tiles);
    SEL_LAUNCH("k_state_flux");
  }
  hipLaunchKernelGGL(k_state_fin, dim3(w.n_list), dim3(kT), 0, st, w);
  SEL_LAUNCH("k_state_fin");
  return 0;
}

int hops_sel_gather_psi(const hops_sel_work_t &w, double *out, const int64_t *off, hipStream_t st) {
  hipLaunchKernelGGL(k_gather_psi, dim3(w.n_list), dim3(kT), 0, st, w, out, off);
  SEL_LAUNCH("k_gather_psi");
  return 0;
}
