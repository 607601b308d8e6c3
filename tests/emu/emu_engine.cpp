// CPU stand-ins for the engine's device entry points in the emulation build
// of the runtime (tests/test_select_emulation.py): Pn, deletion-error sums of
// a simple form, a deterministic toy step that spreads amplitude over the
// hierarchy (it grows the bases; it is not the HOPS equation of motion), the
// Phi permutation and the segment copies.  Also the emulation's thread-local
// work-item ids, barrier and cross-lane buffer.
#include <hip/hip_runtime.h>
#include <complex>
#include <set>
#include "../../include/hops_mi355x.h"
thread_local emu_u3 threadIdx, blockIdx;
std::barrier<> *g_emu_bar = nullptr;
long g_emu_launches = 0;
static int g_lane_buf[1024];
int emu_readlane(int v, int l) {
  g_lane_buf[threadIdx.x] = v;
  __syncthreads();
  const int r = g_lane_buf[l];
  __syncthreads();
  return r;
}
static char g_msg[512];
extern "C" int hops_internal_set_error(const char *msg) {
  snprintf(g_msg, sizeof(g_msg), "%s", msg ? msg : "");
  return -1;
}
extern "C" const char *hops_last_error(void) { return g_msg; }
using cplx = std::complex<double>;
static std::set<int> slots_of(const uint32_t *t, int n) {
  std::set<int> s;
  for (int i = 0; i < n; ++i) s.insert((int)(t[i] >> 20));
  return s;
}
extern "C" {
int hops_build_super(const hops_op_t *, int, void *) { return 0; }
int hops_slots_build(const hops_op_t *, const int32_t *, int, const uint32_t *, int, int, void *) { return 0; }
int hops_copy_segments(const hops_segment_t *s, int n, void *) {
  for (int i = 0; i < n; ++i) memcpy(s[i].dst, s[i].src, (size_t)s[i].bytes);
  return 0;
}
int hops_gather_u64(const uint64_t *const *src, uint64_t *out, int n, void *) {
  for (int i = 0; i < n; ++i) out[i] = *src[i];
  return 0;
}
// toy step: Phi[k, s] += 0.05 * (k + 1)^-1 * Phi[0, (s + k) % n] * i
int hops_slots_step(const hops_op_t *ops, const hops_batch_t *bts, const int32_t *list, int n_list, const uint32_t *,
                    int, const int32_t *, int, int, void *) {
  for (int q = 0; q < n_list; ++q) {
    const int sl = list[q];
    const hops_op_t &op = ops[sl];
    cplx *phi = (cplx *)bts[sl].phi_d;
    const int n = op.n_state, nh = op.n_hier;
    for (int k = nh - 1; k >= 1; --k)
      for (int s = 0; s < n; ++s) phi[(size_t)k * n + s] += cplx(0, 0.05 / (k + 1)) * phi[(s + k) % n];
    // also feed the first-level rows
    for (int s = 0; s < n; ++s) phi[s] *= cplx(std::cos(0.01), std::sin(0.01));
  }
  return 0;
}
int hops_slots_pn(const hops_op_t *ops, const hops_batch_t *bts, const int32_t *list, int n_list, const int64_t *off,
                  void *out, double *nrm, int, int psi_only, void *) {
  for (int q = 0; q < n_list; ++q) {
    const int sl = list[q];
    const hops_op_t &op = ops[sl];
    const cplx *phi = (const cplx *)bts[sl].phi_d;
    cplx *o = (cplx *)out + off[sl];
    const int n = op.n_state;
    if (psi_only) {
      for (int s = 0; s < n; ++s) o[s] = phi[s];
      continue;
    }
    double a = 0;
    for (int s = 0; s < n; ++s) a += std::norm(phi[s]);
    a = std::sqrt(a);
    for (size_t r = 0; r < (size_t)n * op.n_hier; ++r) o[r] = phi[r] / a;
    if (nrm) nrm[sl] = a;
  }
  return 0;
}
int hops_slots_errors(const hops_op_t *ops, const hops_batch_t *, const int32_t *list, int n_list, const uint32_t *,
                      int, const int32_t *, int, int, const int64_t *off, const void *pn, const int64_t *moff,
                      const uint8_t *mask, const int64_t *eoff, double *e, const double *, double, void *) {
  for (int q = 0; q < n_list; ++q) {
    const int sl = list[q];
    const hops_op_t &op = ops[sl];
    const cplx *p = (const cplx *)pn + off[sl];
    const int n = op.n_state, nh = op.n_hier;
    if (!mask) {
      for (int k = 0; k < nh; ++k) {
        double a = 0;
        for (int s = 0; s < n; ++s) a += std::norm(p[(size_t)k * n + s]);
        e[eoff[sl] + k] = 1e-3 * a;
      }
    } else {
      for (int s = 0; s < n; ++s) {
        double a = 0;
        for (int k = 0; k < nh; ++k)
          if (mask[moff[sl] + k]) a += std::norm(p[(size_t)k * n + s]);
        e[eoff[sl] + s] = 1e-3 * a;
      }
    }
  }
  return 0;
}
int hops_slots_permute(const hops_slot_perm_t *perm, const uint32_t *tiles, int n_tiles, void *) {
  for (int sl : slots_of(tiles, n_tiles)) {
    const hops_slot_perm_t &p = perm[sl];
    const cplx *o = (const cplx *)p.old_phi_d;
    cplx *nw = (cplx *)p.new_phi_d;
    for (int j = 0; j < p.nh_new; ++j)
      for (int q = 0; q < p.n_new; ++q) {
        const int a = p.inv_aux_d[j], b = p.inv_state_d[q];
        nw[(size_t)j * p.n_new + q] = (a >= 0 && b >= 0) ? o[(size_t)a * p.n_old + b] * p.scale : cplx(0, 0);
      }
  }
  return 0;
}
}
