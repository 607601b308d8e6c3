// Driver of tests/test_select_emulation.py: two batches of the same
// trajectories (a chain of n sites, two modes per site), one selecting its
// bases with the device kernels (emulated), one on the host threads, stepped
// side by side; every slot's state list, auxiliary list, Phi and records must
// be bitwise the same after every call.
//   select_driver n steps delta_a delta_s markov(0/1) slots
#include <hip/hip_runtime.h>

#include <complex>
#include <cstdio>
#include <vector>

#include "../../include/hops_mi355x.h"
#include "../../include/hops_runtime.h"

extern "C" const char *hops_last_error(void);
extern long g_emu_launches;

namespace {

struct Sys {
  std::vector<double> H, g, w, lval;
  std::vector<int32_t> lom, lptr, lrow, lcol;
  hops_rt_system_desc_t d{};
};

void make_chain(int n, Sys &s) {
  const int nhm = 2 * n;
  s.H.assign(2 * n * n, 0.0);
  for (int i = 0; i < n; ++i) {
    s.H[2 * (i * n + i)] = 100.0 * std::sin(1.3 * i);
    if (i + 1 < n) s.H[2 * (i * n + i + 1)] = s.H[2 * ((i + 1) * n + i)] = 50.0;
  }
  s.g.assign(2 * nhm, 0.0);
  s.w.assign(2 * nhm, 0.0);
  s.lom.assign(nhm, 0);
  for (int m = 0; m < nhm; ++m) {
    s.g[2 * m] = m % 2 ? 200.0 : 5000.0;
    s.g[2 * m + 1] = m % 2 ? 30.0 : -1000.0;
    s.w[2 * m] = m % 2 ? 300.0 : 50.0;
    s.w[2 * m + 1] = m % 2 ? 2.0 : 5.0;
    s.lom[m] = m / 2;
  }
  s.lptr.assign(n + 1, 0);
  s.lrow.assign(n, 0);
  s.lcol.assign(n, 0);
  s.lval.assign(2 * n, 0.0);
  for (int l = 0; l < n; ++l) {
    s.lptr[l] = l;
    s.lrow[l] = s.lcol[l] = l;
    s.lval[2 * l] = 1.0;
  }
  s.lptr[n] = n;
  s.d.n_state = n;
  s.d.n_hmode = nhm;
  s.d.n_lop = n;
  s.d.H = s.H.data();
  s.d.g = s.g.data();
  s.d.w = s.w.data();
  s.d.l_of_mode = s.lom.data();
  s.d.l_ptr = s.lptr.data();
  s.d.l_row = s.lrow.data();
  s.d.l_col = s.lcol.data();
  s.d.l_val = s.lval.data();
}

int fail(const char *what) {
  printf("FAIL %s: %s\n", what, hops_last_error());
  return 1;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 7) return 2;
  const int n = atoi(argv[1]), steps = atoi(argv[2]);
  const double da = atof(argv[3]), ds = atof(argv[4]);
  const int markov = atoi(argv[5]), B = atoi(argv[6]);
  // optional: every slot starts from a STATIC_BASIS (states 0..2 and the
  // first-level auxiliaries of modes 0 and 1) instead of its first selection
  const bool with_static = argc > 7 && atoi(argv[7]) != 0;
  Sys S;
  make_chain(n, S);
  void *sys = hops_rt_system_create(&S.d);
  if (!sys) return fail("system");
  const size_t arena = (size_t)192 << 20;
  std::vector<void *> ar(2);
  std::vector<void *> bat(2);
  std::vector<uint8_t> mk(2 * n, 0);
  for (int m = 0; m < 2 * n; m += 3) mk[m] = 1;
  std::vector<double> z(1 << 12, 0.0);
  for (int v = 0; v < 2; ++v) {
    ar[v] = aligned_alloc(256, arena);
    memset(ar[v], 0x5a, arena);  // stale device memory: reads of unwritten words give wild values
    bat[v] = hops_rt_batch_create(HOPS_PREC_FP64, ar[v], arena, B, 4, nullptr);
    if (!bat[v]) return fail("batch");
    if (hops_rt_batch_option(bat[v], "device_select", v == 0 ? 1 : 0) < 0) return fail("option");
    for (int s = 0; s < B; ++s) {
      std::vector<double> psi0(2 * n, 0.0);
      psi0[2 * ((n / 2 + s) % n)] = 1.0;
      psi0[2 * ((n / 2 + 2 * s + 1) % n) + 1] = 0.5;
      hops_rt_traj_param_t p{};
      p.eom = HOPS_EOM_NORMALIZED_NONLINEAR;
      p.maxhier = 5;
      p.update_step = 1 + (s % 2);
      p.early_integrator = 0;
      p.early_steps = 3;
      p.inchworm_cap = 5;
      p.noise_step = 1;
      p.n_t = 1000;
      p.record = 1;
      p.delta_a = da * (1 + 0.3 * s);
      p.delta_s = ds;
      p.f_discard = s % 2 ? 0.01 : 0.0;
      p.dt = 1.0;
      p.markov = markov ? mk.data() : nullptr;
      const int slot = hops_rt_batch_add(bat[v], sys, &p, psi0.data(), z.data(), nullptr);
      if (slot < 0) return fail("add");
      if (with_static) {
        const int32_t st[3] = {0, 1, 2}, ap[3] = {0, 1, 2}, am[2] = {0, 1}, av[2] = {1, 1};
        if (hops_rt_slot_static_basis(bat[v], slot, 3, st, 2, ap, am, av) < 0) return fail("static basis");
      }
    }
  }
  long long nh_max = 0, rec_total = 0;
  for (int call = 0; call < 3; ++call) {
    for (int v = 0; v < 2; ++v)
      if (hops_rt_batch_step(bat[v], steps, nullptr, 0) < 0) return fail("step");
    for (int s = 0; s < B; ++s) {
      long long i0[9], i1[9];
      hops_rt_slot_info(bat[0], s, i0);
      hops_rt_slot_info(bat[1], s, i1);
      if (i0[0] != i1[0] || i0[1] != i1[1] || i0[4] != i1[4]) {
        printf("FAIL slot %d call %d: n_state %lld/%lld n_hier %lld/%lld records %lld/%lld\n", s, call, i0[0],
               i1[0], i0[1], i1[1], i0[4], i1[4]);
        return 1;
      }
      nh_max = std::max(nh_max, i0[1]);
      const int ns = (int)i0[0], nh = (int)i0[1];
      std::vector<int32_t> st0(ns), st1(ns);
      hops_rt_slot_states(bat[0], s, st0.data());
      hops_rt_slot_states(bat[1], s, st1.data());
      if (st0 != st1) {
        printf("FAIL slot %d call %d: state lists differ\n", s, call);
        return 1;
      }
      std::vector<int32_t> p0(nh + 1), p1(nh + 1), m0(nh * 8), m1(nh * 8), v0(nh * 8), v1(nh * 8);
      hops_rt_slot_aux(bat[0], s, p0.data(), m0.data(), v0.data());
      hops_rt_slot_aux(bat[1], s, p1.data(), m1.data(), v1.data());
      m0.resize(p0[nh]);
      m1.resize(p1[nh]);
      v0.resize(p0[nh]);
      v1.resize(p1[nh]);
      if (p0 != p1 || m0 != m1 || v0 != v1) {
        printf("FAIL slot %d call %d: auxiliary lists differ\n", s, call);
        return 1;
      }
      std::vector<double> f0(2 * (size_t)ns * nh), f1(2 * (size_t)ns * nh);
      hops_rt_slot_phi(bat[0], s, f0.data());
      hops_rt_slot_phi(bat[1], s, f1.data());
      if (memcmp(f0.data(), f1.data(), f0.size() * 8) != 0) {
        printf("FAIL slot %d call %d: Phi differs\n", s, call);
        return 1;
      }
      rec_total += i0[4];
    }
  }
  for (int v = 0; v < 2; ++v) {
    hops_rt_batch_destroy(bat[v]);
    free(ar[v]);
  }
  hops_rt_system_destroy(sys);
  printf("OK slots %d largest hierarchy %lld records %lld kernel launches %ld\n", B, nh_max, rec_total,
         g_emu_launches);
  return 0;
}
