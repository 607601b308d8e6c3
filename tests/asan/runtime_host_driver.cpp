// Host-only exercise of the native runtime (include/hops_runtime.h) for the
// address and undefined-behaviour sanitizers: systems, basis tables, the
// hierarchy and state selection (with the Markovian, Triangular and LongEdge
// flux filters), the threshold helper, and every handle freed.  The device
// entry points of the engine are stubs that fail: the selection path never
// calls them.  Prints "OK <checks>" and exits 0 when every call succeeded.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/hops_mi355x.h"
#include "../../include/hops_runtime.h"

// the engine's device entry points are stubs (engine_stubs.c)
extern "C" const char *stub_last_error(void);

static int g_fail = 0, g_checks = 0;
#define CHECK(cond, ...)                      \
  do {                                        \
    ++g_checks;                               \
    if (!(cond)) {                            \
      std::fprintf(stderr, __VA_ARGS__);      \
      std::fprintf(stderr, " [%s]\n", stub_last_error()); \
      ++g_fail;                               \
    }                                         \
  } while (0)

struct Sys {
  int n, nm;
  std::vector<double> H, g, w, lval, lt;
  std::vector<int32_t> lofm, lptr, lrow, lcol;
};

// chain of n sites, two modes per site on the site projector
static Sys chain(int n, bool ltc, std::mt19937 &rng) {
  Sys s;
  s.n = n;
  s.nm = 2 * n;
  s.H.assign(2 * n * n, 0.0);
  std::normal_distribution<double> nd(0.0, 100.0);
  for (int i = 0; i < n; ++i) {
    s.H[2 * (i * n + i)] = nd(rng);
    if (i + 1 < n) s.H[2 * (i * n + i + 1)] = s.H[2 * ((i + 1) * n + i)] = 50.0;
  }
  for (int i = 0; i < n; ++i) {
    const double gr[2] = {14352.0, 93.0}, gi[2] = {-1750.0, 0.0}, wr[2] = {50.0, 1288.0};
    for (int j = 0; j < 2; ++j) {
      s.g.push_back(gr[j]);
      s.g.push_back(gi[j]);
      s.w.push_back(wr[j]);
      s.w.push_back(0.0);
      s.lofm.push_back(i);
    }
  }
  s.lptr.push_back(0);
  for (int i = 0; i < n; ++i) {
    s.lrow.push_back(i);
    s.lcol.push_back(i);
    s.lval.push_back(1.0);
    s.lval.push_back(0.0);
    s.lptr.push_back((int32_t)s.lrow.size());
    if (ltc) {
      s.lt.push_back(25.0);
      s.lt.push_back(-4.0);
    }
  }
  return s;
}

// auxiliaries of depth <= depth over the given modes, in the reference's
// order: by depth, then the lexicographic sequence of sorted modes
static void aux_list(const std::vector<int> &modes, int depth, std::vector<int32_t> &ptr,
                     std::vector<int32_t> &mode, std::vector<int32_t> &val) {
  ptr.assign(1, 0);
  mode.clear();
  val.clear();
  std::vector<int> seq;
  for (int d = 0; d <= depth; ++d) {
    std::vector<int> idx(d, 0);
    for (;;) {
      // emit the multiset idx as (mode, count) pairs
      for (int i = 0; i < d;) {
        int j = i;
        while (j < d && idx[j] == idx[i]) ++j;
        mode.push_back(modes[idx[i]]);
        val.push_back(j - i);
        i = j;
      }
      ptr.push_back((int32_t)mode.size());
      int p = d - 1;
      while (p >= 0 && idx[p] == (int)modes.size() - 1) --p;
      if (p < 0) break;
      ++idx[p];
      for (int q = p + 1; q < d; ++q) idx[q] = idx[p];
    }
  }
}

int main() {
  std::mt19937 rng(12345);
  std::uniform_real_distribution<double> ud(0.0, 1.0);
  // the threshold helper on edge cases
  {
    double below = -1;
    const double t0 = hops_rt_debug_threshold(nullptr, 0, 1.0, 0.0, &below);
    CHECK(t0 == 0.0, "empty threshold %g", t0);
    std::vector<double> z(17, 0.0);
    hops_rt_debug_threshold(z.data(), (long long)z.size(), 1.0, 0.0, &below);
    std::vector<double> v(5000);
    for (auto &x : v) x = ud(rng) * 1e-6;
    const double t1 = hops_rt_debug_threshold(v.data(), (long long)v.size(), 1e-4, 0.0, &below);
    CHECK(t1 >= 0.0, "threshold %g", t1);
  }
  for (int sys_i = 0; sys_i < 3; ++sys_i) {
    const int n = sys_i == 0 ? 6 : (sys_i == 1 ? 12 : 20);
    Sys s = chain(n, sys_i == 1, rng);
    hops_rt_system_desc_t d{};
    d.n_state = n;
    d.n_hmode = s.nm;
    d.n_lop = n;
    d.H = s.H.data();
    d.g = s.g.data();
    d.w = s.w.data();
    d.l_of_mode = s.lofm.data();
    d.l_ptr = s.lptr.data();
    d.l_row = s.lrow.data();
    d.l_col = s.lcol.data();
    d.l_val = s.lval.data();
    d.lt = s.lt.empty() ? nullptr : s.lt.data();
    void *sys = hops_rt_system_create(&d);
    CHECK(sys != nullptr, "system %d", sys_i);
    if (!sys) continue;
    for (int trial = 0; trial < 4; ++trial) {
      // a contiguous block of states and their modes, depth 1..3
      const int ns = 2 + trial % (n - 1 > 3 ? 3 : n - 1);
      const int s0 = (int)(ud(rng) * (n - ns));
      std::vector<int32_t> states;
      std::vector<int> modes;
      for (int q = 0; q < ns; ++q) {
        states.push_back(s0 + q);
        modes.push_back(2 * (s0 + q));
        modes.push_back(2 * (s0 + q) + 1);
      }
      const int depth = 1 + trial % 3;
      std::vector<int32_t> ap, am, av;
      aux_list(modes, depth, ap, am, av);
      const int nh = (int)ap.size() - 1;
      void *B = hops_rt_debug_basis(sys, 2, ns, states.data(), nh, ap.data(), am.data(), av.data());
      CHECK(B != nullptr, "basis sys %d trial %d", sys_i, trial);
      if (!B) continue;
      int32_t sizes[16] = {0};
      const long long nb = hops_rt_debug_table(B, "sizes", nullptr, 0);
      CHECK(nb > 0 && nb <= (long long)sizeof(sizes), "sizes table %lld", nb);
      if (nb > 0 && nb <= (long long)sizeof(sizes)) hops_rt_debug_table(B, "sizes", sizes, nb);
      CHECK(sizes[1] == nh, "n_hier %d != %d", sizes[1], nh);
      std::vector<double> pn(2 * (size_t)nh * ns), ea(nh), es(ns);
      for (int k = 0; k < nh; ++k) {
        int kd = 0;
        for (int q = ap[k]; q < ap[k + 1]; ++q) kd += av[q];
        double sc = 1.0;
        for (int i = 0; i < kd; ++i) sc *= 0.05;
        for (int q = 0; q < ns; ++q) {
          pn[2 * ((size_t)k * ns + q)] = sc * (ud(rng) - 0.5);
          pn[2 * ((size_t)k * ns + q) + 1] = sc * (ud(rng) - 0.5);
        }
        ea[k] = sc * sc * 1e-6 * ud(rng);
      }
      for (auto &x : es) x = 1e-9 * ud(rng);
      for (int cfg = 0; cfg < 4; ++cfg) {
        hops_rt_traj_param_t p{};
        p.eom = 2;
        p.maxhier = depth + 1;
        p.update_step = 1;
        p.early_steps = 5;
        p.inchworm_cap = 5;
        p.noise_step = 1;
        p.n_t = 10;
        p.record = 1;
        p.delta_a = cfg == 3 ? 0.0 : (cfg == 0 ? 1e-3 : 1e-4);
        p.delta_s = cfg == 2 ? 0.0 : 1e-3;
        p.f_discard = 0.01;
        p.dt = 1.0;
        std::vector<uint8_t> mk(s.nm, 0), fmask(2 * (size_t)s.nm, 0);
        std::vector<int32_t> fkind = {0, 1}, fkmax = {1, 2};
        if (cfg == 1) {
          for (int m = 1; m < s.nm; m += 2) mk[m] = 1;
          p.markov = mk.data();
        }
        if (cfg >= 1) {
          for (int m = 0; m < s.nm; ++m) {
            fmask[m] = (m % 2) == 0;
            fmask[s.nm + m] = (m % 2) == 1;
          }
          p.n_filters = 2;
          p.filter_kind = fkind.data();
          p.filter_kmax = fkmax.data();
          p.filter_mask = fmask.data();
        }
        std::vector<uint8_t> stable(nh, 0);
        const int rc = hops_rt_debug_select(B, &p, pn.data(), ea.data(), 1.0, stable.data());
        CHECK(rc == 0, "select sys %d trial %d cfg %d rc %d", sys_i, trial, cfg, rc);
        if (rc != 0) continue;
        if (p.delta_s > 0) {
          const int rs = hops_rt_debug_select_states(B, es.data());
          CHECK(rs == 0, "select_states rc %d", rs);
        }
        int32_t nso = 0, nao = 0;
        const long long cap = 1 << 22;
        std::vector<int32_t> so(cap), aop(cap), aom(cap), aov(cap);
        const int rr = hops_rt_debug_result(B, &nso, so.data(), &nao, aop.data(), aom.data(), aov.data(), cap);
        CHECK(rr == 0, "result rc %d", rr);
        // the state half runs only with state adaptivity; the batch path
        // keeps the basis states otherwise
        CHECK((p.delta_s <= 0 || nso >= 1) && nao >= 1, "empty result %d %d", nso, nao);
        int nst = 0;
        for (int k = 0; k < nh; ++k) nst += stable[k];
        CHECK(nst >= 1 && nst <= nh, "stable count %d", nst);
      }
      CHECK(hops_rt_debug_free(B) == 0, "free basis");
    }
    CHECK(hops_rt_system_destroy(sys) == 0, "destroy system");
  }
  if (g_fail) {
    std::fprintf(stderr, "FAILED %d of %d checks\n", g_fail, g_checks);
    return 1;
  }
  std::printf("OK %d\n", g_checks);
  return 0;
}
