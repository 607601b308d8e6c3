// hops_runtime.cpp — native host runtime for ragged batches of adaptive HOPS
// trajectories on MI355X (include/hops_runtime.h).
//
// Per batch step the GPU runs the RK4 step of every slot, the derivative of
// the error bound and its per-auxiliary / per-state sums (slot kernels of
// hops_engine.hip); the host worker threads run the set logic of the basis
// update for every slot in parallel (thresholds, boundary auxiliaries, state
// basis, connection tables of the new basis); the GPU then builds the new
// operators and moves every Phi into its new basis in one launch each.
//
// Restated from the reference (MesoHOPS v1.6.0, src/mesohops/):
//   basis/hops_basis.py:117-292      define_basis / update_basis
//   basis/hops_basis.py:294-998      hierarchy and state error bounds
//   basis/hops_basis.py:1000-1151    basis from error list, M2/X2/K2 arrays
//   basis/basis_functions.py:8-41    error threshold
//   basis/hops_fluxfilters.py        flux filters (incl. Markovian)
//   basis/hops_hierarchy.py, hops_aux.py, hops_system.py, hops_modes.py
//                                    auxiliary order, connections, relative
//                                    indexing of states, modes, L-operators
//   trajectory/hops_trajectory.py:253-336, :401-533, :636-714
//                                    initial basis, time loop, INCH_WORM
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <complex>
#include <chrono>
#include <condition_variable>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <functional>
#include <malloc.h>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <type_traits>
#include <unordered_map>
#include <vector>

#include "../../include/hops_mi355x.h"
#include "../../include/hops_runtime.h"
#include "hops_select.h"

extern "C" int hops_internal_set_error(const char *msg);

namespace {

using cplx = std::complex<double>;
constexpr double kHbar = 33357.179 / (2.0 * 3.14159265358979323846);  // util/physical_constants.py:21-28
constexpr int kSlotShift = 20;

int rt_fail(const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  return hops_internal_set_error(buf);
}

#define RT_HIP(expr)                                                          \
  do {                                                                        \
    hipError_t e_ = (expr);                                                   \
    if (e_ != hipSuccess) return rt_fail("%s failed: %s", #expr, hipGetErrorString(e_)); \
  } while (0)
#define RT_TRY(expr)           \
  do {                         \
    int rc_ = (expr);          \
    if (rc_ < 0) return rc_;   \
  } while (0)

// ---------------------------------------------------------------------------
// Auxiliary multi-index: packed (mode << 8 | depth) entries sorted by mode.
// ---------------------------------------------------------------------------
typedef std::u32string Aux;

// complex products without the inf/nan recovery path of std::complex (the
// values here are finite); same formula, same rounding
inline cplx cm(cplx a, cplx b) {
  return cplx(a.real() * b.real() - a.imag() * b.imag(), a.real() * b.imag() + a.imag() * b.real());
}
inline cplx cmi(cplx a) { return cplx(a.imag(), -a.real()); }  // -i a

inline int aux_mode(char32_t e) { return (int)(e >> 8); }
inline int aux_val(char32_t e) { return (int)(e & 0xff); }
inline char32_t aux_pack(int m, int v) { return (char32_t)(((uint32_t)m << 8) | (uint32_t)v); }

int aux_depth(const Aux &a) {
  int d = 0;
  for (char32_t e : a) d += aux_val(e);
  return d;
}
int aux_get(const Aux &a, int m) {
  for (char32_t e : a)
    if (aux_mode(e) == m) return aux_val(e);
  return 0;
}
// a + step e_m (step = +-1); returns false if the depth would go negative
bool aux_step(const Aux &a, int m, int step, Aux &out) {
  out.clear();
  bool done = false;
  for (char32_t e : a) {
    const int mm = aux_mode(e);
    if (!done && mm > m) {
      if (step < 0) return false;
      out.push_back(aux_pack(m, 1));
      done = true;
    }
    if (mm == m) {
      const int v = aux_val(e) + step;
      if (v > 0) out.push_back(aux_pack(m, v));
      done = true;
    } else {
      out.push_back(e);
    }
  }
  if (!done) {
    if (step < 0) return false;
    out.push_back(aux_pack(m, 1));
  }
  return true;
}
// b == a + dir * e_m exactly (dir = +1 or -1): the same (mode, depth)
// entries except mode m, whose depth differs by dir.  Confirms a Zobrist-key
// hit against the auxiliaries themselves.
bool is_step(const Aux &a, const Aux &b, int m, int dir) {
  size_t ia = 0, ib = 0;
  int va = 0, vb = 0;
  while (ia < a.size() || ib < b.size()) {
    const int ma = ia < a.size() ? aux_mode(a[ia]) : INT32_MAX;
    const int mb = ib < b.size() ? aux_mode(b[ib]) : INT32_MAX;
    if (ma == m) {
      va = aux_val(a[ia++]);
      continue;
    }
    if (mb == m) {
      vb = aux_val(b[ib++]);
      continue;
    }
    if (ma != mb || a[ia] != b[ib]) return false;
    ++ia;
    ++ib;
  }
  return vb == va + dir;
}

// Reference order (basis/hops_aux.py:171-201, :460-480): depth, then the
// sorted mode sequence (mode m repeated k_m times) lexicographically.
bool aux_less(const Aux &a, const Aux &b) {
  const int da = aux_depth(a), db = aux_depth(b);
  if (da != db) return da < db;
  size_t ia = 0, ib = 0;
  int ra = a.empty() ? 0 : aux_val(a[0]), rb = b.empty() ? 0 : aux_val(b[0]);
  while (ia < a.size() && ib < b.size()) {
    const int ma = aux_mode(a[ia]), mb = aux_mode(b[ib]);
    if (ma != mb) return ma < mb;
    if (--ra == 0 && ++ia < a.size()) ra = aux_val(a[ia]);
    if (--rb == 0 && ++ib < b.size()) rb = aux_val(b[ib]);
  }
  return false;
}

// ---------------------------------------------------------------------------
// System (basis/system_functions.py:10-171)
// ---------------------------------------------------------------------------
struct LOp {
  std::vector<int> row, col;
  std::vector<cplx> val;
  // the same entries by column (cp/ri/cv) and by row (rp/ci/rv), each in
  // ascending order of the other index, over the full state space
  std::vector<int> cp, ri, rp, ci;
  std::vector<cplx> cv, rv;
  bool off = false;  // an entry off the diagonal (a Peierls-type coupling)
};

struct System {
  int n = 0, n_hm = 0, n_l = 0;
  std::vector<cplx> H;  // n x n
  std::vector<cplx> G, W;
  std::vector<int> l_of_mode;
  std::vector<LOp> L;
  std::vector<cplx> lt;
  std::vector<std::vector<int>> modes_by_state, l_by_state, coupled, dest;
  std::vector<int> site_of_mode;  // the single state of a site-projector mode, or -1
  bool diag_L = true;  // no L-operator has an entry off the diagonal
  std::vector<uint64_t> zob, zob2;  // two independent Zobrist tables [n_hm * 256]
};

// Identity of an auxiliary: two independent 64-bit Zobrist hashes of its
// (mode, depth) entries; a +-e_m step updates them in O(1).
struct AKey {
  uint64_t a = 0, b = 0;
  bool operator==(const AKey &o) const { return a == o.a && b == o.b; }
};
inline AKey key_of(const System &S, const Aux &x) {
  AKey k;
  for (char32_t e : x) {
    k.a ^= S.zob[(size_t)aux_mode(e) * 256 + aux_val(e)];
    k.b ^= S.zob2[(size_t)aux_mode(e) * 256 + aux_val(e)];
  }
  return k;
}
// key of the auxiliary whose depth in mode m changes from v0 to v1
inline AKey key_step(const System &S, AKey k, int m, int v0, int v1) {
  const size_t z0 = (size_t)m * 256 + v0, z1 = (size_t)m * 256 + v1;
  k.a ^= S.zob[z0] ^ S.zob[z1];
  k.b ^= S.zob2[z0] ^ S.zob2[z1];
  return k;
}

// open-addressing map AKey -> auxiliary index
struct KeyTable {
  std::vector<AKey> keys;
  std::vector<int32_t> vals;
  size_t mask = 0;
  static size_t hh(const AKey &k) { return (size_t)(k.a ^ (k.b * 0x9E3779B97F4A7C15ull)); }
  void reset(size_t n) {
    size_t cap = 16;
    while (cap < 2 * n) cap <<= 1;
    keys.assign(cap, AKey());
    vals.assign(cap, -1);
    mask = cap - 1;
  }
  // inserts k -> v; returns -1, or the value already held under key k (which
  // is then kept)
  int insert(const AKey &k, int v) {
    size_t h = hh(k) & mask;
    while (vals[h] >= 0 && !(keys[h] == k)) h = (h + 1) & mask;
    if (vals[h] >= 0) return vals[h];
    keys[h] = k;
    vals[h] = v;
    return -1;
  }
  int find(const AKey &k) const {
    if (vals.empty()) return -1;
    size_t h = hh(k) & mask;
    while (vals[h] >= 0) {
      if (keys[h] == k) return vals[h];
      h = (h + 1) & mask;
    }
    return -1;
  }
};

// ---------------------------------------------------------------------------
// Tables of one basis (hops_op_t, include/hops_mi355x.h) — the same arrays,
// values and order as mesohops_amd/basis.py::BasisTables.
// ---------------------------------------------------------------------------
struct RedL {  // an L-operator reduced to the state list: CSR and CSC
  std::vector<int32_t> rp, ci, cp, ri;
  std::vector<cplx> rv, cv;
};

struct Basis {
  const System *sys = nullptr;
  int eom = 2;
  std::vector<int> states;
  std::vector<Aux> aux;
  KeyTable index;  // auxiliary -> position
  int n = 0, nh = 0, nm = 0, nl = 0, na = 0, n_pat = 0, pat_w = 0, ell_w = 0;
  bool has_lt = false;
  bool has_off = false;  // an L-operator of the basis has entries off the diagonal
  std::vector<int> modes, rel_mode, absL, mode_lrel, mode_lact, act_abs, act_lrel, pos;
  std::vector<RedL> L;
  std::vector<uint64_t> aux_hash, aux_hash2;
  std::vector<int> depth;
  // device tables
  std::vector<int32_t> aux_ptr, aux_mode_t, aux_val_t, up_ptr, up_aux, up_mode, dn_ptr, dn_aux, dn_mode, dn_val,
      lr_ptr, lr_col, lc_ptr, lc_row, mode_lrel32, mode_lact32, mode_abs32, abs_lact, act_lrel32, act_zrow,
      am_ptr, am_mode, la_ptr, la_row, la_col, lq_ptr, lq_row, lq_col, first_aux, first_lact, first_mode,
      pat_ptr, pat_col, pc_ptr, pc_act, pc_kind;
  std::vector<cplx> lr_val, lc_val, mode_g, mode_w, abs_g, abs_w, la_val, lq_val, lt, pat_h, pc_val;
  // destroy the contents of o on this thread
  void swap_out(Basis &o) {
    Basis t(std::move(o));
    o = Basis();
  }
};

int build_basis(const System &S, int eom, std::vector<int> states, std::vector<Aux> aux_in, Basis &B) {
  B = Basis();
  B.sys = &S;
  B.eom = eom;
  B.aux = std::move(aux_in);  // ordered in place: no copy of the auxiliaries
  std::vector<Aux> &aux = B.aux;
  std::sort(states.begin(), states.end());
  states.erase(std::unique(states.begin(), states.end()), states.end());
  {
    // reference order.  A new basis arrives as the kept auxiliaries of the
    // old one (already in order) followed by a few new ones: sort the tail,
    // merge, drop duplicates.
    auto mid = std::is_sorted_until(aux.begin(), aux.end(), aux_less);
    if (mid != aux.end()) {
      std::vector<std::pair<int, int>> key(aux.end() - mid);
      for (size_t i = 0; i < key.size(); ++i) key[i] = {aux_depth(mid[i]), (int)i};
      std::sort(key.begin(), key.end(), [&](const std::pair<int, int> &x, const std::pair<int, int> &y) {
        if (x.first != y.first) return x.first < y.first;
        return aux_less(mid[x.second], mid[y.second]);
      });
      std::vector<Aux> tail;
      tail.reserve(key.size());
      for (auto &k : key) tail.push_back(std::move(mid[k.second]));
      std::move(tail.begin(), tail.end(), mid);
      std::inplace_merge(aux.begin(), mid, aux.end(), aux_less);
    }
    aux.erase(std::unique(aux.begin(), aux.end()), aux.end());
  }
  if (aux.empty() || !aux[0].empty()) return rt_fail("Zero Vector not contained in list_aux!");
  if (states.empty()) return rt_fail("empty state basis");
  B.states = states;
  const int n = (int)states.size(), nh = (int)aux.size();
  B.n = n;
  B.nh = nh;
  B.index.reset(nh);
  B.aux_hash.resize(nh);
  B.aux_hash2.resize(nh);
  B.depth.resize(nh);
  for (int i = 0; i < nh; ++i) {
    const AKey k = key_of(S, aux[i]);
    B.aux_hash[i] = k.a;
    B.aux_hash2[i] = k.b;
    B.depth[i] = aux_depth(aux[i]);
    // the list holds no duplicates, so an equal key is a collision of two
    // different auxiliaries: refused rather than merged
    if (B.index.insert(k, i) >= 0) return rt_fail("two auxiliaries of one basis share a 128-bit Zobrist key");
  }
  // modes: in use by the aux list, or of the states (hops_modes.py:111-171)
  std::vector<char> mflag(S.n_hm, 0);
  for (const Aux &a : aux)
    for (char32_t e : a) {
      if (aux_mode(e) >= S.n_hm) return rt_fail("auxiliary mode %d out of range", aux_mode(e));
      mflag[aux_mode(e)] = 1;
    }
  std::vector<char> in_use = mflag;
  for (int s : states) {
    if (s < 0 || s >= S.n) return rt_fail("state %d out of range", s);
    for (int m : S.modes_by_state[s]) mflag[m] = 1;
  }
  B.rel_mode.assign(S.n_hm, -1);
  for (int m = 0; m < S.n_hm; ++m)
    if (mflag[m]) {
      B.rel_mode[m] = (int)B.modes.size();
      B.modes.push_back(m);
    }
  const int nm = (int)B.modes.size();
  B.nm = nm;
  std::vector<int> relL(S.n_l, -1);
  {
    std::vector<char> lf(S.n_l, 0);
    for (int m : B.modes) lf[S.l_of_mode[m]] = 1;
    for (int l = 0; l < S.n_l; ++l)
      if (lf[l]) {
        relL[l] = (int)B.absL.size();
        B.absL.push_back(l);
      }
  }
  const int nl = (int)B.absL.size();
  B.nl = nl;
  for (int l : B.absL) B.has_off |= S.L[l].off;
  for (int m : B.modes) B.mode_lrel.push_back(relL[S.l_of_mode[m]]);
  {
    std::vector<char> af(S.n_l, 0);
    for (int s : states)
      for (int l : S.l_by_state[s]) af[l] = 1;
    for (int l = 0; l < S.n_l; ++l)
      if (af[l]) {
        if (relL[l] < 0) return rt_fail("an active L-operator without a hierarchy mode");
        B.act_abs.push_back(l);
      }
  }
  const int na = (int)B.act_abs.size();
  B.na = na;
  std::vector<int> lact_of_lrel(nl, -1);
  for (int j = 0; j < na; ++j) {
    B.act_lrel.push_back(relL[B.act_abs[j]]);
    lact_of_lrel[relL[B.act_abs[j]]] = j;
  }
  for (int i = 0; i < nm; ++i) B.mode_lact.push_back(lact_of_lrel[B.mode_lrel[i]]);

  // reduced L-operators (hops_system.py:318-354): duplicates summed, sorted
  B.pos.assign(S.n, -1);
  for (int i = 0; i < n; ++i) B.pos[states[i]] = i;
  B.L.resize(nl);
  for (int j = 0; j < nl; ++j) {
    const LOp &op = S.L[B.absL[j]];
    std::map<std::pair<int, int>, cplx> ent;
    for (size_t q = 0; q < op.row.size(); ++q) {
      const int r = B.pos[op.row[q]], c = B.pos[op.col[q]];
      if (r >= 0 && c >= 0) ent[{r, c}] += op.val[q];
    }
    RedL &R = B.L[j];
    R.rp.assign(n + 1, 0);
    R.cp.assign(n + 1, 0);
    for (auto &kv : ent) {
      R.rp[kv.first.first + 1]++;
      R.cp[kv.first.second + 1]++;
    }
    for (int i = 0; i < n; ++i) {
      R.rp[i + 1] += R.rp[i];
      R.cp[i + 1] += R.cp[i];
    }
    R.ci.resize(ent.size());
    R.rv.resize(ent.size());
    R.ri.resize(ent.size());
    R.cv.resize(ent.size());
    std::vector<int> rf(R.rp.begin(), R.rp.end() - 1), cf(R.cp.begin(), R.cp.end() - 1);
    for (auto &kv : ent) {  // map order: row-major, columns ascending
      const int r = kv.first.first, c = kv.first.second;
      R.ci[rf[r]] = c;
      R.rv[rf[r]++] = kv.second;
    }
    std::vector<std::pair<std::pair<int, int>, cplx>> bycol(ent.begin(), ent.end());
    std::stable_sort(bycol.begin(), bycol.end(), [](const auto &x, const auto &y) {
      return x.first.second != y.first.second ? x.first.second < y.first.second : x.first.first < y.first.first;
    });
    for (auto &kv : bycol) {
      const int r = kv.first.first, c = kv.first.second;
      R.ri[cf[c]] = r;
      R.cv[cf[c]++] = kv.second;
    }
  }
  // aux CSR
  B.aux_ptr.assign(nh + 1, 0);
  for (int i = 0; i < nh; ++i) {
    B.aux_ptr[i + 1] = B.aux_ptr[i] + (int)aux[i].size();
    for (char32_t e : aux[i]) {
      B.aux_mode_t.push_back(aux_mode(e));
      B.aux_val_t.push_back(aux_val(e));
    }
  }
  // connections (hops_hierarchy.py:470-523): j = i + e_m, as flat lists
  // sorted by (aux, relative mode).  The down list comes out in that order
  // (j ascending, the entries of an auxiliary by mode); the up list is
  // placed by a counting sort on i, and each aux's few entries are then
  // ordered by mode.
  {
    std::vector<int32_t> up_i, up_j, up_m;
    up_i.reserve((size_t)nh * 4);
    up_j.reserve((size_t)nh * 4);
    up_m.reserve((size_t)nh * 4);
    B.dn_ptr.assign(nh + 1, 0);
    for (int j = 0; j < nh; ++j) {
      const AKey kj{B.aux_hash[j], B.aux_hash2[j]};
      for (char32_t e : aux[j]) {
        const int m = aux_mode(e), v = aux_val(e);
        const int i = B.index.find(key_step(S, kj, m, v, v - 1));
        if (i < 0) continue;
        if (!is_step(aux[i], aux[j], m, 1)) return rt_fail("Zobrist key collision between auxiliaries");
        up_i.push_back(i);
        up_j.push_back(j);
        up_m.push_back(B.rel_mode[m]);
        B.dn_aux.push_back(i);
        B.dn_mode.push_back(B.rel_mode[m]);
        B.dn_val.push_back(v);
      }
      B.dn_ptr[j + 1] = (int)B.dn_aux.size();
    }
    B.up_ptr.assign(nh + 1, 0);
    for (int32_t i : up_i) B.up_ptr[i + 1]++;
    for (int i = 0; i < nh; ++i) B.up_ptr[i + 1] += B.up_ptr[i];
    B.up_aux.resize(up_i.size());
    B.up_mode.resize(up_i.size());
    std::vector<int32_t> fill(B.up_ptr.begin(), B.up_ptr.end() - 1);
    for (size_t q = 0; q < up_i.size(); ++q) {
      const int32_t at = fill[up_i[q]]++;
      B.up_aux[at] = up_j[q];
      B.up_mode[at] = up_m[q];
    }
    for (int i = 0; i < nh; ++i)
      for (int a = B.up_ptr[i] + 1; a < B.up_ptr[i + 1]; ++a)
        for (int c = a; c > B.up_ptr[i] && B.up_mode[c - 1] > B.up_mode[c]; --c) {
          std::swap(B.up_mode[c - 1], B.up_mode[c]);
          std::swap(B.up_aux[c - 1], B.up_aux[c]);
        }
  }
  // per relative L: row and column CSR tables with running bases
  B.lr_ptr.assign((size_t)nl * (n + 1), 0);
  B.lc_ptr.assign((size_t)nl * (n + 1), 0);
  int base_r = 0, base_c = 0;
  for (int j = 0; j < nl; ++j) {
    const RedL &R = B.L[j];
    for (int i = 0; i <= n; ++i) {
      B.lr_ptr[(size_t)j * (n + 1) + i] = R.rp[i] + base_r;
      B.lc_ptr[(size_t)j * (n + 1) + i] = R.cp[i] + base_c;
    }
    base_r += (int)R.ci.size();
    base_c += (int)R.ri.size();
    B.lr_col.insert(B.lr_col.end(), R.ci.begin(), R.ci.end());
    B.lr_val.insert(B.lr_val.end(), R.rv.begin(), R.rv.end());
    B.lc_row.insert(B.lc_row.end(), R.ri.begin(), R.ri.end());
    B.lc_val.insert(B.lc_val.end(), R.cv.begin(), R.cv.end());
  }
  // ELL width: densest row of K2_kp1 + K2_km1.  Row (i, s) holds one entry
  // per nonzero of row s of L_m for each up connection of i in mode m and of
  // column s for each down connection; the nonzero rows / columns of every
  // reduced L are listed once, so an aux costs its connections, not n of them.
  {
    std::vector<std::vector<std::pair<int, int>>> nzr(nl), nzc(nl);
    for (int j = 0; j < nl; ++j)
      for (int s = 0; s < n; ++s) {
        const RedL &R = B.L[j];
        if (R.rp[s + 1] > R.rp[s]) nzr[j].push_back({s, R.rp[s + 1] - R.rp[s]});
        if (R.cp[s + 1] > R.cp[s]) nzc[j].push_back({s, R.cp[s + 1] - R.cp[s]});
      }
    std::vector<int> cnt(n, 0);
    int w = 0;
    for (int i = 0; i < nh; ++i) {
      for (int u = B.up_ptr[i]; u < B.up_ptr[i + 1]; ++u)
        for (auto &e : nzr[B.mode_lrel[B.up_mode[u]]]) cnt[e.first] += e.second;
      for (int u = B.dn_ptr[i]; u < B.dn_ptr[i + 1]; ++u)
        for (auto &e : nzc[B.mode_lrel[B.dn_mode[u]]]) cnt[e.first] += e.second;
      for (int u = B.up_ptr[i]; u < B.up_ptr[i + 1]; ++u)
        for (auto &e : nzr[B.mode_lrel[B.up_mode[u]]]) {
          w = std::max(w, cnt[e.first]);
          cnt[e.first] = 0;
        }
      for (int u = B.dn_ptr[i]; u < B.dn_ptr[i + 1]; ++u)
        for (auto &e : nzc[B.mode_lrel[B.dn_mode[u]]]) {
          w = std::max(w, cnt[e.first]);
          cnt[e.first] = 0;
        }
    }
    B.ell_w = w;
  }
  // active L-operators: COO of L and L^2, LTC
  for (int j = 0; j < na; ++j) {
    const cplx c = S.lt.empty() ? cplx(0, 0) : S.lt[B.act_abs[j]];
    B.lt.push_back(c);
    if (c != cplx(0, 0)) B.has_lt = true;
  }
  std::vector<std::vector<std::array<int, 2>>> sq_idx(na);
  std::vector<std::vector<cplx>> sq_val(na);
  B.la_ptr.push_back(0);
  B.lq_ptr.push_back(0);
  for (int j = 0; j < na; ++j) {
    const RedL &R = B.L[B.act_lrel[j]];
    for (int r = 0; r < n; ++r)
      for (int q = R.rp[r]; q < R.rp[r + 1]; ++q) {
        B.la_row.push_back(r);
        B.la_col.push_back(R.ci[q]);
        B.la_val.push_back(R.rv[q]);
      }
    B.la_ptr.push_back((int)B.la_row.size());
    // L @ L, zero results dropped, row-major with sorted columns
    for (int r = 0; r < n; ++r) {
      std::map<int, cplx> acc;
      for (int q = R.rp[r]; q < R.rp[r + 1]; ++q) {
        const int c = R.ci[q];
        for (int q2 = R.rp[c]; q2 < R.rp[c + 1]; ++q2) acc[R.ci[q2]] += R.rv[q] * R.rv[q2];
      }
      for (auto &kv : acc)
        if (kv.second != cplx(0, 0)) {
          sq_idx[j].push_back({r, kv.first});
          sq_val[j].push_back(kv.second);
          B.lq_row.push_back(r);
          B.lq_col.push_back(kv.first);
          B.lq_val.push_back(kv.second);
        }
    }
    B.lq_ptr.push_back((int)B.lq_row.size());
  }
  // modes per active L
  {
    std::vector<std::vector<int>> am(na);
    for (int i = 0; i < nm; ++i)
      if (B.mode_lact[i] >= 0) am[B.mode_lact[i]].push_back(i);
    B.am_ptr.push_back(0);
    for (int j = 0; j < na; ++j) {
      B.am_mode.insert(B.am_mode.end(), am[j].begin(), am[j].end());
      B.am_ptr.push_back((int)B.am_mode.size());
    }
  }
  B.abs_lact.assign(S.n_hm, -2);
  for (int i = 0; i < nm; ++i) B.abs_lact[B.modes[i]] = B.mode_lact[i];
  for (int i = 0; i < nm; ++i) {
    const int fi = B.index.find(key_step(S, AKey(), B.modes[i], 0, 1));
    if (fi >= 0 && !is_step(Aux(), aux[fi], B.modes[i], 1)) return rt_fail("Zobrist key collision between auxiliaries");
    if (fi >= 0 && B.mode_lact[i] >= 0) {
      B.first_aux.push_back(fi);
      B.first_lact.push_back(B.mode_lact[i]);
      B.first_mode.push_back(i);
    }
  }
  // system-block pattern: H, active L, (LTC) L^2
  std::vector<std::vector<int>> prow(n);
  for (int r = 0; r < n; ++r)
    for (int c = 0; c < n; ++c)
      if (std::abs(S.H[(size_t)states[r] * S.n + states[c]]) > 0) prow[r].push_back(c);
  for (int j = 0; j < na; ++j) {
    const RedL &R = B.L[B.act_lrel[j]];
    for (int r = 0; r < n; ++r)
      for (int q = R.rp[r]; q < R.rp[r + 1]; ++q)
        if (std::abs(R.rv[q]) > 0) prow[r].push_back(R.ci[q]);
    if (B.has_lt)
      for (size_t q = 0; q < sq_idx[j].size(); ++q)
        if (std::abs(sq_val[j][q]) > 0) prow[sq_idx[j][q][0]].push_back(sq_idx[j][q][1]);
  }
  B.pat_ptr.assign(n + 1, 0);
  for (int r = 0; r < n; ++r) {
    std::sort(prow[r].begin(), prow[r].end());
    prow[r].erase(std::unique(prow[r].begin(), prow[r].end()), prow[r].end());
    B.pat_ptr[r + 1] = B.pat_ptr[r] + (int)prow[r].size();
    B.pat_w = std::max(B.pat_w, (int)prow[r].size());
    for (int c : prow[r]) {
      B.pat_col.push_back(c);
      B.pat_h.push_back(cplx(0, -1) * S.H[(size_t)states[r] * S.n + states[c]]);
    }
  }
  B.n_pat = (int)B.pat_col.size();
  auto pidx = [&](int r, int c) {
    const int *b = B.pat_col.data() + B.pat_ptr[r], *e = B.pat_col.data() + B.pat_ptr[r + 1];
    return (int)(std::lower_bound(b, e, c) - B.pat_col.data());
  };
  std::vector<std::vector<std::tuple<int, int, cplx>>> contrib(B.n_pat);
  for (int j = 0; j < na; ++j) {
    const RedL &R = B.L[B.act_lrel[j]];
    for (int r = 0; r < n; ++r)
      for (int q = R.rp[r]; q < R.rp[r + 1]; ++q) contrib[pidx(r, R.ci[q])].emplace_back(j, 0, R.rv[q]);
    if (B.has_lt)
      for (size_t q = 0; q < sq_idx[j].size(); ++q)
        contrib[pidx(sq_idx[j][q][0], sq_idx[j][q][1])].emplace_back(j, 1, sq_val[j][q]);
  }
  B.pc_ptr.push_back(0);
  for (auto &c : contrib) {
    for (auto &t : c) {
      B.pc_act.push_back(std::get<0>(t));
      B.pc_kind.push_back(std::get<1>(t));
      B.pc_val.push_back(std::get<2>(t));
    }
    B.pc_ptr.push_back((int)B.pc_act.size());
  }
  for (int i = 0; i < nm; ++i) {
    B.mode_lrel32.push_back(B.mode_lrel[i]);
    B.mode_lact32.push_back(B.mode_lact[i]);
    B.mode_abs32.push_back(B.modes[i]);
    B.mode_g.push_back(S.G[B.modes[i]]);
    B.mode_w.push_back(S.W[B.modes[i]]);
  }
  B.abs_g = S.G;
  B.abs_w = S.W;
  for (int j = 0; j < na; ++j) {
    B.act_lrel32.push_back(B.act_lrel[j]);
    B.act_zrow.push_back(B.act_abs[j]);
  }
  return 0;
}

// ---------------------------------------------------------------------------
// Adaptive selection (basis/hops_basis.py:117-182, :294-998)
// ---------------------------------------------------------------------------
struct Params {
  int eom = 2, maxhier = 3, update_step = 1, early_integrator = 0, early_steps = 5, inchworm_cap = 5;
  int noise_step = 1, n_t = 0, it0 = 0, record = 1;
  double delta_a = 0, delta_s = 0, f_discard = 0, dt = 1;
  std::vector<uint8_t> markov;
  struct UpFilter {
    int kind, kmax;  // kind 0: Triangular, 1: LongEdge
    std::vector<uint8_t> mask;  // [n_hmode] filtered modes
  };
  std::vector<UpFilter> filters;
  std::vector<Aux> aux0;
  bool keep_err = false;  // keep the error terms of a selection (inspection)
  // STATIC_BASIS (trajectory/hops_trajectory.py:299-317): joined to the first
  // basis in place of its selection; with early integrator 2
  // (STATIC_STATE_INCHWORM_HIERARCHY) the early steps keep these states
  bool has_static = false;
  std::vector<int> static_states;
  std::vector<Aux> static_aux;
};

// copies the Markovian and Triangular / LongEdge filter parameters
int copy_filters(const hops_rt_traj_param_t *prm, int n_hm, Params &P) {
  if (prm->markov) P.markov.assign(prm->markov, prm->markov + n_hm);
  P.filters.clear();
  if (prm->n_filters < 0) return rt_fail("n_filters < 0");
  if (prm->n_filters > 0 && (!prm->filter_kind || !prm->filter_kmax || !prm->filter_mask))
    return rt_fail("filter arrays missing");
  for (int i = 0; i < prm->n_filters; ++i) {
    if (prm->filter_kind[i] != 0 && prm->filter_kind[i] != 1) return rt_fail("unknown filter kind %d", prm->filter_kind[i]);
    Params::UpFilter f{prm->filter_kind[i], prm->filter_kmax[i],
                       std::vector<uint8_t>(prm->filter_mask + (size_t)i * n_hm, prm->filter_mask + (size_t)(i + 1) * n_hm)};
    P.filters.push_back(std::move(f));
  }
  return 0;
}

double error_threshold(const std::vector<double> &sorted, double max_err, double offset = 0.0) {
  // basis/basis_functions.py:8-41
  if (sorted.empty()) return 0.0;
  double cum = 0;  // cumsum(sorted) + offset, summed in that order
  for (size_t i = 0; i < sorted.size(); ++i) {
    cum += sorted[i];
    if (cum + offset >= max_err) return i > 0 ? sorted[i - 1] : 0.0;
  }
  return sorted[0] + offset < max_err ? INFINITY : 0.0;
}

// ascending sort of non-negative doubles: LSD radix sort of their bit
// patterns (monotonic for x >= 0), 11-bit digits
void sort_nonneg(std::vector<double> &v) {
  const size_t n = v.size();
  if (n < 512) {
    std::sort(v.begin(), v.end());
    return;
  }
  std::vector<uint64_t> a(n), t(n);
  for (size_t i = 0; i < n; ++i) std::memcpy(&a[i], &v[i], 8);
  std::vector<uint32_t> cnt(2048);
  for (int sh = 0; sh < 64; sh += 11) {
    std::fill(cnt.begin(), cnt.end(), 0);
    for (uint64_t x : a) cnt[(x >> sh) & 2047]++;
    uint32_t sum = 0;
    for (auto &c : cnt) {
      const uint32_t c0 = c;
      c = sum;
      sum += c0;
    }
    for (uint64_t x : a) t[cnt[(x >> sh) & 2047]++] = x;
    a.swap(t);
  }
  for (size_t i = 0; i < n; ++i) std::memcpy(&v[i], &a[i], 8);
}

// error_threshold(sort(v), max_err, offset) for non-negative v, sorting only
// a prefix of the sorted order: the values are binned by binary exponent (a
// bin holds a range of values, and the bins are ordered), the bins are walked
// up to the first one where the running sum reaches max_err, and the values
// of those bins, which are exactly the smallest values of v, are sorted and
// scanned as error_threshold scans them.  The binned sum is summed in another
// order than the scan, so the prefix is taken with a margin and widened to the
// whole of v if the scan does not reach max_err inside it.
double prefix_threshold(const std::vector<double> &v, double max_err, double offset) {
  if (v.empty()) return 0.0;
  constexpr int kBins = 2048;
  uint32_t cnt[kBins] = {};
  double sum[kBins] = {};
  for (double x : v) {
    uint64_t u;
    std::memcpy(&u, &x, 8);
    const int e = (int)(u >> 52) & (kBins - 1);
    cnt[e]++;
    sum[e] += x;
  }
  double run = offset;
  int last = kBins - 1;
  for (int e = 0; e < kBins; ++e) {
    run += sum[e];
    if (cnt[e] && run >= max_err * (1.0 - 1e-6)) {
      last = e;
      break;
    }
  }
  std::vector<double> w;
  if (last == kBins - 1) {
    w = v;
  } else {
    size_t m = 0;
    for (int e = 0; e <= last; ++e) m += cnt[e];
    w.reserve(m);
    for (double x : v) {
      uint64_t u;
      std::memcpy(&u, &x, 8);
      if ((int)(u >> 52) <= last) w.push_back(x);
    }
  }
  sort_nonneg(w);
  double cum = 0;
  size_t i = 0;
  for (; i < w.size(); ++i) {
    cum += w[i];
    if (cum + offset >= max_err) break;
  }
  if (i < w.size()) return i > 0 ? w[i - 1] : 0.0;
  if (w.size() == v.size()) return INFINITY;  // the sum of every value stays below max_err
  std::vector<double> all(v);
  sort_nonneg(all);
  return error_threshold(all, max_err, offset);
}

// The same threshold by radix selection, without sorting the prefix.  The
// sorted order's running sum S_i (rounded at each addition, as the reference's
// cumsum) first reaches max_err - offset at some rank i*; the threshold is the
// value of rank i* - 1.  The bins of the value bits (the binary exponent,
// then 11 mantissa bits at a time) are walked to the bin holding rank i*,
// which is refined until it is small and then sorted.  The sums here are
// added in another order.  Two summation orders of m non-negative values
// differ by at most 2 (m - 1) eps times their sum, so:
//  - the total, whose sign against max_err - offset decides whether any
//    crossing exists, by n eps (total);
//  - the partial sums below the crossing, which stay below max_err - offset,
//    by n eps (|max_err| + |offset|); the sum at the crossing adds one value
//    and one rounding, eps of its own size.
// When a comparison lies closer than these margins the result is undecided
// (*ok = false) and the caller sorts instead.
double select_threshold(const std::vector<double> &v, double max_err, double offset, bool *ok) {
  *ok = true;
  const size_t n = v.size();
  if (n == 0) return 0.0;
  if (offset >= max_err) return 0.0;  // x0 + offset >= offset >= max_err
  constexpr double kEps = 2.220446049250313e-16;
  const double tol = (4.0 * (double)n + 16.0) * kEps * (std::abs(offset) + std::abs(max_err));
  auto early = [&](double total, double *thr) {  // true: decided without a crossing
    const double tol_total = (4.0 * (double)n + 16.0) * kEps * (total + std::abs(offset) + std::abs(max_err));
    if (total + offset < max_err - tol_total) {
      *thr = INFINITY;
      return true;
    }
    if (total + offset < max_err + tol_total) {
      *ok = false;
      *thr = 0.0;
      return true;
    }
    return false;
  };
  std::vector<double> cur, nxt;
  double base = 0;          // sum of every value ranked below the current bin
  double below_max = -1.0;  // largest such value (-1: none)
  constexpr int kB = 2048;
  static thread_local std::vector<uint32_t> cnt(kB);
  static thread_local std::vector<double> sum(kB), mx(kB);
  int shift = 52, width = 11;  // exponent first, then the mantissa from the top
  // the first level bins v itself (no copy), and its bin sums give the total
  bool first = n > 48;
  if (!first) {
    double total = 0, thr = 0;
    for (double x : v) total += x;
    if (early(total, &thr)) return thr;
    cur = v;
  }
  while ((first || cur.size() > 48) && shift >= 0) {
    const std::vector<double> &in = first ? v : cur;
    const uint64_t mask = (1ull << width) - 1;
    std::fill(cnt.begin(), cnt.end(), 0u);
    std::fill(sum.begin(), sum.end(), 0.0);
    std::fill(mx.begin(), mx.end(), -1.0);
    for (double x : in) {
      uint64_t u;
      std::memcpy(&u, &x, 8);
      const int e = (int)((u >> shift) & mask);
      cnt[e]++;
      sum[e] += x;
      mx[e] = std::max(mx[e], x);
    }
    if (first) {
      double total = 0, thr = 0;
      for (int e = 0; e <= (int)mask; ++e) total += sum[e];
      if (early(total, &thr)) return thr;
    }
    int hit = -1;
    for (int e = 0; e <= (int)mask; ++e) {
      if (!cnt[e]) continue;
      if (base + sum[e] + offset >= max_err) {
        hit = e;
        break;
      }
      base += sum[e];
      below_max = std::max(below_max, mx[e]);
    }
    if (hit < 0) {
      *ok = false;
      return 0.0;
    }
    nxt.clear();
    nxt.reserve(cnt[hit]);
    for (double x : in) {
      uint64_t u;
      std::memcpy(&u, &x, 8);
      if ((int)((u >> shift) & mask) == hit) nxt.push_back(x);
    }
    cur.swap(nxt);
    first = false;
    if (shift == 0) break;
    width = std::min(11, shift == 52 ? 52 : shift);
    shift = shift == 52 ? 41 : shift - width;
  }
  std::sort(cur.begin(), cur.end());
  double s = base;
  for (size_t j = 0; j < cur.size(); ++j) {
    const double prev = s;
    s += cur[j];
    if (s + offset >= max_err) {
      if (prev + offset >= max_err - tol || s + offset < max_err + tol + 4.0 * kEps * std::abs(s + offset)) {
        *ok = false;
        return 0.0;
      }
      if (j > 0) return cur[j - 1];
      return below_max >= 0 ? below_max : 0.0;
    }
  }
  *ok = false;
  return 0.0;
}

// Sum of f(0) .. f(n-1) in a fixed blocked order: a sequential sum over each
// block of 256 consecutive terms, then a sequential sum of the block sums.
// The device selection (hops_select.hip) adds its sums in the same order, a
// tile of 256 entries per block, so both give the same bits.  A term left out
// of a sum is passed as +0.0, which leaves every partial sum unchanged.
constexpr size_t kSumBlock = 256;
template <class F>
double bsum(size_t n, F f) {
  double tot = 0;
  for (size_t b = 0; b < n; b += kSumBlock) {
    double part = 0;
    const size_t e = std::min(n, b + kSumBlock);
    for (size_t i = b; i < e; ++i) part += f(i);
    tot += part;
  }
  return tot;
}

// error_threshold(sort(v), max_err, offset) and the blocked sum of the values
// <= the threshold (v: non-negative error terms)
double partial_threshold(const std::vector<double> &v, double max_err, double offset, double *below) {
  bool ok = false;
  double thr = select_threshold(v, max_err, offset, &ok);
  if (!ok) thr = prefix_threshold(v, max_err, offset);
  *below = bsum(v.size(), [&](size_t i) { return v[i] <= thr ? v[i] : 0.0; });
  return thr;
}

struct Selection {
  std::vector<int> stable_idx;        // stable auxiliaries (sorted indices)
  std::vector<uint8_t> stable_mask;   // [n_hier]
  std::vector<Aux> aux_new, aux_bound;
  std::vector<int> states_new;
  // kept between the two halves
  std::vector<double> absC2;          // [nh * n] |Pn|^2, aux-major
  std::vector<cplx> X2;               // [nm]
  std::vector<std::vector<std::pair<int, cplx>>> M2;  // per rel mode: (state, L[s,s])
  std::vector<int> sc;
  // L-operators with entries off the diagonal in the basis
  // (hops_basis.py:914-998, :1030-1100); empty without them
  std::vector<double> off2;           // [nm * n] sum over destinations d != s of |L_m[d, s]|^2
  std::vector<int> dsc;               // destination states outside the basis (in sc), ascending
  std::vector<std::vector<std::pair<int, double>>> md2;  // per dsc: (lm * n + i, |L_m[d, s_i]|^2)
  std::vector<cplx> coef;             // per relative L: conj(z1) - i z2 + its noise memory
  // couplings of the outside states sc to the basis states through H2 = -iH
  // + noise operator, as CSR over sc (state, value); with low-temperature
  // operators also through H2 + T_phys (physical wave function) and H2 +
  // T_hier (auxiliaries)
  std::vector<int> h2p, h2s, hpp, hps, hhp, hhs;
  std::vector<cplx> h2v, hpv, hhv;
  std::vector<double> v1, vp, vh;     // [n] sum over d != s of |H2[d, s]|^2 (and with T_phys, T_hier)
  bool lt = false;                    // low-temperature operators of off-diagonal L
  // with Params::keep_err: the error per auxiliary, per basis state and per
  // outside state (state, error) of the last selection
  std::vector<double> err_aux, err_state;
  std::vector<std::pair<int, double>> err_out;
  // a Zobrist-key hit named a different auxiliary than the one looked up
  // (the selection is then invalid and the caller fails)
  bool key_collision = false;
};

std::vector<int> list_sc(const System &S, const Basis &B) {
  std::vector<char> in(S.n, 0), out(S.n, 0);
  for (int s : B.states) in[s] = 1;
  for (int s : B.states) {
    for (int c : S.coupled[s]) out[c] = 1;
    for (int c : S.dest[s]) out[c] = 1;
  }
  std::vector<int> r;
  for (int s = 0; s < S.n; ++s)
    if (out[s] && !in[s]) r.push_back(s);
  return r;
}

// returns false if a Zobrist-key hit named another auxiliary
bool markov_up(const Params &P, const Basis &B, std::vector<uint8_t> &F) {
  // hops_fluxfilters.py:354-397; F[k * nm + lm]
  F.assign((size_t)B.nh * B.nm, 1);
  const int nm = B.nm;
  if (!P.markov.empty())
    for (int lm = 0; lm < nm; ++lm) {
      if (!P.markov[B.modes[lm]]) continue;
      for (int k = 1; k < B.nh; ++k) F[(size_t)k * nm + lm] = 0;
      const int j = B.index.find(key_step(*B.sys, AKey(), B.modes[lm], 0, 1));
      if (j >= 0 && !is_step(Aux(), B.aux[j], B.modes[lm], 1)) return false;
      if (j >= 0)
        for (int m2 = 0; m2 < nm; ++m2) F[(size_t)j * nm + m2] = 0;
    }
  if (P.filters.empty()) return true;
  // Triangular: no flux up along a filtered mode from an auxiliary whose
  // depth in the filtered modes has reached kmax2.  LongEdge: from total
  // depth kmax2 on, an auxiliary with depth in a filtered mode steps up only
  // along its own mode if it is an edge (single-mode) term, and one without
  // steps up along no filtered mode.  The filters of one kind share one mask
  // in order; the kinds multiply (hops_fluxfilters.py:372-501).
  std::vector<uint8_t> G;
  for (int kind = 0; kind < 2; ++kind) {
    bool any = false;
    for (const auto &f : P.filters) any |= f.kind == kind;
    if (!any) continue;
    G.assign((size_t)B.nh * nm, 1);
    for (const auto &f : P.filters) {
      if (f.kind != kind) continue;
      for (int k = 0; k < B.nh; ++k) {
        int fdep = 0;
        for (int q = B.aux_ptr[k]; q < B.aux_ptr[k + 1]; ++q)
          if (f.mask[B.aux_mode_t[q]]) fdep += B.aux_val_t[q];
        uint8_t *g = &G[(size_t)k * nm];
        if (kind == 0) {
          if (fdep >= f.kmax)
            for (int lm = 0; lm < nm; ++lm)
              if (f.mask[B.modes[lm]]) g[lm] = 0;
        } else if (B.depth[k] >= f.kmax) {
          if (fdep > 0) {
            for (int lm = 0; lm < nm; ++lm) g[lm] = 0;
            if (B.aux_ptr[k + 1] - B.aux_ptr[k] == 1) g[B.rel_mode[B.aux_mode_t[B.aux_ptr[k]]]] = 1;
          } else {
            for (int lm = 0; lm < nm; ++lm)
              if (f.mask[B.modes[lm]]) g[lm] = 0;
          }
        }
      }
    }
    for (size_t i = 0; i < F.size(); ++i) F[i] &= G[i];
  }
  return true;
}

// The part of the selection that needs only psi = Pn[:n]: the diagonal
// entries M2 of each mode's reduced L, <L> over psi (X2) and the outside
// states coupled to the basis (sc)
void sel_offdiag(const Basis &B, const cplx *coef, const std::vector<cplx> &avg, Selection &out);

// coef: per relative L of the basis, the noise coefficient of the noise
// operator (noise_coef); used only when the basis has an L-operator with
// entries off the diagonal
void sel_common(const Basis &B, const cplx *pn, Selection &out, const cplx *coef = nullptr) {
  const System &S = *B.sys;
  const int n = B.n, nm = B.nm;
  out = Selection();
  // M2: diagonal entries of each mode's reduced L (basis/hops_basis.py:1030-1151)
  out.M2.resize(nm);
  for (int lm = 0; lm < nm; ++lm) {
    const RedL &R = B.L[B.mode_lrel[lm]];
    for (int s = 0; s < n; ++s)
      for (int q = R.rp[s]; q < R.rp[s + 1]; ++q)
        if (R.ci[q] == s) out.M2[lm].push_back({s, R.rv[q]});
  }
  // <L> over psi = Pn[:n]
  cplx nrm = 0;
  for (int s = 0; s < n; ++s) nrm += cm(std::conj(pn[s]), pn[s]);
  std::vector<cplx> avg(B.nl);
  for (int j = 0; j < B.nl; ++j) {
    const RedL &R = B.L[j];
    cplx a = 0;
    for (int r = 0; r < n; ++r) {
      cplx lv = 0;
      for (int q = R.rp[r]; q < R.rp[r + 1]; ++q) lv += cm(R.rv[q], pn[R.ci[q]]);
      a += cm(std::conj(pn[r]), lv);
    }
    avg[j] = a / nrm;
  }
  out.X2.resize(nm);
  for (int lm = 0; lm < nm; ++lm) out.X2[lm] = avg[B.mode_lrel[lm]];
  out.sc = list_sc(S, B);
  if (B.has_off && coef) sel_offdiag(B, coef, avg, out);
}

// Noise coefficient of each relative L of the basis in the noise operator of
// the state flux: conj(z1) - i z2 in complex64 arithmetic (the noise is
// complex64, as the reference's get_noise returns it) plus the noise memory
// of the L's modes in the basis (hops_basis.py:936-953).  z2 may be NULL
// (noise 2 ZERO); zmem is indexed by absolute mode.
// complex128 noise (c128) is combined in double arithmetic, as the reference
// combines its complex128 noise.
void noise_coef(const Basis &B, const void *z1, const void *z2, bool c128, const cplx *zmem,
                std::vector<cplx> &coef) {
  std::vector<cplx> comp(B.nl, cplx(0, 0));
  for (int lm = 0; lm < B.nm; ++lm) comp[B.mode_lrel[lm]] += zmem[B.modes[lm]];
  coef.resize(B.nl);
  for (int j = 0; j < B.nl; ++j) {
    if (c128) {
      const cplx a = static_cast<const cplx *>(z1)[j], b2 = z2 ? static_cast<const cplx *>(z2)[j] : cplx(0, 0);
      coef[j] = cplx(a.real() + b2.imag(), -a.imag() - b2.real()) + comp[j];
    } else {
      const auto a = static_cast<const std::complex<float> *>(z1)[j];
      const auto b2 = z2 ? static_cast<const std::complex<float> *>(z2)[j] : std::complex<float>(0, 0);
      const float re = a.real() + b2.imag();
      const float im = -a.imag() - b2.real();
      coef[j] = cplx(re, im) + comp[j];
    }
  }
}

// The off-diagonal part of a selection (hops_basis.py:914-998, :1030-1100):
// - off2[lm, i] = sum over the destination states d != s_i of the basis
//   states of |L_m[d, s_i]|^2, and those terms per outside destination d;
// - H2 = -iH + sum_j coef_j L_j over the off-diagonal L of the basis, and
//   with a low-temperature correction on the basis L's T_hier = sum_j
//   conj(c_j) <L_j> L_j and T_phys = T_hier + T_hier^dagger - sum_j c_j L_j^2
//   (off-diagonal L only): their couplings from the basis states to the
//   outside states sc, and the column sums over d != s_i of |H2[d, s_i]|^2
//   (plus T_phys, T_hier).
// Built column by column over the basis states, from the full-space L.
void sel_offdiag(const Basis &B, const cplx *coef, const std::vector<cplx> &avg, Selection &out) {
  const System &S = *B.sys;
  const int n = B.n, nm = B.nm, N = S.n;
  std::vector<char> isdest(N, 0);
  for (int s : B.states)
    for (int c : S.dest[s]) isdest[c] = 1;
  std::vector<int> dpos(N, -1), scpos(N, -1);
  for (size_t q = 0; q < out.sc.size(); ++q) {
    scpos[out.sc[q]] = (int)q;
    if (isdest[out.sc[q]]) {
      dpos[out.sc[q]] = (int)out.dsc.size();
      out.dsc.push_back(out.sc[q]);
    }
  }
  out.off2.assign((size_t)nm * n, 0.0);
  out.md2.assign(out.dsc.size(), {});
  for (int lm = 0; lm < nm; ++lm) {
    const LOp &L = S.L[B.absL[B.mode_lrel[lm]]];
    if (!L.off) continue;
    for (int i = 0; i < n; ++i) {
      const int s = B.states[i];
      double acc = 0;
      for (int q = L.cp[s]; q < L.cp[s + 1]; ++q) {
        const int d = L.ri[q];
        if (d == s || !isdest[d]) continue;
        const double a = std::norm(L.cv[q]);
        acc += a;
        if (dpos[d] >= 0) out.md2[dpos[d]].push_back({lm * n + i, a});
      }
      out.off2[(size_t)lm * n + i] = acc;
    }
  }
  bool any_lt = false;
  if (!S.lt.empty())
    for (int l : B.absL) any_lt |= S.lt[l] != cplx(0, 0);
  out.lt = any_lt;
  const size_t nsc = out.sc.size();
  std::vector<std::vector<std::pair<int, cplx>>> r2(nsc), rp(any_lt ? nsc : 0), rh(any_lt ? nsc : 0);
  out.v1.assign(n, 0.0);
  if (any_lt) {
    out.vp.assign(n, 0.0);
    out.vh.assign(n, 0.0);
  }
  std::vector<cplx> h2(N), th(N), tp(N), sq(N);
  std::vector<char> hit(N, 0);
  std::vector<int> touched;
  auto touch = [&](int d) {
    if (!hit[d]) {
      hit[d] = 1;
      touched.push_back(d);
    }
  };
  for (int i = 0; i < n; ++i) {
    const int s = B.states[i];
    for (int d = 0; d < N; ++d)
      if (S.H[(size_t)d * N + s] != cplx(0, 0)) {
        h2[d] = cmi(S.H[(size_t)d * N + s]);
        touch(d);
      }
    for (int j = 0; j < B.nl; ++j) {
      const LOp &L = S.L[B.absL[j]];
      if (!L.off) continue;
      for (int q = L.cp[s]; q < L.cp[s + 1]; ++q) {
        h2[L.ri[q]] += coef[j] * L.cv[q];
        touch(L.ri[q]);
      }
    }
    if (any_lt) {
      for (int j = 0; j < B.nl; ++j) {
        const LOp &L = S.L[B.absL[j]];
        if (!L.off) continue;
        const cplx c = S.lt[B.absL[j]];
        const cplx a = std::conj(c) * avg[j];
        for (int q = L.cp[s]; q < L.cp[s + 1]; ++q) {  // T_hier[:, s]
          th[L.ri[q]] += a * L.cv[q];
          touch(L.ri[q]);
        }
        const cplx ad = c * std::conj(avg[j]);
        for (int q = L.rp[s]; q < L.rp[s + 1]; ++q) {  // T_hier^dagger[:, s] = conj(T_hier[s, :])
          tp[L.ci[q]] += ad * std::conj(L.rv[q]);
          touch(L.ci[q]);
        }
        for (int q = L.cp[s]; q < L.cp[s + 1]; ++q) {  // (L L)[:, s]
          const int k = L.ri[q];
          for (int q2 = L.cp[k]; q2 < L.cp[k + 1]; ++q2) {
            sq[L.ri[q2]] += L.cv[q2] * L.cv[q];
            touch(L.ri[q2]);
          }
        }
        for (int d : touched) {
          tp[d] -= c * sq[d];
          sq[d] = 0;
        }
      }
    }
    std::sort(touched.begin(), touched.end());
    double v = 0, vp = 0, vh = 0;
    for (int d : touched) {
      const cplx tphys = th[d] + tp[d];  // T_hier + T_hier^dagger - sum c L^2
      if (d != s) {
        v += std::norm(h2[d]);
        if (any_lt) {
          vp += std::norm(h2[d] + tphys);
          vh += std::norm(h2[d] + th[d]);
        }
      }
      const int q = scpos[d];
      if (q >= 0) {
        if (h2[d] != cplx(0, 0)) r2[q].push_back({i, h2[d]});
        if (any_lt) {
          if (h2[d] + tphys != cplx(0, 0)) rp[q].push_back({i, h2[d] + tphys});
          if (h2[d] + th[d] != cplx(0, 0)) rh[q].push_back({i, h2[d] + th[d]});
        }
      }
      h2[d] = th[d] = tp[d] = 0;
      hit[d] = 0;
    }
    touched.clear();
    out.v1[i] = v;
    if (any_lt) {
      out.vp[i] = vp;
      out.vh[i] = vh;
    }
  }
  auto flat = [&](const std::vector<std::vector<std::pair<int, cplx>>> &r, std::vector<int> &ptr,
                  std::vector<int> &st, std::vector<cplx> &val) {
    ptr.assign(r.size() + 1, 0);
    for (size_t q = 0; q < r.size(); ++q) {
      for (auto &e : r[q]) {
        st.push_back(e.first);
        val.push_back(e.second);
      }
      ptr[q + 1] = (int)st.size();
    }
  };
  flat(r2, out.h2p, out.h2s, out.h2v);
  if (any_lt) {
    flat(rp, out.hpp, out.hps, out.hpv);
    flat(rh, out.hhp, out.hhs, out.hhv);
  }
}

// the couplings -iH[sp, s] of each outside state sp (in sc order) to the
// basis states s, as CSR; a zero entry adds an exact zero to the flux sum,
// so it is left out
void sc_couplings(const Basis &B, const std::vector<int> &sc, std::vector<int> &hp, std::vector<int> &hs,
                  std::vector<cplx> &hv) {
  const System &S = *B.sys;
  hp.assign(sc.size() + 1, 0);
  hs.clear();
  hv.clear();
  for (size_t i = 0; i < sc.size(); ++i) {
    for (int s = 0; s < B.n; ++s) {
      const cplx h = S.H[(size_t)sc[i] * S.n + B.states[s]];
      if (h != cplx(0, 0)) {
        hs.push_back(s);
        hv.push_back(cmi(h));
      }
    }
    hp[i + 1] = (int)hs.size();
  }
}

// hierarchy half: stable auxiliaries, boundary auxiliaries, new aux list
void select_hier(const Params &P, const Basis &B, const cplx *pn, const double *e_aux, Selection &out,
                 const cplx *coef = nullptr) {
  const System &S = *B.sys;
  const int n = B.n, nh = B.nh, nm = B.nm;
  if (B.has_off && !coef) {
    // the noise operator needs the noise of the step: a caller without it
    // cannot select for off-diagonal L-operators
    out = Selection();
    out.aux_new = B.aux;
    out.states_new = B.states;
    out.stable_mask.assign(nh, 1);
    out.stable_idx.resize(nh);
    for (int k = 0; k < nh; ++k) out.stable_idx[k] = k;
    return;
  }
  sel_common(B, pn, out, coef);
  const bool off = !out.off2.empty();
  out.absC2.resize((size_t)nh * n);
  for (size_t i = 0; i < (size_t)nh * n; ++i) out.absC2[i] = std::norm(pn[i]);
  if (P.delta_a <= 0) {
    out.aux_new = B.aux;
    out.stable_idx.resize(nh);
    for (int k = 0; k < nh; ++k) out.stable_idx[k] = k;
    out.stable_mask.assign(nh, 1);
    return;
  }
  const double h2 = kHbar * kHbar;
  std::vector<double> err(e_aux, e_aux + nh);
  // flux out of the state basis through H, or with off-diagonal L through H2
  // (-iH plus the noise operator) and the low-temperature operators: T_phys
  // for the physical wave function, T_hier for the auxiliaries
  // (basis_functions_adaptive.py:76-163)
  if (!out.sc.empty()) {
    std::vector<int> hp, hs;
    std::vector<cplx> hv;
    if (!off) sc_couplings(B, out.sc, hp, hs, hv);
    for (int k = 0; k < nh; ++k) {
      const std::vector<int> &cp = !off ? hp : !out.lt ? out.h2p : k == 0 ? out.hpp : out.hhp;
      const std::vector<int> &cs = !off ? hs : !out.lt ? out.h2s : k == 0 ? out.hps : out.hhs;
      const std::vector<cplx> &cv = !off ? hv : !out.lt ? out.h2v : k == 0 ? out.hpv : out.hhv;
      const cplx *pk = pn + (size_t)k * n;
      double acc = 0;
      for (size_t i = 0; i < out.sc.size(); ++i) {
        cplx z = 0;
        for (int q = cp[i]; q < cp[i + 1]; ++q) z += cm(cv[q], pk[cs[q]]);
        z /= kHbar;
        acc += std::norm(z);
      }
      err[k] += acc;
    }
  }
  // flux up / down (basis_functions_adaptive.py:165-390); per-(aux, mode)
  // arrays are aux-major, [k * nm + lm]
  std::vector<uint8_t> kval((size_t)nh * nm, 0);  // K2_aux_by_mode
  for (int k = 0; k < nh; ++k)
    for (int q = B.aux_ptr[k]; q < B.aux_ptr[k + 1]; ++q) kval[(size_t)k * nm + B.rel_mode[B.aux_mode_t[q]]] = B.aux_val_t[q];
  // every entry read below is written first (all of them in the general
  // loop), so the arrays start uninitialised
  std::unique_ptr<double[]> E_up(new double[(size_t)nh * nm]), E_dn(new double[(size_t)nh * nm]);
  std::vector<uint8_t> F_up;
  if (!markov_up(P, B, F_up)) {
    out.key_collision = true;
    return;
  }
  std::vector<double> aw(nm), gw2(nm), nX(nm);
  // the diagonal L entries of each mode as CSR: state, |M2|^2, |M2 - X|^2;
  // with off-diagonal L each state's sum over its destinations d != s of
  // |L_m[d, s]|^2 is added to both (hops_basis.py:1030-1100), and the states
  // with only such entries join the list
  std::vector<int> m2p(nm + 1, 0), m2s;
  std::vector<double> m2u, m2d;
  for (int lm = 0; lm < nm; ++lm) {
    const int m = B.modes[lm];
    aw[lm] = std::abs(S.W[m]);
    gw2[lm] = std::norm(S.G[m] / S.W[m]);
    nX[lm] = std::norm(out.X2[lm]);
    if (!off) {
      for (auto &e : out.M2[lm]) {
        m2s.push_back(e.first);
        m2u.push_back(std::norm(e.second));
        m2d.push_back(std::norm(e.second - out.X2[lm]));
      }
    } else {
      const double *o2 = &out.off2[(size_t)lm * n];
      size_t q = 0;
      for (int s2 = 0; s2 < n; ++s2) {
        cplx dv = 0;
        bool has = false;
        if (q < out.M2[lm].size() && out.M2[lm][q].first == s2) {
          dv = out.M2[lm][q++].second;
          has = true;
        }
        if (!has && o2[s2] == 0) continue;
        m2s.push_back(s2);
        m2u.push_back(std::norm(dv) + o2[s2]);
        m2d.push_back(std::norm(dv - out.X2[lm]) + o2[s2]);
      }
    }
    m2p[lm + 1] = (int)m2s.size();
  }
  // diagonal L-operators give every mode one M2 entry: then the sums below
  // are one product each, and the loop over modes is a plain gather.  Only
  // the terms read later are formed (flux up below the top depth where the
  // filters allow it, flux down where the auxiliary has depth in the mode),
  // and err[k] takes their sums in the same order as the general loop.
  bool single = true;
  for (int lm = 0; lm < nm; ++lm) single &= m2p[lm + 1] - m2p[lm] == 1;
  for (int k = 0; k < nh && single; ++k) {
    const double *c2 = &out.absC2[(size_t)k * n];
    double colsum = 0;
    for (int s2 = 0; s2 < n; ++s2) colsum += c2[s2];
    double *eu = &E_up[(size_t)k * nm], *ed = &E_dn[(size_t)k * nm];
    const uint8_t *kv = &kval[(size_t)k * nm];
    const uint8_t *fu = &F_up[(size_t)k * nm];
    const bool top = B.depth[k] == P.maxhier;
    double acc_up = 0, acc_dn = 0;
    for (int lm = 0; lm < nm; ++lm) {
      const double c = c2[m2s[lm]];
      if (!top && fu[lm]) {
        const double p_up = m2u[lm] * c;
        const double kw = aw[lm] * (1.0 + kv[lm]);
        eu[lm] = (kw * kw * p_up) / h2;
        acc_up += eu[lm];
      }
      if (kv[lm] > 0) {
        double p_dn = m2d[lm] * c;
        if (nX[lm] != 0) p_dn += nX[lm] * (colsum - c);
        ed[lm] = (gw2[lm] * p_dn) / h2;
        acc_dn += ed[lm];
      }
    }
    err[k] += acc_up;
    err[k] += acc_dn;
  }
  for (int k = 0; k < nh && !single; ++k) {
    const double *c2 = &out.absC2[(size_t)k * n];
    double colsum = 0;
    for (int s2 = 0; s2 < n; ++s2) colsum += c2[s2];
    double *eu = &E_up[(size_t)k * nm], *ed = &E_dn[(size_t)k * nm];
    const uint8_t *kv = &kval[(size_t)k * nm];
    for (int lm = 0; lm < nm; ++lm) {
      // sum_s |M2[m,s]|^2 |C2[s,k]|^2 and sum_s |M2[m,s] - X|^2 |C2[s,k]|^2;
      // the states without an M2 entry contribute |X|^2 |C2[s,k]|^2
      double p_up = 0, p_dn = 0, on = 0;
      for (int q = m2p[lm]; q < m2p[lm + 1]; ++q) {
        const double c = c2[m2s[q]];
        p_up += m2u[q] * c;
        p_dn += m2d[q] * c;
        on += c;
      }
      if (nX[lm] != 0) p_dn += nX[lm] * (colsum - on);
      const double kw = aw[lm] * (1.0 + kv[lm]);
      eu[lm] = (kw * kw * p_up) / h2;
      ed[lm] = (gw2[lm] * p_dn) / h2;
    }
  }
  for (int k = 0; k < nh && !single; ++k) {
    const bool top = B.depth[k] == P.maxhier;
    const size_t o = (size_t)k * nm;
    double acc = 0;
    for (int lm = 0; lm < nm; ++lm)
      if (!top && F_up[o + lm]) acc += E_up[o + lm];
    err[k] += acc;
    acc = 0;
    for (int lm = 0; lm < nm; ++lm)
      if (kval[o + lm] > 0) acc += E_dn[o + lm];
    err[k] += acc;
  }
  if (P.keep_err) out.err_aux = err;
  // stable auxiliaries (hops_basis.py:1000-1028)
  const double dsq = P.delta_a * P.delta_a;
  double below = 0;
  const double thr = partial_threshold(err, dsq / 2.0, 0.0, &below);
  out.stable_mask.assign(nh, 0);
  for (int k = 0; k < nh; ++k)
    if (err[k] > thr) {
      out.stable_idx.push_back(k);
      out.stable_mask[k] = 1;
    }
  const double unstable = bsum(nh, [&](size_t k) { return err[k] > thr ? 0.0 : err[k]; });
  const double dbound = dsq - unstable;
  // boundary auxiliaries (hops_basis.py:667-793): flux out of the basis, up
  // (depth < maxhier, k+e_m not in the basis) and down (k-e_m not in the
  // basis), in the reference's insertion order: stable column, flux up by
  // mode, then flux down by mode.  cand bit 0: up candidate, bit 1: down.
  // The candidates' values in this order (nz), with (mode << 1 | down) of
  // each (nzc) and where each stable auxiliary's candidates start (nzo).
  // (per host thread, reused: they keep their capacity between selections)
  static thread_local std::vector<double> nz;
  static thread_local std::vector<uint32_t> nzc;
  nz.clear();
  nzc.clear();
  std::vector<size_t> nzo(out.stable_idx.size() + 1, 0);
  std::vector<uint8_t> cand(nm);
  for (size_t i = 0; i < out.stable_idx.size(); ++i) {
    const int k = out.stable_idx[i];
    const size_t o = (size_t)k * nm;
    nzo[i] = nz.size();
    const bool below_top = B.depth[k] < P.maxhier;
    for (int lm = 0; lm < nm; ++lm)
      cand[lm] = (uint8_t)((below_top && F_up[o + lm] && E_up[o + lm] != 0 ? 1 : 0) |
                           (kval[o + lm] > 0 && E_dn[o + lm] != 0 ? 2 : 0));
    for (int u = B.up_ptr[k]; u < B.up_ptr[k + 1]; ++u) cand[B.up_mode[u]] &= (uint8_t)~1u;
    for (int u = B.dn_ptr[k]; u < B.dn_ptr[k + 1]; ++u) cand[B.dn_mode[u]] &= (uint8_t)~2u;
    for (int lm = 0; lm < nm; ++lm) {
      const uint8_t c = cand[lm];
      if (c & 1) {
        nz.push_back(E_up[o + lm]);
        nzc.push_back((uint32_t)lm << 1);
      }
      if (c & 2) {
        nz.push_back(E_dn[o + lm]);
        nzc.push_back((uint32_t)lm << 1 | 1u);
      }
    }
  }
  nzo[out.stable_idx.size()] = nz.size();
  double discarded = 0;
  const double thr1 = partial_threshold(nz, dbound * P.f_discard * P.f_discard, 0.0, &discarded);
  // accumulate the survivors by target auxiliary (Zobrist key), in the
  // order above, in an open-addressing table that doubles at half load; the
  // slots of the next candidates are prefetched
  struct Cand {
    AKey key;
    int k, lm, dir;
    double v;
  };
  static thread_local std::vector<Cand> cs;
  cs.clear();
  for (size_t i = 0; i < out.stable_idx.size(); ++i) {
    const size_t q0 = nzo[i], q1 = nzo[i + 1];
    if (q0 == q1) continue;
    const int k = out.stable_idx[i];
    const size_t o = (size_t)k * nm;
    const AKey kk{B.aux_hash[k], B.aux_hash2[k]};
    for (uint32_t d = 0; d < 2; ++d) {  // flux up by mode, then flux down by mode
      const int dir = d == 0 ? 1 : -1;
      for (size_t q = q0; q < q1; ++q)
        if ((nzc[q] & 1u) == d && nz[q] > thr1) {
          const int lm = (int)(nzc[q] >> 1);
          const int v0 = kval[o + lm];
          cs.push_back(Cand{key_step(S, kk, B.modes[lm], v0, v0 + dir), k, lm, dir, nz[q]});
        }
    }
  }
  struct Ent {
    uint64_t a;
    int32_t id, pad;
  };
  size_t cap = 1024;
  while (cap < cs.size() / 2) cap <<= 1;
  std::vector<Ent> tab(cap, Ent{0, -1, 0});
  std::vector<uint64_t> keyb;
  struct Rep {
    int k, lm, dir;
  };
  std::vector<Rep> rep;
  std::vector<double> vals;
  auto slot_of = [&](uint64_t a, uint64_t b) {
    size_t h = KeyTable::hh(AKey{a, b}) & (cap - 1);
    while (tab[h].id >= 0 && !(tab[h].a == a && keyb[tab[h].id] == b)) h = (h + 1) & (cap - 1);
    return h;
  };
  Aux t1, t2;
  for (size_t i = 0; i < cs.size(); ++i) {
    if (i + 8 < cs.size()) __builtin_prefetch(&tab[KeyTable::hh(cs[i + 8].key) & (cap - 1)]);
    const Cand &c = cs[i];
    size_t h = slot_of(c.key.a, c.key.b);
    if (tab[h].id >= 0) {
      // a merge: the two candidates must reach the same auxiliary
      const Rep &r = rep[tab[h].id];
      aux_step(B.aux[r.k], B.modes[r.lm], r.dir, t1);
      aux_step(B.aux[c.k], B.modes[c.lm], c.dir, t2);
      if (t1 != t2) {
        out.key_collision = true;
        return;
      }
    }
    if (tab[h].id < 0) {
      if (2 * (vals.size() + 1) > cap) {
        std::vector<Ent> old;
        old.swap(tab);
        cap <<= 1;
        tab.assign(cap, Ent{0, -1, 0});
        for (const Ent &e : old)
          if (e.id >= 0) tab[slot_of(e.a, keyb[e.id])] = e;
        h = slot_of(c.key.a, c.key.b);
      }
      tab[h] = Ent{c.key.a, (int32_t)vals.size(), 0};
      keyb.push_back(c.key.b);
      rep.push_back({c.k, c.lm, c.dir});
      vals.push_back(0.0);
    }
    vals[tab[h].id] += c.v;
  }
  double dummy = 0;
  const double thr2 = partial_threshold(vals, dbound, discarded, &dummy);
  out.aux_new.reserve(out.stable_idx.size() + vals.size());
  for (int k : out.stable_idx) out.aux_new.push_back(B.aux[k]);
  Aux tgt;
  for (size_t i = 0; i < vals.size(); ++i)
    if (vals[i] > thr2) {
      const Rep &cd = rep[i];
      aux_step(B.aux[cd.k], B.modes[cd.lm], cd.dir, tgt);
      out.aux_bound.push_back(tgt);
      out.aux_new.push_back(tgt);
    }
}

// state half (hops_basis.py:294-482, :795-912); e_state[s] = sum over stable
// auxiliaries of |D_stable/hbar + P_stable/dt|^2
void select_states(const Params &P, const Basis &B, const cplx *pn, const double *e_state, Selection &sel) {
  const System &S = *B.sys;
  const int n = B.n, nh = B.nh, nm = B.nm;
  if (P.delta_s <= 0) {
    sel.states_new = B.states;
    return;
  }
  const double h2 = kHbar * kHbar;
  const bool off = !sel.off2.empty();
  std::vector<double> err(e_state, e_state + n);
  // flux through the off-diagonal Hamiltonian (with off-diagonal L: through
  // H2, and with their low-temperature operators through H2 + T_phys for
  // the first stable auxiliary and H2 + T_hier for the others,
  // basis_functions_adaptive.py:893-995)
  const int k_first = sel.stable_idx.empty() ? -1 : sel.stable_idx[0];
  for (int s = 0; s < n; ++s) {
    if (off && sel.lt) {
      if (k_first < 0) continue;
      const double c0 = sel.absC2[(size_t)k_first * n + s];
      const double cr =
          bsum(nh, [&](size_t k) { return sel.stable_mask[k] && (int)k != k_first ? sel.absC2[k * n + s] : 0.0; });
      err[s] += (sel.vp[s] * c0 + sel.vh[s] * cr) / h2;
      continue;
    }
    double v1 = 0;
    if (off) {
      v1 = sel.v1[s];
    } else {
      for (int sp = 0; sp < S.n; ++sp)
        if (sp != B.states[s]) v1 += std::norm(S.H[(size_t)sp * S.n + B.states[s]]);
    }
    const double c1 = bsum(nh, [&](size_t k) { return sel.stable_mask[k] ? sel.absC2[k * n + s] : 0.0; });
    err[s] += v1 * c1 / h2;
  }
  // filters from the boundary auxiliaries (hops_fluxfilters.py:292-352)
  // F_dn[lm, j]: j = a + e_lm for a boundary aux a; F_up[lm, j]: j = a - e_lm
  std::vector<uint8_t> F_dn((size_t)nm * nh, 0), F_up((size_t)nm * nh, 0);
  const bool from_basis = sel.aux_bound.size() * (size_t)nm > B.aux_mode_t.size();
  KeyTable btab;
  if (from_basis) btab.reset(sel.aux_bound.size());
  for (size_t i = 0; i < sel.aux_bound.size(); ++i) {
    const Aux &a = sel.aux_bound[i];
    const AKey ka = key_of(S, a);
    if (from_basis) {
      if (btab.insert(ka, (int)i) >= 0) sel.key_collision = true;
    } else {
      for (int lm = 0; lm < nm; ++lm) {
        const int v = aux_get(a, B.modes[lm]);
        if (v >= 255) continue;
        const int j = B.index.find(key_step(S, ka, B.modes[lm], v, v + 1));
        if (j < 0) continue;
        if (!is_step(a, B.aux[j], B.modes[lm], 1)) {
          sel.key_collision = true;
          return;
        }
        F_dn[(size_t)lm * nh + j] = 1;
      }
    }
    for (char32_t e : a) {
      const int j = B.index.find(key_step(S, ka, aux_mode(e), aux_val(e), aux_val(e) - 1));
      if (j < 0) continue;
      if (!is_step(a, B.aux[j], aux_mode(e), -1)) {
        sel.key_collision = true;
        return;
      }
      F_up[(size_t)B.rel_mode[aux_mode(e)] * nh + j] = 1;
    }
  }
  if (from_basis)  // the same F_dn from the basis side: j - e_m is a boundary aux
    for (int j = 0; j < nh; ++j) {
      const AKey kj{B.aux_hash[j], B.aux_hash2[j]};
      for (int q = B.aux_ptr[j]; q < B.aux_ptr[j + 1]; ++q) {
        const int m = B.aux_mode_t[q], v = B.aux_val_t[q];
        const int i = btab.find(key_step(S, kj, m, v, v - 1));
        if (i < 0) continue;
        if (!is_step(sel.aux_bound[i], B.aux[j], m, 1)) {
          sel.key_collision = true;
          return;
        }
        F_dn[(size_t)B.rel_mode[m] * nh + j] = 1;
      }
    }
  for (int lm = 0; lm < nm; ++lm) {
    const int m = B.modes[lm];
    const double gw2 = std::norm(S.G[m] / S.W[m]);
    const double aw = std::abs(S.W[m]);
    const cplx X = sel.X2[lm];
    // |M2[m,s] - X|^2 and |M2[m,s]|^2 for every state
    std::vector<double> a_dn(n, std::norm(X)), a_up(n, 0.0);
    for (auto &e : sel.M2[lm]) {
      a_dn[e.first] = std::norm(e.second - X);
      a_up[e.first] = std::norm(e.second);
    }
    for (int k : sel.stable_idx) {
      const bool fd = F_dn[(size_t)lm * nh + k], fu = F_up[(size_t)lm * nh + k];
      if (!fd && !fu) continue;
      const double kw = aw * (1.0 + aux_get(B.aux[k], m));
      for (int s = 0; s < n; ++s) {
        const double c2 = sel.absC2[(size_t)k * n + s];
        if (fd) err[s] += (a_dn[s] * gw2) * c2 / h2;
        if (fu) err[s] += (a_up[s] * kw * kw) * c2 / h2;
      }
    }
  }
  // off-diagonal L: the destination sums off2 enter through the filters of
  // the new auxiliary list (hops_fluxfilters.py:292-352 on aux_new;
  // basis_functions_adaptive.py:268-384, :626-758).  fo[p * n_off + o]: bit 0
  // k + e_m in the new list (flux up), bit 1 k - e_m in it (flux down), for
  // stable position p and off-diagonal mode o.
  std::vector<int> offm;
  std::vector<uint8_t> fo;
  if (off) {
    for (int lm = 0; lm < nm; ++lm)
      for (int s = 0; s < n; ++s)
        if (sel.off2[(size_t)lm * n + s] != 0) {
          offm.push_back(lm);
          break;
        }
    KeyTable ntab;
    ntab.reset(sel.aux_bound.size());
    for (size_t i = 0; i < sel.aux_bound.size(); ++i)
      if (ntab.insert(key_of(S, sel.aux_bound[i]), (int)i) >= 0) sel.key_collision = true;
    // membership of base + dir * e_m in the new list, each key hit confirmed
    // against the auxiliaries
    auto in_new = [&](const AKey &key, const Aux &base, int m, int dir) {
      const int j = B.index.find(key);
      if (j >= 0) {
        if (!is_step(base, B.aux[j], m, dir)) sel.key_collision = true;
        return sel.stable_mask[j] != 0;
      }
      const int i = ntab.find(key);
      if (i >= 0 && !is_step(base, sel.aux_bound[i], m, dir)) sel.key_collision = true;
      return i >= 0;
    };
    const size_t no = offm.size();
    fo.assign(sel.stable_idx.size() * no, 0);
    for (size_t p = 0; p < sel.stable_idx.size(); ++p) {
      const int k = sel.stable_idx[p];
      const AKey kk{B.aux_hash[k], B.aux_hash2[k]};
      for (size_t o = 0; o < no; ++o) {
        const int lm = offm[o], m = B.modes[lm];
        const int v = aux_get(B.aux[k], m);
        uint8_t f = 0;
        if (v < 255 && in_new(key_step(S, kk, m, v, v + 1), B.aux[k], m, 1)) f |= 1;
        if (v > 0 && in_new(key_step(S, kk, m, v, v - 1), B.aux[k], m, -1)) f |= 2;
        fo[p * no + o] = f;
        if (!f) continue;
        const double gw2 = std::norm(S.G[m] / S.W[m]);
        const double kw = std::abs(S.W[m]) * (1.0 + v);
        for (int s = 0; s < n; ++s) {
          const double o2 = sel.off2[(size_t)lm * n + s];
          if (o2 == 0) continue;
          const double c2 = sel.absC2[(size_t)k * n + s];
          if (f & 2) err[s] += (o2 * gw2) * c2 / h2;
          if (f & 1) err[s] += (o2 * kw * kw) * c2 / h2;
        }
      }
    }
  }
  if (P.keep_err) sel.err_state = err;
  const double dsq = P.delta_s * P.delta_s;
  std::vector<double> srt(err);
  std::sort(srt.begin(), srt.end());
  const double thr = error_threshold(srt, dsq / 2.0);
  std::vector<int> rel_st, st_stable;
  double unstable = 0;
  for (int s = 0; s < n; ++s) {
    if (err[s] > thr) {
      rel_st.push_back(s);
      st_stable.push_back(B.states[s]);
    } else {
      unstable += err[s];
    }
  }
  const double dbound = dsq - unstable;
  std::vector<int> bound;
  if ((int)rel_st.size() < S.n && !sel.sc.empty()) {
    std::vector<double> E(sel.sc.size(), 0.0);
    std::vector<int> qs;
    std::vector<cplx> qh;
    std::vector<char> is_st(n, 0);
    for (int q : rel_st) is_st[q] = 1;
    std::vector<int> qs2;
    std::vector<cplx> qh2;
    for (size_t i = 0; i < sel.sc.size(); ++i) {
      const int sp = sel.sc[i];
      qs.clear();
      qh.clear();
      if (!off) {
        for (size_t q = 0; q < rel_st.size(); ++q) {  // zero entries add exact zeros: skipped
          const cplx h = S.H[(size_t)sp * S.n + st_stable[q]];
          if (h != cplx(0, 0)) {
            qs.push_back(rel_st[q]);
            qh.push_back(cmi(h));
          }
        }
      } else {
        // H2 (or H2 + T_hier) from the stable states; qs2 / qh2: H2 + T_phys
        // for the first stable auxiliary
        const std::vector<int> &cp = sel.lt ? sel.hhp : sel.h2p;
        const std::vector<int> &cs = sel.lt ? sel.hhs : sel.h2s;
        const std::vector<cplx> &cv = sel.lt ? sel.hhv : sel.h2v;
        for (int q = cp[i]; q < cp[i + 1]; ++q)
          if (is_st[cs[q]]) {
            qs.push_back(cs[q]);
            qh.push_back(cv[q]);
          }
        qs2.clear();
        qh2.clear();
        if (sel.lt)
          for (int q = sel.hpp[i]; q < sel.hpp[i + 1]; ++q)
            if (is_st[sel.hps[q]]) {
              qs2.push_back(sel.hps[q]);
              qh2.push_back(sel.hpv[q]);
            }
      }
      E[i] = bsum(nh, [&](size_t k) {
        if (!sel.stable_mask[k]) return 0.0;
        const bool phys = off && sel.lt && (int)k == k_first;
        const std::vector<int> &ss = phys ? qs2 : qs;
        const std::vector<cplx> &hh = phys ? qh2 : qh;
        cplx z = 0;
        for (size_t q = 0; q < ss.size(); ++q) z += cm(hh[q], pn[k * n + ss[q]] / kHbar);
        return std::norm(z);
      });
    }
    if (off && !sel.dsc.empty()) {
      // flux up and down into each outside destination state from the
      // stable states and auxiliaries (basis_functions_adaptive.py:386-515,
      // :760-891)
      std::vector<int> opos(nm, -1);
      for (size_t o = 0; o < offm.size(); ++o) opos[offm[o]] = (int)o;
      const size_t no = offm.size();
      std::vector<int> scpos(S.n, -1);
      for (size_t q = 0; q < sel.sc.size(); ++q) scpos[sel.sc[q]] = (int)q;
      for (size_t di = 0; di < sel.dsc.size(); ++di) {
        double fu = 0, fd = 0;
        for (size_t p = 0; p < sel.stable_idx.size(); ++p) {
          const int k = sel.stable_idx[p];
          for (const auto &e : sel.md2[di]) {
            const int lm = e.first / n, i = e.first % n;
            if (!is_st[i] || opos[lm] < 0) continue;
            const uint8_t f = fo[p * no + opos[lm]];
            if (!f) continue;
            const int m = B.modes[lm];
            const double c2 = sel.absC2[(size_t)k * n + i];
            if (f & 1) {
              const double kw = std::abs(S.W[m]) * (1.0 + aux_get(B.aux[k], m));
              fu += e.second * (kw * kw) * c2 / h2;
            }
            if (f & 2) fd += e.second * std::norm(S.G[m] / S.W[m]) * c2 / h2;
          }
        }
        E[scpos[sel.dsc[di]]] += fu + fd;
      }
    }
    if (P.keep_err)
      for (size_t i = 0; i < E.size(); ++i) sel.err_out.push_back({sel.sc[i], E[i]});
    std::vector<double> ev;
    std::vector<int> es;
    for (size_t i = 0; i < E.size(); ++i)
      if (E[i] != 0) {
        ev.push_back(E[i]);
        es.push_back(sel.sc[i]);
      }
    if (!ev.empty()) {
      std::vector<double> s2(ev);
      std::sort(s2.begin(), s2.end());
      const double t2 = error_threshold(s2, dbound);
      for (size_t i = 0; i < ev.size(); ++i)
        if (ev[i] > t2) bound.push_back(es[i]);
    }
  }
  std::vector<int> res(st_stable);
  for (int b : bound)
    if (std::find(B.states.begin(), B.states.end(), b) == B.states.end()) res.push_back(b);
  std::sort(res.begin(), res.end());
  res.erase(std::unique(res.begin(), res.end()), res.end());
  sel.states_new = res;
}

// ---------------------------------------------------------------------------
// Thread pool for per-slot host work
// ---------------------------------------------------------------------------
class Pool {
 public:
  explicit Pool(int n) {
    for (int i = 0; i < n; ++i) th_.emplace_back([this] { run(); });
  }
  ~Pool() {
    {
      std::lock_guard<std::mutex> g(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto &t : th_) t.join();
  }
  // fn(i) for i in [0, n); returns after all finished
  void parallel_for(int n, const std::function<void(int)> &fn) {
    if (n <= 0) return;
    if (th_.empty() || n == 1) {
      for (int i = 0; i < n; ++i) fn(i);
      return;
    }
    std::unique_lock<std::mutex> g(mu_);
    fn_ = &fn;
    n_ = n;
    next_.store(0);
    done_ = 0;
    ++gen_;
    cv_.notify_all();
    g.unlock();
    work();
    g.lock();
    done_cv_.wait(g, [&] { return done_ == (int)th_.size(); });
    fn_ = nullptr;
  }

 private:
  void work() {
    for (;;) {
      const int i = next_.fetch_add(1);
      if (i >= n_) break;
      (*fn_)(i);
    }
  }
  void run() {
    long seen = 0;
    for (;;) {
      std::unique_lock<std::mutex> g(mu_);
      cv_.wait(g, [&] { return stop_ || gen_ != seen; });
      if (stop_) return;
      seen = gen_;
      g.unlock();
      work();
      g.lock();
      ++done_;
      done_cv_.notify_all();
    }
  }
  std::vector<std::thread> th_;
  std::mutex mu_;
  std::condition_variable cv_, done_cv_;
  const std::function<void(int)> *fn_ = nullptr;
  std::atomic<int> next_{0};
  int n_ = 0, done_ = 0;
  long gen_ = 0;
  bool stop_ = false;
};

// ---------------------------------------------------------------------------
// Arena: first-fit allocator over caller-owned device memory.  One stream,
// so memory freed after its last use was enqueued can be handed out again.
// ---------------------------------------------------------------------------
class Arena {
 public:
  void init(char *base, size_t bytes) {
    base_ = base;
    size_ = bytes;
    free_.clear();
    free_[0] = bytes;
  }
  long long alloc(size_t bytes) {
    bytes = (bytes + 255) & ~(size_t)255;
    if (bytes == 0) bytes = 256;
    for (auto it = free_.begin(); it != free_.end(); ++it)
      if (it->second >= bytes) {
        const size_t off = it->first, sz = it->second;
        free_.erase(it);
        if (sz > bytes) free_[off + bytes] = sz - bytes;
        used_[off] = bytes;
        return (long long)off;
      }
    return -1;
  }
  // the highest-addressed fit: the staging area lives at the top of the
  // arena so that slot blocks pack from the bottom without holes around it
  long long alloc_top(size_t bytes) {
    bytes = (bytes + 255) & ~(size_t)255;
    if (bytes == 0) bytes = 256;
    for (auto it = free_.rbegin(); it != free_.rend(); ++it)
      if (it->second >= bytes) {
        const size_t off = it->first + it->second - bytes;
        if (it->second == bytes) free_.erase(std::next(it).base());
        else it->second -= bytes;
        used_[off] = bytes;
        return (long long)off;
      }
    return -1;
  }
  void release(long long off) {
    if (off < 0) return;
    auto u = used_.find((size_t)off);
    if (u == used_.end()) return;
    size_t o = u->first, sz = u->second;
    used_.erase(u);
    auto nx = free_.lower_bound(o);
    if (nx != free_.end() && o + sz == nx->first) {
      sz += nx->second;
      nx = free_.erase(nx);
    }
    if (nx != free_.begin()) {
      auto pv = std::prev(nx);
      if (pv->first + pv->second == o) {
        pv->second += sz;
        return;
      }
    }
    free_[o] = sz;
  }
  size_t free_bytes() const {
    size_t s = 0;
    for (auto &kv : free_) s += kv.second;
    return s;
  }
  char *ptr(long long off) const { return base_ + off; }

 private:
  char *base_ = nullptr;
  size_t size_ = 0;
  std::map<size_t, size_t> free_, used_;
};

// ---------------------------------------------------------------------------
// Slots and the batch
// ---------------------------------------------------------------------------
struct Record {
  std::vector<cplx> psi;
  std::vector<int> states;
  int n_hier;
};

struct Slot {
  bool used = false, pending_init = false;
  const System *sys = nullptr;
  Params prm;
  Basis B;
  long long blk = -1, pers = -1;  // arena offsets: per-basis block, persistent block
  size_t blk_bytes = 0;
  hops_op_t op{};
  hops_batch_t bt{};
  const void *z1 = nullptr, *z2 = nullptr;
  // EFFECTIVE_NOISE_INTEGRATION: the RK stages read these window means while
  // the selection keeps the point samples z1 / z2 (integrator_rk.py:115-128,
  // hops_trajectory.py:693-714); NULL: the stages read z1 / z2
  const void *z1s = nullptr, *z2s = nullptr;
  bool c128 = false;  // the noise arrays hold complex128
  int it = 0, early_counter = 0;
  long long steps = 0;
  std::vector<cplx> psi0;
  // device pointers inside the blocks
  void *phi_prev = nullptr;
  double *zmem_prev = nullptr;
  std::vector<Record> rec;
  // per-step scratch
  Selection sel;
  std::vector<cplx> pn;
  std::vector<cplx> zcoef;  // noise coefficients of the noise operator (off-diagonal L only)
  double nrm = 1;
  Basis Bnew;
  bool changed = false;
  double scale = 1;
  // the new basis of the last step did not fit in the arena: the slot was
  // left mid-step (Phi advanced, basis and time not) and must be removed
  bool oom = false;
  // the system's Zobrist tables on the device (a then b, [n_hmode * 256] each)
  const uint64_t *zob_d = nullptr;
};

struct Batch {
  int precision = HOPS_PREC_FP64;
  hipStream_t stream = nullptr;
  Arena arena;
  std::vector<Slot> slots;
  std::unique_ptr<Pool> pool;
  // device descriptor arrays
  long long ops_off = -1, bts_off = -1, perm_off = -1, it_off = -1, dt_off = -1, nrm_off = -1;
  // staging: device scratch region and pinned host buffers
  long long scratch_off = -1;
  size_t scratch_bytes = 0;
  char *h_up = nullptr, *h_down = nullptr;
  size_t h_up_bytes = 0, h_down_bytes = 0;
  std::vector<hops_op_t> ops;
  std::vector<hops_batch_t> bts;
  std::vector<hops_slot_perm_t> perms;
  std::vector<int32_t> its;
  std::vector<double> dts, nrms;
  // accumulated wall time (s) per phase: step, errors on the device,
  // hierarchy selection, state errors + selection, table build, device
  // update, psi fetch; [7]: hierarchy rows advanced by RK4 steps
  double tm[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  // selection on the device (1) or on the host threads (0, default: the
  // device selection has not yet passed on the GPU)
  int sel_device = 0;
  // device work space of the selection: one region per phase
  long long sel_off[4] = {-1, -1, -1, -1};
  size_t sel_bytes[4] = {0, 0, 0, 0};
  // Zobrist tables on the device per system: arena offset, slots using them
  std::map<const System *, std::pair<long long, int>> zob;
};

struct Timer {
  double *acc;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  explicit Timer(double *a) : acc(a) {}
  ~Timer() { *acc += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
};

size_t rsz(size_t x) { return (x + 255) & ~(size_t)255; }

// per-basis block layout
struct Layout {
  size_t total = 0;
  std::vector<std::pair<const void *, size_t>> tabs;  // host source, bytes
  std::vector<size_t> tab_off;
  size_t ell_col, ell_val, k0, phi, phi_prev, ya, yb, acc, y0, coef, coefT;
};

template <typename T>
void add_tab(Layout &L, const std::vector<T> &v) {
  L.tab_off.push_back(L.total);
  L.tabs.push_back({v.data(), v.size() * sizeof(T)});
  L.total += rsz(std::max<size_t>(v.size() * sizeof(T), 16));
}

size_t stage_bytes(int precision) { return precision == HOPS_PREC_FP64 ? 16 : 8; }

Layout layout_of(const Basis &B, int precision) {
  Layout L;
  // order = order of assignment in set_op_pointers
  add_tab(L, B.aux_ptr); add_tab(L, B.aux_mode_t); add_tab(L, B.aux_val_t);
  add_tab(L, B.up_ptr); add_tab(L, B.up_aux); add_tab(L, B.up_mode);
  add_tab(L, B.dn_ptr); add_tab(L, B.dn_aux); add_tab(L, B.dn_mode); add_tab(L, B.dn_val);
  add_tab(L, B.lr_ptr); add_tab(L, B.lr_col); add_tab(L, B.lr_val);
  add_tab(L, B.lc_ptr); add_tab(L, B.lc_row); add_tab(L, B.lc_val);
  add_tab(L, B.mode_lrel32); add_tab(L, B.mode_lact32); add_tab(L, B.mode_abs32);
  add_tab(L, B.mode_g); add_tab(L, B.mode_w); add_tab(L, B.abs_g); add_tab(L, B.abs_w); add_tab(L, B.abs_lact);
  add_tab(L, B.act_lrel32); add_tab(L, B.act_zrow); add_tab(L, B.am_ptr); add_tab(L, B.am_mode);
  add_tab(L, B.la_ptr); add_tab(L, B.la_row); add_tab(L, B.la_col); add_tab(L, B.la_val);
  add_tab(L, B.lq_ptr); add_tab(L, B.lq_row); add_tab(L, B.lq_col); add_tab(L, B.lq_val); add_tab(L, B.lt);
  add_tab(L, B.first_aux); add_tab(L, B.first_lact); add_tab(L, B.first_mode);
  add_tab(L, B.pat_ptr); add_tab(L, B.pat_col); add_tab(L, B.pat_h);
  add_tab(L, B.pc_ptr); add_tab(L, B.pc_act); add_tab(L, B.pc_kind); add_tab(L, B.pc_val);
  const size_t nr = (size_t)B.n * B.nh, es = stage_bytes(precision);
  const size_t w = std::max(B.ell_w, 1);
  auto put = [&](size_t bytes) {
    const size_t o = L.total;
    L.total += rsz(std::max<size_t>(bytes, 16));
    return o;
  };
  L.ell_col = put(w * nr * 4);
  L.ell_val = put(w * nr * es);
  L.k0 = put((size_t)B.nh * 16);
  L.phi = put(nr * 16);
  L.phi_prev = put(nr * 16);
  L.ya = put(nr * es);
  L.yb = put(nr * es);
  L.acc = put(nr * es);
  L.y0 = put(precision == HOPS_PREC_MIXED ? nr * 8 : 16);
  const size_t coef_stride = 4 * (size_t)B.n_pat + 2 * B.nm + 4;
  L.coef = put(coef_stride * 8);
  const size_t rowsT = 2 * (size_t)B.n_pat + B.nm + 1;
  L.coefT = put(rowsT * es);
  return L;
}

size_t pers_bytes(const System &S) {
  const size_t Ma = S.n_hm;
  return rsz(Ma * 16 + 16) * 2 + rsz((4 * Ma + 2) * 8) + rsz(16);
}

void set_op(const Basis &B, char *base, const Layout &L, int precision, hops_op_t &op) {
  op = hops_op_t{};
  op.n_state = B.n;
  op.n_hier = B.nh;
  op.n_mode = B.nm;
  op.n_lop = B.nl;
  op.n_act = B.na;
  op.n_first = (int)B.first_aux.size();
  op.ell_w = B.ell_w;
  op.n_pat = B.n_pat;
  op.n_mode_abs = B.sys->n_hm;
  op.pat_w = B.pat_w;
  size_t t = 0;
  auto P = [&](auto *&field) {
    using T = std::remove_reference_t<decltype(field)>;
    field = reinterpret_cast<T>(base + L.tab_off[t++]);
  };
  P(op.aux_ptr_d); P(op.aux_mode_d); P(op.aux_val_d);
  P(op.up_ptr_d); P(op.up_aux_d); P(op.up_mode_d);
  P(op.dn_ptr_d); P(op.dn_aux_d); P(op.dn_mode_d); P(op.dn_val_d);
  P(op.lr_ptr_d); P(op.lr_col_d); P(op.lr_val_d);
  P(op.lc_ptr_d); P(op.lc_row_d); P(op.lc_val_d);
  P(op.mode_lrel_d); P(op.mode_lact_d); P(op.mode_abs_d);
  P(op.mode_g_d); P(op.mode_w_d); P(op.abs_g_d); P(op.abs_w_d); P(op.abs_lact_d);
  P(op.act_lrel_d); P(op.act_zrow_d); P(op.am_ptr_d); P(op.am_mode_d);
  P(op.la_ptr_d); P(op.la_row_d); P(op.la_col_d); P(op.la_val_d);
  P(op.lq_ptr_d); P(op.lq_row_d); P(op.lq_col_d); P(op.lq_val_d); P(op.lt_d);
  P(op.first_aux_d); P(op.first_lact_d); P(op.first_mode_d);
  P(op.pat_ptr_d); P(op.pat_col_d); P(op.pat_h_d);
  P(op.pc_ptr_d); P(op.pc_act_d); P(op.pc_kind_d); P(op.pc_val_d);
  if (!B.has_lt) op.lt_d = nullptr;
  op.ell_col_d = reinterpret_cast<int32_t *>(base + L.ell_col);
  op.ell_val_d = base + L.ell_val;
  op.k0_d = reinterpret_cast<double *>(base + L.k0);
  (void)precision;
}

// ---------------------------------------------------------------------------
// Staging: a host-pinned upload buffer mirrored by a device scratch region.
// Uploads of one phase go to increasing offsets; the cursor resets at sync.
// ---------------------------------------------------------------------------
struct Stage {
  Batch *b;
  size_t cur = 0, flushed = 0;
  // copy into the pinned buffer; the device copy is issued by flush()
  int push(const void *src, size_t bytes, void **dev) {
    char *h;
    RT_TRY(place(bytes, dev, &h));
    if (bytes) std::memcpy(h, src, bytes);
    return 0;
  }
  // room for bytes in the pinned buffer (filled by the caller) and its device twin
  int place(size_t bytes, void **dev, char **host) {
    const size_t need = cur + rsz(bytes);
    if (need > b->scratch_bytes || need > b->h_up_bytes) return rt_fail("staging area too small (%zu bytes)", need);
    *host = b->h_up + cur;
    *dev = b->arena.ptr(b->scratch_off) + cur;
    cur = need;
    return 0;
  }
  // device-side output region at the current cursor
  int reserve(size_t bytes, void **dev, size_t *off) {
    const size_t need = cur + rsz(bytes);
    if (need > b->scratch_bytes) return rt_fail("staging area too small (%zu bytes)", need);
    *off = cur;
    *dev = b->arena.ptr(b->scratch_off) + cur;
    cur = need;
    return 0;
  }
  // one host-to-device copy of everything placed since the last flush
  int flush() {
    if (cur > flushed)
      RT_HIP(hipMemcpyAsync(b->arena.ptr(b->scratch_off) + flushed, b->h_up + flushed, cur - flushed,
                            hipMemcpyHostToDevice, b->stream));
    flushed = cur;
    return 0;
  }
};

int ensure_staging(Batch &b, size_t bytes) {
  bytes = rsz(bytes) + (1 << 20);
  if (bytes <= b.scratch_bytes && bytes <= b.h_up_bytes && bytes <= b.h_down_bytes) return 0;
  RT_HIP(hipStreamSynchronize(b.stream));
  if (b.scratch_off >= 0) b.arena.release(b.scratch_off);
  const size_t nb = std::max(bytes, b.scratch_bytes * 3 / 2);
  b.scratch_off = b.arena.alloc_top(nb);
  if (b.scratch_off < 0) {
    b.scratch_bytes = 0;
    return rt_fail("arena full: %zu bytes of staging", nb);
  }
  b.scratch_bytes = nb;
  if (b.h_up) RT_HIP(hipHostFree(b.h_up));
  if (b.h_down) RT_HIP(hipHostFree(b.h_down));
  b.h_up = b.h_down = nullptr;
  RT_HIP(hipHostMalloc((void **)&b.h_up, nb, hipHostMallocDefault));
  RT_HIP(hipHostMalloc((void **)&b.h_down, nb, hipHostMallocDefault));
  b.h_up_bytes = b.h_down_bytes = nb;
  return 0;
}

std::vector<uint32_t> tiles_of(const Batch &b, const std::vector<int> &list, bool new_basis = false) {
  std::vector<uint32_t> t;
  for (int s : list) {
    const Basis &B = new_basis ? b.slots[s].Bnew : b.slots[s].B;
    const long nr = (long)B.n * B.nh;
    const long nt = (nr + 255) / 256;
    for (long i = 0; i < nt; ++i) t.push_back(((uint32_t)s << kSlotShift) | (uint32_t)i);
  }
  return t;
}

int shm_of(const Batch &b, const std::vector<int> &list) {
  int m = 0;
  for (int s : list) {
    const Basis &B = b.slots[s].B;
    m = std::max(m, 2 * B.n + 8 * B.na + 2 * B.sys->n_hm + 16);
  }
  return m;
}

// upload descriptor arrays and per-slot scalars
int upload_desc(Batch &b) {
  const size_t S = b.slots.size();
  RT_HIP(hipMemcpyAsync(b.arena.ptr(b.ops_off), b.ops.data(), S * sizeof(hops_op_t), hipMemcpyHostToDevice, b.stream));
  RT_HIP(hipMemcpyAsync(b.arena.ptr(b.bts_off), b.bts.data(), S * sizeof(hops_batch_t), hipMemcpyHostToDevice, b.stream));
  RT_HIP(hipMemcpyAsync(b.arena.ptr(b.it_off), b.its.data(), S * sizeof(int32_t), hipMemcpyHostToDevice, b.stream));
  RT_HIP(hipMemcpyAsync(b.arena.ptr(b.dt_off), b.dts.data(), S * sizeof(double), hipMemcpyHostToDevice, b.stream));
  // pageable sources: make sure the copies have read them before they change
  RT_HIP(hipStreamSynchronize(b.stream));
  return 0;
}

hops_op_t *d_ops(Batch &b) { return reinterpret_cast<hops_op_t *>(b.arena.ptr(b.ops_off)); }
hops_batch_t *d_bts(Batch &b) { return reinterpret_cast<hops_batch_t *>(b.arena.ptr(b.bts_off)); }

void fill_bt(const Batch &b, const Slot &sl, const Basis &B, const Layout &L, char *base, hops_batch_t &bt) {
  bt = hops_batch_t{};
  bt.B = 1;
  bt.precision = b.precision;
  bt.eom = sl.prm.eom;
  bt.n_t = sl.prm.n_t;
  bt.coef_stride = 4 * B.n_pat + 2 * B.nm + 4;
  bt.ws_stride = 4 * B.sys->n_hm + 2;
  bt.store_stride = 0;
  bt.lane_shift = 0;
  bt.dt_over_hbar = sl.prm.dt / kHbar;
  bt.noise_step = sl.prm.noise_step;
  bt.phi_d = base + L.phi;
  bt.y0_d = base + L.y0;
  bt.ya_d = base + L.ya;
  bt.yb_d = base + L.yb;
  bt.acc_d = base + L.acc;
  bt.z1_d = sl.z1s ? sl.z1s : sl.z1;
  bt.z2_d = sl.z1s ? sl.z2s : sl.z2;
  bt.noise_c128 = sl.c128 ? 1 : 0;
  char *pb = b.arena.ptr(sl.pers);
  const size_t Ma = B.sys->n_hm;
  bt.zmem_d = reinterpret_cast<double *>(pb);
  bt.ws_d = reinterpret_cast<double *>(pb + 2 * rsz(Ma * 16 + 16));
  bt.coef_d = reinterpret_cast<double *>(base + L.coef);
  bt.psi_store_d = nullptr;
  bt.tile_d = nullptr;
  bt.n_tiles = 1;
  bt.tile_rows = 256;
  bt.coefT_d = base + L.coefT;
  bt.sigT_d = reinterpret_cast<double *>(pb + 2 * rsz(Ma * 16 + 16) + rsz((4 * Ma + 2) * 8));
  bt.coefT_stride = 1;
}

// allocate the block of a basis, stage its tables in host_tables, point op/bt at it
int place_basis(Batch &b, Slot &sl, Basis &B, long long *blk_out, hops_op_t &op, hops_batch_t &bt,
                std::vector<char> &host_tables) {
  Layout L = layout_of(B, b.precision);
  const long long blk = b.arena.alloc(L.total);
  if (blk < 0) return rt_fail("arena full: a basis of %d states x %d auxiliaries needs %zu bytes", B.n, B.nh, L.total);
  char *base = b.arena.ptr(blk);
  set_op(B, base, L, b.precision, op);
  const size_t tab_bytes = L.tab_off.empty() ? 0 : L.ell_col;
  host_tables.assign(tab_bytes, 0);
  for (size_t i = 0; i < L.tabs.size(); ++i)
    if (L.tabs[i].second) std::memcpy(host_tables.data() + L.tab_off[i], L.tabs[i].first, L.tabs[i].second);
  *blk_out = blk;
  fill_bt(b, sl, B, L, base, bt);
  sl.blk_bytes = L.total;
  return 0;
}

// ---- the step machinery ----------------------------------------------------

// the noise the listed slots' kernels read: the stage noise for an RK step
// (stage = true), the point samples for the selection's derivatives
void point_noise(Batch &b, const std::vector<int> &list, bool stage) {
  for (int s : list) {
    const Slot &sl = b.slots[s];
    const bool eff = stage && sl.z1s;
    b.bts[s].z1_d = eff ? sl.z1s : sl.z1;
    b.bts[s].z2_d = eff ? sl.z2s : sl.z2;
  }
}

int run_step(Batch &b, const std::vector<int> &list) {
  if (list.empty()) return 0;
  Timer tmr(&b.tm[0]);
  for (int s : list) b.its[s] = b.slots[s].it;
  point_noise(b, list, true);
  RT_TRY(upload_desc(b));
  std::vector<uint32_t> tiles = tiles_of(b, list);
  Stage st{&b};
  void *dl, *dt;
  RT_TRY(ensure_staging(b, tiles.size() * 4 + list.size() * 4 + 4096));
  RT_TRY(st.push(list.data(), list.size() * 4, &dl));
  RT_TRY(st.push(tiles.data(), tiles.size() * 4, &dt));
  RT_TRY(st.flush());
  for (int s : list) b.tm[7] += (double)b.slots[s].B.n * b.slots[s].B.nh;
  RT_TRY(hops_slots_step(d_ops(b), d_bts(b), (const int32_t *)dl, (int)list.size(), (const uint32_t *)dt,
                         (int)tiles.size(), (const int32_t *)b.arena.ptr(b.it_off), b.precision, shm_of(b, list),
                         b.stream));
  RT_HIP(hipStreamSynchronize(b.stream));
  return 0;
}

// copy psi (psi_only) of the listed slots to the host
int fetch_psi(Batch &b, const std::vector<int> &list) {
  if (list.empty()) return 0;
  Timer tmr(&b.tm[6]);
  std::vector<int64_t> off(b.slots.size(), 0);
  size_t tot = 0;
  for (int s : list) {
    off[s] = (int64_t)tot;
    tot += b.slots[s].B.n;
  }
  // the kernel reads the slots' descriptors on the device: upload them, as
  // this may run before any step or selection of the call has (a first step
  // whose slots all start from a static basis)
  RT_TRY(upload_desc(b));
  RT_TRY(ensure_staging(b, tot * 16 + off.size() * 8 + list.size() * 4 + 4096));
  Stage st{&b};
  void *dl, *doff, *dout;
  size_t oo;
  RT_TRY(st.push(list.data(), list.size() * 4, &dl));
  RT_TRY(st.push(off.data(), off.size() * 8, &doff));
  RT_TRY(st.reserve(tot * 16, &dout, &oo));
  RT_TRY(st.flush());
  RT_TRY(hops_slots_pn(d_ops(b), d_bts(b), (const int32_t *)dl, (int)list.size(), (const int64_t *)doff, dout,
                       nullptr, b.precision, 1, b.stream));
  RT_HIP(hipMemcpyAsync(b.h_down, dout, tot * 16, hipMemcpyDeviceToHost, b.stream));
  RT_HIP(hipStreamSynchronize(b.stream));
  const cplx *h = reinterpret_cast<const cplx *>(b.h_down);
  for (int s : list) {
    Slot &sl = b.slots[s];
    sl.pn.assign(h + off[s], h + off[s] + sl.B.n);
    sl.nrm = 0;  // marks psi-only content
  }
  return 0;
}

// define_basis for the listed slots on the host threads; dt per slot in b.dts
int define_bases_host(Batch &b, const std::vector<int> &list) {
  if (list.empty()) return 0;
  const size_t S = b.slots.size();
  std::vector<int64_t> off(S, 0), eoff(S, 0);
  size_t tot = 0, etot = 0;
  for (int s : list) {
    off[s] = (int64_t)tot;
    tot += (size_t)b.slots[s].B.n * b.slots[s].B.nh;
    eoff[s] = (int64_t)etot;
    etot += std::max(b.slots[s].B.nh, b.slots[s].B.n);
  }
  auto t_dev = std::chrono::steady_clock::now();
  const std::vector<uint32_t> tiles = tiles_of(b, list);
  for (int s : list) b.its[s] = b.slots[s].it;
  // the noise operator of slots with off-diagonal L-operators needs z1 / z2
  // of the basis L at the step's noise index and the noise memory
  // (hops_basis.py:914-953): one gather of 8-byte words for all of them
  std::vector<const void *> zsrc;
  std::vector<size_t> zat(S, 0);
  for (int s : list) {
    const Slot &sl = b.slots[s];
    if (!sl.B.has_off) continue;
    if (sl.it < 0 || sl.it >= sl.prm.n_t) return rt_fail("noise index %d outside the noise grid", sl.it);
    zat[s] = zsrc.size();
    const size_t zw = sl.c128 ? 16 : 8;  // bytes per noise value
    for (int z = 0; z < (sl.z2 ? 2 : 1); ++z) {
      const char *zb = static_cast<const char *>(z ? sl.z2 : sl.z1);
      for (int l : sl.B.absL)
        for (size_t w = 0; w < zw; w += 8) zsrc.push_back(zb + ((size_t)l * sl.prm.n_t + sl.it) * zw + w);
    }
    const char *zm = reinterpret_cast<const char *>(b.bts[s].zmem_d);
    for (int w = 0; w < 2 * sl.sys->n_hm; ++w) zsrc.push_back(zm + (size_t)w * 8);
  }
  RT_TRY(upload_desc(b));
  RT_TRY(ensure_staging(b, tot * 17 + etot * 16 + S * 8 * 6 + tiles.size() * 8 + list.size() * 8 + zsrc.size() * 16 +
                               65536));
  Stage st{&b};
  void *dl, *dtl, *doff, *deoff, *dpn, *de, *dzs = nullptr, *dz = nullptr;
  size_t opn, oe, oz;
  RT_TRY(st.push(list.data(), list.size() * 4, &dl));
  RT_TRY(st.push(tiles.data(), tiles.size() * 4, &dtl));
  RT_TRY(st.push(off.data(), S * 8, &doff));
  RT_TRY(st.push(eoff.data(), S * 8, &deoff));
  RT_TRY(st.reserve(tot * 16, &dpn, &opn));
  RT_TRY(st.reserve(etot * 8, &de, &oe));
  if (!zsrc.empty()) {
    RT_TRY(st.push(zsrc.data(), zsrc.size() * 8, &dzs));
    RT_TRY(st.reserve(zsrc.size() * 8, &dz, &oz));
  }
  RT_TRY(st.flush());
  if (!zsrc.empty())
    RT_TRY(hops_gather_u64((const uint64_t *const *)dzs, (uint64_t *)dz, (int)zsrc.size(), b.stream));
  double *dnrm = reinterpret_cast<double *>(b.arena.ptr(b.nrm_off));
  const int shm = shm_of(b, list);
  RT_TRY(hops_slots_pn(d_ops(b), d_bts(b), (const int32_t *)dl, (int)list.size(), (const int64_t *)doff, dpn, dnrm,
                       b.precision, 0, b.stream));
  bool any_h = false, any_s = false;
  for (int s : list) {
    any_h |= b.slots[s].prm.delta_a > 0;
    any_s |= b.slots[s].prm.delta_s > 0;
  }
  if (any_h)
    RT_TRY(hops_slots_errors(d_ops(b), d_bts(b), (const int32_t *)dl, (int)list.size(), (const uint32_t *)dtl,
                             (int)tiles.size(), (const int32_t *)b.arena.ptr(b.it_off), b.precision, shm,
                             (const int64_t *)doff, dpn, nullptr, nullptr, (const int64_t *)deoff, (double *)de,
                             (const double *)b.arena.ptr(b.dt_off), kHbar, b.stream));
  RT_HIP(hipMemcpyAsync(b.h_down, dpn, tot * 16, hipMemcpyDeviceToHost, b.stream));
  RT_HIP(hipMemcpyAsync(b.h_down + tot * 16, de, etot * 8, hipMemcpyDeviceToHost, b.stream));
  if (!zsrc.empty())
    RT_HIP(hipMemcpyAsync(b.h_down + tot * 16 + etot * 8, dz, zsrc.size() * 8, hipMemcpyDeviceToHost, b.stream));
  RT_HIP(hipMemcpyAsync(b.nrms.data(), dnrm, S * 8, hipMemcpyDeviceToHost, b.stream));
  RT_HIP(hipStreamSynchronize(b.stream));
  if (!zsrc.empty()) {
    const char *hz = b.h_down + tot * 16 + etot * 8;
    for (int s : list) {
      Slot &sl = b.slots[s];
      if (!sl.B.has_off) continue;
      const size_t zw = sl.c128 ? 16 : 8;
      const char *z1 = hz + zat[s] * 8;
      const char *z2 = sl.z2 ? z1 + (size_t)sl.B.nl * zw : nullptr;
      const auto *zm = reinterpret_cast<const cplx *>(z1 + (size_t)sl.B.nl * zw * (sl.z2 ? 2 : 1));
      noise_coef(sl.B, z1, z2, sl.c128, zm, sl.zcoef);
    }
  }
  // the selection reads Pn and the error sums in place (pinned memory); the
  // slot keeps psi = Pn[:n]
  const cplx *hpn = reinterpret_cast<const cplx *>(b.h_down);
  const double *he = reinterpret_cast<const double *>(b.h_down + tot * 16);
  for (int s : list) {
    Slot &sl = b.slots[s];
    sl.pn.assign(hpn + off[s], hpn + off[s] + sl.B.n);
    sl.nrm = b.nrms[s];
  }
  auto t_h = std::chrono::steady_clock::now();
  b.tm[1] += std::chrono::duration<double>(t_h - t_dev).count();
  b.pool->parallel_for((int)list.size(), [&](int i) {
    Slot &sl = b.slots[list[i]];
    select_hier(sl.prm, sl.B, hpn + off[list[i]], he + eoff[list[i]], sl.sel,
                sl.B.has_off ? sl.zcoef.data() : nullptr);
  });
  auto t_s = std::chrono::steady_clock::now();
  b.tm[2] += std::chrono::duration<double>(t_s - t_h).count();
  Timer tmr_s(&b.tm[3]);
  if (any_s) {
    std::vector<int> ls;
    for (int s : list)
      if (b.slots[s].prm.delta_s > 0) ls.push_back(s);
    std::vector<int64_t> moff(S, 0), soff(S, 0);
    size_t mtot = 0, stot = 0;
    for (int s : ls) {
      moff[s] = (int64_t)mtot;
      mtot += b.slots[s].B.nh;
      soff[s] = (int64_t)stot;
      stot += b.slots[s].B.n;
    }
    std::vector<uint8_t> mask(mtot);
    for (int s : ls) std::memcpy(mask.data() + moff[s], b.slots[s].sel.stable_mask.data(), b.slots[s].B.nh);
    const std::vector<uint32_t> tl = tiles_of(b, ls);
    void *dls, *dtls, *dmoff, *dmask, *dsoff, *des;
    size_t oes;
    RT_TRY(st.push(ls.data(), ls.size() * 4, &dls));
    RT_TRY(st.push(tl.data(), tl.size() * 4, &dtls));
    RT_TRY(st.push(moff.data(), S * 8, &dmoff));
    RT_TRY(st.push(mask.data(), mtot, &dmask));
    RT_TRY(st.push(soff.data(), S * 8, &dsoff));
    RT_TRY(st.reserve(stot * 8, &des, &oes));
    RT_TRY(st.flush());
    RT_TRY(hops_slots_errors(d_ops(b), d_bts(b), (const int32_t *)dls, (int)ls.size(), (const uint32_t *)dtls,
                             (int)tl.size(), (const int32_t *)b.arena.ptr(b.it_off), b.precision, shm_of(b, ls),
                             (const int64_t *)doff, dpn, (const int64_t *)dmoff, (const uint8_t *)dmask,
                             (const int64_t *)dsoff, (double *)des, (const double *)b.arena.ptr(b.dt_off), kHbar,
                             b.stream));
    char *hs_at = b.h_down + tot * 16 + etot * 8;  // after Pn and the aux sums
    RT_HIP(hipMemcpyAsync(hs_at, des, stot * 8, hipMemcpyDeviceToHost, b.stream));
    RT_HIP(hipStreamSynchronize(b.stream));
    const double *hs = reinterpret_cast<const double *>(hs_at);
    b.pool->parallel_for((int)ls.size(), [&](int i) {
      Slot &sl = b.slots[ls[i]];
      select_states(sl.prm, sl.B, hpn + off[ls[i]], hs + soff[ls[i]], sl.sel);
    });
  }
  for (int s : list) {
    Slot &sl = b.slots[s];
    if (sl.prm.delta_s <= 0) sl.sel.states_new = sl.B.states;
  }
  return 0;
}

// ---- selection on the device (hops_select.hip) ------------------------------
//
// The same selection as select_hier / select_states, with the per-auxiliary
// work on the device: the error bound of every auxiliary, the stable set,
// the boundary candidates, the pre-trim, the binning by target auxiliary and
// the per-state sums.  Only psi, a few counts per slot, the stable flags and
// the boundary auxiliaries cross PCIe; the host keeps the per-slot work that
// is O(n_state) or O(boundary) (X2, the coupled outside states, the filters
// of the state half and the state thresholds).

// the slot's system's Zobrist tables on the device, shared by the slots of
// that system in the batch (uploaded on the first device selection); a slot
// whose tables do not fit in the arena keeps the host selection
int ensure_zob(Batch &b, Slot &sl) {
  if (sl.zob_d || !sl.sys) return 0;
  auto it = b.zob.find(sl.sys);
  if (it == b.zob.end()) {
    const System &S = *sl.sys;
    const size_t zb = S.zob.size() * 8;
    const long long zo = b.arena.alloc(2 * zb);
    if (zo < 0) return 0;
    RT_HIP(hipMemcpyAsync(b.arena.ptr(zo), S.zob.data(), zb, hipMemcpyHostToDevice, b.stream));
    RT_HIP(hipMemcpyAsync(b.arena.ptr(zo) + zb, S.zob2.data(), zb, hipMemcpyHostToDevice, b.stream));
    RT_HIP(hipStreamSynchronize(b.stream));  // pageable sources
    it = b.zob.emplace(sl.sys, std::make_pair(zo, 0)).first;
  }
  it->second.second++;
  sl.zob_d = reinterpret_cast<const uint64_t *>(b.arena.ptr(it->second.first));
  return 0;
}

// a slot whose selection runs on the device: hierarchy adaptivity on, no
// Triangular / LongEdge filter, at most one diagonal entry per mode's L
bool dev_eligible(const Slot &sl) {
  if (!(sl.prm.delta_a > 0) || !sl.prm.filters.empty() || !sl.zob_d || sl.B.has_off) return false;
  const Basis &B = sl.B;
  for (int lm = 0; lm < B.nm; ++lm) {
    const RedL &R = B.L[B.mode_lrel[lm]];
    int c = 0;
    for (int s = 0; s < B.n; ++s)
      for (int q = R.rp[s]; q < R.rp[s + 1]; ++q) c += R.ci[q] == s;
    if (c > 1) return false;
  }
  return true;
}

// per relative mode: the terms select_hier forms from S.W, S.G, M2 and X2
void sel_modes(const Params &P, const Basis &B, const Selection &sel, hops_sel_mode_t *md) {
  const System &S = *B.sys;
  for (int lm = 0; lm < B.nm; ++lm) {
    const int m = B.modes[lm];
    hops_sel_mode_t &c = md[lm];
    c.aw = std::abs(S.W[m]);
    c.gw2 = std::norm(S.G[m] / S.W[m]);
    c.nX = std::norm(sel.X2[lm]);
    if (sel.M2[lm].empty()) {
      c.m2s = -1;
      c.m2u = c.m2d = 0.0;
    } else {
      const auto &e = sel.M2[lm][0];
      c.m2s = e.first;
      c.m2u = std::norm(e.second);
      c.m2d = std::norm(e.second - sel.X2[lm]);
    }
    c.markov = !P.markov.empty() && P.markov[m] ? 1 : 0;
  }
}

// the (mode, auxiliary) pairs of the state half's flux filters
// (hops_fluxfilters.py:292-352): F_dn[lm, j] for j = a + e_lm and F_up[lm, j]
// for j = a - e_lm, a a boundary auxiliary, j a stable auxiliary; sorted by
// (lm, j), each as (lm << 2 | F_dn << 1 | F_up, j)
void state_flag_list(const Basis &B, const Selection &sel, std::vector<uint32_t> &out) {
  const System &S = *B.sys;
  std::vector<std::pair<uint64_t, uint32_t>> f;  // ((lm << 32 | j), flag bit)
  for (const Aux &a : sel.aux_bound) {
    const AKey ka = key_of(S, a);
    for (int lm = 0; lm < B.nm; ++lm) {
      const int v = aux_get(a, B.modes[lm]);
      if (v >= 255) continue;
      const int j = B.index.find(key_step(S, ka, B.modes[lm], v, v + 1));
      if (j >= 0 && sel.stable_mask[j]) f.push_back({((uint64_t)lm << 32) | (uint32_t)j, 2u});
    }
    for (char32_t e : a) {
      const int j = B.index.find(key_step(S, ka, aux_mode(e), aux_val(e), aux_val(e) - 1));
      if (j >= 0 && sel.stable_mask[j]) f.push_back({((uint64_t)B.rel_mode[aux_mode(e)] << 32) | (uint32_t)j, 1u});
    }
  }
  std::sort(f.begin(), f.end());
  out.clear();
  for (size_t i = 0; i < f.size();) {
    uint32_t bits = 0;
    size_t e = i;
    for (; e < f.size() && f[e].first == f[i].first; ++e) bits |= f[e].second;
    out.push_back((uint32_t)(f[i].first >> 32) << 2 | bits);
    out.push_back((uint32_t)(f[i].first & 0xffffffffu));
    i = e;
  }
}

// bump layout of a device region: offsets first (base null), pointers after
struct Lay {
  char *base = nullptr;
  size_t cur = 0;
  template <typename T>
  T *take(size_t n) {
    T *p = reinterpret_cast<T *>(base + cur);
    cur += rsz(std::max<size_t>(n * sizeof(T), 16));
    return p;
  }
};

// device region r of at least bytes (its contents are not kept on growth)
int ensure_sel(Batch &b, int r, size_t bytes) {
  if (bytes <= b.sel_bytes[r]) return 0;
  RT_HIP(hipStreamSynchronize(b.stream));
  if (b.sel_off[r] >= 0) b.arena.release(b.sel_off[r]);
  const size_t nb = std::max(rsz(bytes) + (1 << 20), b.sel_bytes[r] * 3 / 2);
  b.sel_off[r] = b.arena.alloc_top(nb);
  if (b.sel_off[r] < 0) {
    b.sel_bytes[r] = 0;
    return HOPS_RT_ARENA_FULL;
  }
  b.sel_bytes[r] = nb;
  return 0;
}

// host<->device copies through the pinned staging buffers; each phase starts
// at offset 0 and ends with a stream synchronisation
struct Xfer {
  Batch &b;
  size_t up = 0, down = 0;
  int put(void *dev, const void *src, size_t bytes) {
    if (!bytes) return 0;
    if (up + rsz(bytes) > b.h_up_bytes) return rt_fail("selection upload exceeds the staging buffer");
    std::memcpy(b.h_up + up, src, bytes);
    RT_HIP(hipMemcpyAsync(dev, b.h_up + up, bytes, hipMemcpyHostToDevice, b.stream));
    up += rsz(bytes);
    return 0;
  }
  int get(const void *dev, size_t bytes, char **host) {
    if (down + rsz(bytes) > b.h_down_bytes) return rt_fail("selection download exceeds the staging buffer");
    *host = b.h_down + down;
    if (bytes) RT_HIP(hipMemcpyAsync(*host, dev, bytes, hipMemcpyDeviceToHost, b.stream));
    down += rsz(std::max<size_t>(bytes, 16));
    return 0;
  }
  int sync() {
    RT_HIP(hipStreamSynchronize(b.stream));
    up = down = 0;
    return 0;
  }
};

template <typename T>
size_t vbytes(const std::vector<T> &v) {
  return v.size() * sizeof(T);
}

int define_bases_dev(Batch &b, const std::vector<int> &list) {
  const int L = (int)list.size();
  const size_t S = b.slots.size();
  auto t_dev = std::chrono::steady_clock::now();
  // ---- layout of the per-auxiliary work --------------------------------------
  std::vector<hops_sel_seg_t> seg(L);
  std::vector<int64_t> off(S, 0), eoff(S, 0), poff(L), ebeg(L), eend(L);
  std::vector<int2> atl;
  size_t tot = 0, etot = 0, htot = 0, ttot = 0, tctot = 0, mtot = 0, stot = 0;
  bool any_s = false;
  for (int i = 0; i < L; ++i) {
    const Slot &sl = b.slots[list[i]];
    const Basis &B = sl.B;
    hops_sel_seg_t &g = seg[i];
    g = hops_sel_seg_t{};
    g.slot = list[i];
    g.n = B.n;
    g.nh = B.nh;
    g.nm = B.nm;
    g.maxhier = sl.prm.maxhier;
    for (uint8_t f : sl.prm.markov) g.markov |= f ? 1 : 0;
    g.n_tiles = (B.nh + 255) / 256;
    g.pn_off = off[list[i]] = (int64_t)tot;
    tot += (size_t)B.n * B.nh;
    g.e_off = eoff[list[i]] = (int64_t)etot;
    etot += std::max(B.nh, B.n);
    g.h_off = ebeg[i] = (int64_t)htot;
    htot += B.nh;
    eend[i] = (int64_t)htot;
    g.t_off = (int64_t)ttot;
    ttot += g.n_tiles;
    g.tc_off = (int64_t)tctot;
    tctot += (size_t)g.n_tiles * B.n;
    g.m_off = (int64_t)mtot;
    mtot += B.nm;
    g.s_off = poff[i] = (int64_t)stot;
    stot += B.n;
    g.dsq = sl.prm.delta_a * sl.prm.delta_a;
    g.f_discard = sl.prm.f_discard;
    g.zob_a = sl.zob_d;
    g.zob_b = sl.zob_d + (size_t)sl.sys->n_hm * 256;
    for (int t = 0; t < g.n_tiles; ++t) atl.push_back(make_int2(i, t));
    any_s |= sl.prm.delta_s > 0;
  }
  // couplings of the outside states (basis only)
  std::vector<std::vector<int>> sc(L), hp(L), hs(L);
  std::vector<std::vector<cplx>> hv(L);
  b.pool->parallel_for(L, [&](int i) {
    const Slot &sl = b.slots[list[i]];
    sc[i] = list_sc(*sl.sys, sl.B);
    sc_couplings(sl.B, sc[i], hp[i], hs[i], hv[i]);
  });
  std::vector<int32_t> scp;
  std::vector<hops_sel_sc_t> sce;
  for (int i = 0; i < L; ++i) {
    seg[i].n_sc = (int)sc[i].size();
    seg[i].sc_off = (int64_t)scp.size();
    seg[i].sce_off = (int64_t)sce.size();
    for (int v : hp[i]) scp.push_back(v);
    for (size_t q = 0; q < hs[i].size(); ++q) sce.push_back(hops_sel_sc_t{hv[i][q].real(), hv[i][q].imag(), hs[i][q], 0});
  }
  const std::vector<uint32_t> tiles = tiles_of(b, list);
  size_t tmp_err = 0;
  RT_TRY(hops_sel_sort_keys(nullptr, &tmp_err, nullptr, nullptr, htot, L, nullptr, nullptr, b.stream));
  // ---- region 0 ------------------------------------------------------------------
  struct R0 {
    int32_t *list;
    uint32_t *tiles;
    int64_t *off, *eoff, *poff, *ebeg, *eend, *gbeg, *gend;
    double *pn, *e, *psi;
    hops_sel_seg_t *segs;
    int2 *atl;
    hops_sel_mode_t *modes;
    int32_t *scp;
    hops_sel_sc_t *sce;
    double *err, *err_sorted, *part_u, *part_c, *c1, *scal;
    uint8_t *flag;
    int32_t *cnt, *scnt;
    int64_t *tot;
    void *tmp;
  } a{};
  auto lay0 = [&](Lay &l) {
    a.list = l.take<int32_t>(L);
    a.tiles = l.take<uint32_t>(tiles.size());
    a.off = l.take<int64_t>(S);
    a.eoff = l.take<int64_t>(S);
    a.poff = l.take<int64_t>(L);
    a.ebeg = l.take<int64_t>(L);
    a.eend = l.take<int64_t>(L);
    a.gbeg = l.take<int64_t>(L);
    a.gend = l.take<int64_t>(L);
    a.pn = l.take<double>(2 * tot);
    a.e = l.take<double>(etot);
    a.psi = l.take<double>(2 * stot);
    a.segs = l.take<hops_sel_seg_t>(L);
    a.atl = l.take<int2>(atl.size());
    a.modes = l.take<hops_sel_mode_t>(mtot);
    a.scp = l.take<int32_t>(scp.size());
    a.sce = l.take<hops_sel_sc_t>(sce.size());
    a.err = l.take<double>(htot);
    a.err_sorted = l.take<double>(htot);
    a.part_u = l.take<double>(ttot);
    a.part_c = l.take<double>(tctot);
    a.c1 = l.take<double>(stot);
    a.scal = l.take<double>(8 * (size_t)L);
    a.flag = l.take<uint8_t>(htot);
    a.cnt = l.take<int32_t>(htot);
    a.scnt = l.take<int32_t>(htot);
    a.tot = l.take<int64_t>(8 * (size_t)L);
    a.tmp = l.take<char>(tmp_err);
  };
  {
    Lay l;
    lay0(l);
    int rc = ensure_sel(b, 0, l.cur);
    if (rc == HOPS_RT_ARENA_FULL) return rc;
    RT_TRY(rc);
    Lay l2;
    l2.base = b.arena.ptr(b.sel_off[0]);
    lay0(l2);
  }
  RT_TRY(ensure_staging(b, std::max({vbytes(seg) + vbytes(atl) + vbytes(scp) + vbytes(sce) + tiles.size() * 4 +
                                         S * 16 + (size_t)L * 48 + mtot * sizeof(hops_sel_mode_t),
                                     stot * 16 + S * 8, htot + (size_t)L * 160}) +
                            (size_t)L * 4096 + 65536));
  hops_sel_work_t w{};
  w.ops = d_ops(b);
  w.segs = a.segs;
  w.pn = a.pn;
  w.e_aux = a.e;
  w.modes = a.modes;
  w.sc_ptr = a.scp;
  w.sc_ent = a.sce;
  w.err = a.err;
  w.err_sorted = a.err_sorted;
  w.flag = a.flag;
  w.cnt = a.cnt;
  w.scnt = a.scnt;
  w.part_u = a.part_u;
  w.part_c = a.part_c;
  w.c1 = a.c1;
  w.scal = a.scal;
  w.tot = a.tot;
  w.seg_beg = a.gbeg;
  w.seg_end = a.gend;
  w.hbar = kHbar;
  w.n_list = L;
  Xfer x{b};
  // ---- phase A: Pn, deletion-flux sums, psi --------------------------------------
  for (int s : list) b.its[s] = b.slots[s].it;
  RT_TRY(upload_desc(b));
  RT_TRY(x.put(a.list, list.data(), list.size() * 4));
  RT_TRY(x.put(a.tiles, tiles.data(), tiles.size() * 4));
  RT_TRY(x.put(a.off, off.data(), S * 8));
  RT_TRY(x.put(a.eoff, eoff.data(), S * 8));
  RT_TRY(x.put(a.poff, poff.data(), (size_t)L * 8));
  RT_TRY(x.put(a.ebeg, ebeg.data(), (size_t)L * 8));
  RT_TRY(x.put(a.eend, eend.data(), (size_t)L * 8));
  RT_TRY(x.put(a.segs, seg.data(), vbytes(seg)));
  RT_TRY(x.put(a.atl, atl.data(), vbytes(atl)));
  RT_TRY(x.put(a.scp, scp.data(), vbytes(scp)));
  RT_TRY(x.put(a.sce, sce.data(), vbytes(sce)));
  double *dnrm = reinterpret_cast<double *>(b.arena.ptr(b.nrm_off));
  const int shm = shm_of(b, list);
  RT_TRY(hops_slots_pn(d_ops(b), d_bts(b), a.list, L, a.off, a.pn, dnrm, b.precision, 0, b.stream));
  RT_TRY(hops_slots_errors(d_ops(b), d_bts(b), a.list, L, a.tiles, (int)tiles.size(),
                           (const int32_t *)b.arena.ptr(b.it_off), b.precision, shm, a.off, a.pn, nullptr, nullptr,
                           a.eoff, a.e, (const double *)b.arena.ptr(b.dt_off), kHbar, b.stream));
  RT_TRY(hops_sel_gather_psi(w, a.psi, a.poff, b.stream));
  char *hpsi, *hnrm;
  RT_TRY(x.get(a.psi, stot * 16, &hpsi));
  RT_TRY(x.get(dnrm, S * 8, &hnrm));
  RT_TRY(x.sync());
  std::vector<hops_sel_mode_t> modes(mtot);
  const cplx *psi = reinterpret_cast<const cplx *>(hpsi);
  for (int i = 0; i < L; ++i) {
    Slot &sl = b.slots[list[i]];
    sl.pn.assign(psi + poff[i], psi + poff[i] + sl.B.n);
    sl.nrm = reinterpret_cast<const double *>(hnrm)[list[i]];
  }
  auto t_h = std::chrono::steady_clock::now();
  b.tm[1] += std::chrono::duration<double>(t_h - t_dev).count();
  b.pool->parallel_for(L, [&](int i) {
    Slot &sl = b.slots[list[i]];
    sel_common(sl.B, sl.pn.data(), sl.sel);
    sel_modes(sl.prm, sl.B, sl.sel, modes.data() + seg[i].m_off);
  });
  // ---- phase B: errors, stable set, candidate counts -------------------------------
  RT_TRY(x.put(a.modes, modes.data(), vbytes(modes)));
  const int2 *datl = a.atl;
  RT_TRY(hops_sel_err(w, datl, (int)atl.size(), b.stream));
  RT_TRY(hops_sel_sort_keys(a.tmp, &tmp_err, a.err, a.err_sorted, htot, L, a.ebeg, a.eend, b.stream));
  RT_TRY(hops_sel_threshold(w, a.err_sorted, 0, b.stream));
  RT_TRY(hops_sel_stable(w, datl, (int)atl.size(), any_s ? 1 : 0, b.stream));
  RT_TRY(hops_sel_fin1(w, any_s ? 1 : 0, b.stream));
  char *htt;
  RT_TRY(x.get(a.tot, (size_t)L * 64, &htt));
  RT_TRY(x.sync());
  std::vector<int64_t> tt(reinterpret_cast<const int64_t *>(htt), reinterpret_cast<const int64_t *>(htt) + 8 * L);
  // ---- phase C: candidates, pre-trim threshold, survivor counts ----------------------
  size_t nztot = 0, nzt = 0;
  std::vector<int2> ntl;
  std::vector<int64_t> nzb(L), nze(L);
  for (int i = 0; i < L; ++i) {
    const int64_t nz = tt[8 * i + TT_NZ];
    seg[i].nz_off = nzb[i] = (int64_t)nztot;
    nztot += nz;
    nze[i] = (int64_t)nztot;
    seg[i].nzt_off = (int64_t)nzt;
    const int64_t nt = (nz + 255) / 256;
    nzt += nt;
    for (int64_t t = 0; t < nt; ++t) ntl.push_back(make_int2(i, (int)t));
  }
  size_t tmp_nz = 0;
  RT_TRY(hops_sel_sort_keys(nullptr, &tmp_nz, nullptr, nullptr, nztot, L, nullptr, nullptr, b.stream));
  struct R1 {
    double *nz, *nzs, *part_d;
    int64_t *nzb, *nze;
    int2 *ntl;
    void *tmp;
  } c{};
  auto lay1 = [&](Lay &l) {
    c.nz = l.take<double>(nztot);
    c.nzs = l.take<double>(nztot);
    c.part_d = l.take<double>(nzt);
    c.nzb = l.take<int64_t>(L);
    c.nze = l.take<int64_t>(L);
    c.ntl = l.take<int2>(ntl.size());
    c.tmp = l.take<char>(tmp_nz);
  };
  {
    Lay l;
    lay1(l);
    int rc = ensure_sel(b, 1, l.cur);
    if (rc == HOPS_RT_ARENA_FULL) return rc;
    RT_TRY(rc);
    Lay l2;
    l2.base = b.arena.ptr(b.sel_off[1]);
    lay1(l2);
  }
  w.nz = c.nz;
  w.nz_sorted = c.nzs;
  w.part_d = c.part_d;
  RT_TRY(ensure_staging(b, vbytes(seg) + vbytes(ntl) + (size_t)L * 96 + 65536));
  RT_TRY(x.put(a.segs, seg.data(), vbytes(seg)));
  RT_TRY(x.put(c.nzb, nzb.data(), (size_t)L * 8));
  RT_TRY(x.put(c.nze, nze.data(), (size_t)L * 8));
  RT_TRY(x.put(c.ntl, ntl.data(), vbytes(ntl)));
  RT_TRY(hops_sel_cand(w, datl, (int)atl.size(), b.stream));
  if (nztot) RT_TRY(hops_sel_sort_keys(c.tmp, &tmp_nz, c.nz, c.nzs, nztot, L, c.nzb, c.nze, b.stream));
  RT_TRY(hops_sel_threshold(w, c.nzs, 1, b.stream));
  RT_TRY(hops_sel_disc(w, c.ntl, (int)ntl.size(), datl, (int)atl.size(), b.stream));
  RT_TRY(hops_sel_fin2(w, b.stream));
  RT_TRY(x.get(a.tot, (size_t)L * 64, &htt));
  RT_TRY(x.sync());
  tt.assign(reinterpret_cast<const int64_t *>(htt), reinterpret_cast<const int64_t *>(htt) + 8 * L);
  // ---- phase D: survivors, binning, boundary threshold -------------------------------
  size_t svtot = 0;
  std::vector<int2> stl;
  std::vector<int64_t> svb(L), sve(L);
  for (int i = 0; i < L; ++i) {
    const int64_t nsv = tt[8 * i + TT_SURV];
    seg[i].sv_off = svb[i] = (int64_t)svtot;
    svtot += nsv;
    sve[i] = (int64_t)svtot;
    for (int64_t t = 0; t < (nsv + 255) / 256; ++t) stl.push_back(make_int2(i, (int)t));
  }
  size_t tmp_sv = 0, tmp_g = 0;
  RT_TRY(hops_sel_sort_pairs(nullptr, &tmp_sv, nullptr, nullptr, nullptr, nullptr, svtot, L, nullptr, nullptr,
                             b.stream));
  RT_TRY(hops_sel_sort_keys(nullptr, &tmp_g, nullptr, nullptr, svtot, L, nullptr, nullptr, b.stream));
  struct R2 {
    uint64_t *ka, *kas, *kb, *desc, *gdesc, *bound;
    uint32_t *pos, *poss;
    double *sval, *gv, *gvals, *gs;
    uint8_t *lead;
    int64_t *svb, *sve, *boff;
    int2 *stl;
    void *tmp;
  } d{};
  auto lay2 = [&](Lay &l) {
    d.ka = l.take<uint64_t>(svtot);
    d.kas = l.take<uint64_t>(svtot);
    d.kb = l.take<uint64_t>(svtot);
    d.desc = l.take<uint64_t>(svtot);
    d.gdesc = l.take<uint64_t>(svtot);
    d.bound = l.take<uint64_t>(svtot);
    d.pos = l.take<uint32_t>(svtot);
    d.poss = l.take<uint32_t>(svtot);
    d.sval = l.take<double>(svtot);
    d.gv = l.take<double>(svtot);
    d.gvals = l.take<double>(svtot);
    d.gs = l.take<double>(svtot);
    d.lead = l.take<uint8_t>(svtot);
    d.svb = l.take<int64_t>(L);
    d.sve = l.take<int64_t>(L);
    d.boff = l.take<int64_t>((size_t)L + 1);
    d.stl = l.take<int2>(stl.size());
    d.tmp = l.take<char>(std::max(tmp_sv, tmp_g));
  };
  {
    Lay l;
    lay2(l);
    int rc = ensure_sel(b, 2, l.cur);
    if (rc == HOPS_RT_ARENA_FULL) return rc;
    RT_TRY(rc);
    Lay l2;
    l2.base = b.arena.ptr(b.sel_off[2]);
    lay2(l2);
  }
  w.ka = d.ka;
  w.ka_sorted = d.kas;
  w.kb = d.kb;
  w.desc = d.desc;
  w.pos = d.pos;
  w.pos_sorted = d.poss;
  w.sval = d.sval;
  w.lead = d.lead;
  w.gv = d.gv;
  w.gvals = d.gvals;
  w.gsorted = d.gs;
  w.gdesc = d.gdesc;
  w.bound = d.bound;
  w.bound_off = d.boff;
  RT_TRY(ensure_staging(b, std::max(vbytes(seg) + vbytes(stl) + (size_t)L * 96, htot + (size_t)L * 200) + 65536));
  RT_TRY(x.put(a.segs, seg.data(), vbytes(seg)));
  RT_TRY(x.put(d.svb, svb.data(), (size_t)L * 8));
  RT_TRY(x.put(d.sve, sve.data(), (size_t)L * 8));
  RT_TRY(x.put(d.stl, stl.data(), vbytes(stl)));
  RT_TRY(hops_sel_surv(w, datl, (int)atl.size(), b.stream));
  if (svtot) {
    size_t tb = std::max(tmp_sv, tmp_g);
    RT_TRY(hops_sel_sort_pairs(d.tmp, &tb, d.ka, d.kas, d.pos, d.poss, svtot, L, d.svb, d.sve, b.stream));
  }
  RT_TRY(hops_sel_group(w, d.stl, (int)stl.size(), b.stream));
  RT_TRY(hops_sel_fin3(w, b.stream));
  if (svtot) {
    size_t tb = std::max(tmp_sv, tmp_g);
    RT_TRY(hops_sel_sort_keys(d.tmp, &tb, d.gvals, d.gs, svtot, L, a.gbeg, a.gend, b.stream));
  }
  RT_TRY(hops_sel_threshold(w, d.gs, 2, b.stream));
  RT_TRY(hops_sel_bound(w, b.stream));
  char *hboff, *hflag;
  RT_TRY(x.get(d.boff, ((size_t)L + 1) * 8, &hboff));
  RT_TRY(x.get(a.flag, htot, &hflag));
  RT_TRY(x.sync());
  const std::vector<int64_t> boff(reinterpret_cast<const int64_t *>(hboff),
                                  reinterpret_cast<const int64_t *>(hboff) + L + 1);
  std::vector<uint8_t> flags(reinterpret_cast<const uint8_t *>(hflag), reinterpret_cast<const uint8_t *>(hflag) + htot);
  RT_TRY(ensure_staging(b, (size_t)boff[L] * 8 + 65536));
  char *hb;
  RT_TRY(x.get(d.bound, (size_t)boff[L] * 8, &hb));
  RT_TRY(x.sync());
  const uint64_t *bd = reinterpret_cast<const uint64_t *>(hb);
  b.pool->parallel_for(L, [&](int i) {
    Slot &sl = b.slots[list[i]];
    const Basis &B = sl.B;
    Selection &out = sl.sel;
    out.stable_mask.assign(flags.begin() + seg[i].h_off, flags.begin() + seg[i].h_off + B.nh);
    out.stable_idx.clear();
    for (int k = 0; k < B.nh; ++k)
      if (out.stable_mask[k]) out.stable_idx.push_back(k);
    out.aux_new.clear();
    out.aux_new.reserve(out.stable_idx.size() + (size_t)(boff[i + 1] - boff[i]));
    for (int k : out.stable_idx) out.aux_new.push_back(B.aux[k]);
    out.aux_bound.clear();
    Aux tgt;
    for (int64_t q = boff[i]; q < boff[i + 1]; ++q) {
      const uint64_t dsc = bd[q];
      const int k = (int)(dsc & 0xffffffffu), lm = (int)((dsc >> 32) & 0xffff);
      aux_step(B.aux[k], B.modes[lm], (dsc >> 48) & 1 ? 1 : -1, tgt);
      out.aux_bound.push_back(tgt);
      out.aux_new.push_back(tgt);
    }
  });
  auto t_s = std::chrono::steady_clock::now();
  b.tm[2] += std::chrono::duration<double>(t_s - t_h).count();
  // ---- state half ------------------------------------------------------------------
  Timer tmr_s(&b.tm[3]);
  if (any_s) {
    std::vector<int> ls, li;  // slots, list indices
    for (int i = 0; i < L; ++i)
      if (b.slots[list[i]].prm.delta_s > 0) {
        ls.push_back(list[i]);
        li.push_back(i);
      }
    std::vector<int64_t> moff(S, 0), soff(S, 0);
    for (int i : li) {
      moff[list[i]] = seg[i].h_off;
      soff[list[i]] = seg[i].s_off;
    }
    std::vector<std::vector<uint32_t>> fls(L);
    std::vector<double> v1(stot, 0.0);
    b.pool->parallel_for((int)li.size(), [&](int q) {
      const int i = li[q];
      const Slot &sl = b.slots[list[i]];
      const Basis &B = sl.B;
      const System &Sy = *B.sys;
      state_flag_list(B, sl.sel, fls[i]);
      for (int s = 0; s < B.n; ++s) {
        double v = 0;
        for (int sp = 0; sp < Sy.n; ++sp)
          if (sp != B.states[s]) v += std::norm(Sy.H[(size_t)sp * Sy.n + B.states[s]]);
        v1[seg[i].s_off + s] = v;
      }
    });
    std::vector<uint32_t> flat;
    for (int i = 0; i < L; ++i) {
      seg[i].fl_off = (int64_t)flat.size() / 2;
      seg[i].n_fl = (int)fls[i].size() / 2;
      flat.insert(flat.end(), fls[i].begin(), fls[i].end());
    }
    const std::vector<uint32_t> stiles = tiles_of(b, ls);
    struct R3 {
      int32_t *ls;
      uint32_t *stiles;
      int64_t *moff, *soff;
      double *es, *v1, *err_s;
      uint32_t *fl;
    } e{};
    auto lay3 = [&](Lay &l) {
      e.ls = l.take<int32_t>(ls.size());
      e.stiles = l.take<uint32_t>(stiles.size());
      e.moff = l.take<int64_t>(S);
      e.soff = l.take<int64_t>(S);
      e.es = l.take<double>(stot);
      e.v1 = l.take<double>(stot);
      e.err_s = l.take<double>(stot);
      e.fl = l.take<uint32_t>(flat.size());
    };
    {
      Lay l;
      lay3(l);
      int rc = ensure_sel(b, 3, l.cur);
      if (rc == HOPS_RT_ARENA_FULL) return rc;
      RT_TRY(rc);
      Lay l2;
      l2.base = b.arena.ptr(b.sel_off[3]);
      lay3(l2);
    }
    RT_TRY(ensure_staging(b, vbytes(seg) + vbytes(flat) + stot * 24 + S * 16 + stiles.size() * 4 +
                                 ls.size() * 4 + 65536));
    RT_TRY(x.put(a.segs, seg.data(), vbytes(seg)));
    RT_TRY(x.put(e.ls, ls.data(), ls.size() * 4));
    RT_TRY(x.put(e.stiles, stiles.data(), stiles.size() * 4));
    RT_TRY(x.put(e.moff, moff.data(), S * 8));
    RT_TRY(x.put(e.soff, soff.data(), S * 8));
    RT_TRY(x.put(e.v1, v1.data(), vbytes(v1)));
    RT_TRY(x.put(e.fl, flat.data(), vbytes(flat)));
    RT_TRY(hops_slots_errors(d_ops(b), d_bts(b), e.ls, (int)ls.size(), e.stiles, (int)stiles.size(),
                             (const int32_t *)b.arena.ptr(b.it_off), b.precision, shm_of(b, ls), a.off, a.pn, e.moff,
                             a.flag, e.soff, e.es, (const double *)b.arena.ptr(b.dt_off), kHbar, b.stream));
    w.fl = e.fl;
    w.err_s = e.err_s;
    w.e_state = e.es;
    w.v1 = e.v1;
    RT_TRY(hops_sel_state_err(w, b.stream));
    char *hes;
    RT_TRY(x.get(e.err_s, stot * 8, &hes));
    RT_TRY(x.sync());
    const std::vector<double> err_s(reinterpret_cast<const double *>(hes), reinterpret_cast<const double *>(hes) + stot);
    // state thresholds (host: O(n_state) per slot), then the flux into the
    // outside coupled states from the stable states
    std::vector<std::vector<int>> rel_st(L), st_stable(L);
    std::vector<double> dbound(L, 0.0);
    std::vector<std::vector<hops_sel_sc_t>> qe(L);
    std::vector<int> nq(L, 0), nsc(L, 0);
    b.pool->parallel_for((int)li.size(), [&](int q) {
      const int i = li[q];
      const Slot &sl = b.slots[list[i]];
      const Basis &B = sl.B;
      const System &Sy = *B.sys;
      const int n = B.n;
      std::vector<double> srt(err_s.begin() + seg[i].s_off, err_s.begin() + seg[i].s_off + n);
      const std::vector<double> er(srt);
      std::sort(srt.begin(), srt.end());
      const double dsq = sl.prm.delta_s * sl.prm.delta_s;
      const double thr = error_threshold(srt, dsq / 2.0);
      double unstable = 0;
      for (int s = 0; s < n; ++s) {
        if (er[s] > thr) {
          rel_st[i].push_back(s);
          st_stable[i].push_back(B.states[s]);
        } else {
          unstable += er[s];
        }
      }
      dbound[i] = dsq - unstable;
      if ((int)rel_st[i].size() < Sy.n && !sl.sel.sc.empty()) {
        std::vector<std::vector<hops_sel_sc_t>> per(sl.sel.sc.size());
        size_t mx = 1;
        for (size_t ii = 0; ii < sl.sel.sc.size(); ++ii) {
          const int sp = sl.sel.sc[ii];
          for (size_t qq = 0; qq < rel_st[i].size(); ++qq) {
            const cplx h = Sy.H[(size_t)sp * Sy.n + st_stable[i][qq]];
            if (h != cplx(0, 0)) {
              const cplx mh = cmi(h);
              per[ii].push_back(hops_sel_sc_t{mh.real(), mh.imag(), rel_st[i][qq], 0});
            }
          }
          mx = std::max(mx, per[ii].size());
        }
        nq[i] = (int)mx;
        nsc[i] = (int)sl.sel.sc.size();
        qe[i].assign(sl.sel.sc.size() * mx, hops_sel_sc_t{0, 0, -1, 0});
        for (size_t ii = 0; ii < per.size(); ++ii)
          std::copy(per[ii].begin(), per[ii].end(), qe[i].begin() + ii * mx);
      }
    });
    std::vector<hops_sel_sc_t> qflat;
    size_t tetot = 0, esctot = 0;
    for (int i = 0; i < L; ++i) {
      seg[i].n_sc = nsc[i];
      seg[i].n_q = nq[i];
      seg[i].q_off = (int64_t)qflat.size();
      qflat.insert(qflat.end(), qe[i].begin(), qe[i].end());
      seg[i].te_off = (int64_t)tetot;
      tetot += (size_t)seg[i].n_tiles * nsc[i];
      seg[i].esc_off = (int64_t)esctot;
      esctot += nsc[i];
    }
    if (esctot) {
      struct R4 {
        hops_sel_sc_t *q;
        double *pe, *esc;
      } f{};
      auto lay4 = [&](Lay &l) {
        f.q = l.take<hops_sel_sc_t>(qflat.size());
        f.pe = l.take<double>(tetot);
        f.esc = l.take<double>(esctot);
      };
      // region 3 holds the state arrays above; region 1 (candidates) is
      // free again and takes these
      {
        Lay l;
        lay4(l);
        int rc = ensure_sel(b, 1, l.cur);
        if (rc == HOPS_RT_ARENA_FULL) return rc;
        RT_TRY(rc);
        Lay l2;
        l2.base = b.arena.ptr(b.sel_off[1]);
        lay4(l2);
      }
      RT_TRY(ensure_staging(b, vbytes(seg) + vbytes(qflat) + esctot * 8 + 65536));
      RT_TRY(x.put(a.segs, seg.data(), vbytes(seg)));
      RT_TRY(x.put(f.q, qflat.data(), vbytes(qflat)));
      w.q_ent = f.q;
      w.part_e = f.pe;
      w.e_sc = f.esc;
      RT_TRY(hops_sel_state_flux(w, datl, (int)atl.size(), b.stream));
      char *hesc;
      RT_TRY(x.get(f.esc, esctot * 8, &hesc));
      RT_TRY(x.sync());
      const double *Esc = reinterpret_cast<const double *>(hesc);
      for (int i : li) {
        Slot &sl = b.slots[list[i]];
        std::vector<int> bound;
        if (nsc[i]) {
          std::vector<double> ev;
          std::vector<int> es;
          for (int ii = 0; ii < nsc[i]; ++ii)
            if (Esc[seg[i].esc_off + ii] != 0) {
              ev.push_back(Esc[seg[i].esc_off + ii]);
              es.push_back(sl.sel.sc[ii]);
            }
          if (!ev.empty()) {
            std::vector<double> s2(ev);
            std::sort(s2.begin(), s2.end());
            const double t2 = error_threshold(s2, dbound[i]);
            for (size_t q = 0; q < ev.size(); ++q)
              if (ev[q] > t2) bound.push_back(es[q]);
          }
        }
        std::vector<int> res(st_stable[i]);
        for (int bs : bound)
          if (std::find(sl.B.states.begin(), sl.B.states.end(), bs) == sl.B.states.end()) res.push_back(bs);
        std::sort(res.begin(), res.end());
        res.erase(std::unique(res.begin(), res.end()), res.end());
        sl.sel.states_new = res;
      }
    } else {
      for (int i : li) {
        std::vector<int> res(st_stable[i]);
        std::sort(res.begin(), res.end());
        res.erase(std::unique(res.begin(), res.end()), res.end());
        b.slots[list[i]].sel.states_new = res;
      }
    }
  }
  for (int s : list) {
    Slot &sl = b.slots[s];
    if (sl.prm.delta_s <= 0) sl.sel.states_new = sl.B.states;
  }
  return 0;
}

// define_basis for the listed slots: on the device where the slot's
// selection runs there, on the host threads otherwise (and for every slot
// when the arena cannot hold the device work space)
int define_bases(Batch &b, const std::vector<int> &list) {
  if (list.empty()) return 0;
  point_noise(b, list, false);
  std::vector<int> dev, host;
  if (b.sel_device)
    for (int s : list) RT_TRY(ensure_zob(b, b.slots[s]));
  for (int s : list) (b.sel_device && dev_eligible(b.slots[s]) ? dev : host).push_back(s);
  if (!dev.empty()) {
    const int rc = define_bases_dev(b, dev);
    if (rc == HOPS_RT_ARENA_FULL)
      host.insert(host.end(), dev.begin(), dev.end());
    else
      RT_TRY(rc);
  }
  RT_TRY(define_bases_host(b, host));
  for (int s : list)
    if (b.slots[s].sel.key_collision)
      return rt_fail("slot %d: two auxiliaries share a 128-bit Zobrist key; the selection is refused", s);
  return 0;
}

bool same_basis(const Basis &B, const std::vector<int> &states, const std::vector<Aux> &aux) {
  if (states.size() != B.states.size() || aux.size() != B.aux.size()) return false;
  if (!std::equal(states.begin(), states.end(), B.states.begin())) return false;
  for (const Aux &a : aux) {
    const int j = B.index.find(key_of(*B.sys, a));
    if (j < 0 || B.aux[j] != a) return false;
  }
  return true;
}

// move the listed slots to the bases (states, aux) given in tgt_states/tgt_aux;
// scale: per slot scale of the moved Phi
int update_bases(Batch &b, const std::vector<int> &list, std::vector<std::vector<int>> &tgt_states,
                 std::vector<std::vector<Aux>> &tgt_aux, const std::vector<double> &scale) {
  std::vector<int> ch;
  std::vector<int> idx_of(b.slots.size(), -1);
  for (size_t i = 0; i < list.size(); ++i) {
    Slot &sl = b.slots[list[i]];
    std::sort(tgt_states[i].begin(), tgt_states[i].end());
    sl.changed = !same_basis(sl.B, tgt_states[i], tgt_aux[i]);
    if (sl.changed) {
      idx_of[list[i]] = (int)i;
      ch.push_back(list[i]);
    }
  }
  if (ch.empty()) return 0;
  std::vector<int> rcs(ch.size(), 0);
  {
    Timer tmr(&b.tm[4]);
    b.pool->parallel_for((int)ch.size(), [&](int i) {
      Slot &sl = b.slots[ch[i]];
      const int j = idx_of[ch[i]];
      rcs[i] = build_basis(*sl.sys, sl.prm.eom, std::move(tgt_states[j]), std::move(tgt_aux[j]), sl.Bnew);
    });
  }
  for (int rc : rcs)
    if (rc < 0) return rc;
  // blocks, descriptors, staging layout (serial: the arena is not shared).
  // A new basis that does not fit marks its slot out of memory; the others
  // are placed as usual.
  Timer tmr_u(&b.tm[5]);
  std::vector<long long> nblk;
  std::vector<Layout> lays;
  {
    std::vector<int> fit;
    for (int s : ch) {
      Slot &sl = b.slots[s];
      Layout L = layout_of(sl.Bnew, b.precision);
      const long long blk = b.arena.alloc(L.total);
      if (blk < 0) {
        sl.oom = true;
        sl.changed = false;
        Basis().swap_out(sl.Bnew);
      } else {
        fit.push_back(s);
        nblk.push_back(blk);
        lays.push_back(std::move(L));
      }
    }
    ch.swap(fit);
    if (ch.empty()) return 0;
  }
  std::vector<hops_op_t> nop(ch.size());
  std::vector<hops_batch_t> nbt(ch.size());
  std::vector<size_t> tab_bytes(ch.size()), tab_off(ch.size()), inv_off(ch.size()), prev_off(ch.size());
  size_t stage_bytes_needed = 0;
  for (size_t i = 0; i < ch.size(); ++i) {
    Slot &sl = b.slots[ch[i]];
    const Layout &L = lays[i];
    char *base = b.arena.ptr(nblk[i]);
    set_op(sl.Bnew, base, L, b.precision, nop[i]);
    fill_bt(b, sl, sl.Bnew, L, base, nbt[i]);
    tab_bytes[i] = L.tab_off.empty() ? 0 : L.ell_col;
    prev_off[i] = L.phi_prev;
    stage_bytes_needed += rsz(tab_bytes[i]) + rsz((size_t)(sl.Bnew.nh + sl.Bnew.n) * 4);
  }
  const std::vector<uint32_t> tiles = tiles_of(b, ch, true);
  RT_TRY(ensure_staging(b, stage_bytes_needed + tiles.size() * 4 + ch.size() * (4 + sizeof(hops_segment_t)) +
                               b.slots.size() * sizeof(hops_slot_perm_t) + 65536));
  Stage st{&b};
  std::vector<char *> htab(ch.size()), hinv(ch.size());
  std::vector<void *> dtab(ch.size()), dinv(ch.size());
  for (size_t i = 0; i < ch.size(); ++i) {
    const Basis &Bn = b.slots[ch[i]].Bnew;
    RT_TRY(st.place(tab_bytes[i], &dtab[i], &htab[i]));
    RT_TRY(st.place((size_t)(Bn.nh + Bn.n) * 4, &dinv[i], &hinv[i]));
  }
  // fill tables and inverse maps in parallel, straight into pinned memory;
  // each key hit of the inverse map is confirmed against the auxiliaries
  std::vector<uint8_t> collide(ch.size(), 0);
  b.pool->parallel_for((int)ch.size(), [&](int i) {
    Slot &sl = b.slots[ch[i]];
    const Layout &L = lays[i];
    for (size_t t = 0; t < L.tabs.size(); ++t)
      if (L.tabs[t].second) std::memcpy(htab[i] + L.tab_off[t], L.tabs[t].first, L.tabs[t].second);
    const Basis &Bo = sl.B, &Bn = sl.Bnew;
    int32_t *inv = reinterpret_cast<int32_t *>(hinv[i]);
    for (int j = 0; j < Bn.nh; ++j) {
      const int io = Bo.index.find(AKey{Bn.aux_hash[j], Bn.aux_hash2[j]});
      inv[j] = io >= 0 && Bo.aux[io] == Bn.aux[j] ? io : -1;
      if (io >= 0 && inv[j] < 0) collide[i] = 1;
    }
    for (int q = 0; q < Bn.n; ++q) inv[Bn.nh + q] = Bo.pos[Bn.states[q]];
  });
  for (uint8_t c : collide)
    if (c) return rt_fail("two auxiliaries of an old and a new basis share a 128-bit Zobrist key");
  std::vector<hops_segment_t> segs;
  for (size_t i = 0; i < ch.size(); ++i) {
    Slot &sl = b.slots[ch[i]];
    if (tab_bytes[i]) segs.push_back({dtab[i], b.arena.ptr(nblk[i]), (long long)tab_bytes[i]});
    hops_slot_perm_t &p = b.perms[ch[i]];
    p.old_phi_d = sl.bt.phi_d;
    p.new_phi_d = nbt[i].phi_d;
    p.inv_aux_d = (const int32_t *)dinv[i];
    p.inv_state_d = (const int32_t *)dinv[i] + sl.Bnew.nh;
    p.n_old = sl.B.n;
    p.n_new = sl.Bnew.n;
    p.nh_new = sl.Bnew.nh;
    p.scale = scale[idx_of[ch[i]]];
    b.ops[ch[i]] = nop[i];
  }
  void *dsegs, *dperm, *dl, *dt;
  RT_TRY(st.push(segs.data(), segs.size() * sizeof(hops_segment_t), &dsegs));
  RT_TRY(st.push(b.perms.data(), b.slots.size() * sizeof(hops_slot_perm_t), &dperm));
  RT_TRY(st.push(ch.data(), ch.size() * 4, &dl));
  RT_TRY(st.push(tiles.data(), tiles.size() * 4, &dt));
  RT_TRY(st.flush());
  RT_TRY(hops_copy_segments((const hops_segment_t *)dsegs, (int)segs.size(), b.stream));
  RT_HIP(hipMemcpyAsync(b.arena.ptr(b.ops_off), b.ops.data(), b.slots.size() * sizeof(hops_op_t),
                        hipMemcpyHostToDevice, b.stream));
  RT_TRY(hops_slots_build(d_ops(b), (const int32_t *)dl, (int)ch.size(), (const uint32_t *)dt, (int)tiles.size(),
                          b.precision, b.stream));
  RT_TRY(hops_slots_permute((const hops_slot_perm_t *)dperm, (const uint32_t *)dt, (int)tiles.size(), b.stream));
  RT_HIP(hipStreamSynchronize(b.stream));
  // retire the old bases on the worker threads (many small frees each)
  std::vector<Basis> old(ch.size());
  for (size_t i = 0; i < ch.size(); ++i) old[i] = std::move(b.slots[ch[i]].B);
  for (size_t i = 0; i < ch.size(); ++i) {
    Slot &sl = b.slots[ch[i]];
    b.arena.release(sl.blk);
    sl.blk = nblk[i];
    sl.blk_bytes = lays[i].total;
    sl.B = std::move(sl.Bnew);
    sl.Bnew = Basis();
    sl.op = nop[i];
    sl.bt = nbt[i];
    b.bts[ch[i]] = nbt[i];
    sl.phi_prev = b.arena.ptr(sl.blk) + prev_off[i];
  }
  b.pool->parallel_for((int)old.size(), [&](int i) { Basis().swap_out(old[i]); });
  return 0;
}

// psi of a slot after its basis update, from the Pn of its define step
void record_step(Slot &sl, const std::vector<int> &old_states, const std::vector<cplx> &psi_old) {
  if (!sl.prm.record) return;
  Record r;
  r.states = sl.B.states;
  r.n_hier = sl.B.nh;
  r.psi.assign(sl.B.n, cplx(0, 0));
  for (int s = 0; s < sl.B.n; ++s) {
    const auto it = std::lower_bound(old_states.begin(), old_states.end(), sl.B.states[s]);
    if (it != old_states.end() && *it == sl.B.states[s]) r.psi[s] = sl.scale * psi_old[it - old_states.begin()];
  }
  sl.rec.push_back(std::move(r));
}

// scale of update_basis: |psi_old| / |psi_old on the states kept|
double keep_scale(const Slot &sl, const std::vector<int> &new_states, const std::vector<cplx> &psi) {
  double a = 0, k = 0;
  for (int s = 0; s < sl.B.n; ++s) {
    const double v = std::norm(psi[s]);
    a += v;
    if (std::binary_search(new_states.begin(), new_states.end(), sl.B.states[s])) k += v;
  }
  return std::sqrt(a) / std::sqrt(k);
}

// device-to-device copies of one phase, run as one segment-copy launch
// (every region is a 256-byte aligned block, sizes rounded to 16 bytes fit)
struct Copies {
  std::vector<hops_segment_t> segs;
  void add(void *dst, const void *src, size_t bytes) {
    segs.push_back({src, dst, (long long)((bytes + 15) & ~(size_t)15)});
  }
  int run(Batch &b) {
    if (segs.empty()) return 0;
    RT_TRY(ensure_staging(b, segs.size() * sizeof(hops_segment_t) + 4096));
    Stage st{&b};
    void *d;
    RT_TRY(st.push(segs.data(), segs.size() * sizeof(hops_segment_t), &d));
    RT_TRY(st.flush());
    RT_TRY(hops_copy_segments((const hops_segment_t *)d, (int)segs.size(), b.stream));
    RT_HIP(hipStreamSynchronize(b.stream));
    segs.clear();
    return 0;
  }
};

// initial basis update of newly added slots (trajectory/hops_trajectory.py:279-317)
int init_pending(Batch &b) {
  std::vector<int> list;
  for (size_t s = 0; s < b.slots.size(); ++s)
    if (b.slots[s].used && b.slots[s].pending_init) list.push_back((int)s);
  if (list.empty()) return 0;
  std::vector<int> adapt, sel_list, stat;
  for (int s : list) {
    b.dts[s] = 1.0;
    Slot &sl = b.slots[s];
    if (sl.prm.delta_a > 0 || sl.prm.delta_s > 0) {
      adapt.push_back(s);
      (sl.prm.has_static ? stat : sel_list).push_back(s);
    }
  }
  RT_TRY(define_bases(b, sel_list));
  // STATIC_BASIS: the first basis joins the static lists instead of
  // selecting (trajectory/hops_trajectory.py:303-317); psi as it is
  RT_TRY(fetch_psi(b, stat));
  for (int s : stat) {
    Slot &sl = b.slots[s];
    std::vector<int> st(sl.B.states);
    st.insert(st.end(), sl.prm.static_states.begin(), sl.prm.static_states.end());
    std::sort(st.begin(), st.end());
    st.erase(std::unique(st.begin(), st.end()), st.end());
    sl.sel.states_new = st;
    std::vector<Aux> ax(sl.B.aux);
    ax.insert(ax.end(), sl.prm.static_aux.begin(), sl.prm.static_aux.end());
    sl.sel.aux_new = std::move(ax);
    sl.nrm = 1.0;  // psi = Pn[:n] as fetched
  }
  std::vector<std::vector<int>> ts;
  std::vector<std::vector<Aux>> ta;
  std::vector<double> sc;
  std::vector<std::vector<int>> old_states;
  std::vector<std::vector<cplx>> old_psi;
  for (int s : adapt) {
    Slot &sl = b.slots[s];
    ts.push_back(sl.sel.states_new);
    ta.push_back(std::move(sl.sel.aux_new));  // not read again before the next selection
    std::vector<cplx> psi(sl.pn.begin(), sl.pn.begin() + sl.B.n);
    for (auto &v : psi) v *= sl.nrm;
    std::vector<int> sn(sl.sel.states_new);
    std::sort(sn.begin(), sn.end());
    sc.push_back(keep_scale(sl, sn, psi));
    old_states.push_back(sl.B.states);
    old_psi.push_back(psi);
  }
  RT_TRY(update_bases(b, adapt, ts, ta, sc));
  for (size_t i = 0; i < adapt.size(); ++i) {
    Slot &sl = b.slots[adapt[i]];
    if (sl.oom) continue;
    if (!sl.changed) sl.scale = 1.0;
    else sl.scale = sc[i];
    record_step(sl, old_states[i], old_psi[i]);
  }
  for (int s : list) {
    Slot &sl = b.slots[s];
    if (!(sl.prm.delta_a > 0 || sl.prm.delta_s > 0) && sl.prm.record) {
      Record r;
      r.states = sl.B.states;
      r.n_hier = sl.B.nh;
      r.psi.assign(sl.B.n, 0);
      for (int q = 0; q < sl.B.n; ++q) r.psi[q] = sl.psi0[sl.B.states[q]];
      sl.rec.push_back(std::move(r));
    }
    sl.pending_init = false;
  }
  return 0;
}

// status of a step in which some new bases did not fit in the arena: the
// other slots completed it; the listed out-of-memory slots are left mid-step
int arena_full(const Batch &b, const std::vector<int> &list) {
  std::string ids;
  for (int s : list)
    if (b.slots[s].oom) ids += (ids.empty() ? "" : ",") + std::to_string(s);
  rt_fail("arena full: the new bases of slots %s do not fit; remove them (the other slots completed the step)",
          ids.c_str());
  return HOPS_RT_ARENA_FULL;
}

// index_t: step index within this call (the update schedule restarts with every
// propagate call, trajectory/hops_trajectory.py:456, :511)
int batch_step_once(Batch &b, const std::vector<int> &list, int index_t) {
  // early INCH_WORM slots keep the state before the step
  std::vector<int> early, upd, plain;
  Copies cp;
  for (int s : list) {
    Slot &sl = b.slots[s];
    const bool adaptive = sl.prm.delta_a > 0 || sl.prm.delta_s > 0;
    const bool is_early = adaptive && sl.early_counter < sl.prm.early_steps;
    if (is_early && (sl.prm.early_integrator == 0 || sl.prm.early_integrator == 2)) {
      early.push_back(s);
      cp.add(sl.phi_prev, sl.bt.phi_d, (size_t)sl.B.n * sl.B.nh * 16);
      cp.add(sl.zmem_prev, sl.bt.zmem_d, (size_t)sl.sys->n_hm * 16);
    }
    const int us = sl.prm.update_step > 0 ? sl.prm.update_step : 1;
    if (adaptive && (is_early || (index_t + 1) % us == 0)) {
      if (!(is_early && sl.prm.early_integrator == 1)) upd.push_back(s);
    } else {
      plain.push_back(s);
    }
  }
  RT_TRY(cp.run(b));
  RT_TRY(run_step(b, list));
  for (int s : upd) b.dts[s] = b.slots[s].prm.dt;
  RT_TRY(define_bases(b, upd));
  // INCH_WORM (trajectory/hops_trajectory.py:469-510, :636-691)
  std::vector<int> redo;
  std::vector<char> is_early(b.slots.size(), 0);
  for (int s : early) is_early[s] = 1;
  // STATIC_STATE_INCHWORM_HIERARCHY: the early steps keep the static states
  // (hops_trajectory.py:478-491)
  auto pin_states = [&](const std::vector<int> &ls) {
    for (int s : ls) {
      Slot &sl = b.slots[s];
      if (is_early[s] && sl.prm.early_integrator == 2) sl.sel.states_new = sl.prm.static_states;
    }
  };
  pin_states(upd);
  for (int s : upd) {
    Slot &sl = b.slots[s];
    if (is_early[s] && !same_basis(sl.B, sl.sel.states_new, sl.sel.aux_new)) redo.push_back(s);
  }
  std::vector<int> tries(b.slots.size(), 0);
  while (!redo.empty()) {
    std::vector<std::vector<int>> ts;
    std::vector<std::vector<Aux>> ta;
    std::vector<double> sc;
    for (int s : redo) {
      Slot &sl = b.slots[s];
      std::vector<int> su(sl.sel.states_new);
      su.insert(su.end(), sl.B.states.begin(), sl.B.states.end());
      std::sort(su.begin(), su.end());
      su.erase(std::unique(su.begin(), su.end()), su.end());
      std::vector<Aux> au(sl.sel.aux_new);
      au.insert(au.end(), sl.B.aux.begin(), sl.B.aux.end());
      ts.push_back(su);
      ta.push_back(au);
      sc.push_back(1.0);  // every old state is kept: |psi| unchanged
      cp.add(sl.bt.phi_d, sl.phi_prev, (size_t)sl.B.n * sl.B.nh * 16);
      cp.add(sl.bt.zmem_d, sl.zmem_prev, (size_t)sl.sys->n_hm * 16);
    }
    RT_TRY(cp.run(b));
    RT_TRY(update_bases(b, redo, ts, ta, sc));
    std::vector<int> live;
    for (int s : redo)
      if (!b.slots[s].oom) live.push_back(s);
    redo.swap(live);
    for (int s : redo) {
      Slot &sl = b.slots[s];
      cp.add(sl.phi_prev, sl.bt.phi_d, (size_t)sl.B.n * sl.B.nh * 16);
    }
    RT_TRY(cp.run(b));
    RT_TRY(run_step(b, redo));
    RT_TRY(define_bases(b, redo));
    pin_states(redo);
    std::vector<int> nxt;
    for (int s : redo) {
      Slot &sl = b.slots[s];
      if (sl.oom || ++tries[s] >= sl.prm.inchworm_cap) continue;
      if (!same_basis(sl.B, sl.sel.states_new, sl.sel.aux_new)) nxt.push_back(s);
    }
    redo.swap(nxt);
  }
  // final update of every defining slot
  std::vector<std::vector<int>> ts;
  std::vector<std::vector<Aux>> ta;
  std::vector<double> sc;
  std::vector<std::vector<int>> old_states;
  std::vector<std::vector<cplx>> old_psi;
  {
    std::vector<int> live;
    for (int s : upd)
      if (!b.slots[s].oom) live.push_back(s);
    upd.swap(live);
  }
  for (int s : upd) {
    Slot &sl = b.slots[s];
    ts.push_back(sl.sel.states_new);
    ta.push_back(std::move(sl.sel.aux_new));  // not read again before the next selection
    std::vector<cplx> psi(sl.pn.begin(), sl.pn.begin() + sl.B.n);
    for (auto &v : psi) v *= sl.nrm;
    std::vector<int> sn(sl.sel.states_new);
    std::sort(sn.begin(), sn.end());
    sc.push_back(keep_scale(sl, sn, psi));
    old_states.push_back(sl.B.states);
    old_psi.push_back(psi);
  }
  RT_TRY(update_bases(b, upd, ts, ta, sc));
  for (size_t i = 0; i < upd.size(); ++i) {
    Slot &sl = b.slots[upd[i]];
    if (sl.oom) continue;
    sl.scale = sl.changed ? sc[i] : 1.0;
    record_step(sl, old_states[i], old_psi[i]);
  }
  // psi of the slots without a basis update this step
  std::vector<int> rec_plain;
  for (int s : plain)
    if (b.slots[s].prm.record) rec_plain.push_back(s);
  std::vector<int> early_static;
  for (int s : list) {
    Slot &sl = b.slots[s];
    const bool adaptive = sl.prm.delta_a > 0 || sl.prm.delta_s > 0;
    if (adaptive && sl.early_counter < sl.prm.early_steps && sl.prm.early_integrator == 1 && sl.prm.record)
      rec_plain.push_back(s);
  }
  RT_TRY(fetch_psi(b, rec_plain));
  for (int s : rec_plain) {
    Slot &sl = b.slots[s];
    Record r;
    r.states = sl.B.states;
    r.n_hier = sl.B.nh;
    r.psi = sl.pn;
    sl.rec.push_back(std::move(r));
  }
  int n_oom = 0;
  for (int s : list) {
    Slot &sl = b.slots[s];
    if (sl.oom) {
      ++n_oom;
      continue;
    }
    const bool adaptive = sl.prm.delta_a > 0 || sl.prm.delta_s > 0;
    if (adaptive && sl.early_counter < sl.prm.early_steps) sl.early_counter++;
    sl.steps++;
    sl.it += 2 * sl.prm.noise_step;
  }
  if (n_oom) return arena_full(b, list);
  return 0;
}

// The selection and table building allocate and free buffers of several MB
// per slot and step.  glibc hands blocks above its mmap threshold back to the
// kernel when they are freed, and trims the heap top above its trim
// threshold, so every step pays the page faults of those buffers again;
// raising both keeps freed blocks in the process heap for reuse.
// drop a slot's use of its system's device Zobrist tables
void release_zob(Batch &b, Slot &sl) {
  if (!sl.zob_d) return;
  auto it = b.zob.find(sl.sys);
  if (it != b.zob.end() && --it->second.second == 0) {
    b.arena.release(it->second.first);
    b.zob.erase(it);
  }
  sl.zob_d = nullptr;
}

void keep_heap() {
  static std::once_flag once;
  std::call_once(once, [] {
    mallopt(M_MMAP_THRESHOLD, 32 << 20);  // the largest value glibc accepts
    mallopt(M_TRIM_THRESHOLD, 1 << 30);
  });
}

int check_batch(void *p) {
  if (!p) return rt_fail("batch is NULL");
  return 0;
}

}  // namespace

extern "C" {

void *hops_rt_system_create(const hops_rt_system_desc_t *d) {
  keep_heap();
  if (!d || d->n_state <= 0 || d->n_hmode <= 0 || d->n_lop <= 0 || !d->H || !d->g || !d->w || !d->l_of_mode ||
      !d->l_ptr || !d->l_row || !d->l_col || !d->l_val) {
    rt_fail("hops_rt_system_create: bad description");
    return nullptr;
  }
  auto *S = new System();
  S->n = d->n_state;
  S->n_hm = d->n_hmode;
  S->n_l = d->n_lop;
  S->H.resize((size_t)S->n * S->n);
  for (size_t i = 0; i < S->H.size(); ++i) S->H[i] = cplx(d->H[2 * i], d->H[2 * i + 1]);
  for (int m = 0; m < S->n_hm; ++m) {
    S->G.push_back(cplx(d->g[2 * m], d->g[2 * m + 1]));
    S->W.push_back(cplx(d->w[2 * m], d->w[2 * m + 1]));
    if (d->l_of_mode[m] < 0 || d->l_of_mode[m] >= S->n_l) {
      delete S;
      rt_fail("mode %d has L-operator index %d", m, d->l_of_mode[m]);
      return nullptr;
    }
    S->l_of_mode.push_back(d->l_of_mode[m]);
  }
  S->L.resize(S->n_l);
  for (int l = 0; l < S->n_l; ++l)
    for (int q = d->l_ptr[l]; q < d->l_ptr[l + 1]; ++q) {
      const int r = d->l_row[q], c = d->l_col[q];
      if (r < 0 || r >= S->n || c < 0 || c >= S->n) {
        delete S;
        rt_fail("L-operator %d entry (%d, %d) out of range", l, r, c);
        return nullptr;
      }
      S->L[l].row.push_back(r);
      S->L[l].col.push_back(c);
      S->L[l].val.push_back(cplx(d->l_val[2 * q], d->l_val[2 * q + 1]));
      if (r != c) {
        S->diag_L = false;
        S->L[l].off = true;
      }
    }
  for (LOp &L : S->L) {
    // by column and by row, duplicates summed
    std::vector<std::tuple<int, int, cplx>> e;
    for (size_t q = 0; q < L.row.size(); ++q) e.emplace_back(L.col[q], L.row[q], L.val[q]);
    std::sort(e.begin(), e.end(), [](const auto &x, const auto &y) {
      return std::get<0>(x) != std::get<0>(y) ? std::get<0>(x) < std::get<0>(y) : std::get<1>(x) < std::get<1>(y);
    });
    L.cp.assign(S->n + 1, 0);
    for (size_t q = 0; q < e.size(); ++q) {
      const auto [c, r, v] = e[q];
      if (!L.ri.empty() && q > 0 && std::get<0>(e[q - 1]) == c && std::get<1>(e[q - 1]) == r) {
        L.cv.back() += v;
        continue;
      }
      L.ri.push_back(r);
      L.cv.push_back(v);
      L.cp[c + 1]++;
    }
    for (int c = 0; c < S->n; ++c) L.cp[c + 1] += L.cp[c];
    L.rp.assign(S->n + 1, 0);
    for (int c = 0; c < S->n; ++c)
      for (int q = L.cp[c]; q < L.cp[c + 1]; ++q) L.rp[L.ri[q] + 1]++;
    for (int r = 0; r < S->n; ++r) L.rp[r + 1] += L.rp[r];
    L.ci.resize(L.ri.size());
    L.rv.resize(L.ri.size());
    std::vector<int> fill(L.rp.begin(), L.rp.end() - 1);
    for (int c = 0; c < S->n; ++c)  // columns ascending: each row's entries come out by ascending column
      for (int q = L.cp[c]; q < L.cp[c + 1]; ++q) {
        L.ci[fill[L.ri[q]]] = c;
        L.rv[fill[L.ri[q]]++] = L.cv[q];
      }
  }
  if (d->lt)
    for (int l = 0; l < S->n_l; ++l) S->lt.push_back(cplx(d->lt[2 * l], d->lt[2 * l + 1]));
  S->modes_by_state.resize(S->n);
  S->l_by_state.resize(S->n);
  S->coupled.resize(S->n);
  S->dest.resize(S->n);
  for (int l = 0; l < S->n_l; ++l) {
    std::vector<int> touched(S->L[l].row);
    touched.insert(touched.end(), S->L[l].col.begin(), S->L[l].col.end());
    std::sort(touched.begin(), touched.end());
    touched.erase(std::unique(touched.begin(), touched.end()), touched.end());
    for (int s : touched) S->l_by_state[s].push_back(l);
    for (size_t q = 0; q < S->L[l].row.size(); ++q) S->dest[S->L[l].row[q]].push_back(S->L[l].col[q]);
  }
  for (int m = 0; m < S->n_hm; ++m)
    for (int s = 0; s < S->n; ++s)
      if (std::find(S->l_by_state[s].begin(), S->l_by_state[s].end(), S->l_of_mode[m]) != S->l_by_state[s].end())
        S->modes_by_state[s].push_back(m);
  for (int r = 0; r < S->n; ++r)
    for (int c = 0; c < S->n; ++c)
      if (S->H[(size_t)r * S->n + c] != cplx(0, 0)) S->coupled[r].push_back(c);
  uint64_t x = 0x9E3779B97F4A7C15ull;
  S->zob.resize((size_t)S->n_hm * 256);
  S->zob2.resize((size_t)S->n_hm * 256);
  for (auto *tab : {&S->zob, &S->zob2})
    for (auto &z : *tab) {  // splitmix64
      x += 0x9E3779B97F4A7C15ull;
      uint64_t v = x;
      v = (v ^ (v >> 30)) * 0xBF58476D1CE4E5B9ull;
      v = (v ^ (v >> 27)) * 0x94D049BB133111EBull;
      z = v ^ (v >> 31);
    }
  for (int m = 0; m < S->n_hm; ++m) S->zob[(size_t)m * 256] = S->zob2[(size_t)m * 256] = 0;  // depth 0: no key
  return S;
}

int hops_rt_system_destroy(void *sys) {
  delete static_cast<System *>(sys);
  return 0;
}

void *hops_rt_batch_create(int precision, void *arena_d, size_t arena_bytes, int max_slots, int n_threads,
                           void *stream) {
  if (!arena_d || arena_bytes < (1 << 20) || max_slots <= 0 || max_slots > (1 << (32 - kSlotShift))) {
    rt_fail("hops_rt_batch_create: bad arguments");
    return nullptr;
  }
  if (precision != HOPS_PREC_FP64 && precision != HOPS_PREC_MIXED) {
    rt_fail("unknown precision %d", precision);
    return nullptr;
  }
  if (((uintptr_t)arena_d & 255) != 0) {
    rt_fail("arena not 256-byte aligned");
    return nullptr;
  }
  keep_heap();
  auto *b = new Batch();
  b->precision = precision;
  b->stream = (hipStream_t)stream;
  b->arena.init((char *)arena_d, arena_bytes);
  b->slots.resize(max_slots);
  b->ops.assign(max_slots, hops_op_t{});
  b->bts.assign(max_slots, hops_batch_t{});
  b->perms.assign(max_slots, hops_slot_perm_t{});
  b->its.assign(max_slots, 0);
  b->dts.assign(max_slots, 1.0);
  b->nrms.assign(max_slots, 1.0);
  b->ops_off = b->arena.alloc(max_slots * sizeof(hops_op_t));
  b->bts_off = b->arena.alloc(max_slots * sizeof(hops_batch_t));
  b->perm_off = b->arena.alloc(max_slots * sizeof(hops_slot_perm_t));
  b->it_off = b->arena.alloc(max_slots * 4);
  b->dt_off = b->arena.alloc(max_slots * 8);
  b->nrm_off = b->arena.alloc(max_slots * 8);
  if (b->ops_off < 0 || b->bts_off < 0 || b->perm_off < 0 || b->it_off < 0 || b->dt_off < 0 || b->nrm_off < 0 ||
      ensure_staging(*b, 4 << 20) < 0) {
    rt_fail("arena too small for the batch descriptors");
    delete b;
    return nullptr;
  }
  const int nt = n_threads > 0 ? n_threads : 16;
  b->pool.reset(new Pool(nt - 1));
  return b;
}

int hops_rt_batch_destroy(void *p) {
  auto *b = static_cast<Batch *>(p);
  if (!b) return 0;
  (void)hipStreamSynchronize(b->stream);
  if (b->h_up) (void)hipHostFree(b->h_up);
  if (b->h_down) (void)hipHostFree(b->h_down);
  delete b;
  return 0;
}

int hops_rt_batch_add(void *p, void *sys, const hops_rt_traj_param_t *prm, const double *psi0, const void *z1_d,
                      const void *z2_d) {
  RT_TRY(check_batch(p));
  auto &b = *static_cast<Batch *>(p);
  if (!sys || !prm || !psi0 || !z1_d) return rt_fail("hops_rt_batch_add: NULL argument");
  const System &S = *static_cast<System *>(sys);
  int id = -1;
  for (size_t s = 0; s < b.slots.size(); ++s)
    if (!b.slots[s].used) {
      id = (int)s;
      break;
    }
  if (id < 0) return rt_fail("no free slot (max %zu)", b.slots.size());
  Slot &sl = b.slots[id];
  sl = Slot();
  sl.sys = &S;
  Params &P = sl.prm;
  P.eom = prm->eom;
  P.maxhier = std::min(prm->maxhier, 255);
  P.update_step = prm->update_step;
  P.early_integrator = prm->early_integrator;
  if (P.early_integrator < 0 || P.early_integrator > 2) return rt_fail("unknown early integrator %d", P.early_integrator);
  P.early_steps = prm->early_steps;
  P.inchworm_cap = prm->inchworm_cap;
  P.noise_step = prm->noise_step;
  P.n_t = prm->n_t;
  P.it0 = prm->it0;
  P.record = prm->record;
  P.delta_a = prm->delta_a;
  P.delta_s = prm->delta_s;
  P.f_discard = prm->f_discard;
  P.dt = prm->dt;
  if (P.eom < 0 || P.eom > 2) return rt_fail("unknown eom %d", P.eom);
  if (P.noise_step <= 0 || P.n_t <= 0 || P.it0 < 0 || P.it0 >= P.n_t) return rt_fail("bad noise grid");
  RT_TRY(copy_filters(prm, S.n_hm, P));
  if (P.delta_a <= 0) {
    if (prm->n_aux0 <= 0 || !prm->aux0_ptr) return rt_fail("non-adaptive hierarchy needs aux0");
    for (int i = 0; i < prm->n_aux0; ++i) {
      Aux a;
      for (int q = prm->aux0_ptr[i]; q < prm->aux0_ptr[i + 1]; ++q) a.push_back(aux_pack(prm->aux0_mode[q], prm->aux0_val[q]));
      std::sort(a.begin(), a.end());
      P.aux0.push_back(a);
    }
  }
  std::vector<int> states;
  sl.psi0.resize(S.n);
  for (int s = 0; s < S.n; ++s) {
    sl.psi0[s] = cplx(psi0[2 * s], psi0[2 * s + 1]);
    if (P.delta_s <= 0 || std::abs(sl.psi0[s]) > 0) states.push_back(s);
  }
  std::vector<Aux> aux = P.delta_a > 0 ? std::vector<Aux>{Aux()} : P.aux0;
  sl.z1 = z1_d;
  sl.z2 = z2_d;
  sl.c128 = prm->noise_c128 != 0;
  sl.it = P.it0;
  RT_TRY(build_basis(S, P.eom, states, aux, sl.B));
  sl.pers = b.arena.alloc(pers_bytes(S));
  if (sl.pers < 0) return rt_fail("arena full");

  std::vector<char> tabs;
  hops_op_t op;
  hops_batch_t bt;
  if (int rc = place_basis(b, sl, sl.B, &sl.blk, op, bt, tabs); rc < 0) {
    b.arena.release(sl.pers);
    return rc;
  }
  sl.op = op;
  sl.bt = bt;
  b.ops[id] = op;
  b.bts[id] = bt;
  char *base = b.arena.ptr(sl.blk);
  Layout L = layout_of(sl.B, b.precision);
  sl.phi_prev = base + L.phi_prev;
  sl.zmem_prev = reinterpret_cast<double *>(b.arena.ptr(sl.pers) + rsz(S.n_hm * 16 + 16));
  RT_HIP(hipMemcpyAsync(base, tabs.data(), tabs.size(), hipMemcpyHostToDevice, b.stream));
  std::vector<cplx> phi0((size_t)sl.B.n * sl.B.nh, cplx(0, 0));
  for (int q = 0; q < sl.B.n; ++q) phi0[q] = sl.psi0[sl.B.states[q]];
  RT_HIP(hipMemcpyAsync(bt.phi_d, phi0.data(), phi0.size() * 16, hipMemcpyHostToDevice, b.stream));
  RT_HIP(hipMemsetAsync(b.arena.ptr(sl.pers), 0, pers_bytes(S), b.stream));
  // super-operator of the initial basis
  std::vector<int> one{id};
  RT_HIP(hipMemcpyAsync(b.arena.ptr(b.ops_off) + id * sizeof(hops_op_t), &b.ops[id], sizeof(hops_op_t),
                        hipMemcpyHostToDevice, b.stream));
  RT_TRY(hops_build_super(&b.ops[id], b.precision, b.stream));
  RT_HIP(hipStreamSynchronize(b.stream));
  sl.used = true;
  sl.pending_init = true;
  return id;
}

int hops_rt_batch_remove(void *p, int slot) {
  RT_TRY(check_batch(p));
  auto &b = *static_cast<Batch *>(p);
  if (slot < 0 || slot >= (int)b.slots.size() || !b.slots[slot].used) return rt_fail("no slot %d", slot);
  RT_HIP(hipStreamSynchronize(b.stream));
  Slot &sl = b.slots[slot];
  b.arena.release(sl.blk);
  b.arena.release(sl.pers);
  release_zob(b, sl);
  sl = Slot();
  b.ops[slot] = hops_op_t{};
  b.bts[slot] = hops_batch_t{};
  return 0;
}

int hops_rt_slot_noise(void *p, int slot, const void *z1_d, const void *z2_d, int n_t, int it, int noise_step,
                       int c128) {
  RT_TRY(check_batch(p));
  auto &b = *static_cast<Batch *>(p);
  if (slot < 0 || slot >= (int)b.slots.size() || !b.slots[slot].used) return rt_fail("no slot %d", slot);
  if (!z1_d || n_t <= 0 || it < 0 || it >= n_t || noise_step <= 0) return rt_fail("hops_rt_slot_noise: bad noise");
  Slot &sl = b.slots[slot];
  sl.z1 = z1_d;
  sl.z2 = z2_d;
  sl.c128 = c128 != 0;
  sl.it = it;
  sl.prm.n_t = n_t;
  sl.prm.noise_step = noise_step;
  for (hops_batch_t *bt : {&sl.bt, &b.bts[slot]}) {
    bt->z1_d = z1_d;
    bt->z2_d = z2_d;
    bt->n_t = n_t;
    bt->noise_step = noise_step;
    bt->noise_c128 = sl.c128 ? 1 : 0;
  }
  return 0;
}

int hops_rt_slot_stage_noise(void *p, int slot, const void *z1_d, const void *z2_d) {
  RT_TRY(check_batch(p));
  auto &b = *static_cast<Batch *>(p);
  if (slot < 0 || slot >= (int)b.slots.size() || !b.slots[slot].used) return rt_fail("no slot %d", slot);
  Slot &sl = b.slots[slot];
  if (z1_d && (sl.z2 != nullptr) != (z2_d != nullptr)) return rt_fail("stage noise 2 must match the slot's noise 2");
  sl.z1s = z1_d;
  sl.z2s = z1_d ? z2_d : nullptr;
  sl.bt.z1_d = b.bts[slot].z1_d = sl.z1s ? sl.z1s : sl.z1;
  sl.bt.z2_d = b.bts[slot].z2_d = sl.z1s ? sl.z2s : sl.z2;
  return 0;
}

int hops_rt_slot_static_basis(void *p, int slot, int n_states, const int32_t *states, int n_aux,
                              const int32_t *aux_ptr, const int32_t *aux_mode, const int32_t *aux_val) {
  RT_TRY(check_batch(p));
  auto &b = *static_cast<Batch *>(p);
  if (slot < 0 || slot >= (int)b.slots.size() || !b.slots[slot].used) return rt_fail("no slot %d", slot);
  Slot &sl = b.slots[slot];
  if (!sl.pending_init) return rt_fail("a static basis is set before the slot's first step");
  if (n_states < 0 || n_aux < 0 || (n_states && !states) || (n_aux && (!aux_ptr || !aux_mode || !aux_val)))
    return rt_fail("hops_rt_slot_static_basis: bad arguments");
  const System &S = *sl.sys;
  Params &P = sl.prm;
  P.static_states.clear();
  for (int i = 0; i < n_states; ++i) {
    if (states[i] < 0 || states[i] >= S.n) return rt_fail("static state %d out of range", states[i]);
    P.static_states.push_back(states[i]);
  }
  std::sort(P.static_states.begin(), P.static_states.end());
  P.static_states.erase(std::unique(P.static_states.begin(), P.static_states.end()), P.static_states.end());
  P.static_aux.clear();
  for (int i = 0; i < n_aux; ++i) {
    Aux a;
    for (int q = aux_ptr[i]; q < aux_ptr[i + 1]; ++q) {
      if (aux_mode[q] < 0 || aux_mode[q] >= S.n_hm || aux_val[q] < 1 || aux_val[q] > 255)
        return rt_fail("static auxiliary %d: bad (mode, depth) entry", i);
      a.push_back(aux_pack(aux_mode[q], aux_val[q]));
    }
    std::sort(a.begin(), a.end());
    P.static_aux.push_back(a);
  }
  P.has_static = true;
  return 0;
}

long long hops_rt_slot_estimate(void *p, void *sys, int n_state, int n_hier) {
  if (!p || !sys) return rt_fail("NULL argument");
  auto &b = *static_cast<Batch *>(p);
  const System &S = *static_cast<System *>(sys);
  const size_t nr = (size_t)n_state * n_hier, es = stage_bytes(b.precision);
  const size_t ell_w = 2 * (size_t)std::max(1, std::min(S.n_hm, 2 * n_state));  // up + down per row
  const size_t tables = (size_t)n_hier * (16 + 8 * ell_w) + (size_t)n_state * n_state * 48 + 4096;
  return (long long)(pers_bytes(S) + tables + ell_w * nr * (4 + es) + nr * (16 * 2 + 3 * es + 8) + n_hier * 16);
}

int hops_rt_batch_option(void *p, const char *name, long long value) {
  RT_TRY(check_batch(p));
  auto &b = *static_cast<Batch *>(p);
  if (!name) return rt_fail("hops_rt_batch_option: NULL name");
  if (std::strcmp(name, "device_select") == 0) {
    b.sel_device = value != 0;
    return 0;
  }
  return rt_fail("unknown batch option '%s'", name);
}

int hops_rt_batch_timers(void *p, double *out) {
  RT_TRY(check_batch(p));
  auto &b = *static_cast<Batch *>(p);
  for (int i = 0; i < 8; ++i) out[i] = b.tm[i];
  return 0;
}

long long hops_rt_arena_free(void *p) {
  if (!p) return rt_fail("batch is NULL");
  return (long long)static_cast<Batch *>(p)->arena.free_bytes();
}

int hops_rt_batch_step(void *p, int n_steps, const int32_t *slots_or_null, int n_slots) {
  RT_TRY(check_batch(p));
  auto &b = *static_cast<Batch *>(p);
  for (size_t s = 0; s < b.slots.size(); ++s)
    if (b.slots[s].used && b.slots[s].oom) return rt_fail("slot %d ran out of arena memory: remove it first", (int)s);
  RT_TRY(init_pending(b));
  {
    std::vector<int> all;
    for (size_t s = 0; s < b.slots.size(); ++s)
      if (b.slots[s].used) all.push_back((int)s);
    for (int s : all)
      if (b.slots[s].oom) return arena_full(b, all);
  }
  std::vector<int> list;
  if (slots_or_null) {
    for (int i = 0; i < n_slots; ++i) {
      const int s = slots_or_null[i];
      if (s < 0 || s >= (int)b.slots.size() || !b.slots[s].used) return rt_fail("no slot %d", s);
      list.push_back(s);
    }
  } else {
    for (size_t s = 0; s < b.slots.size(); ++s)
      if (b.slots[s].used) list.push_back((int)s);
  }
  for (int s : list) {
    const Slot &sl = b.slots[s];
    if (sl.it + (long)n_steps * 2 * sl.prm.noise_step >= sl.prm.n_t)
      return rt_fail("slot %d: trajectory times longer than the noise (index %ld >= %d)", s,
                     sl.it + (long)n_steps * 2 * sl.prm.noise_step, sl.prm.n_t);
  }
  for (int k = 0; k < n_steps; ++k) RT_TRY(batch_step_once(b, list, k));
  return 0;
}

int hops_rt_slot_info(void *p, int slot, long long *info) {
  RT_TRY(check_batch(p));
  auto &b = *static_cast<Batch *>(p);
  if (slot < 0 || slot >= (int)b.slots.size() || !info) return rt_fail("bad slot %d", slot);
  const Slot &sl = b.slots[slot];
  info[0] = sl.B.n;
  info[1] = sl.B.nh;
  info[2] = sl.steps;
  info[3] = sl.it;
  info[4] = (long long)sl.rec.size();
  info[5] = (long long)(sl.blk_bytes + (sl.sys ? pers_bytes(*sl.sys) : 0));
  info[6] = std::max(0, sl.prm.early_steps - sl.early_counter);
  info[7] = sl.used;
  info[8] = sl.oom;
  return 0;
}

int hops_rt_slot_states(void *p, int slot, int32_t *out) {
  RT_TRY(check_batch(p));
  auto &b = *static_cast<Batch *>(p);
  if (slot < 0 || slot >= (int)b.slots.size() || !b.slots[slot].used) return rt_fail("no slot %d", slot);
  for (int i = 0; i < b.slots[slot].B.n; ++i) out[i] = b.slots[slot].B.states[i];
  return 0;
}

int hops_rt_slot_aux(void *p, int slot, int32_t *ptr, int32_t *mode, int32_t *val) {
  RT_TRY(check_batch(p));
  auto &b = *static_cast<Batch *>(p);
  if (slot < 0 || slot >= (int)b.slots.size() || !b.slots[slot].used) return rt_fail("no slot %d", slot);
  const Basis &B = b.slots[slot].B;
  std::copy(B.aux_ptr.begin(), B.aux_ptr.end(), ptr);
  std::copy(B.aux_mode_t.begin(), B.aux_mode_t.end(), mode);
  std::copy(B.aux_val_t.begin(), B.aux_val_t.end(), val);
  return 0;
}

int hops_rt_slot_phi(void *p, int slot, double *out) {
  RT_TRY(check_batch(p));
  auto &b = *static_cast<Batch *>(p);
  if (slot < 0 || slot >= (int)b.slots.size() || !b.slots[slot].used) return rt_fail("no slot %d", slot);
  const Slot &sl = b.slots[slot];
  RT_HIP(hipMemcpyAsync(out, sl.bt.phi_d, (size_t)sl.B.n * sl.B.nh * 16, hipMemcpyDeviceToHost, b.stream));
  RT_HIP(hipStreamSynchronize(b.stream));
  return 0;
}

int hops_rt_slot_zmem(void *p, int slot, double *out) {
  RT_TRY(check_batch(p));
  auto &b = *static_cast<Batch *>(p);
  if (slot < 0 || slot >= (int)b.slots.size() || !b.slots[slot].used) return rt_fail("no slot %d", slot);
  const Slot &sl = b.slots[slot];
  RT_HIP(hipMemcpyAsync(out, sl.bt.zmem_d, (size_t)sl.sys->n_hm * 16, hipMemcpyDeviceToHost, b.stream));
  RT_HIP(hipStreamSynchronize(b.stream));
  return 0;
}

int hops_rt_slot_record(void *p, int slot, int idx, double *psi_out, int32_t *states_out, int32_t *n_state_out,
                        int32_t *n_hier_out) {
  RT_TRY(check_batch(p));
  auto &b = *static_cast<Batch *>(p);
  if (slot < 0 || slot >= (int)b.slots.size()) return rt_fail("bad slot %d", slot);
  const Slot &sl = b.slots[slot];
  if (idx < 0 || idx >= (int)sl.rec.size()) return rt_fail("slot %d has no record %d", slot, idx);
  const Record &r = sl.rec[idx];
  if (n_state_out) *n_state_out = (int32_t)r.states.size();
  if (n_hier_out) *n_hier_out = r.n_hier;
  if (psi_out)
    for (size_t i = 0; i < r.psi.size(); ++i) {
      psi_out[2 * i] = r.psi[i].real();
      psi_out[2 * i + 1] = r.psi[i].imag();
    }
  if (states_out) std::copy(r.states.begin(), r.states.end(), states_out);
  return 0;
}

int hops_rt_slot_clear_records(void *p, int slot) {
  RT_TRY(check_batch(p));
  auto &b = *static_cast<Batch *>(p);
  if (slot < 0 || slot >= (int)b.slots.size()) return rt_fail("bad slot %d", slot);
  b.slots[slot].rec.clear();
  return 0;
}

// ---- host-only hooks -------------------------------------------------------
struct DebugBasis {
  Basis B;
  Selection sel;
  Params P;
  std::vector<cplx> pn;
  std::vector<cplx> coef;  // noise coefficients (hops_rt_debug_noise)
};

void *hops_rt_debug_basis(void *sys, int eom, int n_states, const int32_t *states, int n_aux, const int32_t *aux_ptr,
                          const int32_t *amode, const int32_t *aval) {
  if (!sys) return nullptr;
  auto *d = new DebugBasis();
  std::vector<int> st(states, states + n_states);
  std::vector<Aux> aux;
  for (int i = 0; i < n_aux; ++i) {
    Aux a;
    for (int q = aux_ptr[i]; q < aux_ptr[i + 1]; ++q) a.push_back(aux_pack(amode[q], aval[q]));
    std::sort(a.begin(), a.end());
    aux.push_back(a);
  }
  if (build_basis(*static_cast<System *>(sys), eom, st, aux, d->B) < 0) {
    delete d;
    return nullptr;
  }
  return d;
}

long long hops_rt_debug_table(void *p, const char *name, void *out, long long cap) {
  if (!p || !name) return rt_fail("NULL argument");
  const Basis &B = static_cast<DebugBasis *>(p)->B;
  std::string nm(name);
  auto give = [&](const void *src, size_t bytes) -> long long {
    if (out && (long long)bytes <= cap) std::memcpy(out, src, bytes);
    return (long long)bytes;
  };
#define TAB(key, vec) \
  if (nm == key) return give(B.vec.data(), B.vec.size() * sizeof(B.vec[0]));
  TAB("aux_ptr", aux_ptr) TAB("aux_mode", aux_mode_t) TAB("aux_val", aux_val_t) TAB("up_ptr", up_ptr)
  TAB("up_aux", up_aux) TAB("up_mode", up_mode) TAB("dn_ptr", dn_ptr) TAB("dn_aux", dn_aux)
  TAB("dn_mode", dn_mode) TAB("dn_val", dn_val) TAB("lr_ptr", lr_ptr) TAB("lr_col", lr_col)
  TAB("lr_val", lr_val) TAB("lc_ptr", lc_ptr) TAB("lc_row", lc_row) TAB("lc_val", lc_val)
  TAB("mode_lrel", mode_lrel32) TAB("mode_lact", mode_lact32) TAB("mode_abs", mode_abs32)
  TAB("mode_g", mode_g) TAB("mode_w", mode_w) TAB("abs_g", abs_g) TAB("abs_w", abs_w)
  TAB("abs_lact", abs_lact) TAB("act_lrel", act_lrel32) TAB("act_zrow", act_zrow) TAB("am_ptr", am_ptr)
  TAB("am_mode", am_mode) TAB("la_ptr", la_ptr) TAB("la_row", la_row) TAB("la_col", la_col)
  TAB("la_val", la_val) TAB("lq_ptr", lq_ptr) TAB("lq_row", lq_row) TAB("lq_col", lq_col)
  TAB("lq_val", lq_val) TAB("lt", lt) TAB("first_aux", first_aux) TAB("first_lact", first_lact)
  TAB("first_mode", first_mode) TAB("pat_ptr", pat_ptr) TAB("pat_col", pat_col) TAB("pat_h", pat_h)
  TAB("pc_ptr", pc_ptr) TAB("pc_act", pc_act) TAB("pc_kind", pc_kind) TAB("pc_val", pc_val)
#undef TAB
  if (nm == "sizes") {
    const int32_t v[10] = {B.n, B.nh, B.nm, B.nl, B.na, (int32_t)B.first_aux.size(), B.ell_w, B.n_pat,
                           B.sys->n_hm, B.pat_w};
    return give(v, sizeof(v));
  }
  if (nm == "has_lt") {
    const int32_t v = B.has_lt;
    return give(&v, sizeof(v));
  }
  return rt_fail("unknown table %s", name);
}

int hops_rt_debug_select(void *p, const hops_rt_traj_param_t *prm, const double *pn, const double *e_aux, double dt,
                         uint8_t *stable_out) {
  if (!p || !prm || !pn || !e_aux) return rt_fail("NULL argument");
  auto *d = static_cast<DebugBasis *>(p);
  d->P = Params();
  d->P.eom = prm->eom;
  d->P.maxhier = prm->maxhier;
  d->P.delta_a = prm->delta_a;
  d->P.delta_s = prm->delta_s;
  d->P.f_discard = prm->f_discard;
  if (int rc = copy_filters(prm, d->B.sys->n_hm, d->P)) return rc;
  (void)dt;
  const size_t N = (size_t)d->B.n * d->B.nh;
  d->pn.resize(N);
  for (size_t i = 0; i < N; ++i) d->pn[i] = cplx(pn[2 * i], pn[2 * i + 1]);
  d->P.keep_err = true;
  if (d->B.has_off && d->coef.empty()) return rt_fail("off-diagonal L-operators: hops_rt_debug_noise first");
  select_hier(d->P, d->B, d->pn.data(), e_aux, d->sel, d->coef.empty() ? nullptr : d->coef.data());
  if (d->sel.key_collision) return rt_fail("two auxiliaries share a 128-bit Zobrist key");
  if (stable_out) std::copy(d->sel.stable_mask.begin(), d->sel.stable_mask.end(), stable_out);
  return 0;
}

int hops_rt_debug_set_zobrist(void *sys, int mode, int depth, unsigned long long a, unsigned long long b) {
  if (!sys) return rt_fail("NULL argument");
  auto *S = static_cast<System *>(sys);
  if (mode < 0 || mode >= S->n_hm || depth < 1 || depth > 255) return rt_fail("bad (mode, depth)");
  S->zob[(size_t)mode * 256 + depth] = a;
  S->zob2[(size_t)mode * 256 + depth] = b;
  return 0;
}

int hops_rt_debug_noise(void *p, const float *z1, const float *z2, const double *zmem) {
  if (!p || !z1 || !zmem) return rt_fail("NULL argument");
  auto *d = static_cast<DebugBasis *>(p);
  std::vector<cplx> zm(d->B.sys->n_hm);
  for (size_t m = 0; m < zm.size(); ++m) zm[m] = cplx(zmem[2 * m], zmem[2 * m + 1]);
  noise_coef(d->B, z1, z2, false, zm.data(), d->coef);
  return 0;
}

long long hops_rt_debug_errors(void *p, int which, double *out, long long cap) {
  if (!p) return rt_fail("NULL argument");
  const Selection &s = static_cast<DebugBasis *>(p)->sel;
  std::vector<double> v;
  if (which == 0) {
    v = s.err_aux;
  } else if (which == 1) {
    v = s.err_state;
  } else if (which == 2) {
    for (auto &e : s.err_out) {
      v.push_back(e.first);
      v.push_back(e.second);
    }
  } else {
    return rt_fail("unknown error table %d", which);
  }
  if (out && (long long)v.size() <= cap) std::copy(v.begin(), v.end(), out);
  return (long long)v.size();
}

int hops_rt_debug_select_states(void *p, const double *e_state) {
  if (!p) return rt_fail("NULL argument");
  auto *d = static_cast<DebugBasis *>(p);
  if (d->P.delta_s > 0 && !e_state) return rt_fail("state errors missing");
  select_states(d->P, d->B, d->pn.data(), e_state, d->sel);
  if (d->sel.key_collision) return rt_fail("two auxiliaries share a 128-bit Zobrist key");
  return 0;
}

int hops_rt_debug_result(void *p, int32_t *n_states, int32_t *states_out, int32_t *n_aux, int32_t *aux_ptr,
                         int32_t *amode, int32_t *aval, long long cap) {
  if (!p) return rt_fail("NULL argument");
  auto *d = static_cast<DebugBasis *>(p);
  const Selection &s = d->sel;
  *n_states = (int32_t)s.states_new.size();
  *n_aux = (int32_t)s.aux_new.size();
  long long ent = 0;
  for (const Aux &a : s.aux_new) ent += (long long)a.size();
  if ((long long)s.states_new.size() > cap || (long long)s.aux_new.size() + 1 > cap || ent > cap) return 1;
  std::copy(s.states_new.begin(), s.states_new.end(), states_out);
  aux_ptr[0] = 0;
  int q = 0;
  for (size_t i = 0; i < s.aux_new.size(); ++i) {
    for (char32_t e : s.aux_new[i]) {
      amode[q] = aux_mode(e);
      aval[q] = aux_val(e);
      ++q;
    }
    aux_ptr[i + 1] = q;
  }
  return 0;
}

int hops_rt_debug_free(void *p) {
  delete static_cast<DebugBasis *>(p);
  return 0;
}

double hops_rt_debug_threshold(const double *v, long long n, double max_err, double offset, double *below) {
  std::vector<double> w(v, v + (n > 0 ? n : 0));
  double b = 0;
  const double thr = partial_threshold(w, max_err, offset, &b);
  if (below) *below = b;
  return thr;
}

}  // extern "C"
