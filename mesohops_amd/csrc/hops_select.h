// hops_select.h — device side of the adaptive basis selection for ragged
// batches (internal to libhops_mi355x.so; the host orchestration is in
// hops_runtime.cpp).
//
// Restated from the reference (MesoHOPS v1.6.0, src/mesohops/):
//   basis/hops_basis.py:484-558      hierarchy basis: stable set, boundary
//   basis/hops_basis.py:560-665      hier_stable_error (flux up / down, state
//                                    flux out of the basis)
//   basis/hops_basis.py:667-793      boundary auxiliaries, f_discard pre-trim
//   basis/hops_basis.py:795-912      state_stable_error
//   basis/basis_functions.py:8-41    error threshold
//   basis/basis_functions_adaptive.py:76-390   flux terms
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

#include "../../include/hops_mi355x.h"

// One relative mode of a slot's basis: |w|, |g/w|^2, |M2|^2 and |M2 - <L>|^2
// of its single diagonal L entry (state m2s, -1: none in the state basis),
// |<L>|^2, Markovian-filter flag.
struct hops_sel_mode_t {
  double aw, gw2, m2u, m2d, nX;
  int32_t m2s, markov;
};

// One entry -iH[sp, s] of the coupling of an outside state sp to basis state s.
struct hops_sel_sc_t {
  double re, im;
  int32_t s, pad;
};

// One listed slot.  Offsets are element offsets into the batch-wide arrays
// of hops_sel_work_t.
struct hops_sel_seg_t {
  int32_t slot, n, nh, nm;
  int32_t maxhier, n_sc, markov, n_tiles;  // n_tiles = ceil(nh / 256)
  int64_t pn_off;   // Pn rows (complex)
  int64_t e_off;    // deletion-flux sums per auxiliary
  int64_t h_off;    // per-auxiliary arrays [nh]
  int64_t t_off;    // per-tile partial sums [n_tiles]
  int64_t tc_off;   // per-tile, per-state partial sums [n_tiles * n]
  int64_t m_off;    // modes [nm]
  int64_t sc_off;   // sc CSR pointers [n_sc + 1]
  int64_t sce_off;  // sc entries
  int64_t s_off;    // per-state arrays [n]
  int64_t nz_off;   // candidate segment (set after the candidate count)
  int64_t sv_off;   // survivor segment (set after the survivor count)
  int64_t nzt_off;  // first candidate tile of the slot (partial sums)
  int64_t fl_off;   // flagged (mode, auxiliary) list of the state half
  int32_t n_fl, n_q;  // flagged entries; couplings per sc state (padded)
  int64_t q_off;    // boundary-state coupling entries [n_sc * n_q]
  int64_t te_off;   // per-tile boundary-state flux partials [n_tiles * n_sc]
  int64_t esc_off;  // boundary-state flux sums [n_sc]
  double dsq, f_discard;
  const uint64_t *zob_a, *zob_b;  // Zobrist tables [n_hmode * 256]
};

// Per-slot scalars written by the kernels (scal[8 * i + j]) and counts
// (tot[8 * i + j]) of listed slot i.  A candidate (k, relative mode, step)
// is packed as desc = k | lm << 32 | (step > 0) << 48.
enum { SC_THR = 0, SC_UNSTABLE, SC_DBOUND, SC_THR1, SC_DISC, SC_THR2 };
enum { TT_STABLE = 0, TT_NZ, TT_SURV, TT_LEAD, TT_BOUND };

struct hops_sel_work_t {
  const hops_op_t *ops;       // per slot (device)
  const hops_sel_seg_t *segs; // per listed slot
  const double *pn;           // complex Pn rows
  const double *e_aux;
  const hops_sel_mode_t *modes;
  const int32_t *sc_ptr;
  const hops_sel_sc_t *sc_ent;
  double *err, *err_sorted;   // [sum nh]
  uint8_t *flag;              // [sum nh] stable auxiliaries
  int32_t *cnt, *scnt;        // [sum nh] candidates / survivors per aux, then offsets
  double *part_u, *part_c, *c1;
  double *scal;
  int64_t *tot;
  int64_t *seg_beg, *seg_end; // [n_list] segment bounds for the sorts
  // candidates
  double *nz, *nz_sorted, *part_d;
  // survivors
  uint64_t *ka, *ka_sorted, *kb, *desc;
  uint32_t *pos, *pos_sorted;
  double *sval;
  uint8_t *lead;
  double *gv, *gvals, *gsorted;
  uint64_t *gdesc;
  uint64_t *bound;            // descs of the boundary auxiliaries, slot i at bound_off[i]
  int64_t *bound_off;         // [n_list + 1]
  // state half
  const uint32_t *fl;         // flagged (mode << 2 | fd << 1 | fu, aux) pairs: 2 words each
  double *err_s;              // [sum n]
  const double *e_state;      // [sum n] sums of the deletion flux over stable auxiliaries
  const double *v1;           // [sum n]
  const hops_sel_sc_t *q_ent; // boundary-state couplings
  double *part_e, *e_sc;      // boundary-state flux partials and sums
  double hbar;
  int n_list;
};

// Launchers (C linkage: the host-only sanitizer build of the runtime links
// stand-ins for them).  tiles: (list index, tile) pairs over the auxiliaries
// (or candidate / survivor entries) of each listed slot, 256 per tile.
extern "C" {
int hops_sel_err(const hops_sel_work_t &w, const int2 *tiles, int n_tiles, hipStream_t st);
int hops_sel_stable(const hops_sel_work_t &w, const int2 *tiles, int n_tiles, int want_c1, hipStream_t st);
int hops_sel_fin1(const hops_sel_work_t &w, int want_c1, hipStream_t st);
int hops_sel_cand(const hops_sel_work_t &w, const int2 *tiles, int n_tiles, hipStream_t st);
int hops_sel_disc(const hops_sel_work_t &w, const int2 *nz_tiles, int n_nz_tiles, const int2 *tiles, int n_tiles,
                  hipStream_t st);
int hops_sel_fin2(const hops_sel_work_t &w, hipStream_t st);
int hops_sel_surv(const hops_sel_work_t &w, const int2 *tiles, int n_tiles, hipStream_t st);
int hops_sel_group(const hops_sel_work_t &w, const int2 *sv_tiles, int n_sv_tiles, hipStream_t st);
int hops_sel_fin3(const hops_sel_work_t &w, hipStream_t st);
int hops_sel_bound(const hops_sel_work_t &w, hipStream_t st);
// threshold of the sorted segments (kind 0: stable, 1: pre-trim, 2: boundary)
int hops_sel_threshold(const hops_sel_work_t &w, const double *sorted, int kind, hipStream_t st);
// segmented ascending sorts (rocPRIM); tmp == nullptr: *tmp_bytes gets the size
int hops_sel_sort_keys(void *tmp, size_t *tmp_bytes, const double *in, double *out, size_t size, int segs,
                       const int64_t *beg, const int64_t *end, hipStream_t st);
int hops_sel_sort_pairs(void *tmp, size_t *tmp_bytes, const uint64_t *kin, uint64_t *kout, const uint32_t *vin,
                        uint32_t *vout, size_t size, int segs, const int64_t *beg, const int64_t *end,
                        hipStream_t st);
// state half
int hops_sel_state_err(const hops_sel_work_t &w, hipStream_t st);
int hops_sel_state_flux(const hops_sel_work_t &w, const int2 *tiles, int n_tiles, hipStream_t st);
// copy n rows of complex Pn of each listed slot to out + off[i]
int hops_sel_gather_psi(const hops_sel_work_t &w, double *out, const int64_t *off, hipStream_t st);
}  // extern "C"
