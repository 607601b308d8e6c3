/*
 * hops_runtime.h — native host runtime of the MI355X HOPS engine: ragged
 * batches of adaptive trajectories (each with its own basis and, if wanted,
 * its own system), advanced together on one GPU.
 *
 * The runtime owns the basis bookkeeping the reference does in Python per
 * trajectory, and drives the slot kernels of hops_mi355x.h:
 *   hops_rt_batch_step   trajectory/hops_trajectory.py:401-533 (time loop with
 *                        basis updates), :636-714 (INCH_WORM early steps)
 *   (inside)             basis/hops_basis.py:117-292 define_basis /
 *                        update_basis, :294-998 error bounds,
 *                        basis/basis_functions.py:8-41 thresholds,
 *                        basis/hops_fluxfilters.py flux filters,
 *                        basis/hops_hierarchy.py, hops_aux.py, hops_system.py,
 *                        hops_modes.py (auxiliary order, connections, relative
 *                        indexing), eom/eom_hops_ksuper.py:387-668 (operator of
 *                        each new basis, rebuilt on the device)
 *
 * Device memory comes from one caller-allocated arena; the runtime never
 * calls hipMalloc for trajectory data, so admission by HBM footprint is a
 * question of arena space (hops_rt_arena_free).  All work runs on the stream
 * given at creation.  Functions return 0 (or a non-negative value) on
 * success and a negative status on failure; hops_last_error() has the text.
 */
#ifndef HOPS_RUNTIME_H
#define HOPS_RUNTIME_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* A system (basis/system_functions.py:10-171): Hamiltonian, hierarchy modes
   and their unique L-operators (COO per operator, zeros removed), and the
   low-temperature correction coefficient of each L-operator. */
typedef struct hops_rt_system_desc {
  int32_t n_state, n_hmode, n_lop, pad;
  const double *H;          /* complex [n_state][n_state] row-major          */
  const double *g, *w;      /* complex [n_hmode]                             */
  const int32_t *l_of_mode; /* [n_hmode] L-operator index of each mode       */
  const int32_t *l_ptr;     /* [n_lop + 1] CSR over operators of COO entries */
  const int32_t *l_row, *l_col;
  const double *l_val;      /* complex                                       */
  const double *lt;         /* complex [n_lop] LTC coefficients, or NULL     */
} hops_rt_system_desc_t;

/* Per-trajectory parameters (eom/hops_eom.py:20-41, hops_trajectory.py:34-50,
   :338-399).  delta_a / delta_s = 0 turns that half of the adaptivity off;
   with delta_a = 0 the hierarchy is aux0 (CSR over auxiliaries of absolute
   mode / depth entries, first entry the zero auxiliary). */
typedef struct hops_rt_traj_param {
  int32_t eom, maxhier, update_step, early_integrator; /* 0 INCH_WORM, 1 STATIC,
                                                          2 STATIC_STATE_INCHWORM_HIERARCHY */
  int32_t early_steps, inchworm_cap, noise_step, n_t;
  int32_t it0, record, n_aux0;
  int32_t noise_c128;      /* 1: z1_d / z2_d hold complex128 (interpolated or
                              complex128 PRE_CALCULATED noise), else complex64 */
  double delta_a, delta_s, f_discard, dt;
  const uint8_t *markov;   /* [n_hmode] Markovian filter flags, or NULL */
  const int32_t *aux0_ptr, *aux0_mode, *aux0_val;
  /* Triangular (kind 0) and LongEdge (kind 1) static filters on the flux up
     of the adaptive hierarchy (basis/hops_fluxfilters.py:372-501), in order:
     kind, kmax2 and a [n_hmode] mask of filtered modes each */
  int32_t n_filters, pad2;
  const int32_t *filter_kind, *filter_kmax;
  const uint8_t *filter_mask; /* [n_filters][n_hmode] */
} hops_rt_traj_param_t;

void *hops_rt_system_create(const hops_rt_system_desc_t *desc);
int hops_rt_system_destroy(void *sys);

/* A ragged batch over arena_d[0, arena_bytes) (caller-owned device memory,
   256-byte aligned).  precision: HOPS_PREC_FP64 or HOPS_PREC_MIXED.
   n_threads: host worker threads for the basis bookkeeping (<= 0: 16). */
void *hops_rt_batch_create(int precision, void *arena_d, size_t arena_bytes, int max_slots, int n_threads,
                           void *stream);
int hops_rt_batch_destroy(void *batch);

/* Admit one trajectory: initial state psi0 (complex [n_state] host), its
   noise z1_d / z2_d (complex64 [n_lop][n_t] device, caller-owned, z2_d may be
   NULL), parameters prm.  Builds the initial basis and its first update
   (trajectory/hops_trajectory.py:253-336).  Returns the slot id. */
int hops_rt_batch_add(void *batch, void *sys, const hops_rt_traj_param_t *prm, const double *psi0,
                      const void *z1_d, const void *z2_d);
/* Release a slot and its arena memory. */
int hops_rt_batch_remove(void *batch, int slot);
/* Points a slot at new noise arrays for its next steps: z1_d / z2_d
   ([n_lop][n_t], complex64 or with c128 complex128, caller-owned), the
   noise index it of the slot's current time and the index step noise_step
   per half RK4 step.  Used per propagate call for interpolated noise, whose
   values are sampled on the call's own half-step grid
   (noise/hops_noise.py:379-385); no counterpart in the reference. */
int hops_rt_slot_noise(void *batch, int slot, const void *z1_d, const void *z2_d, int n_t, int it,
                       int noise_step, int c128);
/* EFFECTIVE_NOISE_INTEGRATION (integrator/integrator_rk.py:115-128): the
   noise the slot's RK stages read (window means, same layout as z1_d / z2_d
   of hops_rt_batch_add); the selection keeps reading the point samples, as
   the reference's error bound does (trajectory/hops_trajectory.py:693-714).
   z1_d NULL: the stages read the point samples again.  No counterpart in the
   reference's interfaces. */
int hops_rt_slot_stage_noise(void *batch, int slot, const void *z1_d, const void *z2_d);
/* STATIC_BASIS (trajectory/hops_trajectory.py:299-317, :478-491) of a slot,
   before its first step: states and auxiliaries (CSR of absolute mode /
   depth entries) joined to its first basis in place of the first selection;
   the early integrator 2 keeps these states through the early steps. */
int hops_rt_slot_static_basis(void *batch, int slot, int n_states, const int32_t *states, int n_aux,
                              const int32_t *aux_ptr, const int32_t *aux_mode, const int32_t *aux_val);
/* Device bytes a trajectory would take at the given basis size, for
   admission by HBM footprint. */
long long hops_rt_slot_estimate(void *batch, void *sys, int n_state, int n_hier);
long long hops_rt_arena_free(void *batch);
/* Accumulated wall seconds per phase of hops_rt_batch_step (no counterpart
   in the reference): [0] RK4 steps, [1] error sums on the device, [2] host
   hierarchy selection, [3] state errors and selection, [4] table build of
   new bases, [5] device update (operators, Phi moves), [6] psi fetch;
   [7] (a count, not seconds) hierarchy rows advanced by RK4 steps. */
int hops_rt_batch_timers(void *batch, double *out /*[8]*/);

/* Batch options (no counterpart in the reference).  "device_select": 1 runs
   the adaptive basis selection of the eligible slots on the device
   (hierarchy adaptivity on, no Triangular / LongEdge filter, at most one
   diagonal entry per mode's L-operator in the state basis); 0 (default) runs
   every selection on the host threads.  The device form is experimental: it
   equals the host selection bit for bit in a CPU emulation of its kernels,
   but has not passed on the GPU. */
int hops_rt_batch_option(void *batch, const char *name, long long value);

/* Advance every active slot (slots_or_null == NULL) or the listed slots by
   n_steps steps of their own dt.  Returns HOPS_RT_ARENA_FULL when the new
   basis of some slots did not fit in the arena: those slots are marked out
   of memory (info[8]) and left mid-step, every other slot completed the step,
   and the call stops there.  Remove the marked slots before the next call
   (a serving loop requeues their requests). */
#define HOPS_RT_ARENA_FULL (-2)
int hops_rt_batch_step(void *batch, int n_steps, const int32_t *slots_or_null, int n_slots);

/* Read-backs.  info[0..8] = n_state, n_hier, steps taken, noise index,
   recorded steps, device bytes, early steps left, active, out of memory. */
int hops_rt_slot_info(void *batch, int slot, long long *info);
int hops_rt_slot_states(void *batch, int slot, int32_t *states_out);
int hops_rt_slot_aux(void *batch, int slot, int32_t *ptr_out, int32_t *mode_out, int32_t *val_out);
int hops_rt_slot_phi(void *batch, int slot, double *phi_out);     /* complex [n_state*n_hier] */
int hops_rt_slot_zmem(void *batch, int slot, double *zmem_out);   /* complex [n_hmode]        */
/* Recorded step idx (0 = initial state): time index, psi over the basis
   states, the states, n_hier. */
int hops_rt_slot_record(void *batch, int slot, int idx, double *psi_out, int32_t *states_out,
                        int32_t *n_state_out, int32_t *n_hier_out);
int hops_rt_slot_clear_records(void *batch, int slot);

/* Host-only hooks used by the tests (no GPU): build the tables of one basis,
   read one table, run the hierarchy / state selection on given inputs. */
void *hops_rt_debug_basis(void *sys, int eom, int n_states, const int32_t *states, int n_aux,
                          const int32_t *aux_ptr, const int32_t *aux_mode, const int32_t *aux_val);
long long hops_rt_debug_table(void *basis, const char *name, void *out, long long cap_bytes);
int hops_rt_debug_select(void *basis, const hops_rt_traj_param_t *prm, const double *pn, const double *e_aux,
                         double dt, uint8_t *stable_out);
int hops_rt_debug_select_states(void *basis, const double *e_state);
/* The noise of the step for a basis with off-diagonal L-operators, before
   hops_rt_debug_select: z1 / z2 (complex64 [n_lop of the basis], z2 may be
   NULL) of the basis L-operators and the noise memory (complex128 [n_hmode]).
   The error terms of the last selection: which 0 per auxiliary, 1 per basis
   state, 2 (state, error) pairs of the outside states; returns the count of
   doubles, copied to out when it fits in cap. */
int hops_rt_debug_noise(void *basis, const float *z1, const float *z2, const double *zmem);
/* Overwrites the two Zobrist words of (mode, depth) of a system, so that a
   test can make two auxiliaries share a key: every key hit is confirmed
   against the auxiliaries, and a collision fails the basis or selection. */
int hops_rt_debug_set_zobrist(void *sys, int mode, int depth, unsigned long long a, unsigned long long b);
long long hops_rt_debug_errors(void *basis, int which, double *out, long long cap);
int hops_rt_debug_result(void *basis, int32_t *n_states, int32_t *states_out, int32_t *n_aux, int32_t *aux_ptr,
                         int32_t *aux_mode, int32_t *aux_val, long long cap);
int hops_rt_debug_free(void *basis);
/* error_threshold(sort(v), max_err, offset) of basis/basis_functions.py:8-41
   as the selection computes it (only the needed prefix of v is sorted), and
   the sum of the values <= it. */
double hops_rt_debug_threshold(const double *v, long long n, double max_err, double offset, double *below);

#ifdef __cplusplus
}
#endif
#endif
