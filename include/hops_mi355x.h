/*
 * hops_mi355x.h — C ABI of the MI355X HOPS propagation engine.
 *
 * All pointers named *_d are device pointers (HBM) owned by the caller; the
 * library never frees caller memory.  Host pointers are copied during the call.
 * Every entry point returns 0 on success and a negative status on failure;
 * hops_last_error() returns a thread-local message for the last failure.
 *
 * Reference interfaces replaced (MesoHOPS v1.6.0, src/mesohops/...):
 *   hops_build_super       eom/eom_hops_ksuper.py:575-668 calculate_ksuper (and
 *                          :387-572 update_ksuper: the device rebuilds the
 *                          operator of the new basis instead of permuting it)
 *   hops_propagate         trajectory/hops_trajectory.py:456-533 (time loop),
 *                          integrator/integrator_rk.py:10-134 (RK4 + noise
 *                          lookup), eom/hops_eom.py:196-542 (dsystem_dt),
 *                          eom/eom_functions.py:11-309, and
 *                          trajectory/hops_trajectory.py:616-634 (normalize)
 *   hops_noise_filter      noise/hops_noise.py:630-716 _construct_correlated_noise
 *   hops_noise_embedding   noise/hops_noise.py:644-667 (circulant embedding
 *                          spectrum of the correlation function)
 */
#ifndef HOPS_MI355X_H
#define HOPS_MI355X_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* precision modes of hops_propagate */
#define HOPS_PREC_FP64 0  /* complex128 stages, complex128 state          */
#define HOPS_PREC_MIXED 1 /* complex64 stages/SpMV, complex128 state carry */

/* equation-of-motion kinds (eom/hops_eom.py:95-110) */
#define HOPS_EOM_LINEAR 0
#define HOPS_EOM_NONLINEAR 1
#define HOPS_EOM_NORMALIZED_NONLINEAR 2

/*
 * Operator of one basis (one state list x one auxiliary list).  Built by the
 * host from the auxiliary multi-indices; the ELL super-operator arrays are
 * filled on the device by hops_build_super.  Complex values are interleaved
 * (re, im) doubles unless noted.  Row r of the hierarchy vector is
 * r = i_aux * n_state + i_state (eom/hops_eom.py:386-391).
 */
typedef struct hops_op {
  int32_t n_state, n_hier, n_mode, n_lop, n_act, n_first, ell_w, n_pat;
  int32_t n_mode_abs; /* number of absolute hierarchy modes (z_mem length) */
  int32_t pat_w;      /* most pattern entries in one state row            */
  /* auxiliary multi-indices, CSR over auxiliaries: absolute mode, depth */
  const int32_t *aux_ptr_d, *aux_mode_d, *aux_val_d;
  /* connections: up (k -> k+e_m) and down (k -> k-e_m, with k_m of aux k) */
  const int32_t *up_ptr_d, *up_aux_d, *up_mode_d;
  const int32_t *dn_ptr_d, *dn_aux_d, *dn_mode_d, *dn_val_d;
  /* relative L-operators: row CSR and column CSR, [n_lop*(n_state+1)] ptrs */
  const int32_t *lr_ptr_d, *lr_col_d;
  const double *lr_val_d;
  const int32_t *lc_ptr_d, *lc_row_d;
  const double *lc_val_d;
  /* relative modes: L index (relative), active L index or -1, absolute mode,
     g and w (complex) */
  const int32_t *mode_lrel_d, *mode_lact_d, *mode_abs_d;
  const double *mode_g_d, *mode_w_d;
  /* absolute modes: g, w and the active L index of the mode (-2: mode not
     in the basis, -1: in the basis with an inactive L-operator) */
  const double *abs_g_d, *abs_w_d;
  const int32_t *abs_lact_d;
  /* active L-operators: relative index, relative modes sharing the operator
     (CSR over active L), COO of L and of L^2, LTC coefficient */
  const int32_t *act_lrel_d;
  /* noise row of each active L-operator in the z1/z2 arrays of one
     trajectory; NULL: row j holds active L-operator j */
  const int32_t *act_zrow_d;
  const int32_t *am_ptr_d, *am_mode_d;
  const int32_t *la_ptr_d, *la_row_d, *la_col_d;
  const double *la_val_d;
  const int32_t *lq_ptr_d, *lq_row_d, *lq_col_d;
  const double *lq_val_d;
  const double *lt_d;
  /* first-level auxiliaries e_m: aux index, active L index, relative mode */
  const int32_t *first_aux_d, *first_lact_d, *first_mode_d;
  /* system-block pattern (union of H, L and L^2) over the state rows */
  const int32_t *pat_ptr_d, *pat_col_d;
  const double *pat_h_d; /* -iH at each pattern entry */
  const int32_t *pc_ptr_d, *pc_act_d, *pc_kind_d;
  const double *pc_val_d;
  /* outputs of hops_build_super (caller-allocated) */
  int32_t *ell_col_d; /* [ell_w * n_rows], column-major            */
  void *ell_val_d;    /* [ell_w * n_rows] complex, fp64 or fp32     */
  double *k0_d;       /* [n_hier] complex: -sum_m k_m w_m           */
  /* dense system block (no counterpart in the reference's tables: the
     reference multiplies (-1j*H) @ Phi, eom/hops_eom.py:389-391, 526).  NULL:
     -iH is in the pattern above.  Otherwise -iH over the state basis is
     applied on the matrix cores (hops_dense_h) and the pattern holds the L
     and L^2 entries only: two planes (real, imaginary) of hd_n x hd_n values
     in stage precision, row-major, zero beyond n_state; hd_n a multiple of
     16. */
  const void *hd_d;
  int32_t hd_n, hd_pad;
} hops_op_t;

/*
 * Batch of B trajectories sharing one operator (one basis).  The state pools
 * (phi, y0, ya, yb, acc) hold G = ceil(B / T) groups of T = 2^lane_shift
 * trajectories interleaved row by row: element (b, r) is at index
 * ((b >> lane_shift) * n_rows + r) * T + (b & (T - 1)); each pool has
 * G * T * n_rows elements.  The tile list names (group, tile) pairs, a tile
 * being 256 / T rows of one group.
 */
typedef struct hops_batch {
  int32_t B, precision, eom, n_t;  /* n_t = noise time points           */
  int32_t coef_stride, ws_stride;  /* per-trajectory doubles            */
  int32_t store_stride, lane_shift; /* psi slots per trajectory; log2 T */
  double dt_over_hbar;             /* tau / hbar                        */
  int32_t noise_step;              /* noise points per tau/2            */
  int32_t sigma_raw;               /* 1: the first step of the call takes
                                      Phi as it is (sigma = 1) instead of
                                      normalising it, as the reference steps
                                      a state an operator has just changed */
  void *phi_d;   /* complex128 [G*T*n_rows] state carry                 */
  void *y0_d;    /* complex64 [G*T*n_rows] (mixed only)                 */
  void *ya_d;    /* stage input ping   (complex R)                      */
  void *yb_d;    /* stage input pong   (complex R)                      */
  void *acc_d;   /* RK4 accumulator    (complex R)                      */
  const void *z1_d; /* complex64 (or complex128, noise_c128) [B][n_act][n_t]
                       noise 1 (active rows)                              */
  const void *z2_d; /* same layout, noise 2, may be NULL                  */
  double *zmem_d;   /* complex128 [B][n_mode_abs] noise memory          */
  double *ws_d;     /* per-trajectory workspace, ws_stride doubles      */
  double *coef_d;   /* per-trajectory stage coefficients                */
  double *psi_store_d; /* complex128 [B][store_stride][n_state] or NULL */
  const int32_t *tile_d; /* [n_tiles] group*65536+tile, or -1 idle      */
  int32_t n_tiles, tile_rows;
  /* stage coefficients in stage precision, trajectory-fast: complex R
     [(2 n_pat + n_mode + 1) rows][coefT_stride >= B]: rows A, A0, zc, nc;
     sigT_d: double [coefT_stride] scale of Phi in the current step */
  void *coefT_d;
  double *sigT_d;
  int32_t coefT_stride;
  int32_t noise_c128; /* 1: z1_d / z2_d hold complex128 instead of complex64 */
  /* order in which the operator-stationary stage kernel visits its tiles of
     512 >> lane_shift rows (a permutation of the tiles), or NULL for tile
     order; any order gives the same results */
  const int32_t *tile_order_d;
  /* (-iH) Y of the current stage input (complex R, the pool layout), when
     the operator has a dense system block (hd_d), else NULL */
  void *hy_d;
} hops_batch_t;

const char *hops_last_error(void);

/* ABI version of this library (bumped on any change of the structs). */
int hops_version(void);

/* Per-launch timing of hops_propagate (no counterpart in the reference; used
   by the benchmark).  enable(1) clears the counters and records a HIP event
   on the launch stream after every kernel; read() returns the summed
   milliseconds and launch counts for [0] the stage prologue, [1] the fused
   derivative/RK4 stage kernel, [2] other kernels, [3] the dense system block
   (-iH Y on the matrix cores).  The switch and counters belong to the calling
   host thread: they cover the calls that thread makes. */
int hops_profile_enable(int on);
int hops_profile_read(double *ms_out /*[4]*/, long long *n_out /*[4]*/);

/* Engine options (no counterpart in the reference).  "stationary_stage":
   1 (default) runs the operator-stationary stage kernel for batches of at
   least 8 trajectories with ell_w <= 16; 0 forces the streaming kernel.
   "chunk_traj": trajectories per chunk propagated through all steps before
   the next (0, default: the whole batch).  "nt_stream": nontemporal hint
   level 0..4 of the stationary kernel.  "groups_q": trajectory-group split
   of the stationary kernel (0: automatic).  "vec_lanes": 1 (default), 2 or
   4 adjacent trajectory lanes per thread of the stationary kernel (reduced
   to the interleave and to a divisor of the batch size).  "lds_stage": 1
   runs the LDS-staged form of the stationary kernel where its tiles fit.
   "fuse_prologue": 1 (default) computes the four stage prologues of a LINEAR
   step in one launch (stream-ordered scratch for the tables of stages 1-3).
   "graph_steps": K (default 16) RK4 steps of hops_propagate captured in one
   HIP graph and replayed for steps 1, 2, ... (0, or profiling enabled, or a
   call with fewer than 4 K + 1 steps: every kernel launched eagerly).
   "resident_rows": bases of at most this many rows per trajectory (default
   1024) run a whole hops_propagate call in one launch, one workgroup per
   trajectory looping over the steps (0: never).
   Options are process-wide library state: set them from one thread while
   no other call is in progress.  No option changes any result. */
int hops_set_option(const char *name, int value);

/* Fill op->ell_col_d/ell_val_d/k0_d from the connection tables.
   precision selects the ELL value type (complex128 or complex64). */
int hops_build_super(const hops_op_t *op, int precision, void *stream);

/* Advance every trajectory of the batch by n_steps RK4 steps of size tau
   starting at noise index it0 (time t0 = it0 * TAU), storing psi after each
   step into psi_store_d slot (store0 + step) when psi_store_d != NULL.
   On return phi_d holds the (normalised, for NORMALIZED NONLINEAR) state. */
int hops_propagate(const hops_op_t *op, const hops_batch_t *batch, int it0,
                   int n_steps, int store0, void *stream);

/* Largest scratch hops_propagate_noalloc uses: the coefficient tables of
   the stages of a fused LINEAR prologue, 3 * 16 * coefT_stride *
   (2 n_pat + n_mode + 1) + 3 * 8 * B * coef_stride bytes in FP64
   (complex64 tables in mixed precision take half of the first term). */
#define HOPS_NOALLOC_SCRATCH_BYTES (32u << 20)

/* hops_propagate without any device allocation, for callers that cannot
   allocate (code under HIP graph capture, tight serving loops, harnesses
   that forbid hipMalloc).  It has no counterpart in the reference.  Same
   arguments, kernels and results as hops_propagate, except that:
   - the stage-coefficient scratch is a library-owned device array of
     HOPS_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends; a
     batch that needs more fails with an error;
   - steps are launched directly: no HIP graph is built and no profiler
     event is recorded.
   Every call on one GPU shares that GPU's array: calls that overlap on one
   GPU (on different streams or host threads) race.  Calls on one stream
   are ordered and safe. */
int hops_propagate_noalloc(const hops_op_t *op, const hops_batch_t *batch, int it0,
                           int n_steps, int store0, void *stream);

/* hbar dY/dt for every trajectory of the batch at the input pool y_d
   (complex128, same layout as phi_d), with the batch's current z-memory and
   the noise at index tidx; output d_d (complex128, same layout).  FP64
   batches only.  Replaces the extra dsystem_dt calls of the adaptive error
   estimate (basis/basis_functions_adaptive.py:10-74). */
int hops_derivative(const hops_op_t *op, const hops_batch_t *bt, const void *y_d, void *d_d, int tidx,
                    void *stream);

/* dst[dst_idx[i]] = scale * src[src_idx[i]] for i < n (complex128): moves
   the stable entries of the hierarchy vector into a new basis and rescales
   them (basis/hops_basis.py:214-274). */
int hops_permute_c128(const void *src_d, void *dst_d, const int32_t *src_idx_d, const int32_t *dst_idx_d,
                      int n, double scale, void *stream);

/* An operator on the system states applied to every auxiliary of the
   hierarchy vector (complex128): phi_out[k, s] = sum_s' op[s, s'] phi_in[k, s']
   for B trajectories in the interleaved layout of hops_batch_t (element
   (b, r) at ((b >> lane_shift) * n_state * n_hier + r) * 2^lane_shift +
   (b & (2^lane_shift - 1)), r = k * n_state + s).  op_d: dense row-major
   [n_state][n_state] complex128 in device memory; phi_in_d and phi_out_d must
   not overlap.  Replaces the matrix product of HopsTrajectory._operator
   (trajectory/hops_trajectory.py:540-577). */
int hops_apply_state_op(const void *phi_in_d, void *phi_out_d, const void *op_d, int n_state, int n_hier, int B,
                        int lane_shift, void *stream);

/* Circulant-embedding spectrum s_w[l][j] = max(0, Re FFT(embed(tilde c_l)))
   for n_lop correlation functions of n_t points (FFT length 2(n_t-1)).
   ang_d, mag_d: float32 [n_lop][n_t] phase and modulus of the complex64
   correlation function as NumPy's np.angle/np.abs round them (the embedding
   is ill-conditioned in its near-zero bins, so these roundings are inputs,
   noise/hops_noise.py:653-656); s_w_d: double [n_lop][2(n_t-1)];
   nu_d: double [n_lop] rotation frequency of each row. */
int hops_noise_embedding(const float *ang_d, const float *mag_d, int n_lop, int n_t,
                         double *s_w_d, double *nu_d, void *stream);

/* Correlated noise for R rows: out[r] = c64( unrotate( IFFT( FFT(z[r]) *
   sqrt(s_w[lop(r)]/2) )[:n_t] ) ).  z_d: complex64 [R][2(n_t-1)];
   row_lop_d: int32 [R] row -> embedding row; out_d: complex64 [R][n_t].
   The embedding and both filters may be called from several host threads
   on different streams at once: transforms longer than the LDS holds work
   in library-owned scratch rows kept per stream, and the shared Bluestein
   plans are published only once written. */
int hops_noise_filter(const void *z_d, int R, int n_t, const double *s_w_d,
                      const double *nu_d, const int32_t *row_lop_d, void *out_d,
                      void *stream);

/* The same filter with the complex128 values before the complex64 cast
   (out_d: complex128 [R][n_t]): the reference builds its cubic interpolation
   on these (noise/hops_noise.py:236-243). */
int hops_noise_filter_c128(const void *z_d, int R, int n_t, const double *s_w_d, const double *nu_d,
                           const int32_t *row_lop_d, void *out_d, void *stream);

/* Cubic interpolation of R complex128 rows y[n] given on the grid i*h, with
   not-a-knot ends (scipy.interpolate.interp1d(kind="cubic"), which the
   reference builds for INTERPOLATE = True, noise/hops_noise.py:242-243,
   259-261, 379-385), evaluated at t0 + j*dt for j < n_out into out_d
   (complex128 [R][n_out]).  cp_d: double [>= n - 4] pivots c_0 = 1/4,
   c_k = 1/(4 - c_{k-1}); work_d: complex128 [R][n]; n >= 4. */
int hops_noise_spline(const void *y_d, int R, int n, double h, double t0, double dt, int n_out,
                      const double *cp_d, void *work_d, void *out_d, void *stream);

/* Moving average of the noise for EFFECTIVE_NOISE_INTEGRATION
   (integrator/integrator_rk.py:115-128): out[r][i] = mean(z[r][i .. i+w)),
   complex64 [R][n_t] in and out, w = (tau/2)/TAU samples (1..64), summed in
   NumPy's order for complex64 means; 0 where the window passes the end.  The
   RK4 stages then read out[] at the usual indices it + {0, w, w, 2w}. */
int hops_noise_window_mean(const void *z_d, void *out_d, int R, int n_t, int w, void *stream);

/* ---------------------------------------------------------------------------
 * Ragged batches ("slots"): trajectories with different bases, and different
 * systems, advanced by one launch per stage.  Slot i has its own operator
 * ops_d[i] and a B = 1, lane_shift = 0 batch descriptor bts_d[i] (both arrays
 * in device memory); list_d names the slots of a call and tiles_d their
 * 256-row tiles (slot << 20 | tile, every tile of every listed slot once).
 * it_d[i]: noise index of slot i at the start of its current step.
 * shm_doubles: max over the listed slots of 2 n_state + 8 n_act + 2 n_mode_abs
 * + 16.  Together these replace the per-trajectory Python loop
 * (trajectory/hops_trajectory.py:456-533) for a batch of adaptive
 * trajectories whose bases differ.
 * ------------------------------------------------------------------------- */

/* Phi move of one slot into its new basis (basis/hops_basis.py:214-274):
   new row (j, s) = scale * old row (inv_aux[j], inv_state[s]), or 0 when
   either index is -1. */
typedef struct hops_slot_perm {
  const void *old_phi_d; /* complex128 [nh_old * n_old] */
  void *new_phi_d;       /* complex128 [nh_new * n_new] */
  const int32_t *inv_aux_d, *inv_state_d;
  int32_t n_old, n_new, nh_new, pad;
  double scale;
} hops_slot_perm_t;

/* One RK4 step (integrator/integrator_rk.py:10-75, eom/hops_eom.py:196-542)
   of every listed slot, then the normalisation of NORMALIZED NONLINEAR
   slots (trajectory/hops_trajectory.py:616-634).  One precision per call. */
int hops_slots_step(const hops_op_t *ops_d, const hops_batch_t *bts_d, const int32_t *list_d, int n_list,
                    const uint32_t *tiles_d, int n_tiles, const int32_t *it_d, int precision,
                    int shm_doubles, void *stream);

/* Pn = Phi / |psi| of every listed slot (basis/hops_basis.py:157): complex128
   into out_d + off_d[slot] (elements), stage precision into the slot's ya
   pool, |psi| into nrm_d[slot].  psi_only != 0: copy psi = Phi[:n_state] to
   out_d + off_d[slot] and nothing else. */
int hops_slots_pn(const hops_op_t *ops_d, const hops_batch_t *bts_d, const int32_t *list_d, int n_list,
                  const int64_t *off_d, void *out_d, double *nrm_d, int precision, int psi_only,
                  void *stream);

/* Derivative part of the adaptive error bound (basis_functions_adaptive.py:
   10-74; hops_basis.py:560-665, :795-912).  With mask_d == NULL: D at Pn
   (from hops_slots_pn, pn_d + off_d[slot]) and e_d + eoff_d[slot] gets, per
   auxiliary k, sum_s |D[k,s]/hbar + Pn[k,s]/dt[slot]|^2.  With mask_d:
   Pn restricted to the auxiliaries k with mask_d[moff_d[slot] + k] != 0,
   and per state s the same sum over those auxiliaries. */
int hops_slots_errors(const hops_op_t *ops_d, const hops_batch_t *bts_d, const int32_t *list_d, int n_list,
                      const uint32_t *tiles_d, int n_tiles, const int32_t *it_d, int precision,
                      int shm_doubles, const int64_t *off_d, const void *pn_d, const int64_t *moff_d,
                      const uint8_t *mask_d, const int64_t *eoff_d, double *e_d, const double *dt_d,
                      double hbar, void *stream);

/* hops_build_super for every listed slot (eom/eom_hops_ksuper.py:387-668). */
int hops_slots_build(const hops_op_t *ops_d, const int32_t *list_d, int n_list, const uint32_t *tiles_d,
                     int n_tiles, int precision, void *stream);

/* Byte copies from a staging area into per-slot device blocks, one launch for
   all segments (no counterpart in the reference: it uploads the tables of the
   new bases of a whole batch at once).  bytes: multiple of 16, src and dst
   16-byte aligned. */
typedef struct hops_segment {
  const void *src;
  void *dst;
  long long bytes;
} hops_segment_t;
int hops_copy_segments(const hops_segment_t *segs_d, int n, void *stream);

/* Gathers n 8-byte words: out_d[i] = *src_d[i] (no counterpart in the
   reference: it collects, for the adaptive selection of a whole batch in
   one launch, the noise values and noise memory that the noise operator of
   off-diagonal L-operators needs, hops_basis.py:914-953).  src_d and out_d
   are device arrays; every src_d[i] is an 8-byte aligned device address. */
int hops_gather_u64(const uint64_t *const *src_d, uint64_t *out_d, int n, void *stream);

/* Phi of the new bases; tiles run over the new rows of each slot and index
   perm_d by slot. */
int hops_slots_permute(const hops_slot_perm_t *perm_d, const uint32_t *tiles_d, int n_tiles, void *stream);

#ifdef __cplusplus
}
#endif
#endif
