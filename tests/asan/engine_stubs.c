/* Stand-ins for the device entry points of libhops_mi355x.so when the native
   runtime is built for the host alone (the sanitizer test): each fails, and
   the host-only selection path calls none of them.  The runtime reports its
   errors through hops_internal_set_error, kept here for the driver. */
#include <stdio.h>
#include <string.h>

static char g_msg[512];

int hops_internal_set_error(const char *msg) {
  snprintf(g_msg, sizeof(g_msg), "%s", msg ? msg : "");
  return -1;
}
const char *stub_last_error(void) { return g_msg; }

int hops_build_super(void) { return -1; }
int hops_slots_step(void) { return -1; }
int hops_slots_pn(void) { return -1; }
int hops_slots_errors(void) { return -1; }
int hops_slots_build(void) { return -1; }
int hops_slots_permute(void) { return -1; }
int hops_copy_segments(void) { return -1; }
int hops_gather_u64(void) { return -1; }
int hops_sel_err(void) { return -1; }
int hops_sel_stable(void) { return -1; }
int hops_sel_fin1(void) { return -1; }
int hops_sel_cand(void) { return -1; }
int hops_sel_disc(void) { return -1; }
int hops_sel_fin2(void) { return -1; }
int hops_sel_surv(void) { return -1; }
int hops_sel_group(void) { return -1; }
int hops_sel_fin3(void) { return -1; }
int hops_sel_bound(void) { return -1; }
int hops_sel_threshold(void) { return -1; }
int hops_sel_sort_keys(void) { return -1; }
int hops_sel_sort_pairs(void) { return -1; }
int hops_sel_state_err(void) { return -1; }
int hops_sel_state_flux(void) { return -1; }
int hops_sel_gather_psi(void) { return -1; }
