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
