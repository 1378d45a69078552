/* err.c -- error reporting of the C ABI.  The reference prints
 * "Error at file:line" and returns EXIT_FAILURE (translation.h:64-126) or
 * calls exit() (octree.c:569-571); here every entry point returns a negative
 * code and keeps a per-thread message. */
#include <stdarg.h>
#include <stdio.h>

#include "../fmb_internal.h"

static _Thread_local char g_msg[512] = "";

int fmb_set_error(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_msg, sizeof(g_msg), fmt, ap);
  va_end(ap);
  return code;
}

const char *fmb_last_error(void) { return g_msg; }

int fmb_version(void) { return 10000; }
