/* host_services.c -- the host-provided services of gadget_defines.h:11-32.
 *
 * The reference library imports hubble_function, terminate, message and the
 * mymalloc/myfree pair from the N-body code.  Here they are weak references:
 * a host that defines them (Gadget, or a test host library loaded with
 * RTLD_GLOBAL) is used; otherwise terminate prints and exits with the code,
 * message prints when KN_VERBOSE is set, allocation uses malloc, and the Hubble
 * function must be registered with kn_set_hubble_function.  A background set
 * with kn_set_background takes precedence over both forms of hubble_function.
 */
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "kn_internal.h"

extern double hubble_function(double a) __attribute__((weak));
extern void terminate(int ierr, const char *fmt, ...) __attribute__((weak));
extern void message(int ierr, const char *fmt, ...) __attribute__((weak));
extern void *mymalloc_fullinfo(const char *string, size_t size, const char *func, const char *file, int line)
    __attribute__((weak));
extern void myfree_fullinfo(void *ptr, const char *func, const char *file, int line) __attribute__((weak));

static double (*g_registered_hubble)(double);
static double (*g_ctx_hubble)(double, void *);
static void *g_ctx_hubble_ctx;
static char g_last_error[512];

void kn_set_hubble_function(double (*fn)(double a)) { g_registered_hubble = fn; }

void kn_set_hubble_ctx(double (*fn)(double a, void *ctx), void *ctx)
{
    g_ctx_hubble = fn;
    g_ctx_hubble_ctx = ctx;
}

double kn_hubble(double a)
{
    if (g_ctx_hubble) return g_ctx_hubble(a, g_ctx_hubble_ctx);
    if (g_registered_hubble) return g_registered_hubble(a);
    if (hubble_function) return hubble_function(a);
    kn_terminate(KN_ERR_NOHUBBLE, "no hubble_function: define it or call kn_set_hubble_function\n");
}

void kn_terminate(int code, const char *fmt, ...)
{
    char buf[1024];
    va_list va;
    va_start(va, fmt);
    vsnprintf(buf, sizeof buf, fmt, va);
    va_end(va);
    if (terminate) {
        terminate(code, "%s", buf);
    }
    fprintf(stderr, "kspace_neutrinos_hip: fatal error %d: %s", code, buf);
    fflush(stderr);
    exit(code);
}

void kn_message(int level, const char *fmt, ...)
{
    char buf[1024];
    va_list va;
    va_start(va, fmt);
    vsnprintf(buf, sizeof buf, fmt, va);
    va_end(va);
    if (message) {
        message(level, "%s", buf);
        return;
    }
    if (getenv("KN_VERBOSE")) fputs(buf, stdout);
}

void *kn_malloc(const char *name, size_t n)
{
    void *p;
    if (mymalloc_fullinfo) p = mymalloc_fullinfo(name, n, __func__, __FILE__, __LINE__);
    else p = malloc(n);
    if (!p && n) kn_terminate(2016, "could not allocate %zu bytes for %s\n", n, name);
    return p;
}

void kn_free(void *p)
{
    if (!p) return;
    if (myfree_fullinfo) myfree_fullinfo(p, __func__, __FILE__, __LINE__);
    else free(p);
}

void kn_set_last_error(const char *fmt, ...)
{
    va_list va;
    va_start(va, fmt);
    vsnprintf(g_last_error, sizeof g_last_error, fmt, va);
    va_end(va);
}

const char *kn_last_error(void) { return g_last_error; }
