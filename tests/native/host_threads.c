/* Host-side stress of libhasty_hip's entry points that need no GPU, from several threads
 * at once, for a ThreadSanitizer / AddressSanitizer build of the library's host code:
 * options (set / get), the LLR shift sequence and workspace queries (the shared row-split
 * table), error paths and the per-thread last-error message. Exits 0 when every check
 * holds; the sanitizer reports races and memory errors on stderr. */
#include <pthread.h>
#include <stdio.h>
#include <string.h>

#include "../../include/hasty_hip.h"

#define NTHREADS 8
#define ROUNDS 2000

static int failures = 0;
static pthread_mutex_t fail_mu = PTHREAD_MUTEX_INITIALIZER;

static void fail(const char* what)
{
    pthread_mutex_lock(&fail_mu);
    ++failures;
    fprintf(stderr, "check failed: %s\n", what);
    pthread_mutex_unlock(&fail_mu);
}

static void* worker(void* arg)
{
    const int id = (int)(long)arg;
    const int64_t shape[3] = {16, 16, 16};
    const int64_t dims5[5] = {20 + id, 1, 64, 64, 64};
    int32_t shifts[3 * 4];
    for (int r = 0; r < ROUNDS; ++r) {
        /* options: every thread writes valid values and reads them back as some valid value */
        if (hasty_set_option("llr_max_sweeps", 1 + (r + id) % 15) != 0) fail("set llr_max_sweeps");
        int64_t v = -1;
        if (hasty_get_option("llr_max_sweeps", &v) != 0 || v < 1 || v > 15) fail("get llr_max_sweeps");
        if (hasty_set_option("sense_zconv_batch", 1 + (r % 16)) != 0) fail("set sense_zconv_batch");
        /* the shift sequence is the same on every call */
        if (hasty_llr_shifts(4, shape, shifts) != 0) fail("llr shifts");
        if (shifts[0] != 0 || shifts[4] != 0 || shifts[8] != 0) fail("first shift is zero");
        /* workspace queries read the shared row-split table */
        if (hasty_llr_workspace_size(dims5, shape, shape) == 0) fail("llr workspace size");
        if (hasty_llr_blocks_workspace_size(dims5, shape, 1 + r % 64) == 0) fail("llr blocks workspace size");
        /* an invalid call sets this thread's message and no other thread's */
        if (hasty_set_option("no_such_option", 1) != 1) fail("unknown option accepted");
        if (strstr(hasty_last_error(), "no_such_option") == NULL) fail("thread-local error message");
    }
    return NULL;
}

int main(void)
{
    pthread_t th[NTHREADS];
    for (long i = 0; i < NTHREADS; ++i) pthread_create(&th[i], NULL, worker, (void*)i);
    for (int i = 0; i < NTHREADS; ++i) pthread_join(th[i], NULL);
    hasty_set_option("llr_max_sweeps", 15);
    hasty_set_option("sense_zconv_batch", 16);
    printf("host_threads: %d threads x %d rounds, %d failures\n", NTHREADS, ROUNDS, failures);
    return failures ? 1 : 0;
}
