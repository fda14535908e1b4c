/* Sanitizer driver for the oracle (test infrastructure only): builds with
 * fpntt_oracle.c under AddressSanitizer + UndefinedBehaviorSanitizer, or
 * ThreadSanitizer, and runs products on several threads at once (the plan
 * cache is shared, as the reference's is, src/fpntt.jl:273-293), each checked
 * against a schoolbook product.  Exit status 0 iff every product matched;
 * the sanitizers report (and fail the run on) any memory or race error. */
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

int or_mul(uint64_t *r, const uint64_t *a, int64_t m, const uint64_t *b, int64_t n);
int or_sqr(uint64_t *r, const uint64_t *a, int64_t n);

typedef unsigned __int128 u128;

static uint64_t splitmix(uint64_t *s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static void schoolbook(uint64_t *r, const uint64_t *a, int64_t m, const uint64_t *b, int64_t n) {
    memset(r, 0, sizeof(uint64_t) * (size_t)(m + n));
    for (int64_t i = 0; i < m; i++) {
        uint64_t carry = 0;
        for (int64_t j = 0; j < n; j++) {
            u128 t = (u128)a[i] * b[j] + r[i + j] + carry;
            r[i + j] = (uint64_t)t;
            carry = (uint64_t)(t >> 64);
        }
        r[i + n] = carry;
    }
}

/* operand styles of test/test_fpntt.jl:120-131: uniform, all-ones, single high bit */
static void fill(uint64_t *a, int64_t n, int style, uint64_t *seed) {
    for (int64_t i = 0; i < n; i++) a[i] = style == 1 ? ~0ull : style == 2 ? 0 : splitmix(seed);
    if (style == 2) a[n - 1] = 1ull << 63;
    if (a[n - 1] == 0) a[n - 1] = 1;
}

static int failures = 0;
static pthread_mutex_t mu = PTHREAD_MUTEX_INITIALIZER;

static void *worker(void *arg) {
    uint64_t seed = (uint64_t)(uintptr_t)arg * 0x51ED2701ull + 7;
    /* sizes reaching lengths m * 2^k for m in {1, 3, 5, 7, 9, 15, 21, 35, 45, 63}, and b = 49 at 1 limb */
    static const int64_t sizes[][2] = {{1, 1}, {3, 2}, {40, 40}, {152, 152}, {300, 100}, {700, 400},
                                       {1000, 999}, {1500, 64}, {2100, 1300}};
    const int ns = (int)(sizeof sizes / sizeof sizes[0]);
    for (int it = 0; it < 2 * ns; it++) {
        const int64_t m = sizes[it % ns][0], n = sizes[it % ns][1];
        const int style = (int)(((uintptr_t)arg + (uint64_t)it) % 3);
        uint64_t *a = malloc(8 * (size_t)m), *b = malloc(8 * (size_t)n);
        uint64_t *r = malloc(8 * (size_t)(m + n)), *w = malloc(8 * (size_t)(m + n));
        uint64_t *s = malloc(16 * (size_t)m), *ws = malloc(16 * (size_t)m);
        fill(a, m, style, &seed);
        fill(b, n, (style + 1) % 3, &seed);
        int bad = or_mul(r, a, m, b, n) != 0;
        schoolbook(w, a, m, b, n);
        bad |= memcmp(r, w, 8 * (size_t)(m + n)) != 0;
        bad |= or_sqr(s, a, m) != 0;
        schoolbook(ws, a, m, a, m);
        bad |= memcmp(s, ws, 16 * (size_t)m) != 0;
        if (bad) {
            pthread_mutex_lock(&mu);
            failures++;
            fprintf(stderr, "mismatch: %lld x %lld style %d\n", (long long)m, (long long)n, style);
            pthread_mutex_unlock(&mu);
        }
        free(a); free(b); free(r); free(w); free(s); free(ws);
    }
    return NULL;
}

int main(void) {
    enum { NT = 4 };
    pthread_t th[NT];
    for (int i = 0; i < NT; i++) pthread_create(&th[i], NULL, worker, (void *)(uintptr_t)(i + 1));
    for (int i = 0; i < NT; i++) pthread_join(th[i], NULL);
    printf("sanitized oracle: %d mismatches\n", failures);
    return failures ? 1 : 0;
}
