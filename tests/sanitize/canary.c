/* A known heap overflow: the sanitizer run first checks that AddressSanitizer
 * reports this one, so a clean run of the suite means something. */
#include <stdlib.h>

int canary_overflow(int n) {
    int* a = (int*)malloc((size_t)n * sizeof(int));
    for (int i = 0; i < n; ++i) a[i] = i;
    int s = a[n]; /* one past the end */
    free(a);
    return s;
}
