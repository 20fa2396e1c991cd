/* Checks that fma(fma(-q, 1000, a), RN(1/1000), q) with q = a * RN(1/1000) equals
 * the correctly rounded a / 1000.0 for every 32-bit integer a (the kernels' form
 * of the depth box mean's first division).  Prints the number of mismatches. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>

int main(void) {
    const double y = 1.0 / 1000.0;
    long long bad = 0;
    for (int64_t a = INT32_MIN; a <= INT32_MAX; a++) {
        const double ad = (double)a;
        const double q = ad * y;
        const double q1 = fma(fma(-q, 1000.0, ad), y, q);
        if (q1 != ad / 1000.0) bad++;
    }
    printf("%lld\n", bad);
    return 0;
}
