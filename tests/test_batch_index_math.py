"""The batch kernel's float-reciprocal index arithmetic (cave_amd/csrc/batch.hip:
div_ceil_small, divmod_rcp) against exact integer division.

The kernel divides with v_rcp_f32 (within 1 ulp of 1/b) and corrects the
estimate with selects, relying on this bound: for a < 2^22 the float product
a * rcp(b) is within 0.75 / b of a / b, so its truncation (or ceiling) is the
result, one less or one more.  This test restates those few lines in C and
checks them over a < 2^22 (every a below 100,000, then a stride, and the
last 5,000 below 2^22) for divisors from 1 to 2^19, with the reciprocal
exact and one ulp off either way.  The device code itself is checked end to
end by the GPU tests, whose batch results must equal the oracle's.
"""
import os
import subprocess
import tempfile

SRC = r"""
#include <math.h>
#include <stdint.h>
#include <stdio.h>
static float rcp(uint32_t b, int ulp) {
    float r = 1.0f / (float)b;
    return ulp > 0 ? nextafterf(r, 2.0f) : (ulp < 0 ? nextafterf(r, 0.0f) : r);
}
static uint32_t div_ceil_small(uint32_t a, uint32_t b, int ulp) {
    uint32_t q = (uint32_t)ceilf((float)a * rcp(b, ulp));
    q = q * b < a ? q + 1u : q;
    q = (q > 0u && (q - 1u) * b >= a) ? q - 1u : q;
    return q;
}
static uint32_t divmod_rcp(uint32_t a, uint32_t b, int ulp, uint32_t *rem) {
    const uint32_t q = (uint32_t)((float)a * rcp(b, ulp));
    const int32_t r = (int32_t)(a - q * b), bi = (int32_t)b;
    *rem = (uint32_t)(r < 0 ? r + bi : (r >= bi ? r - bi : r));
    return r < 0 ? q - 1u : (r >= bi ? q + 1u : q);
}
static long check(uint32_t a, uint32_t b) {
    long bad = 0;
    for (int u = -1; u <= 1; u++) {
        uint32_t rem, q = divmod_rcp(a, b, u, &rem);
        bad += div_ceil_small(a, b, u) != a / b + (a % b != 0u);
        bad += q != a / b || rem != a % b;
    }
    return bad;
}
int main(void) {
    static const uint32_t bs[] = {1, 2, 3, 5, 7, 9, 11, 13, 17, 31, 33, 49, 63, 64, 65, 97, 127, 255, 257,
                                  1000, 1023, 4095, 4096, 4097, 65535, 65537, 524287, 1u << 19};
    long bad = 0, n = 0;
    for (unsigned k = 0; k < sizeof bs / sizeof *bs; k++) {
        for (uint32_t a = 0; a < (1u << 22); a += a < 100000u ? 1u : 97u, n++) bad += check(a, bs[k]);
        for (uint32_t a = (1u << 22) - 5000u; a < (1u << 22); a++, n++) bad += check(a, bs[k]);
    }
    printf("%ld %ld\n", n, bad);
    return bad != 0;
}
"""


def test_float_reciprocal_division_is_exact_below_2_22():
    with tempfile.TemporaryDirectory() as tmp:
        c = os.path.join(tmp, "idx.c")
        exe = os.path.join(tmp, "idx")
        with open(c, "w") as f:
            f.write(SRC)
        # -ffp-contract=off: a * rcp stays one rounded product, as v_mul_f32 computes it
        subprocess.check_call(["gcc", "-O2", "-ffp-contract=off", "-o", exe, c, "-lm"])
        out = subprocess.run([exe], capture_output=True, text=True)
        n, bad = (int(v) for v in out.stdout.split())
        assert n > 1_000_000 and bad == 0 and out.returncode == 0
