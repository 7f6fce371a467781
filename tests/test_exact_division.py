"""The RK4 kernels replace k/6, k/3 and (len a)/|a| by FMA-corrected products
(Markstein's theorem: z = RN(1/d) within half an ulp, q = RN(k z) within one ulp,
r = k - d q exact, then RN(q + r z) = RN(k/d)), written -fma(-r, z, -q) so that a
zero keeps its sign.  This checks the same formulas, compiled for the CPU with C's
correctly rounded fma(), against IEEE division on random operands over a wide
exponent range and on operands next to multiples of d (the hard cases).  The GPU's
v_fma_f64 is the same correctly rounded operation.  (A 2.2e9-operand run of this check
found no mismatch; the test runs a smaller sample.)
"""
import ctypes
import os
import shutil
import subprocess
import tempfile

import pytest

SRC = r"""
#include <math.h>
#include <stdint.h>
#include <string.h>
static uint64_t s = 0x9E3779B97F4A7C15ull;
static inline uint64_t nx(void) { s ^= s << 13; s ^= s >> 7; s ^= s << 17; return s; }
static inline double bits(uint64_t u) { double d; memcpy(&d, &u, 8); return d; }
static inline int same(double a, double b) { return memcmp(&a, &b, 8) == 0; }
/* a + k / d as in the kernels' add_div (with a = 0: the quotient itself) */
static inline double q_const(double k, double d, double z)
{
    const double q = k * z, r = fma(-q, d, k);
    return -fma(-r, z, -q);
}
/* (len a) / nrm as in the kernels' renorm, with y = 1 / nrm */
static inline double q_var(double n, double nrm, double y)
{
    const double q = n * y, r = fma(-q, nrm, n);
    return -fma(-r, y, -q);
}
long check(long n)
{
    long bad = 0;
    const double ds[2] = {3.0, 6.0};
    for (int j = 0; j < 2; j++) {
        const double d = ds[j], z = 1.0 / d;
        for (long i = 0; i < n; i++) {
            uint64_t u = nx();
            if (i & 1) u = (u & 0x800FFFFFFFFFFFFFull) | ((uint64_t)(1023 - 200 + nx() % 401) << 52);
            const double x = bits(u);
            if (!isfinite(x) || (fabs(x) < 0x1p-1000 && x != 0)) continue;
            if (!same(q_const(x, d, z), x / d)) bad++;
            const double m = bits((nx() & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull) * ldexp(1.0, (int)(nx() % 60) - 30);
            const double xm = d * m, xs[3] = {xm, nextafter(xm, 0.0), nextafter(xm, 1e300)};
            for (int t = 0; t < 3; t++) if (!same(q_const(xs[t], d, z), xs[t] / d)) bad++;
        }
        if (!same(q_const(0.0, d, z), 0.0) || !same(q_const(-0.0, d, z), -0.0)) bad++;
    }
    for (long i = 0; i < n; i++) {
        const double nrm = bits((nx() & 0x000FFFFFFFFFFFFFull) | ((uint64_t)(1023 - 4 + nx() % 8) << 52));
        double a = bits((nx() & 0x800FFFFFFFFFFFFFull) | ((uint64_t)(1023 - 60 + nx() % 64) << 52));
        if (i % 7 == 0) a = nrm * bits((nx() & 0x800FFFFFFFFFFFFFull) | 0x3FE0000000000000ull);
        if (!same(q_var(a, nrm, 1.0 / nrm), a / nrm)) bad++;
    }
    return bad;
}
"""


def test_fma_corrected_division_equals_ieee_division():
    cc = shutil.which("gcc")
    if cc is None:
        pytest.skip("gcc not available")
    with tempfile.TemporaryDirectory() as d:
        src, lib = os.path.join(d, "c.c"), os.path.join(d, "c.so")
        with open(src, "w") as f:
            f.write(SRC)
        subprocess.run([cc, "-O2", "-ffp-contract=off", "-fPIC", "-shared", src, "-o", lib, "-lm"], check=True)
        h = ctypes.CDLL(lib)
        h.check.restype = ctypes.c_long
        h.check.argtypes = [ctypes.c_long]
        assert h.check(2_000_000) == 0
