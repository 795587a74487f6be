"""The arithmetic of the complex64 GEMM on the bf16 matrix cores, emulated on
the CPU: every FP32 operand value v is split into three bf16 parts (each the
round-to-nearest-even bf16 of what the previous parts leave), and a real
product keeps the six partial products hi.hi, hi.mid, mid.hi, hi.lo, mid.mid,
lo.hi.  The split represents v to 2^-24 |v|, and the six products give x y to
a few units of 2^-24 |x y| -- FP32 accuracy, which the GPU kernel's own tests
then check against the oracle's matvec."""

import numpy as np

from bench import product_form


def bf16_rne(x):
    """FP32 -> nearest bf16 (ties to even), returned as FP32."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def split3(v):
    v = np.asarray(v, dtype=np.float32)
    h = bf16_rne(v)
    r1 = (v - h).astype(np.float32)  # exact: h carries the leading 8 bits of v
    m = bf16_rne(r1)
    r2 = (r1 - m).astype(np.float32)
    return h, m, bf16_rne(r2)


def test_split_represents_fp32_values():
    rng = np.random.default_rng(3)
    v = (rng.standard_normal(200_000) * np.exp2(rng.integers(-40, 40, 200_000))).astype(np.float32)
    h, m, l = split3(v)
    for part in (h, m, l):  # every part is a bf16 value
        assert np.all(part.view(np.uint32) & 0xFFFF == 0)
    # residuals exact in FP32
    assert np.array_equal((v - h).astype(np.float32).astype(np.float64), v.astype(np.float64) - h)
    err = np.abs(v.astype(np.float64) - (h.astype(np.float64) + m + l))
    assert np.all(err <= 2.0 ** -24 * np.abs(v.astype(np.float64)))


def test_six_partial_products_keep_fp32_accuracy():
    rng = np.random.default_rng(5)
    x = rng.standard_normal(100_000).astype(np.float32)
    y = rng.standard_normal(100_000).astype(np.float32)
    xh, xm, xl = (p.astype(np.float64) for p in split3(x))
    yh, ym, yl = (p.astype(np.float64) for p in split3(y))
    six = xl * yh + xm * ym + xh * yl + xm * yh + xh * ym + xh * yh
    exact = x.astype(np.float64) * y.astype(np.float64)
    rel = np.abs(six - exact) / np.abs(exact)
    assert rel.max() <= 4 * 2.0 ** -24
    # dropping the smallest products loses FP32 accuracy: the terms are needed
    four = xm * yh + xh * ym + xh * yh + xm * ym
    assert (np.abs(four - exact) / np.abs(exact)).max() > 2.0 ** -20


def test_bench_reads_the_product_form_from_the_tuner_table():
    t = ["fp32 op00 step A chi_l 1024 chi_r 1024 d 2 w 5/5 chains 1: bk 32 bf16x6 split 1 3M (512.4 us, default 564.9 us, 29 candidates)",
         "fp32 op00 step C chi_l 1024 chi_r 1024 d 2 w 5/5 chains 1: bk 32 bf16x6 split 1 3M (503.3 us, default 571.1 us, 29 candidates)",
         "fp64 op00 step A chi_l 1024 chi_r 1024 d 2 w 5/5 chains 1: bk 16 split 0 3M default (1048.7 us, default 1065.8 us, 25 candidates)",
         "fp64 op00 step C chi_l 1024 chi_r 1024 d 2 w 5/5 chains 1: bk 16 split 2 4M (1.0 us, default 1.1 us, 25 candidates)"]
    assert product_form(t, "fp32", 1024, 2) == "bf16x6"
    assert product_form(t, "fp64", 1024, 2) == "3M+4M"
    assert product_form(t[:3], "fp64", 1024, 2) == "3M"
    assert product_form([], "fp32", 1024, 2) == "3M"
