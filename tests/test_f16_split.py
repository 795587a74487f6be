"""The arithmetic of the fp16 plane form of the complex64 matvec, emulated on
the CPU: each matrix is multiplied by the power of two 2^(14 - e) (e the binary
exponent of its largest |re|, |im|, |re + im|, so every scaled value is at most
2^14 and far from the fp16 limit 65504), every scaled FP32 value is split into
two fp16 parts (hi the round-to-nearest fp16 of v, lo that of the exact FP32
residual), and a real product keeps the three partial products hi.hi, hi.lo,
lo.hi.  The split represents v to 2^-22 |v| where both parts are fp16
normal numbers (|v| >= 2^-3, i.e. within 2^17 of the largest value) and to
2^-25 absolute below; the three products give x y to a few units of
2^-23 |x y|: FP32 accuracy, as the bf16 form's six products do.  The GPU
kernel's own tests then check the matvec against the oracle."""

import numpy as np


def scale_for(m):
    """The planes' scale for the largest value m (as the device computes it)."""
    if m == 0:
        return 1.0
    e = np.frexp(np.float32(m))[1]  # m = f 2^e, f in [0.5, 1)
    return float(2.0 ** min(60, max(-60, 14 - int(e))))


def split2(v):
    v = np.asarray(v, dtype=np.float32)
    h = v.astype(np.float16)
    r = (v - h.astype(np.float32)).astype(np.float32)  # exact in FP32
    return h.astype(np.float64), r.astype(np.float16).astype(np.float64)


def test_scale_is_a_power_of_two_keeping_values_below_2_14():
    rng = np.random.default_rng(11)
    for _ in range(200):
        m = float(np.abs(rng.standard_normal()) * 10.0 ** rng.uniform(-12, 12))
        s = scale_for(m)
        assert np.frexp(s)[0] == 0.5  # a power of two
        assert m * s <= 2.0 ** 14
        assert m * s > 2.0 ** 13  # and the largest one at least half of it
    assert scale_for(0.0) == 1.0


def test_split_represents_scaled_values():
    rng = np.random.default_rng(3)
    v = rng.standard_normal(200_000) * np.exp2(rng.integers(-30, 0, 200_000))
    s = scale_for(np.abs(v).max())
    w = (v * s).astype(np.float32)
    h, lo = split2(w)
    w64 = w.astype(np.float64)
    err = np.abs(w64 - (h + lo))
    # both parts are fp16 normal numbers from |w| >= 2^-3 on (the residual is
    # about 2^-11 |w|): relative accuracy 2^-22 there, absolute 2^-25 (the
    # subnormal spacing of lo) below, i.e. 2^-39 of the largest value
    normal = np.abs(w64) >= 2.0 ** -3
    assert np.all(err[normal] <= 2.0 ** -22 * np.abs(w64[normal]))
    assert np.all(err[~normal] <= 2.0 ** -25)
    assert np.all(np.abs(h) <= 2.0 ** 14)


def test_three_partial_products_keep_fp32_accuracy():
    rng = np.random.default_rng(5)
    x = rng.standard_normal(100_000).astype(np.float32) * np.float32(2.0 ** 12)
    y = rng.standard_normal(100_000).astype(np.float32) * np.float32(2.0 ** 12)
    xh, xl = split2(x)
    yh, yl = split2(y)
    three = xl * yh + xh * yl + xh * yh
    exact = x.astype(np.float64) * y.astype(np.float64)
    big = np.abs(x) >= 2.0 ** -2
    big &= np.abs(y) >= 2.0 ** -2
    rel = np.abs(three - exact)[big] / np.abs(exact[big])
    assert rel.max() <= 2.0 ** -21
    # hi.hi alone is fp16 accuracy: the cross terms are needed
    assert (np.abs(xh * yh - exact)[big] / np.abs(exact[big])).max() > 2.0 ** -13


def test_gemm_with_scaled_fp16_parts_matches_fp32_accuracy():
    """A 3M complex product through the scaled two-part split, FP32 partial
    sums per K stage of 32, against the FP64 product: within a few 1e-7 of the
    product's norm, on operands spanning twelve decades."""
    rng = np.random.default_rng(9)
    n = 96
    a = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) * np.exp2(rng.integers(-20, 20, (n, 1)))
    b = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) * 1e-3

    def parts(m):
        s = scale_for(max(np.abs(m.real).max(), np.abs(m.imag).max(), np.abs(m.real + m.imag).max()))
        comps = [m.real * s, m.imag * s, (m.real + m.imag) * s]
        return [split2(c.astype(np.float32)) for c in comps], s

    pa, sa = parts(a)
    pb, sb = parts(b)
    t = []
    for (ah, al), (bh, bl) in zip(pa, pb):
        acc = np.zeros((n, n), np.float32)
        for k0 in range(0, n, 32):
            ks = slice(k0, k0 + 32)
            part = al[:, ks] @ bh[ks] + ah[:, ks] @ bl[ks] + ah[:, ks] @ bh[ks]
            acc = (acc + part.astype(np.float32)).astype(np.float32)
        t.append(acc.astype(np.float64))
    c = ((t[0] - t[1]) + 1j * (t[2] - t[0] - t[1])) / (sa * sb)
    exact = a @ b
    assert np.linalg.norm(c - exact) / np.linalg.norm(exact) < 1e-6
