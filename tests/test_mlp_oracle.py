"""Pin the MLP oracle: its FP64 backward pass against central finite
differences of its own loss, and its BF16 rounding against exact cases."""
import numpy as np

from oracle.mlp_oracle import KIN, KOUT, NIN, NOUT, MlpOracle, param_count, to_bf16


def random_params(rng, scale=0.3):
    p = []
    for i, k in zip(KIN, KOUT):
        w = np.zeros((i, k))
        ri = NIN if i == 16 else i
        rk = NOUT if k == 64 else k
        w[:ri, :rk] = rng.normal(0, scale / np.sqrt(ri), (ri, rk))
        b = np.zeros(k)
        b[:rk] = rng.normal(0, 0.1, rk)
        p.append(np.concatenate([w.ravel(), b]))
    return np.concatenate(p)


def test_bf16_rounding_exact_cases():
    # representable values are unchanged; halfway cases round to even
    assert to_bf16(1.0) == 1.0 and to_bf16(-2.5) == -2.5
    one_ulp = 2.0 ** -7  # bf16 spacing at 1.0
    assert to_bf16(1.0 + one_ulp / 2) == 1.0            # tie -> even (1.0)
    assert to_bf16(1.0 + 3 * one_ulp / 2) == 1.0 + 2 * one_ulp  # tie -> even
    assert to_bf16(1.0 + 0.75 * one_ulp) == 1.0 + one_ulp


def test_fp64_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = random_params(rng)
    x = rng.normal(0, 1, (7, NIN))
    t = rng.normal(0, 1, (7, NOUT))
    loss, g = MlpOracle(p, bf16=False).gradients(x, t)
    assert g.shape == (param_count(),)
    # check a spread of entries in every layer (padding entries have zero gradient)
    idx, o = [], 0
    for i, k in zip(KIN, KOUT):
        ri = NIN if i == 16 else i
        rk = NOUT if k == 64 else k
        for _ in range(6):
            a, b = rng.integers(ri), rng.integers(rk)
            idx.append(o + a * k + b)
        idx.append(o + i * k + rng.integers(rk))  # a bias
        pad = o + (i - 1) * k + (k - 1)           # a padding weight when padded
        if i == 16 or k == 64:
            assert g[pad] == 0.0
        o += i * k + k
    h = 1e-6
    for j in idx:
        pp, pm = p.copy(), p.copy()
        pp[j] += h
        pm[j] -= h
        lp, _ = MlpOracle(pp, bf16=False).gradients(x, t)
        lm, _ = MlpOracle(pm, bf16=False).gradients(x, t)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[j]) <= 1e-6 * max(1.0, abs(fd)), (j, fd, g[j])


def test_silu_and_derivative_point_values():
    """silu(z) = z sigmoid(z) and its derivative at points with known values:
    silu(0) = 0, silu'(0) = 1/2; odd-part identity silu(z) - silu(-z) = z;
    silu'(z) + silu'(-z) = 1; ln 3 gives sigmoid = 3/4 exactly; large |z|
    tends to z (resp. 0) with derivative 1 (resp. 0)."""
    from oracle.mlp_oracle import dsilu, silu
    assert silu(0.0) == 0.0 and dsilu(0.0) == 0.5
    z = np.log(3.0)  # sigmoid(ln 3) = 3/4
    assert abs(silu(z) - 0.75 * z) <= 1e-15
    assert abs(dsilu(z) - 0.75 * (1.0 + z * 0.25)) <= 1e-15
    zs = np.linspace(-6.0, 6.0, 49)
    np.testing.assert_allclose(silu(zs) - silu(-zs), zs, rtol=0, atol=1e-14)
    np.testing.assert_allclose(dsilu(zs) + dsilu(-zs), 1.0, rtol=0, atol=1e-14)
    assert abs(silu(40.0) - 40.0) <= 1e-12 and abs(silu(-40.0)) <= 1e-15
    assert abs(dsilu(40.0) - 1.0) <= 1e-15 and abs(dsilu(-40.0)) <= 1e-15
    # the derivative is that of silu: central differences at sample points
    h = 1e-6
    np.testing.assert_allclose(dsilu(zs), (silu(zs + h) - silu(zs - h)) / (2 * h), rtol=0, atol=1e-8)
    # the minimum of silu: silu'(z*) = 0 where 1 + z* + exp(z*) = 0, z* = -1 - W(1/e)
    zstar = -1.2784645427610738
    assert abs(dsilu(zstar)) <= 1e-14
