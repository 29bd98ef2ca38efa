"""The closure perceptron's CPU restatement (oracle/pll_closure_oracle.py),
pinned without the GPU: BF16 rounding against hand-worked values, and the
back propagation against central finite differences of the loss in FP64."""
import numpy as np

from oracle import pll_closure_oracle as mo


def test_bf16_rounding_known_values():
    one = np.float32(1.0)
    ulp = 2.0 ** -7  # BF16 spacing at 1 (8 significant bits)
    cases = {
        1.0: 1.0,
        1.0 + ulp: 1.0 + ulp,
        1.0 + ulp / 2: 1.0,               # tie -> even (mantissa ...0)
        1.0 + 3 * ulp / 2: 1.0 + 2 * ulp,  # tie -> even (mantissa ...10)
        1.0 + ulp / 2 + 2.0 ** -20: 1.0 + ulp,
        -2.5: -2.5,
        3.0e38: 226 / 128 * 2.0 ** 127,  # 3e38 = 1.7633 * 2^127 -> 226/128
    }
    for x, want in cases.items():
        got = float(mo.bf16(np.array([x], dtype=np.float32))[0])
        assert abs(got - want) <= 1e-3 * abs(want), (x, got, want)
    assert np.isinf(mo.bf16(np.array([np.inf], dtype=np.float32))[0])
    assert float(mo.bf16(np.array([one], dtype=np.float32))[0]) == 1.0


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = rng.normal(0, 0.3, mo.N_PARAM)
    p[mo.O_W1:mo.O_B1].reshape(16, 64)[10:] = 0.0
    x = rng.normal(0, 1, (37, 10))
    y = np.abs(rng.normal(1, 0.5, 37))
    g, _ = mo.gradient(p, x, y, bf16_round=False)
    # every parameter block, a few entries each
    idx = list(rng.choice(np.arange(mo.O_W1, mo.O_W1 + 10 * 64), 6)) + \
        list(rng.choice(np.arange(mo.O_B1, mo.O_W2), 4)) + \
        list(rng.choice(np.arange(mo.O_W2, mo.O_B2), 6)) + \
        list(rng.choice(np.arange(mo.O_B2, mo.O_W3), 4)) + \
        list(rng.choice(np.arange(mo.O_W3, mo.O_B3), 4)) + [mo.O_B3]
    h = 1e-6
    for i in idx:
        q = p.copy()
        q[i] += h
        lp = mo.loss(q, x, y, bf16_round=False)
        q[i] -= 2 * h
        lm = mo.loss(q, x, y, bf16_round=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])
    # padded input rows carry no gradient
    assert not g[mo.O_W1 + 10 * 64:mo.O_B1].any()


def test_bf16_forward_close_to_exact():
    rng = np.random.default_rng(4)
    p = rng.normal(0, 0.3, mo.N_PARAM)
    x = rng.normal(0, 1, (200, 10))
    a = mo.predict(p, x, bf16_round=True)
    b = mo.predict(p, x, bf16_round=False)
    assert np.max(np.abs(a - b)) < 0.05 * np.max(np.abs(b))
