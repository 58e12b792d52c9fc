"""CPU checks of arithmetic shortcuts the march kernels take, restated in
numpy float32 (the kernels' own arithmetic):

* the chunk's sample positions t0 + i * step (one fma) equal the sequential
  float additions t += step whenever the kernel's exactness test
  (sum_chain_exact in render.hip) passes;
* the 8-lane sum by quad permutes and a half-row mirror equals the
  xor-shuffle tree 1, 2, 4 bit for bit.
"""
import numpy as np


def _bits(x: np.float32) -> int:
    return int(np.array([x], np.float32).view(np.uint32)[0])


def _step_ulp_max(step: np.float32) -> int:
    # step = m * 2^(e_step - 150): a whole multiple of 2^(e - 150) exactly
    # when e <= e_step + trailing zeros of the 24-bit significand m.
    b = _bits(step)
    m = (b & 0x7FFFFF) | 0x800000
    return ((b >> 23) & 0xFF) + ((m & -m).bit_length() - 1)


def _sum_chain_exact(t0: np.float32, step: np.float32) -> bool:
    # The kernel's integer form: exponent field of t0 in 1..253 (positive
    # normal), step a multiple of t0's ulp, the rounded end point in t0's
    # binade.
    e = _bits(t0) >> 23
    t_end = np.float32(np.float64(31) * np.float64(step) + np.float64(t0))  # fma: one rounding
    return (0 < e < 254) and e <= _step_ulp_max(step) and (_bits(t_end) >> 23) == e


def _step_is_ulp_multiple(t0: np.float32, step: np.float32) -> bool:
    # Direct restatement: step / ulp(t0) is an integer (ulp = 2^(e - 150)).
    e = (_bits(t0) >> 23) & 0xFF
    q = np.float64(step) / np.ldexp(1.0, e - 150)
    return bool(q == np.floor(q))


def test_chunk_positions_exact_shortcut():
    g = np.random.default_rng(0)
    n_fast = 0
    steps = [0.5, 0.25, 1.0, 0.125, 0.1, 0.3, 0.75, 2.0]
    for _ in range(20000):
        step = np.float32(steps[g.integers(len(steps))])
        t0 = np.float32(g.uniform(0, 1200) if g.random() < 0.8 else g.uniform(0, 4))
        if not _sum_chain_exact(t0, step):
            continue
        n_fast += 1
        t = t0
        for i in range(32):
            fma = np.float32(np.float64(i) * np.float64(step) + np.float64(t0))
            assert t == fma, (t0, step, i)
            t = np.float32(t + step)
    assert n_fast > 10000  # the shortcut applies to most chunks at step 0.5


def test_step_ulp_max_matches_direct_division():
    g = np.random.default_rng(2)
    for step in [0.5, 0.25, 1.0, 0.125, 0.1, 0.3, 0.75, 2.0, 1e-3, 3.0, 0.7]:
        step = np.float32(step)
        for _ in range(300):
            t0 = np.float32(g.uniform(1e-3, 5000))
            e = _bits(t0) >> 23
            assert (e <= _step_ulp_max(step)) == _step_is_ulp_multiple(t0, step), (t0, step)


def test_shortcut_rejects_rounding_chains():
    # Crossing a power of two, or a step with bits below the ulp, must not pass.
    assert not _sum_chain_exact(np.float32(511.9), np.float32(0.5))
    assert not _sum_chain_exact(np.float32(3.0), np.float32(0.1))
    assert not _sum_chain_exact(np.float32(0.0), np.float32(0.5))
    assert not _sum_chain_exact(np.float32(-3.0), np.float32(0.5))
    assert not _sum_chain_exact(np.float32(np.inf), np.float32(0.5))
    assert not _sum_chain_exact(np.float32(np.nan), np.float32(0.5))
    assert _sum_chain_exact(np.float32(300.25), np.float32(0.5))


def test_dpp_sum8_equals_xor_tree():
    g = np.random.default_rng(1)
    v = g.normal(size=(1000, 8)).astype(np.float32) * np.float32(1e3) ** g.integers(
        -2, 3, size=(1000, 1)).astype(np.float32)
    idx = np.arange(8)

    def xor_tree(x):
        for m in (1, 2, 4):
            x = (x + x[:, idx ^ m]).astype(np.float32)
        return x

    def dpp(x):
        x = (x + x[:, idx ^ 1]).astype(np.float32)    # quad_perm [1,0,3,2]
        x = (x + x[:, idx ^ 2]).astype(np.float32)    # quad_perm [2,3,0,1]
        x = (x + x[:, 7 - idx]).astype(np.float32)    # row_half_mirror
        return x

    assert np.array_equal(xor_tree(v), dpp(v))
