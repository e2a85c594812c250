"""CPU: the oracle's e4m3 rounding, tile scaling and FP8-tail model.

Pins, in order:
  * e4m3_round against torch.float8_e4m3fn (PyTorch's conversion to the OCP
    E4M3 format) on a million values over the whole finite range, and against
    the format's own code table: every finite code is a fixed point, every
    midpoint between neighbouring codes goes to the one with an even code;
  * the tile exponent rule: the largest e with max * 2^e <= 448;
  * gemm_fp8: exact on operands that e4m3 holds exactly; a planted truncation
    in place of round-to-nearest-even is caught;
  * expm_fp8tail_model: with min_degree above the degree it is the float64
    Paterson-Stockmeyer evaluation (oracle.expm_taylor's algorithm), and its
    error against oracle.expm is of the size the 4-bit significand predicts.
"""
import numpy as np
import pytest
import torch

import oracle


def _torch_e4m3(x):
    return torch.tensor(np.asarray(x, dtype=np.float64)).float().to(torch.float8_e4m3fn) \
        .to(torch.float64).numpy()


def test_e4m3_round_matches_torch_float8():
    rng = np.random.default_rng(0)
    x = rng.standard_normal(1_000_000) * np.exp(rng.uniform(-14, 6.5, 1_000_000))
    x = np.clip(x, -448.0, 448.0)
    assert np.array_equal(oracle.e4m3_round(x), _torch_e4m3(x))


def test_e4m3_code_table():
    codes = torch.arange(256, dtype=torch.uint8).view(torch.float8_e4m3fn).to(torch.float64).numpy()
    finite = np.sort(np.unique(codes[np.isfinite(codes)]))
    assert finite.max() == 448.0 and finite.min() == -448.0
    assert len(finite) == 253  # 254 finite codes, +0 and -0 equal
    assert np.array_equal(oracle.e4m3_round(finite), finite)
    pos = finite[finite >= 0]
    mids = (pos[:-1] + pos[1:]) / 2
    got = oracle.e4m3_round(mids)
    # ties to the neighbour whose code is even (its last fraction bit is 0)
    code_of = {float(v): i for i, v in enumerate(codes) if np.isfinite(v) and v >= 0 and i < 128}
    for m, lo, hi, g in zip(mids, pos[:-1], pos[1:], got):
        want = lo if code_of[float(lo)] % 2 == 0 else hi
        assert g == want, (m, lo, hi, g)
    # known answers of the format: 1 = 0x38, 0.5 = 0x30, 3 = 0x44, 2^-9 = 0x01
    assert [code_of[v] for v in (1.0, 0.5, 3.0, 2.0 ** -9, 448.0)] == [0x38, 0x30, 0x44, 0x01, 0x7e]
    with pytest.raises(ValueError):
        oracle.e4m3_round([449.0])


def test_tile_exponent_rule():
    for mx in (448.0, 449.0, 224.0, 224.0001, 1.0, 3.0, 1e-300, 7.5e5):
        e = oracle.fp8_tile_exponent(mx)
        assert np.ldexp(mx, e) <= 448.0 < np.ldexp(mx, e + 1)
    assert oracle.fp8_tile_exponent(0.0) == 0
    rng = np.random.default_rng(1)
    x = rng.standard_normal((300, 200)) * np.exp2(rng.integers(-20, 20, (300, 1)))
    q, ex = oracle.fp8_quantize(x)
    assert ex.shape == (3, 2)
    for ti in range(3):
        for tj in range(2):
            blk = x[ti * 128:(ti + 1) * 128, tj * 128:(tj + 1) * 128]
            mx = np.abs(np.ldexp(blk, ex[ti, tj])).max()
            assert 224.0 < mx <= 448.0
    # 3 fraction bits: relative rounding at most 2^-4 for normal values
    big = np.abs(x) > np.ldexp(2.0 ** -6, -ex.repeat(128, 0).repeat(128, 1)[:300, :200])
    assert np.all(np.abs(q - x)[big] <= 2.0 ** -4 * np.abs(x)[big])


def test_gemm_fp8_exact_on_representable_operands():
    rng = np.random.default_rng(2)
    a = rng.integers(-15, 16, (130, 260)).astype(np.float64) * 2.0 ** -7
    b = rng.integers(-15, 16, (260, 70)).astype(np.float64) * 2.0 ** 9
    assert np.array_equal(oracle.gemm_fp8(a, b), a @ b)


def test_gemm_fp8_catches_truncation():
    """a truncating quantizer (a plausible mistake) moves the product far beyond
    the FP32 summation bound the GPU test allows"""
    rng = np.random.default_rng(3)
    a = rng.standard_normal((128, 256))
    b = rng.standard_normal((256, 128))
    good = oracle.gemm_fp8(a, b)

    def trunc_q(x):
        e = oracle.fp8_tile_exponent(np.abs(x).max())
        y = np.ldexp(x, e)
        _, ex = np.frexp(y)
        q = np.ldexp(1.0, np.maximum(ex - 1, -6) - 3)
        return np.ldexp(np.trunc(y / q) * q, -e)

    bad = np.hstack([trunc_q(a[:, :128]), trunc_q(a[:, 128:])]) @ \
        np.vstack([trunc_q(b[:128]), trunc_q(b[128:])])
    bound = 2 * (128 + 2) * 2.0 ** -24 * (np.abs(a) @ np.abs(b))
    assert np.abs(bad - good).max() > 100 * bound.max()


@pytest.mark.parametrize("norm,degree,s", [(1.0, 15, 0), (16.0, 15, 4), (0.3, 8, 0)])
def test_fp8tail_model_reduces_to_float64_ps(norm, degree, s):
    rng = np.random.default_rng(4)
    n = 40
    a = rng.standard_normal((n, n))
    a *= norm / np.abs(a).sum(0).max()
    exact_ps = oracle.expm_fp8tail_model(a, degree, s, min_degree=degree + 1)
    coeffs = oracle.inv_factorials()[:degree + 1]
    want = oracle.paterson_stockmeyer(coeffs, np.ldexp(a, -s))
    for _ in range(s):
        want = want @ want
    assert np.linalg.norm(exact_ps - want) <= 1e-13 * np.linalg.norm(want)


def test_fp8tail_model_error_size():
    """degree-4-and-up terms through e4m3: the error sits between 2^-12 and 2^-4
    of the size of those terms, far above float64 and far below O(1)"""
    rng = np.random.default_rng(5)
    n = 128
    a = rng.standard_normal((n, n)) / np.sqrt(n)
    a *= 1.0 / np.abs(a).sum(0).max()
    ref = oracle.expm(a)
    model = oracle.expm_fp8tail_model(a, 15, 0)
    err = np.linalg.norm(model - ref) / np.linalg.norm(ref)
    tail = sum(np.linalg.matrix_power(a, k) / np.prod(np.arange(1, k + 1), dtype=float)
               for k in range(4, 16))
    rel_tail = np.linalg.norm(tail) / np.linalg.norm(ref)
    assert 2.0 ** -12 * rel_tail < err < 2.0 ** -4 * rel_tail
