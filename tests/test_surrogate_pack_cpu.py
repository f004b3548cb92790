"""Host packing of the surrogate networks for the fused MFMA inference
kernel, checked by emulating the kernel's operand flow in numpy.

The emulation follows the documented v_mfma_f32_32x32x16_bf16 lane layout:
A lane l element t = A[l % 32][slot (l // 32, t)], B lane l element t =
B[slot (l // 32, t)][l % 32], D lane l item v = D[8 (v // 4) + 4 (l // 32) +
v % 4][l % 32]; the kernel turns items 8p..8p+7 of a D tile into the next B
operand.  Reference: the plain numpy forward pass with BF16 weights and BF16
hi + lo activations."""

import numpy as np

from pyphasefield_amd.surrogate import _bf16, _f2bf_bits, _split, pack_phase


def _bits_to_f(bits):
    return (bits.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def _mfma(A, B, D):
    """A, B: [64, 8] lane operands; D: [64, 16] lane accumulators."""
    Am = np.zeros((32, 2, 8))
    Bm = np.zeros((2, 8, 32))
    for l in range(64):
        Am[l % 32, l // 32] = A[l]
        Bm[l // 32, :, l % 32] = B[l]
    full = np.einsum("iut,utj->ij", Am, Bm)
    out = D.copy()
    for l in range(64):
        for v in range(16):
            out[l, v] += full[8 * (v // 4) + 4 * (l // 32) + v % 4, l % 32]
    return out


def _bias_tile(b, row0):
    D = np.zeros((64, 16))
    for l in range(64):
        for v in range(16):
            D[l, v] = b[row0 + 8 * (v // 4) + 4 * (l // 32) + v % 4]
    return D


def _relu_split(D, p):
    v = np.maximum(D[:, 8 * p:8 * p + 8], 0).astype(np.float32).astype(np.float64)  # FP32 accumulators
    hi = _bf16(v)
    return hi, _bf16(v - hi)


def _emulate(words, b1, b2, b3, X):
    """words: [14*64, 8] BF16 bits; X: [3, 32] normalised inputs of 32 cells."""
    W = _bits_to_f(words).reshape(14, 64, 8)
    xin = np.zeros((64, 8))
    xs = X.astype(np.float32).astype(np.float64)
    hi = _bf16(xs)
    lo = _bf16(xs - hi)
    xin[:32, 0:3] = hi.T
    xin[:32, 3:6] = lo.T
    xin[:32, 6:8] = 1.0  # constant inputs of the packed layer-1 bias (hi, lo)
    h = [_mfma(W[hh], xin, np.zeros((64, 16))) for hh in range(2)]
    packs = [_relu_split(h[kb // 2], kb % 2) for kb in range(4)]
    z = []
    for m in range(2):
        acc = _bias_tile(b2, 32 * m)
        for kb in range(4):
            acc = _mfma(W[2 + 4 * m + kb], packs[kb][0], acc)
            acc = _mfma(W[2 + 4 * m + kb], packs[kb][1], acc)
        z.append(acc)
    packs = [_relu_split(z[kb // 2], kb % 2) for kb in range(4)]
    y = np.zeros((64, 16))
    y[:32, 0:3] = b3[:3]
    for kb in range(4):
        y = _mfma(W[10 + kb], packs[kb][0], y)
        y = _mfma(W[10 + kb], packs[kb][1], y)
    return y[:32, 0:3].T


def test_packed_layout_reproduces_forward():
    rng = np.random.default_rng(0)
    W1 = rng.normal(0, 0.8, (64, 3))
    W2 = rng.normal(0, 0.2, (64, 64))
    W3 = rng.normal(0, 0.2, (3, 64))
    b1, b2, b3 = rng.normal(0, 0.1, 64), rng.normal(0, 0.1, 64), rng.normal(0, 0.1, 4)
    X = rng.uniform(-1, 1, (3, 32))
    words = pack_phase(W1, W2, W3, b1)
    assert words.shape == (14 * 64, 8)
    got = _emulate(words, b1, b2, b3, X)
    # the layer-1 bias travels as a BF16 hi + lo pair
    H1 = np.maximum(_bf16(W1) @ _split(X) + _split(b1)[:, None], 0)
    H2 = np.maximum(_bf16(W2) @ _split(H1) + b2[:, None], 0)
    want = _bf16(W3) @ _split(H2) + b3[:3, None]
    # equal up to FP32 rounding of the accumulators and summation order
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-6 * np.abs(want).max())
    # a transposed hidden layer is caught
    bad = _emulate(pack_phase(W1, W2.T, W3, b1), b1, b2, b3, X)
    assert np.abs(bad - want).max() > 1e-3
    # so is a layer-1 bias left out of the packed words
    nob = _emulate(pack_phase(W1, W2, W3, np.zeros(64)), b1, b2, b3, X)
    assert np.abs(nob - want).max() > 1e-3


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 3.0e-3], np.float32)
    got = _bits_to_f(_f2bf_bits(x))
    # ties go to the even mantissa; 1 + 2^-8 is halfway between 1 and 1 + 2^-7
    assert got[0] == 1.0 and got[1] == 1.0 and got[2] == 1.0 + 2 ** -6
    assert got[3] == -2.5
    assert abs(got[4] - 3.0e-3) <= 3.0e-3 * 2 ** -8
