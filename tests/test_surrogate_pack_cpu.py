"""Host packing of the surrogate networks for the fused MFMA inference
kernel, checked by emulating the kernel's operand flow in numpy.

The emulation follows the documented v_mfma_f32_32x32x16_f16 lane layout:
A lane l element t = A[l % 32][slot (l // 32, t)], B lane l element t =
B[slot (l // 32, t)][l % 32], D lane l item v = D[8 (v // 4) + 4 (l // 32) +
v % 4][l % 32], and for layer 3 that of the 16-block 4x4x4 MFMA (lane 4b + j,
item r += sum_k A[lane 4b + r][k] B[lane 4b + j][k]); the kernel turns items
8p..8p+7 of a D tile into the next B operand, either clamped to [0, 1] and
rounded to FP16 (default) or as the FP16 pair hi + lo of the ReLU.  Reference: the plain numpy forward pass of
the unscaled network with FP16 weights."""

import numpy as np

from pyphasefield_amd.surrogate import _f16, _f2h_bits, _split16, activation_scales, pack_phase


def _bits_to_f(bits):
    return bits.astype(np.uint16).view(np.float16).astype(np.float64)


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


def _mfma4(A, B, D):
    """16-block 4x4x4 MFMA on [64, 4] lane operands and [64, 4] accumulators:
    lane 4b + j, item r += sum_k A[4b + r][k] B[4b + j][k] (the layout
    measured on gfx950)."""
    out = D.copy()
    for l in range(64):
        b0 = l - l % 4
        for r in range(4):
            out[l, r] += float(np.dot(A[b0 + r], B[l]))
    return out


def _bias_tile(b, row0):
    D = np.zeros((64, 16))
    for l in range(64):
        for v in range(16):
            D[l, v] = b[row0 + 8 * (v // 4) + 4 * (l // 32) + v % 4]
    return D


def _act(D, p, split):
    """next-layer B operand(s) from items 8p..8p+7 of FP32 accumulators"""
    v = D[:, 8 * p:8 * p + 8].astype(np.float32).astype(np.float64)
    if split:
        r = np.maximum(v, 0)
        hi = _f16(r)
        return [hi, _f16(r - hi)]
    return [_f16(np.clip(v, 0.0, 1.0))]


def _emulate(words, bias, X, split=False):
    """words: [14*64, 8] FP16 bits; bias: [132] FP32; X: [3, 32] normalised
    inputs of 32 cells.  Returns the three outputs and the largest scaled
    hidden pre-activation seen."""
    W = _bits_to_f(words).reshape(14, 64, 8)
    b2, b3 = bias[64:128].astype(np.float64), bias[128:131].astype(np.float64)
    xin = np.zeros((64, 8))
    xs = X.astype(np.float32).astype(np.float64)
    hi = _f16(xs)
    lo = _f16(xs - hi)
    xin[:32, 0:3] = hi.T
    xin[:32, 3:6] = lo.T
    xin[:32, 6:8] = 1.0  # constant inputs of the packed layer-1 bias (hi, lo)
    h = [_mfma(W[hh], xin, np.zeros((64, 16))) for hh in range(2)]
    peak = max(float(np.max(x)) for x in h)
    packs = [_act(h[kb // 2], kb % 2, split) for kb in range(4)]
    z = []
    for m in range(2):
        acc = _bias_tile(b2, 32 * m)
        for kb in range(4):
            for op in packs[kb]:
                acc = _mfma(W[2 + 4 * m + kb], op, acc)
        z.append(acc)
    peak = max(peak, max(float(np.max(x)) for x in z))
    packs = [_act(z[kb // 2], kb % 2, split) for kb in range(4)]
    y = np.zeros((64, 4))
    for kb in range(4):
        for op in packs[kb]:
            for e in range(2):
                y = _mfma4(W[10 + kb][:, 4 * e:4 * e + 4], op[:, 4 * e:4 * e + 4], y)
    # the two lane halves hold the two halves of the hidden rows of each cell
    out = y[:32, 0:3] + y[32:, 0:3] + b3
    return out.T, peak


def _net(seed=0):
    rng = np.random.default_rng(seed)
    W1 = rng.normal(0, 0.8, (64, 3))
    W2 = rng.normal(0, 0.2, (64, 64))
    W3 = rng.normal(0, 0.2, (3, 64))
    b1, b2, b3 = rng.normal(0, 0.1, 64), rng.normal(0, 0.1, 64), rng.normal(0, 0.1, 4)
    return W1, W2, W3, b1, b2, b3


def _reference(W1, W2, W3, b1, b2, b3, X, split):
    """unscaled network with FP16 weights (scaling by powers of two commutes
    with FP16 rounding) and the kernel's activation rounding"""
    H1 = np.maximum(_f16(W1) @ _split16(X) + _split16(b1)[:, None], 0)
    H1 = _split16(H1) if split else _f16(H1)
    H2 = np.maximum(_f16(W2) @ H1 + b2[:, None], 0)
    H2 = _split16(H2) if split else _f16(H2)
    return _f16(W3) @ H2 + b3[:3, None]


def test_packed_layout_reproduces_forward():
    W1, W2, W3, b1, b2, b3 = _net()
    X = np.random.default_rng(1).uniform(-1, 1, (3, 32))
    words, bias = pack_phase(W1, W2, W3, b1, b2, b3)
    assert words.shape == (14 * 64, 8) and bias.shape == (132,)
    for split in (False, True):
        got, peak = _emulate(words, bias, X, split)
        assert peak <= 1.0  # every hidden value inside the clamp range
        want = _reference(W1, W2, W3, b1, b2, b3, X, split)
        if split:
            # pairs: equal up to FP32 rounding of the accumulators and order
            np.testing.assert_allclose(got, want, rtol=0, atol=1e-6 * np.abs(want).max())
        else:
            # FP16 activations: scaling by powers of two keeps the FP16 grid
            # relative, so scaled and unscaled roundings differ only where a
            # scaled activation falls into FP16's subnormal range
            np.testing.assert_allclose(got, want, rtol=0, atol=1e-4 * np.abs(want).max())
            # and against exact arithmetic: FP16 rounding of weights and activations
            exact = W3 @ np.maximum(W2 @ np.maximum(W1 @ X + b1[:, None], 0) + b2[:, None], 0) + b3[:3, None]
            np.testing.assert_allclose(got, exact, rtol=0, atol=2e-3 * np.abs(exact).max())
    # a transposed hidden layer is caught
    bad, _ = _emulate(*pack_phase(W1, W2.T, W3, b1, b2, b3), X)
    want = _reference(W1, W2, W3, b1, b2, b3, X, False)
    assert np.abs(bad - want).max() > 1e-2 * np.abs(want).max()
    # so is a layer-1 bias left out of the packed words
    nob, _ = _emulate(*pack_phase(W1, W2, W3, np.zeros(64), b2, b3), X)
    assert np.abs(nob - want).max() > 1e-2 * np.abs(want).max()


def test_activation_scales_bound_the_box():
    """Layer 1 is linear in x, so its largest pre-activation over the box
    [-1, 1]^3 sits at a corner: the bound is tight there.  Layer 2's bound
    holds for every sampled input; the scales are powers of two."""
    W1, W2, W3, b1, b2, b3 = _net(2)
    s1, s2 = activation_scales(W1, b1, W2, b2, x_bound=1.0, margin=1.0)
    assert np.log2(s1) == np.round(np.log2(s1)) and np.log2(s2) == np.round(np.log2(s2))
    corners = np.array(np.meshgrid([-1, 1], [-1, 1], [-1, 1])).reshape(3, -1)
    z1_max = (W1 @ corners + b1[:, None]).max()
    assert s1 / 2 < z1_max <= s1
    X = np.concatenate([corners, np.random.default_rng(3).uniform(-1, 1, (3, 20000))], axis=1)
    H1 = np.maximum(W1 @ X + b1[:, None], 0)
    z2_max = (W2 @ H1 + b2[:, None]).max()
    assert z2_max <= s2
    # a slip that drops the absolute value of the weights is caught
    slip = np.maximum(b1 + W1.sum(axis=1), 0).max()
    assert slip < z1_max


def test_scaled_network_equals_unscaled_in_exact_arithmetic():
    """ReLU(z) = s clamp(z / s, 0, 1) whenever z <= s, so the scaled layers
    compute the same function (checked without any FP16 rounding)."""
    W1, W2, W3, b1, b2, b3 = _net(4)
    s1, s2 = activation_scales(W1, b1, W2, b2)
    X = np.random.default_rng(5).uniform(-1, 1, (3, 5000))
    A1 = np.clip((W1 / s1) @ X + (b1 / s1)[:, None], 0, 1)
    A2 = np.clip((W2 * s1 / s2) @ A1 + (b2 / s2)[:, None], 0, 1)
    Y = (W3 * s2) @ A2 + b3[:3, None]
    want = W3 @ np.maximum(W2 @ np.maximum(W1 @ X + b1[:, None], 0) + b2[:, None], 0) + b3[:3, None]
    np.testing.assert_allclose(Y, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())
    # too small a layer-2 scale clips and is caught
    A2c = np.clip((W2 * s1 / (s2 / 8)) @ A1 + (b2 / (s2 / 8))[:, None], 0, 1)
    assert np.abs((W3 * s2 / 8) @ A2c + b3[:3, None] - want).max() > 1e-3 * np.abs(want).max()


def test_fp16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -11, 1.0 + 3 * 2 ** -11, -2.5, 3.0e-3, 1e-6], np.float32)
    got = _bits_to_f(_f2h_bits(x))
    # ties go to the even mantissa; 1 + 2^-11 is halfway between 1 and 1 + 2^-10
    assert got[0] == 1.0 and got[1] == 1.0 and got[2] == 1.0 + 2 ** -9
    assert got[3] == -2.5
    assert abs(got[4] - 3.0e-3) <= 3.0e-3 * 2 ** -11
    # subnormal range: absolute spacing 2^-24
    assert abs(got[5] - 1e-6) <= 2 ** -25
