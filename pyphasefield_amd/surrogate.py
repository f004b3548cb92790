"""Learned surrogate for the N-component engine's Gibbs-energy evaluation.

Two 3-64-64-3 ReLU networks (one per phase) map (T, c1, c2) to the
non-ideal part of the Gibbs energy and its derivatives along e_l - e_N; the
ideal-mixing term R T sum c ln c stays analytic in the kernel (its log
singularity at c -> 0 is what an MLP fits worst, and it cancels in the
driving force).  The solid network's first output is G_S - G_L directly, so
the driving force is not the small difference of two large predictions.

Everything runs on the GPU through libppf_hip.so: training samples come from
the plain evaluation (``ppf_thermo_eval``), forward and backward passes are
MFMA GEMMs with FP32 accumulation (``ppf_mlp_gemm``: FP16 weights and
activation pairs forward, as inference runs them; BF16 backward) plus the
project's elementwise and Adam kernels; inference is the fused FP16-MFMA
kernel ``ppf_ncomp_thermo_mlp``, which falls back to the plain FP64
evaluation for cells outside the trained range.  torch only provides device
memory.
"""

import numpy as np

from . import _lib
from . import _lib_ext

HIDDEN = 64


def _f2bf_bits(a):
    """FP32 -> BF16 bits, round to nearest even (as the device does)."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def _f2h_bits(a):
    """FP32 -> FP16 bits, round to nearest even (as the device does)."""
    return np.ascontiguousarray(np.asarray(a, np.float64).astype(np.float32).astype(np.float16)).view(np.uint16)


def _slot_rows(h, p):
    """Hidden row held by slot (u, t) of K block kb = 2h + p: an array [2, 8]."""
    u = np.arange(2)[:, None]
    t = np.arange(8)[None, :]
    return 32 * h + 16 * p + 8 * (t // 4) + 4 * u + t % 4


def activation_scales(W1, b1, W2, b2, x_bound=1.001, margin=1.01):
    """Powers of two s1, s2 at least the interval upper bounds of the layer-1
    and layer-2 pre-activations over inputs |x_i| <= x_bound (the trained box
    is normalised to [-1, 1]; the bound leaves room for FP32 rounding of the
    normalisation and the margin for FP16 rounding of the weights).  With
    layer 1 scaled by 1/s1 and layer 2 by 1/s2, every hidden value of an
    accepted cell lies in [0, 1], where ReLU equals a clamp to [0, 1]."""
    W1, b1, W2, b2 = (np.asarray(a, np.float64) for a in (W1, b1, W2, b2))
    u1 = np.maximum(b1 + x_bound * np.abs(W1).sum(axis=1), 0.0)   # a1_j in [0, u1_j]
    u2 = np.maximum(b2 + np.maximum(W2, 0.0) @ u1, 0.0)          # z2_i <= u2_i
    pow2 = lambda u: float(2.0 ** np.ceil(np.log2(max(margin * float(u.max()), 2.0 ** -10))))
    return pow2(u1), pow2(u2)


def pack_phase(W1, W2, W3, b1, b2, b3):
    """FP16 A-operand lane layout and FP32 biases of ``ppf_ncomp_thermo_mlp``
    for one network: (words [14*64, 8] FP16 bits, bias [132] FP32).

    32x32x16 MFMA A operand: lane l, element t holds A[l % 32][slot (l // 32, t)];
    slot (u, t) of K block kb = 2h + p is hidden row 32h + 16p + 8(t//4) + 4u + t%4
    (the row the previous layer's D tile leaves in that lane and element).
    Sets of 64 lanes x 8: layer 1 (2 hidden halves; slots (0, 0..5) carry the
    inputs T, c1, c2, T, c1, c2 for the hi and lo parts, slots (0, 6..7) the
    layer-1 bias as an FP16 hi + lo pair against two constant-1 inputs, so the
    bias needs no accumulator tile), layer 2 (2 output halves x 4 K blocks),
    layer 3 (4 K blocks as 4x4x4-MFMA A operands: lane l, element t holds the
    weight of output row l % 4, zero for row 3, from the hidden row of slot
    (l // 32, t)).  Scaled by the powers
    of two of ``activation_scales``: layer 1 by 1/s1, layer 2's weights by
    s1/s2 and its bias by 1/s2, layer 3's weights by s2 (the network's outputs
    are unchanged; its hidden values become clamp-able to [0, 1]).
    bias: b1 (unused by the kernel), b2 / s2, b3, 0.
    """
    s1, s2 = activation_scales(W1, b1, W2, b2)
    W1 = np.asarray(W1, np.float64) / s1
    W2 = np.asarray(W2, np.float64) * (s1 / s2)
    W3 = np.asarray(W3, np.float64) * s2
    lanes = np.arange(64)
    r = lanes % 32
    u = lanes // 32
    sets = []
    col_of_t = [0, 1, 2, 0, 1, 2]
    b1f = (np.asarray(b1, np.float64) / s1).astype(np.float32).astype(np.float64)
    b1_hi = _f16(b1f)
    b1_lo = b1f - b1_hi  # rounded to FP16 with the rest of the set
    for h in range(2):
        blk = np.zeros((64, 8), np.float64)
        for t, col in enumerate(col_of_t):
            blk[u == 0, t] = W1[32 * h + r[u == 0], col]
        blk[u == 0, 6] = b1_hi[32 * h + r[u == 0]]
        blk[u == 0, 7] = b1_lo[32 * h + r[u == 0]]
        sets.append(blk)
    for m in range(2):
        for kb in range(4):
            rows = _slot_rows(kb // 2, kb % 2)      # [2, 8]
            sets.append(W2[32 * m + r[:, None], rows[u]])  # [64, 8]
    # layer 3 as the A operands of the 16-block 4x4x4 MFMA: lane l supplies
    # output row l % 4 (row 3 zero) for the hidden rows its block's B lanes
    # hold, those of half l // 32
    orow = lanes % 4
    for kb in range(4):
        rows = _slot_rows(kb // 2, kb % 2)
        blk = np.zeros((64, 8), np.float64)
        ok = orow < 3
        blk[ok] = W3[orow[ok][:, None], rows[u[ok]]]
        sets.append(blk)
    bias = np.zeros(132, np.float32)
    bias[:64] = b1
    bias[64:128] = np.asarray(b2, np.float64) / s2
    bias[128:131] = np.asarray(b3, np.float64)[:3]
    return _f2h_bits(np.concatenate(sets, axis=0)), bias


class PhaseNet:
    """FP32 master parameters of one 3-64-64-3 network, flat on the device."""

    SHAPES = [("W1", (HIDDEN, 3)), ("b1", (HIDDEN,)), ("W2", (HIDDEN, HIDDEN)), ("b2", (HIDDEN,)),
              ("W3", (3, HIDDEN)), ("b3", (3,))]

    def __init__(self, device, rng):
        import torch
        n = sum(int(np.prod(s)) for _, s in self.SHAPES)
        host = np.zeros(n, np.float32)
        off = 0
        for name, shp in self.SHAPES:
            size = int(np.prod(shp))
            if name.startswith("W"):
                fan_in = shp[1]
                host[off:off + size] = rng.normal(0, np.sqrt(2.0 / fan_in), size)
            off += size
        self.flat = torch.from_numpy(host).to(device)
        self.grad = torch.zeros_like(self.flat)
        self.m = torch.zeros_like(self.flat)
        self.v = torch.zeros_like(self.flat)
        self.view = {}
        self.gview = {}
        off = 0
        for name, shp in self.SHAPES:
            size = int(np.prod(shp))
            self.view[name] = self.flat[off:off + size].view(*shp)
            self.gview[name] = self.grad[off:off + size].view(*shp)
            off += size

    def numpy(self):
        return {k: v.detach().cpu().numpy().astype(np.float64) for k, v in self.view.items()}


class SurrogateThermo:
    """Trainable surrogate of a ThermoModel (3 components) on a (T, c) box."""

    def __init__(self, model, T_range, c_lo=0.02, seed=0, device="cuda", split2=False, split3=False):
        import torch
        if len(model.components) != 3:
            raise ValueError("the surrogate kernel is built for 3 components")
        self.model = model
        self.T_lo, self.T_hi = float(T_range[0]), float(T_range[1])
        self.c_lo = float(c_lo)
        self.device = torch.device(device)
        self.rng = np.random.default_rng(seed)
        self.nets = [PhaseNet(self.device, self.rng), PhaseNet(self.device, self.rng)]  # liquid, solid
        lo = np.array([self.T_lo, self.c_lo, self.c_lo])
        hi = np.array([self.T_hi, 1 - 2 * self.c_lo, 1 - 2 * self.c_lo])
        self.in_lo = lo
        self.in_scale = 2.0 / (hi - lo)
        self.out_mu = np.zeros((2, 3))
        self.out_sigma = np.ones((2, 3))
        self.history = []
        self._packed = None
        # inference: the inputs of layers 2 / 3 as FP16 hi + lo pairs (True)
        # or FP16 only (default: FP16 activations add a few J/mol of rounding
        # scatter and halve the MFMA and conversion work; the training GEMMs
        # always use pairs)
        self.split2 = bool(split2)
        self.split3 = bool(split3)

    # --------------------------------------------------------- samples --
    def sample_inputs(self, n):
        T = self.rng.uniform(self.T_lo, self.T_hi, n)
        c = self.c_lo + (1 - 3 * self.c_lo) * self.rng.dirichlet([1.0, 1.0, 1.0], n)
        return T, c[:, :2].copy()

    def plain_targets(self, T, c):
        """Plain FP64 evaluation on the GPU (ppf_thermo_eval), converted to the
        networks' targets: liquid (G_ni, dG_ni), solid (G_S - G_L, dG_S,ni)."""
        import torch
        L = _lib.lib()
        n = len(T)
        Td = torch.from_numpy(np.ascontiguousarray(T, np.float64)).to(self.device)
        cd = torch.from_numpy(np.ascontiguousarray(c, np.float64)).to(self.device)
        out = torch.empty(n * 6, dtype=torch.float64, device=self.device)
        _lib.check(L.ppf_thermo_eval(self.model.to_ctypes(), n, _lib.ptr(Td), _lib.ptr(cd), _lib.ptr(out),
                                     _lib.stream_handle()), "thermo_eval")
        o = out.cpu().numpy().reshape(n, 6)
        R = self.model.liquid.R
        cs = [c[:, 0], c[:, 1], 1 - c[:, 0] - c[:, 1]]
        lnc = [np.log(np.maximum(x, 1e-10)) for x in cs]
        Gid = R * T * sum(x * l for x, l in zip(cs, lnc))
        did = [R * T * (lnc[l] - lnc[2]) for l in range(2)]
        yL = np.stack([o[:, 0] - Gid, o[:, 1] - did[0], o[:, 2] - did[1]], axis=1)
        yS = np.stack([o[:, 3] - o[:, 0], o[:, 4] - did[0], o[:, 5] - did[1]], axis=1)
        return yL, yS

    def normalise_inputs(self, T, c):
        x = np.stack([T, c[:, 0], c[:, 1]], axis=0)
        return ((x - self.in_lo[:, None]) * self.in_scale[:, None] - 1.0).astype(np.float32)

    # ---------------------------------------------------- forward/back --
    def _gemm(self, M, N, K, A, sai, sak, B, sbk, sbj, C, alpha=1.0, beta=0.0, ksplit=1, split_b=1):
        L = _lib.lib()
        if ksplit > 1:
            C.zero_()
        _lib.check(L.ppf_mlp_gemm(M, N, K, _lib.ptr(A), sai, sak, _lib.ptr(B), sbk, sbj, _lib.ptr(C), C.stride(0),
                                  alpha, beta, ksplit, split_b, _lib.stream_handle()), "mlp_gemm")

    def forward(self, net, X, keep=None):
        """X: device [3 x B] normalised inputs -> Y [3 x B] (normalised outputs)."""
        import torch
        L = _lib.lib()
        s = _lib.stream_handle()
        B = X.shape[1]
        v = net.view
        Z1 = torch.empty(HIDDEN, B, device=self.device)
        # forward products in the inference kernel's formats: FP16 weights,
        # FP16 hi + lo activations (training sees the weights it will run with)
        self._gemm(HIDDEN, B, 3, v["W1"], 3, 1, X, B, 1, Z1, split_b=2)
        H1 = torch.empty_like(Z1)
        _lib.check(L.ppf_mlp_bias_act(_lib.ptr(Z1), _lib.ptr(v["b1"]), _lib.ptr(H1), HIDDEN, B, 1, s), "bias_act")
        Z2 = torch.empty_like(Z1)
        self._gemm(HIDDEN, B, HIDDEN, v["W2"], HIDDEN, 1, H1, B, 1, Z2, split_b=2)
        H2 = torch.empty_like(Z1)
        _lib.check(L.ppf_mlp_bias_act(_lib.ptr(Z2), _lib.ptr(v["b2"]), _lib.ptr(H2), HIDDEN, B, 1, s), "bias_act")
        Z3 = torch.empty(3, B, device=self.device)
        self._gemm(3, B, HIDDEN, v["W3"], HIDDEN, 1, H2, B, 1, Z3, split_b=2)
        Y = torch.empty_like(Z3)
        _lib.check(L.ppf_mlp_bias_act(_lib.ptr(Z3), _lib.ptr(v["b3"]), _lib.ptr(Y), 3, B, 0, s), "bias_act")
        if keep is not None:
            keep.update(X=X, H1=H1, H2=H2)
        return Y

    def backward(self, net, dY, keep):
        """Gradients of the loss w.r.t. the parameters into net.grad."""
        import torch
        L = _lib.lib()
        s = _lib.stream_handle()
        X, H1, H2 = keep["X"], keep["H1"], keep["H2"]
        B = X.shape[1]
        v, gv = net.view, net.gview
        ks = max(1, B // 512)
        self._gemm(3, HIDDEN, B, dY, B, 1, H2, 1, B, gv["W3"], ksplit=ks)
        _lib.check(L.ppf_mlp_rowsum(_lib.ptr(dY), _lib.ptr(gv["b3"]), 3, B, s), "rowsum")
        dH2 = torch.empty(HIDDEN, B, device=self.device)
        self._gemm(HIDDEN, B, 3, v["W3"], 1, HIDDEN, dY, B, 1, dH2)
        dZ2 = torch.empty_like(dH2)
        _lib.check(L.ppf_mlp_relu_bwd(_lib.ptr(dH2), _lib.ptr(H2), _lib.ptr(dZ2), HIDDEN * B, s), "relu_bwd")
        self._gemm(HIDDEN, HIDDEN, B, dZ2, B, 1, H1, 1, B, gv["W2"], ksplit=ks)
        _lib.check(L.ppf_mlp_rowsum(_lib.ptr(dZ2), _lib.ptr(gv["b2"]), HIDDEN, B, s), "rowsum")
        dH1 = torch.empty_like(dH2)
        self._gemm(HIDDEN, B, HIDDEN, v["W2"], 1, HIDDEN, dZ2, B, 1, dH1)
        dZ1 = torch.empty_like(dH2)
        _lib.check(L.ppf_mlp_relu_bwd(_lib.ptr(dH1), _lib.ptr(H1), _lib.ptr(dZ1), HIDDEN * B, s), "relu_bwd")
        self._gemm(HIDDEN, 3, B, dZ1, B, 1, X, 1, B, gv["W1"], ksplit=ks)
        _lib.check(L.ppf_mlp_rowsum(_lib.ptr(dZ1), _lib.ptr(gv["b1"]), HIDDEN, B, s), "rowsum")

    def loss_and_grad(self, Y, T):
        import torch
        L = _lib.lib()
        n = Y.numel()
        dY = torch.empty_like(Y)
        part = torch.empty(int(L.ppf_mlp_mse_partials(n)), device=self.device)
        _lib.check(L.ppf_mlp_mse(_lib.ptr(Y), _lib.ptr(T), _lib.ptr(dY), _lib.ptr(part), n, _lib.stream_handle()),
                   "mse")
        return part, dY

    # ------------------------------------------------------------ train --
    def train(self, n_samples=1_000_000, epochs=20, batch=1024, lr=2e-3, lr_final=5e-5, verbose=False):
        """Adam on normalised MSE; the learning rate decays geometrically."""
        import torch
        L = _lib.lib()
        T, c = self.sample_inputs(n_samples)
        yL, yS = self.plain_targets(T, c)
        for p, y in enumerate((yL, yS)):
            self.out_mu[p] = y.mean(axis=0)
            self.out_sigma[p] = y.std(axis=0) + 1e-12
        X = torch.from_numpy(self.normalise_inputs(T, c)).to(self.device)
        Ys = [torch.from_numpy(((y - self.out_mu[p]) / self.out_sigma[p]).T.astype(np.float32).copy()).to(self.device)
              for p, y in enumerate((yL, yS))]
        steps_per_epoch = n_samples // batch
        total = epochs * steps_per_epoch
        gamma = (lr_final / lr) ** (1.0 / max(1, total - 1))
        s = _lib.stream_handle()
        for p, net in enumerate(self.nets):
            t = 0
            for ep in range(epochs):
                perm = torch.from_numpy(self.rng.permutation(n_samples)).to(self.device)
                losses = []
                for b in range(steps_per_epoch):
                    idx = perm[b * batch:(b + 1) * batch]
                    Xb = X[:, idx].contiguous()
                    Tb = Ys[p][:, idx].contiguous()
                    keep = {}
                    Y = self.forward(net, Xb, keep)
                    part, dY = self.loss_and_grad(Y, Tb)
                    self.backward(net, dY, keep)
                    t += 1
                    rate = lr * gamma ** (t - 1)
                    _lib.check(L.ppf_mlp_adam(_lib.ptr(net.flat), _lib.ptr(net.grad), _lib.ptr(net.m),
                                              _lib.ptr(net.v), net.flat.numel(), rate, 0.9, 0.999, 1e-8, t, s),
                               "adam")
                    if b % 32 == 0:
                        losses.append(part)
                ep_loss = float(torch.stack([q.sum() for q in losses]).mean()) / (3 * batch)
                self.history.append((p, ep, ep_loss))
                if verbose:
                    print("phase %d epoch %d loss %.3e" % (p, ep, ep_loss), flush=True)
        self._packed = None
        return self.history

    # -------------------------------------------------------- evaluate --
    def predict(self, T, c):
        """Physical outputs (liquid, solid) of the trained networks (GPU forward)."""
        import torch
        X = torch.from_numpy(self.normalise_inputs(T, c)).to(self.device)
        out = []
        for p, net in enumerate(self.nets):
            Y = self.forward(net, X).cpu().numpy().T.astype(np.float64)
            out.append(Y * self.out_sigma[p] + self.out_mu[p])
        return out

    # ------------------------------------------------------ save / load --
    def save(self, path):
        """Trained parameters and normalisation to an .npz file (no pickles)."""
        np.savez(path, T_range=np.array([self.T_lo, self.T_hi]), c_lo=np.array(self.c_lo), in_lo=self.in_lo,
                 in_scale=self.in_scale, out_mu=self.out_mu, out_sigma=self.out_sigma,
                 liquid=self.nets[0].flat.cpu().numpy(), solid=self.nets[1].flat.cpu().numpy(),
                 components=np.array(list(self.model.components)))

    @classmethod
    def load(cls, path, model, device="cuda"):
        """Surrogate of ``model`` from :meth:`save` output; the components must match."""
        import torch
        z = np.load(path, allow_pickle=False)
        if [str(c) for c in z["components"]] != list(model.components):
            raise ValueError("surrogate %s was trained for components %s, not %s"
                             % (path, list(z["components"]), list(model.components)))
        sur = cls(model, tuple(z["T_range"]), c_lo=float(z["c_lo"]), device=device)
        sur.in_lo, sur.in_scale = z["in_lo"], z["in_scale"]
        sur.out_mu, sur.out_sigma = z["out_mu"], z["out_sigma"]
        for net, key in zip(sur.nets, ("liquid", "solid")):
            flat = z[key]
            if flat.shape != tuple(net.flat.shape):
                raise ValueError("surrogate %s: %s network has %d parameters, expected %d"
                                 % (path, key, flat.size, net.flat.numel()))
            net.flat.copy_(torch.from_numpy(flat.astype(np.float32)))
        return sur

    # ------------------------------------------------------- inference --
    def packed(self):
        """Device buffers + ppf_mlp_net descriptor for ppf_ncomp_thermo_mlp."""
        import torch
        if self._packed is not None:
            return self._packed
        net = _lib_ext.MlpNet()
        keep = []
        for p, pn in enumerate(self.nets):
            h = pn.numpy()
            words, bias = pack_phase(h["W1"], h["W2"], h["W3"], h["b1"], h["b2"], h["b3"])
            wdev = torch.from_numpy(words.view(np.int16).copy()).to(self.device)
            bdev = torch.from_numpy(bias).to(self.device)
            keep += [wdev, bdev]
            net.weights[p] = wdev.data_ptr()
            net.bias[p] = bdev.data_ptr()
            for q in range(3):
                net.out_mu[p][q] = float(self.out_mu[p][q])
                net.out_sigma[p][q] = float(self.out_sigma[p][q])
        for q in range(3):
            net.in_lo[q] = float(self.in_lo[q])
            net.in_scale[q] = float(self.in_scale[q])
        net.T_lo, net.T_hi, net.c_lo = self.T_lo, self.T_hi, self.c_lo
        # per-step-parity counts of cells handed to the FP64 fallback
        marks = torch.zeros(2, dtype=torch.int32, device=self.device)
        keep.append(marks)
        net.marks = marks.data_ptr()
        net.flags = ((_lib_ext.PPF_MLP_SPLIT2 if self.split2 else 0) |
                     (_lib_ext.PPF_MLP_SPLIT3 if self.split3 else 0))
        self._packed = (net, keep)
        return self._packed


def _bf16(a):
    bits = _f2bf_bits(np.asarray(a, np.float64).astype(np.float32)).astype(np.uint32) << 16
    return bits.view(np.float32).astype(np.float64)


def _f16(a):
    return np.asarray(a, np.float64).astype(np.float32).astype(np.float16).astype(np.float64)


def _split(a):
    """the BF16 pair hi + lo the device feeds for an FP32 operand"""
    a32 = np.asarray(a, np.float64).astype(np.float32).astype(np.float64)
    hi = _bf16(a32)
    return hi + _bf16(a32 - hi)


def _split16(a):
    """the FP16 pair hi + lo the device feeds for an FP32 operand"""
    a32 = np.asarray(a, np.float64).astype(np.float32).astype(np.float64)
    hi = _f16(a32)
    return hi + _f16(a32 - hi)


def numpy_forward_backward(params, X, Yt, rounded=False):
    """FP64 numpy MLP (checker for the GPU gradients); rounded=True reproduces
    the device operand rounding: forward products with FP16 weights and FP16
    hi + lo inputs, backward products with the A operand (a weight or the
    left factor of a gradient product) in BF16 and the B operand as a BF16
    hi + lo pair."""
    fa = _f16 if rounded else (lambda a: a)
    fb = _split16 if rounded else (lambda a: a)
    ra = _bf16 if rounded else (lambda a: a)
    rb = _split if rounded else (lambda a: a)

    W1, b1, W2, b2, W3, b3 = (params[k] for k in ("W1", "b1", "W2", "b2", "W3", "b3"))
    Z1 = fa(W1) @ fb(X) + b1[:, None]
    H1 = np.maximum(Z1, 0)
    Z2 = fa(W2) @ fb(H1) + b2[:, None]
    H2 = np.maximum(Z2, 0)
    Y = fa(W3) @ fb(H2) + b3[:, None]
    n = Y.size
    dY = 2 * (Y - Yt) / n
    g = {"W3": ra(dY) @ rb(H2).T, "b3": dY.sum(1)}
    dH2 = ra(W3).T @ rb(dY)
    dZ2 = dH2 * (H2 > 0)
    g["W2"] = ra(dZ2) @ rb(H1).T
    g["b2"] = dZ2.sum(1)
    dH1 = ra(W2).T @ rb(dZ2)
    dZ1 = dH1 * (H1 > 0)
    g["W1"] = ra(dZ1) @ rb(X).T
    g["b1"] = dZ1.sum(1)
    return Y, float(np.sum((Y - Yt) ** 2) / n), g
