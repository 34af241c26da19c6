"""Learned gradient preconditioner for the L-BFGS inversion.

A 4-level 3D convolutional encoder-decoder (U-Net with additive skips,
1.31 M parameters) maps a shear-modulus search direction, sampled on the
element grid, to the shape of the model error it comes from.  It replaces the
scaled identity H0 = (y.s)/(y.y) of the coupled L-BFGS two-loop recursion
(LBFGS_Coupled.F90:95-104) for the mu field: r_mu = H0 P(q_mu), with P
normalised so that <P(q), q> = <q, q>.  The inversion keeps the plain step as
its fallback (the line search judges the preconditioned step first).

Everything numeric runs in the HIP kernels of ``csrc/hr_nn.hip`` (BF16
operands, FP32 accumulation on the matrix cores, FP32 activations, Adam on FP32
master weights): torch only provides device memory and the stream.  Training
pairs (raw mu gradient, true mu error) come from the FP64 solver on random
layered models with Gaussian inclusions.

The reference has no learned component; the interface point is the H0
scaling of LBFGS_Coupled.F90:95-104.
"""
from __future__ import annotations

import ctypes as C
import math
import time

import numpy as np

from . import native
from .mesh import BoxMesh, Model, vs_vp_rho_to_lame

# (name, cin, cout, level): encoder, bottleneck, decoder (additive skips), output
LAYERS = (
    ("e0a", 1, 16, 0), ("e0b", 16, 16, 0),
    ("e1a", 16, 32, 1), ("e1b", 32, 32, 1),
    ("e2a", 32, 64, 2), ("e2b", 64, 64, 2),
    ("ba", 64, 128, 3), ("bb", 128, 128, 3),
    ("d2a", 128, 64, 2), ("d2b", 64, 64, 2),
    ("d1a", 64, 32, 1), ("d1b", 32, 32, 1),
    ("d0a", 32, 16, 0), ("d0b", 16, 16, 0),
    ("out", 16, 1, 0),
)
LEVELS = 4
SLOPE = 0.1          # leaky ReLU


def n_params():
    return sum(ci * co * 27 + co for _, ci, co, _ in LAYERS)


# ----------------------------------------------------------------- C ABI
_F = C.c_float
_I = C.c_int
_LL = C.c_longlong
_P = C.c_void_p


def lib():
    """The HIP library with the hr_nn_* signatures set."""
    L = native.load()
    if getattr(L, "_nn_ready", False):
        return L
    L.hr_nn_conv3d.argtypes = [_P, _P, _P, _P, _I, _I, _I, _I, _I, _I, _I, _P]
    L.hr_nn_conv3d_wgrad.argtypes = [_P, _P, _P, _I, _P, _P, _I, _I, _I, _I, _I, _I, _P]
    L.hr_nn_transpose_weights.argtypes = [_P, _P, _I, _I, _P]
    L.hr_nn_leaky_relu.argtypes = [_P, _LL, _F, _P]
    L.hr_nn_leaky_relu_grad.argtypes = [_P, _P, _LL, _F, _P]
    L.hr_nn_avgpool2.argtypes = [_P, _P, _LL, _I, _I, _I, _P]
    L.hr_nn_avgpool2_grad.argtypes = [_P, _P, _LL, _I, _I, _I, _I, _P]
    L.hr_nn_upsample2.argtypes = [_P, _P, _LL, _I, _I, _I, _P]
    L.hr_nn_upsample2_grad.argtypes = [_P, _P, _LL, _I, _I, _I, _P]
    L.hr_nn_mse.argtypes = [_P, _P, _LL, _P, _P, _P, _P]
    L.hr_nn_adam.argtypes = [_P, _P, _P, _P, _P, _LL, _F, _F, _F, _F, _I, _P]
    L.hr_nn_to_bf16.argtypes = [_P, _P, _LL, _P]
    for name in native.NN_EXPORTS:
        getattr(L, name).restype = C.c_int
    L._nn_ready = True
    return L


def _check(rc, what):
    if rc != 0:
        raise native.HrError(f"{what} failed (code {rc})")


# ----------------------------------------------------------------- network
class ConvNet:
    """The encoder-decoder on one device.  Inputs are float32 volumes
    (NB, 1, D, H, W) with D, H, W divisible by 8; outputs have the same shape."""

    def __init__(self, seed=0, device="cuda:0"):
        import torch
        self.torch = torch
        self.dev = torch.device(device)
        self.L = lib()
        rng = np.random.default_rng(seed)
        self.woff, self.boff = [], []
        nw = nb = 0
        for _, ci, co, _ in LAYERS:
            self.woff.append(nw)
            self.boff.append(nb)
            nw += ci * co * 27
            nb += co
        w = np.empty(nw, np.float32)
        for li, (name, ci, co, _) in enumerate(LAYERS):
            # He initialisation for the leaky ReLU; the output layer starts small
            std = math.sqrt(2.0 / ((1.0 + SLOPE ** 2) * ci * 27)) * (0.1 if name == "out" else 1.0)
            w[self.woff[li]:self.woff[li] + ci * co * 27] = rng.normal(0.0, std, ci * co * 27)
        f32 = dict(dtype=torch.float32, device=self.dev)
        self.w = torch.as_tensor(w, **f32)
        self.b = torch.zeros(nb, **f32)
        self.wb = torch.empty(nw, dtype=torch.int16, device=self.dev)     # BF16 bits
        self.wt = torch.empty(nw, dtype=torch.int16, device=self.dev)     # transposed BF16 bits
        self.gw = torch.zeros(nw, **f32)
        self.gb = torch.zeros(nb, **f32)
        self.mw, self.vw = torch.zeros(nw, **f32), torch.zeros(nw, **f32)
        self.mb, self.vb = torch.zeros(nb, **f32), torch.zeros(nb, **f32)
        self.part = torch.zeros(256, **f32)
        self.loss_dev = torch.zeros(1, **f32)
        self.step = 0
        self._sync_weights()

    # ------------------------------------------------------------ helpers
    @property
    def stream(self):
        return self.torch.cuda.current_stream(self.dev).cuda_stream

    def _sync_weights(self):
        _check(self.L.hr_nn_to_bf16(self.w.data_ptr(), self.wb.data_ptr(), self.w.numel(), self.stream), "to_bf16")
        self._transpose()

    def _transpose(self):
        for li, (_, ci, co, _) in enumerate(LAYERS):
            o = 2 * self.woff[li]
            _check(self.L.hr_nn_transpose_weights(self.wb.data_ptr() + o, self.wt.data_ptr() + o, ci, co, self.stream),
                   "transpose_weights")

    def _empty(self, nb, c, dims):
        return self.torch.empty((nb, c) + tuple(dims), dtype=self.torch.float32, device=self.dev)

    def _conv(self, li, x, out, dims, accumulate=False):
        _, ci, co, _ = LAYERS[li]
        nb = x.shape[0]
        _check(self.L.hr_nn_conv3d(x.data_ptr(), self.wb.data_ptr() + 2 * self.woff[li],
                                   self.b.data_ptr() + 4 * self.boff[li], out.data_ptr(), ci, co, nb, *dims,
                                   int(accumulate), self.stream), f"conv {LAYERS[li][0]}")

    def _conv_dx(self, li, dy, dims):
        """Input gradient of layer li: the convolution with the mirrored, transposed weights."""
        _, ci, co, _ = LAYERS[li]
        nb = dy.shape[0]
        dx = self._empty(nb, ci, dims)
        _check(self.L.hr_nn_conv3d(dy.data_ptr(), self.wt.data_ptr() + 2 * self.woff[li], None, dx.data_ptr(), co, ci,
                                   nb, *dims, 0, self.stream), f"conv_dx {LAYERS[li][0]}")
        return dx

    def _wgrad(self, li, x, dy, dims):
        _, ci, co, _ = LAYERS[li]
        nb = x.shape[0]
        ncol = nb * dims[0] * dims[1] * dims[2]
        tiles = -(-ci * 27 // 64) * -(-co // 16)
        nslice = max(1, min(-(-1024 // tiles), ncol // 64))
        part = self.torch.empty(nslice * co * ci * 27, dtype=self.torch.float32, device=self.dev)
        _check(self.L.hr_nn_conv3d_wgrad(x.data_ptr(), dy.data_ptr(), part.data_ptr(), nslice,
                                         self.gw.data_ptr() + 4 * self.woff[li], self.gb.data_ptr() + 4 * self.boff[li],
                                         ci, co, nb, *dims, self.stream), f"wgrad {LAYERS[li][0]}")

    def _act(self, x):
        _check(self.L.hr_nn_leaky_relu(x.data_ptr(), x.numel(), SLOPE, self.stream), "leaky_relu")
        return x

    def _act_grad(self, y, dy):
        _check(self.L.hr_nn_leaky_relu_grad(y.data_ptr(), dy.data_ptr(), dy.numel(), SLOPE, self.stream),
               "leaky_relu_grad")
        return dy

    def _pool(self, x, dims):
        nb, c = x.shape[:2]
        out = self._empty(nb, c, [d // 2 for d in dims])
        _check(self.L.hr_nn_avgpool2(x.data_ptr(), out.data_ptr(), nb * c, *dims, self.stream), "avgpool2")
        return out

    def _pool_grad(self, dy, dx, dims):
        nb, c = dx.shape[:2]
        _check(self.L.hr_nn_avgpool2_grad(dy.data_ptr(), dx.data_ptr(), nb * c, *dims, 1, self.stream),
               "avgpool2_grad")

    def _up(self, x, dims):
        nb, c = x.shape[:2]
        out = self._empty(nb, c, dims)
        _check(self.L.hr_nn_upsample2(x.data_ptr(), out.data_ptr(), nb * c, *dims, self.stream), "upsample2")
        return out

    def _up_grad(self, dy, dims):
        nb, c = dy.shape[:2]
        dx = self._empty(nb, c, [d // 2 for d in dims])
        _check(self.L.hr_nn_upsample2_grad(dy.data_ptr(), dx.data_ptr(), nb * c, *dims, self.stream),
               "upsample2_grad")
        return dx

    # ------------------------------------------------------------ passes
    def forward(self, x, keep=False):
        """y = net(x); keep: retain the activations for :meth:`backward`."""
        if x.dtype != self.torch.float32 or not x.is_contiguous() or x.dim() != 5 or x.shape[1] != 1:
            raise ValueError("input must be a contiguous float32 tensor (NB, 1, D, H, W)")
        dims0 = list(x.shape[2:])
        if any(d % (1 << (LEVELS - 1)) for d in dims0):
            raise ValueError(f"volume {dims0} is not divisible by {1 << (LEVELS - 1)}")
        nb = x.shape[0]
        dims = [[d >> lev for d in dims0] for lev in range(LEVELS)]
        a = {"x": x}
        enc_in = x
        for lev in range(LEVELS):            # encoder and bottleneck: layers 2 lev, 2 lev + 1
            la, lb = 2 * lev, 2 * lev + 1
            h = self._empty(nb, LAYERS[la][2], dims[lev])
            self._conv(la, enc_in, h, dims[lev])
            self._act(h)
            e = self._empty(nb, LAYERS[lb][2], dims[lev])
            self._conv(lb, h, e, dims[lev])
            self._act(e)
            a[f"h{lev}"], a[f"e{lev}"] = h, e
            if lev < LEVELS - 1:
                enc_in = self._pool(e, dims[lev])
                a[f"p{lev + 1}"] = enc_in
        dec = a[f"e{LEVELS - 1}"]
        for lev in range(LEVELS - 2, -1, -1):   # decoder: layers 8 + 2 (2 - lev), +1
            la = 2 * LEVELS + 2 * (LEVELS - 2 - lev)
            lb = la + 1
            u = self._up(dec, dims[lev])
            z = a[f"e{lev}"].clone()           # skip: conv(u) + e
            self._conv(la, u, z, dims[lev], accumulate=True)
            self._act(z)
            d = self._empty(nb, LAYERS[lb][2], dims[lev])
            self._conv(lb, z, d, dims[lev])
            self._act(d)
            a[f"u{lev}"], a[f"z{lev}"], a[f"d{lev}"] = u, z, d
            dec = d
        y = self._empty(nb, 1, dims[0])
        self._conv(len(LAYERS) - 1, dec, y, dims[0])
        if keep:
            self._acts, self._dims = a, dims
        return y

    def backward(self, dy):
        """Parameter gradients (into gw, gb) of the last kept forward for dL/dy."""
        a, dims = self._acts, self._dims
        lo = len(LAYERS) - 1
        self._wgrad(lo, a["d0"], dy, dims[0])
        g = self._conv_dx(lo, dy, dims[0])
        ge = {}
        for lev in range(LEVELS - 1):            # decoder, finest level first
            la = 2 * LEVELS + 2 * (LEVELS - 2 - lev)
            lb = la + 1
            self._act_grad(a[f"d{lev}"], g)
            self._wgrad(lb, a[f"z{lev}"], g, dims[lev])
            g = self._conv_dx(lb, g, dims[lev])
            self._act_grad(a[f"z{lev}"], g)
            self._wgrad(la, a[f"u{lev}"], g, dims[lev])
            ge[lev] = g                           # the skip's gradient into e_lev
            gu = self._conv_dx(la, g, dims[lev])
            g = self._up_grad(gu, dims[lev])      # gradient of the coarser decoder output
        # bottleneck output e3 receives g; then the encoder, coarsest level first
        for lev in range(LEVELS - 1, -1, -1):
            la, lb = 2 * lev, 2 * lev + 1
            if lev < LEVELS - 1:
                g = ge[lev]
                self._pool_grad(gp, g, dims[lev])   # + pooling path
            self._act_grad(a[f"e{lev}"], g)
            self._wgrad(lb, a[f"h{lev}"], g, dims[lev])
            g = self._conv_dx(lb, g, dims[lev])
            self._act_grad(a[f"h{lev}"], g)
            xin = a["x"] if lev == 0 else a[f"p{lev}"]
            self._wgrad(la, xin, g, dims[lev])
            if lev > 0:
                gp = self._conv_dx(la, g, dims[lev])
        self._acts = None

    def loss(self, y, t, grad=True):
        """mean (y - t)^2 (device scalar) and, when grad, dL/dy."""
        dy = self.torch.empty_like(y) if grad else None
        _check(self.L.hr_nn_mse(y.data_ptr(), t.data_ptr(), y.numel(), dy.data_ptr() if grad else None,
                                self.part.data_ptr(), self.loss_dev.data_ptr(), self.stream), "mse")
        return self.loss_dev, dy

    def adam(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.step += 1
        s = self.stream
        _check(self.L.hr_nn_adam(self.w.data_ptr(), self.gw.data_ptr(), self.mw.data_ptr(), self.vw.data_ptr(),
                                 self.wb.data_ptr(), self.w.numel(), lr, b1, b2, eps, self.step, s), "adam")
        _check(self.L.hr_nn_adam(self.b.data_ptr(), self.gb.data_ptr(), self.mb.data_ptr(), self.vb.data_ptr(), None,
                                 self.b.numel(), lr, b1, b2, eps, self.step, s), "adam")
        self._transpose()

    # ------------------------------------------------------------ state
    def state(self):
        return {"w": self.w.cpu().numpy(), "b": self.b.cpu().numpy()}

    def load_state(self, st):
        torch = self.torch
        w, b = np.asarray(st["w"], np.float32), np.asarray(st["b"], np.float32)
        if w.shape != tuple(self.w.shape) or b.shape != tuple(self.b.shape):
            raise ValueError("weights do not match the layer table")
        self.w.copy_(torch.as_tensor(w))
        self.b.copy_(torch.as_tensor(b))
        self._sync_weights()

    def save(self, path):
        np.savez(path, **self.state())

    def load(self, path):
        with np.load(path) as f:
            self.load_state({"w": f["w"], "b": f["b"]})

    def infer(self, vol):
        """Host volume (D, H, W) -> host volume, one sample."""
        torch = self.torch
        x = torch.as_tensor(np.ascontiguousarray(vol, np.float32)[None, None], device=self.dev)
        y = self.forward(x)
        return y[0, 0].cpu().numpy()


# ------------------------------------------------------------ element-grid maps
def nodal_to_voxels(mesh: BoxMesh, v):
    """Element-grid volume (nez, ney, nex): mean of each element's 27 lattice nodes."""
    a = np.asarray(v, np.float64).reshape(mesh.NZ, mesh.NY, mesh.NX)
    s = np.zeros((mesh.nez, mesh.ney, mesh.nex))
    for c in range(3):
        for b in range(3):
            for x in range(3):
                s += a[c:c + 2 * mesh.nez:2, b:b + 2 * mesh.ney:2, x:x + 2 * mesh.nex:2]
    return s / 27.0


def voxels_to_nodal(mesh: BoxMesh, vox):
    """Nodal field: each lattice node takes the mean of the elements that hold it."""
    vox = np.asarray(vox, np.float64).reshape(mesh.nez, mesh.ney, mesh.nex)
    acc = np.zeros((mesh.NZ, mesh.NY, mesh.NX))
    cnt = np.zeros_like(acc)
    for c in range(3):
        for b in range(3):
            for x in range(3):
                sl = (slice(c, c + 2 * mesh.nez, 2), slice(b, b + 2 * mesh.ney, 2), slice(x, x + 2 * mesh.nex, 2))
                acc[sl] += vox
                cnt[sl] += 1.0
    return (acc / cnt).reshape(-1)


def rd_voxel_mask(mesh: BoxMesh):
    """1 on regular-domain elements, 0 in the PML, as an element-grid volume."""
    return mesh.element_is_rd().reshape(mesh.nez, mesh.ney, mesh.nex).astype(np.float64)


def normalise(vol, mask):
    v = vol * mask
    s = math.sqrt(float((v * v).sum()) / max(float(mask.sum()), 1.0))
    return (v / s if s > 0 else v), s


# ------------------------------------------------------------ training data
def random_case(mesh: BoxMesh, rng, vp_over_vs=2.0):
    """(true, start) models: 2-4 horizontal layers (Vs increasing with depth,
    150-400 m/s, rho 1900-2100) and, in the true model only, 1-3 Gaussian Vs
    inclusions of +-15..30 % with radii of 2-5 element edges inside the
    regular domain."""
    xyz = mesh.xyz()
    p = xyz[mesh.material_map()]
    z = p[:, 2]
    depth = (mesh.nez - mesh.rd_lo[2]) * mesh.h[2]
    nl = int(rng.integers(2, 5))
    vs_l = np.sort(rng.uniform(150.0, 400.0, nl))
    rho_l = np.sort(rng.uniform(1900.0, 2100.0, nl))
    zi = -np.sort(rng.uniform(0.1, 0.9, nl - 1)) * depth        # descending
    vs = np.full(mesh.nnode, vs_l[-1])
    rho = np.full(mesh.nnode, rho_l[-1])
    for zc, v, r in zip(zi[::-1], vs_l[:-1][::-1], rho_l[:-1][::-1]):
        vs = np.where(z >= zc, v, vs)
        rho = np.where(z >= zc, r, rho)
    lo = np.array([mesh.xmin + mesh.rd_lo[0] * mesh.h[0], mesh.ymin + mesh.rd_lo[1] * mesh.h[1], -depth])
    hi = np.array([mesh.xmin + mesh.rd_hi[0] * mesh.h[0], mesh.ymin + mesh.rd_hi[1] * mesh.h[1], 0.0])
    fac = np.ones(mesh.nnode)
    for _ in range(int(rng.integers(1, 4))):
        amp = rng.uniform(0.15, 0.30) * rng.choice([-1.0, 1.0])
        rad = rng.uniform(2.0, 5.0) * mesh.h[0]
        c = lo + rng.uniform(0.2, 0.8, 3) * (hi - lo)
        fac = fac + amp * np.exp(-((p - c) ** 2).sum(axis=1) / rad ** 2)
    ez = np.arange(mesh.nez)
    zc_el = mesh.zmin + (ez + 0.5) * mesh.h[2]
    rho_z = np.full(mesh.nez, rho_l[-1])
    for zc, r in zip(zi[::-1], rho_l[:-1][::-1]):
        rho_z = np.where(zc_el >= zc, r, rho_z)
    rho_el = np.repeat(rho_z, mesh.nex * mesh.ney) * 1.0e-6

    def model(f):
        lam, mu, _ = vs_vp_rho_to_lame(vs * f, vp_over_vs * vs * f, rho)
        return Model(lam, mu, rho_el)

    return model(fac), model(np.ones(mesh.nnode))


def training_mesh(nex=32, ney=32, nez=16, npml=4):
    return BoxMesh(nex, ney, nez, npml=(npml,) * 5)


def training_survey(mesh: BoxMesh, nstep, dt=1.0e-4, fr=30.0, nshot_side=2, nrec_side=12):
    """Vertical point loads on a surface grid, receivers on a surface grid over the regular domain."""
    from . import fwi as F
    hx = 0.5 * (mesh.rd_hi[0] - mesh.rd_lo[0]) * mesh.h[0]
    hy = 0.5 * (mesh.rd_hi[1] - mesh.rd_lo[1]) * mesh.h[1]
    r = lambda n, h: np.round(np.linspace(-h, h, n) * 2.0) / 2.0
    rec = np.array([mesh.surface_node(x, y) for y in r(nrec_side, hy) for x in r(nrec_side, hx)], dtype=np.int64)
    src = np.array([mesh.surface_node(x, y) for y in r(nshot_side, 0.5 * hy) for x in r(nshot_side, 0.5 * hx)],
                   dtype=np.int64)
    return F.survey_of(dict(mesh=mesh, src_nodes=src, rec_nodes=rec, fr=fr, dt=dt, nstep=nstep))


def generate_pairs(n, seed=0, mesh=None, nstep=900, log=None):
    """n pairs (normalised raw mu gradient, normalised true mu error) on the
    element grid of ``mesh`` (float32 arrays (n, nez, ney, nex)).  The gradient
    is the inversion's first-iteration gradient (adjoint + control pass, 1/M_g)
    of the start model against traces of the true model, all shots batched."""
    from . import fwi as F
    from . import solver as S
    mesh = mesh or training_mesh()
    rng = np.random.default_rng(seed)
    sv = training_survey(mesh, nstep)
    ctx = S.Context(mesh)
    mask = rd_voxel_mask(mesh)
    X = np.zeros((n, mesh.nez, mesh.ney, mesh.nex), np.float32)
    T = np.zeros_like(X)
    t0 = time.perf_counter()
    try:
        for k in range(n):
            true, start = random_case(mesh, rng)
            prob = F.HipProblem(ctx, sv, true.rho)
            sv.dis_meas = prob.observe(true.lam, true.mu)
            prob.forward_misfit(start.lam, start.mu)
            _, g_m, _, _ = prob.gradient(1, start.lam, start.mu)
            X[k] = normalise(nodal_to_voxels(mesh, g_m), mask)[0]
            T[k] = normalise(nodal_to_voxels(mesh, start.mu - true.mu), mask)[0]
            if log and (k + 1) % max(1, n // 10) == 0:
                log(f"pairs: {k + 1}/{n} ({time.perf_counter() - t0:.1f} s)")
    finally:
        ctx.close()
    return X, T


def cosine(a, b):
    a, b = np.asarray(a, np.float64).ravel(), np.asarray(b, np.float64).ravel()
    d = np.linalg.norm(a) * np.linalg.norm(b)
    return float(a @ b / d) if d > 0 else 0.0


def train(net: ConvNet, X, T, epochs=40, batch=8, lr=1e-3, seed=0, log=None, log_every=1):
    """Adam on the mean squared error; returns the per-epoch mean loss."""
    torch = net.torch
    rng = np.random.default_rng(seed)
    Xd = torch.as_tensor(X[:, None], device=net.dev)
    Td = torch.as_tensor(T[:, None], device=net.dev)
    n = X.shape[0]
    hist = []
    for ep in range(epochs):
        perm = rng.permutation(n)
        tot = torch.zeros(1, dtype=torch.float32, device=net.dev)
        nb = 0
        for i in range(0, n - batch + 1, batch):
            idx = torch.as_tensor(perm[i:i + batch], device=net.dev)
            x, t = Xd[idx].contiguous(), Td[idx].contiguous()
            y = net.forward(x, keep=True)
            lossd, dy = net.loss(y, t)
            tot += lossd
            net.backward(dy)
            # cosine decay of the step size
            net.adam(lr * 0.5 * (1.0 + math.cos(math.pi * (ep + i / n) / epochs)))
            nb += 1
        hist.append(float(tot.item()) / max(nb, 1))
        if log and (ep == 0 or (ep + 1) % log_every == 0 or ep + 1 == epochs):
            log(f"train: epoch {ep + 1}/{epochs} loss {hist[-1]:.4f}")
    return hist


def evaluate(net: ConvNet, X, T):
    """Mean cosine similarity to the true error of the raw gradient and of the network output."""
    raw, learned = [], []
    for x, t in zip(X, T):
        raw.append(cosine(x, t))
        learned.append(cosine(net.infer(x), t))
    return float(np.mean(raw)), float(np.mean(learned))


# ------------------------------------------------------------ L-BFGS hook
class LearnedH0:
    """The map q_mu -> P(q_mu) used in place of the identity inside the scaled
    H0 of the L-BFGS recursion: element-grid sample, normalisation, network,
    back to the nodes (regular domain only), scaled so that <P(q), q> = <q, q>.
    Returns None (plain H0) when the result is not a descent-compatible
    direction (<P(q), q> <= 0) or the input is zero."""

    def __init__(self, net: ConvNet, mesh: BoxMesh, restriction):
        self.net, self.mesh = net, mesh
        self.restriction = np.asarray(restriction, np.float64)
        self.mask = rd_voxel_mask(mesh)
        self.calls = self.used = 0

    def __call__(self, q):
        self.calls += 1
        q = np.asarray(q, np.float64)
        x, s = normalise(nodal_to_voxels(self.mesh, q), self.mask)
        if s == 0.0:
            return None
        y = self.net.infer(x) * self.mask
        r = voxels_to_nodal(self.mesh, y) * self.restriction
        c = float(r @ q)
        if not (c > 0.0 and math.isfinite(c)):
            return None
        self.used += 1
        return r * (float(q @ q) / c)
