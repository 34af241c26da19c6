"""HIP kernels of the learned preconditioner (hirensi_amd/csrc/hr_nn.hip,
include/hirensi_nn.h) against a plain PyTorch FP32 reference on the CPU.

The kernels round both matrix operands to BF16 (nearest even) and accumulate in
FP32, so the reference rounds the same operands (weights and the convolution
input; for the weight gradient also the output gradient) and computes in FP32:
the remaining difference is FP32 summation order, bounded here by 1e-5
relative to the largest output.  The whole network's forward is checked
operation by operation from its own intermediate activations at the same bar,
and end to end at 2e-2 relative L2 (an FP32 difference can move an activation
across a BF16 rounding boundary, and at the 4-voxel bottleneck that moves a
whole sample); its parameter gradients, where the kernels also round every
layer's output gradient to BF16 and autograd does not, at 3e-2 relative L2 per
layer.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
F = torch.nn.functional


def _L():
    from hirensi_amd import precond as PC
    return PC.lib()


def bf(x):
    """BF16 rounding (nearest even), kept in FP32."""
    return x.to(torch.bfloat16).to(torch.float32)


def bits(x):
    """BF16 bit patterns of an FP32 tensor, as int16."""
    return x.to(torch.bfloat16).view(torch.int16)


def dev(x):
    return x.contiguous().to("cuda:0")


def _stream():
    return torch.cuda.current_stream().cuda_stream


def _rel(a, b):
    return float((a - b).abs().max() / max(float(b.abs().max()), 1e-30))


CONV_CASES = [
    # cin, cout, NB, (D, H, W)
    (1, 16, 1, (8, 8, 8)),
    (16, 1, 2, (4, 6, 10)),
    (3, 5, 1, (5, 7, 3)),
    (32, 64, 2, (4, 4, 8)),
    (128, 64, 1, (2, 4, 4)),
]


@pytest.mark.parametrize("cin,cout,nb,dims", CONV_CASES)
def test_conv3d_forward_and_accumulate(cin, cout, nb, dims):
    print(f"conv3d forward cin={cin} cout={cout} nb={nb} dims={dims}", flush=True)
    g = torch.Generator().manual_seed(cin * 100 + cout)
    x = torch.randn((nb, cin) + dims, generator=g)
    w = torch.randn(cout, cin * 27, generator=g) / np.sqrt(cin * 27)
    b = torch.randn(cout, generator=g)
    ref = F.conv3d(bf(x), bf(w).view(cout, cin, 3, 3, 3), b, padding=1)
    L = _L()
    xd, wd, bd = dev(x), dev(bits(w)), dev(b)
    out = torch.zeros((nb, cout) + dims, device="cuda:0")
    assert L.hr_nn_conv3d(xd.data_ptr(), wd.data_ptr(), bd.data_ptr(), out.data_ptr(), cin, cout, nb, *dims, 0,
                          _stream()) == 0
    got = out.cpu()
    assert _rel(got, ref) < 1e-5
    # accumulate onto existing values, no bias
    base = torch.randn((nb, cout) + dims, generator=g)
    out2 = dev(base)
    assert L.hr_nn_conv3d(xd.data_ptr(), wd.data_ptr(), None, out2.data_ptr(), cin, cout, nb, *dims, 1,
                          _stream()) == 0
    ref2 = base + F.conv3d(bf(x), bf(w).view(cout, cin, 3, 3, 3), None, padding=1)
    assert _rel(out2.cpu(), ref2) < 1e-5


@pytest.mark.parametrize("cin,cout,nb,dims", CONV_CASES[:4])
def test_conv3d_weight_and_input_gradients(cin, cout, nb, dims):
    print(f"conv3d gradients cin={cin} cout={cout} nb={nb} dims={dims}", flush=True)
    g = torch.Generator().manual_seed(7 + cin + cout)
    x = torch.randn((nb, cin) + dims, generator=g)
    w = torch.randn(cout, cin * 27, generator=g) / np.sqrt(cin * 27)
    dy = torch.randn((nb, cout) + dims, generator=g)
    W5 = bf(w).view(cout, cin, 3, 3, 3)
    ref_dw = torch.nn.grad.conv3d_weight(bf(x), W5.shape, bf(dy), padding=1).reshape(cout, cin * 27)
    ref_db = dy.sum(dim=(0, 2, 3, 4))
    ref_dx = torch.nn.grad.conv3d_input(x.shape, W5, bf(dy), padding=1)
    L = _L()
    xd, dyd = dev(x), dev(dy)
    dw = torch.zeros(cout * cin * 27, device="cuda:0")
    db = torch.zeros(cout, device="cuda:0")
    for nslice in (1, 7):
        part = torch.empty(nslice * cout * cin * 27, device="cuda:0")
        assert L.hr_nn_conv3d_wgrad(xd.data_ptr(), dyd.data_ptr(), part.data_ptr(), nslice, dw.data_ptr(),
                                    db.data_ptr(), cin, cout, nb, *dims, _stream()) == 0
        assert _rel(dw.cpu().view(cout, cin * 27), ref_dw) < 1e-5
        assert _rel(db.cpu(), ref_db) < 1e-5
    # input gradient: hr_nn_conv3d of dy with the transposed, mirrored weights
    wt = torch.empty(cin * cout * 27, dtype=torch.int16, device="cuda:0")
    assert L.hr_nn_transpose_weights(dev(bits(w)).data_ptr(), wt.data_ptr(), cin, cout, _stream()) == 0
    dx = torch.zeros((nb, cin) + dims, device="cuda:0")
    assert L.hr_nn_conv3d(dyd.data_ptr(), wt.data_ptr(), None, dx.data_ptr(), cout, cin, nb, *dims, 0,
                          _stream()) == 0
    assert _rel(dx.cpu(), ref_dx) < 1e-5


def test_pool_upsample_activation_loss_and_adam():
    print("pool / upsample / leaky relu / mse / adam", flush=True)
    L = _L()
    s = _stream()
    g = torch.Generator().manual_seed(3)
    C_, dims = 6, (4, 6, 8)
    x = torch.randn((C_,) + dims, generator=g)
    xd = dev(x)
    p = torch.empty((C_,) + tuple(d // 2 for d in dims), device="cuda:0")
    assert L.hr_nn_avgpool2(xd.data_ptr(), p.data_ptr(), C_, *dims, s) == 0
    assert _rel(p.cpu(), F.avg_pool3d(x[None], 2)[0]) < 1e-6
    u = torch.empty((C_,) + dims, device="cuda:0")
    assert L.hr_nn_upsample2(p.data_ptr(), u.data_ptr(), C_, *dims, s) == 0
    assert torch.equal(u.cpu(), F.interpolate(p.cpu()[None], scale_factor=2, mode="nearest")[0])
    # backward of both against autograd
    xr = x.clone().requires_grad_(True)
    dy = torch.randn((C_,) + tuple(d // 2 for d in dims), generator=g)
    F.avg_pool3d(xr[None], 2)[0].backward(dy)
    gx = dev(torch.ones_like(x))              # accumulate onto ones
    assert L.hr_nn_avgpool2_grad(dev(dy).data_ptr(), gx.data_ptr(), C_, *dims, 1, s) == 0
    assert _rel(gx.cpu(), xr.grad + 1.0) < 1e-6
    pr = torch.randn(dy.shape, generator=g).requires_grad_(True)
    dU = torch.randn((C_,) + dims, generator=g)
    F.interpolate(pr[None], scale_factor=2, mode="nearest")[0].backward(dU)
    gp = torch.empty(dy.shape, device="cuda:0")
    assert L.hr_nn_upsample2_grad(dev(dU).data_ptr(), gp.data_ptr(), C_, *dims, s) == 0
    assert _rel(gp.cpu(), pr.grad) < 1e-6
    # leaky ReLU and its backward (exact)
    y = dev(x)
    assert L.hr_nn_leaky_relu(y.data_ptr(), y.numel(), 0.1, s) == 0
    assert torch.equal(y.cpu(), F.leaky_relu(x, 0.1))
    d = dev(torch.ones_like(x))
    assert L.hr_nn_leaky_relu_grad(y.data_ptr(), d.data_ptr(), d.numel(), 0.1, s) == 0
    assert torch.equal(d.cpu(), torch.where(x > 0, 1.0, 0.1).to(torch.float32))
    # mean squared error and its gradient
    t = torch.randn(x.shape, generator=g)
    grad = torch.empty_like(xd)
    part = torch.empty(256, device="cuda:0")
    loss = torch.empty(1, device="cuda:0")
    assert L.hr_nn_mse(xd.data_ptr(), dev(t).data_ptr(), x.numel(), grad.data_ptr(), part.data_ptr(),
                       loss.data_ptr(), s) == 0
    assert abs(float(loss.item()) - float(((x - t) ** 2).mean())) < 1e-5 * float(((x - t) ** 2).mean())
    assert _rel(grad.cpu(), 2.0 * (x - t) / x.numel()) < 1e-6
    # two Adam steps against the update rule in float64
    n = 1000
    w0 = torch.randn(n, generator=g)
    gs = [torch.randn(n, generator=g) for _ in range(2)]
    w, m, v = dev(w0), torch.zeros(n, device="cuda:0"), torch.zeros(n, device="cuda:0")
    wb = torch.empty(n, dtype=torch.int16, device="cuda:0")
    wr, mr, vr = w0.double(), torch.zeros(n, dtype=torch.float64), torch.zeros(n, dtype=torch.float64)
    for k, gk in enumerate(gs, start=1):
        assert L.hr_nn_adam(w.data_ptr(), dev(gk).data_ptr(), m.data_ptr(), v.data_ptr(), wb.data_ptr(), n, 1e-2, 0.9,
                            0.999, 1e-8, k, s) == 0
        gd = gk.double()
        mr = 0.9 * mr + 0.1 * gd
        vr = 0.999 * vr + 0.001 * gd * gd
        wr = wr - 1e-2 * (mr / (1 - 0.9 ** k)) / ((vr / (1 - 0.999 ** k)).sqrt() + 1e-8)
    assert float((w.cpu().double() - wr).abs().max()) < 1e-5
    assert torch.equal(wb.cpu(), bits(w.cpu()))
    # argument errors
    assert L.hr_nn_avgpool2(xd.data_ptr(), p.data_ptr(), C_, 3, 6, 8, s) != 0
    assert L.hr_nn_conv3d(None, None, None, None, 1, 1, 1, 1, 1, 1, 0, s) != 0


def _ref_net(net, x, round_inputs=True):
    """Torch CPU restatement of the layer graph with the network's BF16 weights."""
    from hirensi_amd import precond as PC
    w = bf(net.w.cpu())
    b = net.b.cpu()
    r = bf if round_inputs else (lambda t: t)

    def conv(li, t):
        _, ci, co, _ = PC.LAYERS[li]
        o = net.woff[li]
        return F.conv3d(r(t), w[o:o + ci * co * 27].view(co, ci, 3, 3, 3), b[net.boff[li]:net.boff[li] + co],
                        padding=1)

    act = lambda t: F.leaky_relu(t, PC.SLOPE)
    e, h = [], x
    for lev in range(PC.LEVELS):
        h = act(conv(2 * lev + 1, act(conv(2 * lev, h))))
        e.append(h)
        if lev < PC.LEVELS - 1:
            h = F.avg_pool3d(h, 2)
    d = e[-1]
    for lev in range(PC.LEVELS - 2, -1, -1):
        la = 2 * PC.LEVELS + 2 * (PC.LEVELS - 2 - lev)
        u = F.interpolate(d, scale_factor=2, mode="nearest")
        d = act(conv(la + 1, act(conv(la, u) + e[lev])))
    return conv(len(PC.LAYERS) - 1, d)


def test_network_forward_and_parameter_gradients():
    from hirensi_amd import precond as PC
    print("network forward/backward nb=2 dims=(8, 16, 16)", flush=True)
    net = PC.ConvNet(seed=5)
    # non-zero biases and a non-small output layer, so that every term matters
    g = torch.Generator().manual_seed(9)
    net.b.copy_(0.1 * torch.randn(net.b.shape, generator=g))
    o = net.woff[-1]
    net.w[o:].mul_(10.0)
    net._sync_weights()
    x = torch.randn((2, 1, 8, 16, 16), generator=g)
    t = torch.randn(x.shape, generator=g)
    y = net.forward(dev(x), keep=True)
    acts = {k: v.cpu() for k, v in net._acts.items()}
    # every operation of the layer graph, from the network's own inputs to it (tight)
    w = bf(net.w.cpu())
    b = net.b.cpu()

    def conv(li, tt):
        _, ci, co, _ = PC.LAYERS[li]
        o_ = net.woff[li]
        return F.conv3d(bf(tt), w[o_:o_ + ci * co * 27].view(co, ci, 3, 3, 3), b[net.boff[li]:net.boff[li] + co],
                        padding=1)
    act = lambda tt: F.leaky_relu(tt, PC.SLOPE)
    checks = []
    for lev in range(PC.LEVELS):
        xin = acts["x"] if lev == 0 else acts[f"p{lev}"]
        checks.append((f"h{lev}", act(conv(2 * lev, xin))))
        checks.append((f"e{lev}", act(conv(2 * lev + 1, acts[f"h{lev}"]))))
        if lev < PC.LEVELS - 1:
            checks.append((f"p{lev + 1}", F.avg_pool3d(acts[f"e{lev}"], 2)))
    for lev in range(PC.LEVELS - 2, -1, -1):
        la = 2 * PC.LEVELS + 2 * (PC.LEVELS - 2 - lev)
        coarse = acts[f"e{PC.LEVELS - 1}"] if lev == PC.LEVELS - 2 else acts[f"d{lev + 1}"]
        checks.append((f"u{lev}", F.interpolate(coarse, scale_factor=2, mode="nearest")))
        checks.append((f"z{lev}", act(conv(la, acts[f"u{lev}"]) + acts[f"e{lev}"])))
        checks.append((f"d{lev}", act(conv(la + 1, acts[f"z{lev}"]))))
    checks.append(("y", conv(len(PC.LAYERS) - 1, acts["d0"])))
    acts["y"] = y.cpu()
    assert len(checks) == 3 * PC.LEVELS - 1 + 3 * (PC.LEVELS - 1) + 1
    for name, ref in checks:
        assert _rel(acts[name], ref) < 1e-5, name
    # end to end against the reference's own forward (its rounding decisions differ where
    # FP32 sums differ; at the 4-voxel bottleneck one flip moves a whole sample by ~0.3 %)
    ref = _ref_net(net, x)
    e2e = float((y.cpu() - ref).norm() / ref.norm())
    print(f"  end-to-end forward rel L2 {e2e:.2e}", flush=True)
    assert e2e < 2e-2
    loss, dy = net.loss(y, dev(t))
    assert abs(float(loss.item()) - float(((y.cpu() - t) ** 2).mean())) < 1e-5 * float(((y.cpu() - t) ** 2).mean())
    net.backward(dy)
    torch.cuda.synchronize()
    # autograd of the reference (weights as leaves, BF16-valued; inputs rounded as in the kernels)
    wl = bf(net.w.cpu()).requires_grad_(True)
    bl = net.b.cpu().clone().requires_grad_(True)

    def conv_leaf(li, tt):
        _, ci, co, _ = PC.LAYERS[li]
        o_ = net.woff[li]
        return F.conv3d(bf(tt), wl[o_:o_ + ci * co * 27].view(co, ci, 3, 3, 3), bl[net.boff[li]:net.boff[li] + co],
                        padding=1)
    e, h = [], x
    for lev in range(PC.LEVELS):
        h = act(conv_leaf(2 * lev + 1, act(conv_leaf(2 * lev, h))))
        e.append(h)
        if lev < PC.LEVELS - 1:
            h = F.avg_pool3d(h, 2)
    d = e[-1]
    for lev in range(PC.LEVELS - 2, -1, -1):
        la = 2 * PC.LEVELS + 2 * (PC.LEVELS - 2 - lev)
        d = act(conv_leaf(la + 1, act(conv_leaf(la, F.interpolate(d, scale_factor=2, mode="nearest")) + e[lev])))
    yr = conv_leaf(len(PC.LAYERS) - 1, d)
    ((yr - t) ** 2).mean().backward()
    gw, gb = net.gw.cpu(), net.gb.cpu()
    for li, (name, ci, co, _) in enumerate(PC.LAYERS):
        a, bnd = net.woff[li], net.woff[li] + ci * co * 27
        rw = wl.grad[a:bnd]
        err = float((gw[a:bnd] - rw).norm() / rw.norm())
        rb = bl.grad[net.boff[li]:net.boff[li] + co]
        errb = float((gb[net.boff[li]:net.boff[li] + co] - rb).norm() / rb.norm())
        print(f"  {name}: weight grad rel L2 {err:.2e}, bias grad {errb:.2e}", flush=True)
        assert err < 3e-2 and errb < 3e-2, name


def test_training_reduces_the_loss():
    from hirensi_amd import precond as PC
    print("training: 80 Adam steps on a smoothing map, nb=4 dims=(8, 8, 8)", flush=True)
    rng = np.random.default_rng(0)
    X = rng.standard_normal((16, 8, 8, 8)).astype(np.float32)
    # target: a fixed linear smoothing of the input (learnable by the convolutions)
    T = np.zeros_like(X)
    for ax in (1, 2, 3):
        T += np.roll(X, 1, axis=ax) + np.roll(X, -1, axis=ax)
    T = (T / 6.0).astype(np.float32)
    net = PC.ConvNet(seed=1)
    hist = PC.train(net, X, T, epochs=20, batch=4, lr=2e-3)
    print("  losses", [round(h, 4) for h in hist], flush=True)
    assert np.isfinite(hist).all() and hist[-1] < 0.6 * hist[0]
    # save / load round trip reproduces the outputs
    st = net.state()
    net2 = PC.ConvNet(seed=2)
    net2.load_state(st)
    assert np.array_equal(net.infer(X[0]), net2.infer(X[0]))


def test_solver_made_pairs_training_and_learned_h0_in_an_inversion():
    """The whole learned-preconditioner path on a small mesh: training pairs from
    the FP64 solver (first-iteration gradient of a layered start model, true mu
    error), a few epochs, the H0 map's normalisation, and inversion iterations
    that use it."""
    from hirensi_amd import fwi as F
    from hirensi_amd import precond as PC
    from hirensi_amd import solver as S
    from hirensi_amd.mesh import BoxMesh
    mesh = BoxMesh(16, 16, 8, npml=(2,) * 5)
    print(f"pairs + training + inversion on {mesh.nex}x{mesh.ney}x{mesh.nez}, neq={mesh.neq}", flush=True)
    X, T = PC.generate_pairs(4, seed=3, mesh=mesh, nstep=150)
    mask = PC.rd_voxel_mask(mesh)
    assert X.shape == (4, mesh.nez, mesh.ney, mesh.nex) and np.isfinite(X).all() and np.isfinite(T).all()
    for k in range(4):
        assert abs(float(np.sqrt((X[k].astype(np.float64) ** 2).sum() / mask.sum())) - 1.0) < 1e-5
        assert np.all(X[k][mask == 0] == 0) and np.all(T[k][mask == 0] == 0)
    net = PC.ConvNet(seed=4)
    hist = PC.train(net, X, T, epochs=3, batch=2, lr=1e-3)
    assert np.isfinite(hist).all()
    # an inversion on a case of the same family, with the map as H0
    rng = np.random.default_rng(7)
    true, start = PC.random_case(mesh, rng)
    sv = PC.training_survey(mesh, 150)
    ctx = S.Context(mesh)
    sv.dis_meas = F.HipProblem(ctx, sv, true.rho).observe(true.lam, true.mu)
    _, restriction = S.gradient_mass(ctx)
    h0 = PC.LearnedH0(net, mesh, restriction)
    q = np.random.default_rng(1).standard_normal(mesh.nnode) * restriction
    r = h0(q)
    if r is not None:
        assert abs(float(r @ q) - float(q @ q)) < 1e-9 * float(q @ q)
    params = F.InversionParams(m_lbfgs=3, bound_lambda_lower=10.0, bound_lambda_upper=2000.0, bound_mu_lower=5.0,
                               bound_mu_upper=1000.0, fac_max_update_lambda=5.0, fac_max_update_mu=5.0)
    inv = F.Inversion(F.HipProblem(ctx, sv, start.rho, params), start.lam, start.mu, restriction, params, h0_map=h0)
    recs = [inv.iterate() for _ in range(2)]
    for rec in recs:
        assert rec["misfit"] <= rec["misfit_pre"] and ("learned_h0" in rec and "fallback" in rec)
    assert h0.calls >= 3
    ctx.close()
