"""The closure network's BF16 MFMA kernels (csrc/cnn.hip) through the C-ABI
against the oracle (oracle/cnn_oracle.py).

The bf16-emulating oracle rounds at the GPU's storage points, so only the
FP32-vs-FP64 accumulation order differs: prediction and loss within
relative L2 1e-3, gradients per layer within 5e-3.  Against the plain FP64
oracle the bf16 storage itself shows: prediction within 3e-2, gradients
within 6e-2."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _dev(a, dtype):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(device="cuda", dtype=dtype)


def _case(cin, cout, n, H, W, seed):
    from oracle import cnn_oracle as co
    rng = np.random.default_rng(seed)
    p = co.init_params(cin, cout, seed, last_scale=1.0)
    p[co.nparams(cin, cout) - cout:] = 0.1 * rng.standard_normal(cout)       # last-layer biases
    x = rng.standard_normal((n, H, W, cin)).astype(np.float32)
    t = rng.standard_normal((n, H, W, cout)).astype(np.float32)
    return p, x, t


def _blocks(v, cin, cout):
    from oracle import cnn_oracle as co
    out, off = [], 0
    for i, o in co.shapes(cin, cout):
        out.append(v[off:off + o * 9 * i])
        out.append(v[off + o * 9 * i:off + o * 9 * i + o])
        off += o * 9 * i + o
    return out


# 16 x 16 tiles (32 x 48, 256 x 256), 8 x 32 tiles (24 x 64), and the
# per-pixel kernel where neither tile divides the image (20 x 48)
@pytest.mark.parametrize("cin,cout,n,H,W", [(16, 8, 2, 32, 48), (32, 16, 1, 24, 64), (16, 8, 1, 256, 256),
                                            (16, 8, 1, 20, 48)])
def test_forward_backward_adam_vs_oracle(cin, cout, n, H, W):
    import torch
    from oracle import cnn_oracle as co
    from qgn_fortran_amd.cnn import CNN, to_bf16_bits
    print("cnn", cin, cout, n, H, W, flush=True)
    p, x, t = _case(cin, cout, n, H, W, seed=cin + H)
    net = CNN(cin, cout, p)
    xd = _dev(to_bf16_bits(x), torch.int16)
    td = _dev(t, torch.float32)
    yd = torch.empty((n, H, W, cout), device="cuda", dtype=torch.float32)
    net.forward(xd, yd)
    y = yd.cpu().numpy()
    loss_e, g_e, pred_e = co.loss_and_grads(p, x, t, cin, cout, bf16_points=True)
    loss_f, g_f, pred_f = co.loss_and_grads(p, x, t, cin, cout, bf16_points=False)
    e1, e2 = rel_l2(y, pred_e), rel_l2(y, pred_f)
    print(f"pred rel L2 vs bf16 oracle {e1:.2e}, vs fp64 oracle {e2:.2e}", flush=True)
    assert e1 < 1e-3 and e2 < 3e-2
    loss = net.backward(td)
    assert abs(loss - loss_e) < 1e-3 * loss_e
    g = net.grads()
    for k, (a, b, c) in enumerate(zip(_blocks(g, cin, cout), _blocks(g_e, cin, cout), _blocks(g_f, cin, cout))):
        ea, eb = rel_l2(a, b), rel_l2(a, c)
        print(f"grad block {k}: rel L2 vs bf16 oracle {ea:.2e}, vs fp64 {eb:.2e}", flush=True)
        assert ea < 5e-3 and eb < 6e-2, k
    # one Adam step against the oracle's Adam with the GPU's gradients
    net.adam(1e-3)
    ref, _, _ = co.adam(p.astype(np.float64), g.astype(np.float64), np.zeros(p.size), np.zeros(p.size), 1, 1e-3)
    assert np.max(np.abs(net.params() - ref)) < 1e-6
    net.close()


def test_training_reduces_loss_on_a_teacher():
    """A student network learns a random teacher network's map (targets from
    the FP64 oracle): 300 Adam steps on the GPU cut the loss 5-fold."""
    import torch
    from oracle import cnn_oracle as co
    from qgn_fortran_amd.cnn import CNN, to_bf16_bits
    cin, cout, n, H, W = 16, 8, 4, 32, 32
    rng = np.random.default_rng(7)
    x = rng.standard_normal((n, H, W, cin)).astype(np.float32)
    teacher = co.init_params(cin, cout, 101, last_scale=1.0)
    t, _ = co.forward(teacher, x, cin, cout)
    net = CNN(cin, cout, co.init_params(cin, cout, 5, last_scale=0.1))
    xd = _dev(to_bf16_bits(x), torch.int16)
    td = _dev(t.astype(np.float32), torch.float32)
    yd = torch.empty((n, H, W, cout), device="cuda", dtype=torch.float32)
    losses = [net.train_step(xd, td, yd, 2e-3) for _ in range(300)]
    print(f"loss {losses[0]:.4e} -> {losses[-1]:.4e}", flush=True)
    assert losses[-1] < 0.2 * losses[0]
    net.close()


@pytest.mark.parametrize("precision,nz,n", [("fp64", 8, 64), ("mixed", 8, 64), ("fp64", 16, 64), ("mixed", 16, 64),
                                            ("fp64", 8, 256), ("mixed", 8, 256)])
def test_closure_forcing_in_rhs(precision, nz, n):
    """qgn_set_closure: GetRHS with the network attached minus GetRHS without
    it is r2c(out_scale * net(in_scale * (q, psi))) on the kept modes, with q
    and psi the physical fields of the state (the oracle's WriteQP fields,
    QGN_Module.f90:353-361) and the network the bf16-emulating oracle; at
    64^2 and at the learned-closure run's 256^2 x 8."""
    from conftest import make_case
    from oracle import cnn_oracle as co
    from oracle import qgn_oracle as orc
    from qgn_fortran_amd import QGN
    from qgn_fortran_amd.cnn import CNN
    case = make_case(n, n, nz, k_peak=6, precision=precision)
    om = orc.Model(case["po"], case["H"], case["Modes"], case["EVals"], case["uBar"], case["qyBar"])
    qh = om.initialize(case["q0"])
    qp, pp = om.get_psi_phys(qh)
    # per-channel input scales that differ, so a channel routed to the wrong
    # place would show
    s_in = np.concatenate([np.linspace(0.8, 1.2, nz) / np.std(qp), np.linspace(1.1, 0.9, nz) / np.std(pp)])
    base = om.get_rhs(qh)
    s_out = np.linspace(0.2, 0.4, nz) * np.sqrt(np.mean(om.c2r(base) ** 2)) / (n * n)
    params = co.init_params(2 * nz, nz, seed=3, last_scale=1.0)
    x = np.concatenate([qp, pp], axis=0).transpose(1, 2, 0)[None] * s_in      # (1, H, W, 2 nz)
    y, _ = co.forward(params, x.astype(np.float32), 2 * nz, nz, bf16_points=True)
    F = orc.dealias(orc.rfft2(y[0].transpose(2, 0, 1) * s_out[:, None, None]), n, n)
    gm = QGN(case["pp"], case["H"], case["Modes"], case["EVals"], case["uBar"], case["qyBar"])
    r0 = gm.get_rhs(qh)
    net = CNN(2 * nz, nz, params)
    gm.set_closure(net, s_in, s_out)
    r1 = gm.get_rhs(qh)
    err = rel_l2(r1 - r0, F)
    print(precision, nz, n, "closure forcing rel L2", err, "forcing/RHS", np.linalg.norm(F) / np.linalg.norm(base), flush=True)
    assert err < (2e-3 if precision == "fp64" else 5e-3)
    # the closure inside the steppers: a step with it differs from a step without
    gm.set_state(qh)
    t, dt = gm.time_step(0.0, 600.0)
    s1 = gm.get_state()
    gm.set_closure(None)
    g2 = QGN(case["pp"], case["H"], case["Modes"], case["EVals"], case["uBar"], case["qyBar"])
    g2.set_state(qh)
    g2.time_step(0.0, 600.0)
    assert rel_l2(s1, g2.get_state()) > 1e-9
    gm.close()
    g2.close()
    net.close()


def test_closure_fp64_row_limit():
    """The FP64 path's closure row passes hold 8 rows of nx/2 FP64 complex
    values in LDS, so qgn_set_closure rejects nx > 1024 there with an error
    (and the mixed path accepts it)."""
    from conftest import make_case
    from oracle import cnn_oracle as co
    from qgn_fortran_amd import QGN
    from qgn_fortran_amd.cnn import CNN
    net = CNN(16, 8, co.init_params(16, 8, seed=1))
    for precision in ("fp64", "mixed"):
        case = make_case(2048, 64, 8, precision=precision)
        gm = QGN(case["pp"], case["H"], case["Modes"], case["EVals"], case["uBar"], case["qyBar"])
        if precision == "fp64":
            with pytest.raises(RuntimeError, match="nx <= 1024"):
                gm.set_closure(net, np.ones(16), np.ones(8))
        else:
            gm.set_closure(net, np.ones(16), np.ones(8))
            gm.set_closure(None)
        gm.close()
    net.close()


def test_weight_gradients_are_reproducible():
    """Each workgroup of the weight-gradient kernel writes its own partial
    sums and a second kernel adds them in a fixed order, so the same
    backward pass twice gives bit-identical gradients (and a training run
    is reproducible)."""
    import torch
    from oracle import cnn_oracle as co
    from qgn_fortran_amd.cnn import CNN, to_bf16_bits
    rng = np.random.default_rng(3)
    for cin, cout, n, H, W in ((16, 8, 2, 32, 48), (32, 16, 3, 40, 32)):
        x = rng.standard_normal((n, H, W, cin)).astype(np.float32)
        t = rng.standard_normal((n, H, W, cout)).astype(np.float32)
        xd = torch.from_numpy(to_bf16_bits(x)).cuda()
        td = torch.from_numpy(t).cuda()
        y = torch.empty((n, H, W, cout), device="cuda", dtype=torch.float32)
        net = CNN(cin, cout, co.init_params(cin, cout, seed=2, last_scale=1.0))
        grads = []
        for _ in range(3):
            net.forward(xd, y)
            net.backward(td)
            grads.append(net.grads())
        assert np.any(grads[0] != 0)
        assert all(np.array_equal(grads[0], gr) for gr in grads[1:])
        net.close()


def test_inputs_written_by_torch_just_before_the_call():
    """The network's kernels run on the library's own stream: inputs that a
    torch kernel is still writing when the call is made (an index_select of
    a batch, as the closure's training does) must be complete before the
    library reads them.  Forward on a freshly gathered batch equals forward
    on the same batch after a full synchronisation, every time."""
    import torch
    from oracle import cnn_oracle as co
    from qgn_fortran_amd.cnn import CNN, to_bf16_bits
    cin, cout, n, H, W = 16, 8, 10, 256, 256
    rng = np.random.default_rng(11)
    pool = _dev(to_bf16_bits(rng.standard_normal((3 * n, H, W, cin)).astype(np.float32)), torch.int16)
    net = CNN(cin, cout, co.init_params(cin, cout, 9, last_scale=1.0))
    y1 = torch.empty((n, H, W, cout), device="cuda", dtype=torch.float32)
    y2 = torch.empty_like(y1)
    for trial in range(4):
        sel = torch.from_numpy(np.sort(rng.permutation(3 * n)[:n])).cuda()
        xs = pool.index_select(0, sel)        # queued on torch's stream
        net.forward(xs, y1)
        torch.cuda.synchronize()
        net.forward(xs, y2)
        assert torch.equal(y1, y2), trial
    net.close()


def test_closure_network_lifetime_is_enforced():
    """A model with the network attached, and an ensemble whose captured
    graph ran it, use the network's activation buffers: while either holds
    the network, a pass on more pixels and qgn_cnn_destroy return an error
    instead of freeing them; qgn_destroy refuses a context an ensemble
    holds; backward after the closure's own forward (inside an RHS) is
    refused.  Only the error paths run: nothing touches freed memory."""
    import torch
    from oracle import cnn_oracle as co
    from qgn_fortran_amd import Batch, Ensemble, configs
    from qgn_fortran_amd._lib import QGN_EINVAL, QgnError
    from qgn_fortran_amd.cnn import CNN, to_bf16_bits
    group = configs.ensemble96(sizes=(64,), per_size=1, nz=8, seed=3, scheme="ARK43")[0]
    net = CNN(16, 8, co.init_params(16, 8, seed=4, last_scale=0.05))
    s_in, s_out = np.full(16, 1e5), np.full(8, 1e-12)
    b = Batch(group.params, group.members)
    qh = b.initialize(group.q0)
    b.set_closure(net, s_in, s_out)                  # holds 64 x 64 pixels
    rng = np.random.default_rng(0)
    big = _dev(to_bf16_bits(rng.standard_normal((1, 128, 128, 16))), torch.int16)
    ybig = torch.empty((1, 128, 128, 8), device="cuda", dtype=torch.float32)
    with pytest.raises(QgnError, match="reallocate") as e:
        net.forward(big, ybig)
    assert e.value.code == QGN_EINVAL
    with pytest.raises(QgnError, match="attached"):
        net.close()
    # a user forward within the held size, then an RHS (the closure's forward): backward refuses
    x = _dev(to_bf16_bits(rng.standard_normal((1, 64, 64, 16))), torch.int16)
    y = torch.empty((1, 64, 64, 8), device="cuda", dtype=torch.float32)
    t = torch.zeros_like(y)
    net.forward(x, y)
    net.backward(t)
    b.get_rhs(qh)
    with pytest.raises(QgnError, match="closure"):
        net.backward(t)
    net.forward(x, y)
    net.backward(t)
    # an ensemble keeps its hold after the model detaches
    b.set_time(0.0, 1000.0)
    ens = Ensemble([b])
    ens.step(1)
    b.times()
    b.set_closure(None)
    with pytest.raises(QgnError, match="ensemble"):
        net.close()
    with pytest.raises(QgnError, match="ensemble"):
        b.close()
    with pytest.raises(QgnError, match="reallocate"):
        net.forward(big, ybig)
    ens.close()
    net.forward(big, ybig)                           # free again: the buffers may grow
    b.close()
    net.close()
    with pytest.raises(ValueError, match="int16"):
        CNN(16, 8, co.init_params(16, 8, seed=4)).forward(x.float(), y)
