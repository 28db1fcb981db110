"""The learned closure's kernels on the GPU (BF16 operands on the matrix
cores, FP32 accumulation) against the FP64 reference network
(oracle/cnn_ref.py, itself pinned by finite differences), and the closure
inside a coarse simulation.

Bars: BF16 keeps 8 significant bits (relative rounding 2^-9 = 2e-3).
Against a reference that rounds the inputs, the weights and each stored
activation to BF16 the same way, the forward output is held to 2e-3
relative L2 (what remains is FP32 accumulation order and tanh); against
the exact FP64 network to 3e-2.  Gradients (each layer's weight block and
the whole vector) to 3e-2 relative L2 of the FP64 gradient, with cosine
similarity above 0.999."""
import numpy as np
import pytest

import pllcart3d_amd as P
from oracle import cnn_ref as R
from pllcart3d_amd import closure as CL

pytestmark = pytest.mark.gpu


def bf16(a):
    """Round to BF16 (nearest even) and back, in numpy."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def forward_bf16(x, Pd):
    """The reference network with the GPU's storage roundings."""
    Q = {k: bf16(v) if k.startswith("W") else v for k, v in Pd.items()}
    a1 = bf16(np.tanh(R.conv(bf16(x), Q["W1"]) + Q["b1"]))
    a2 = bf16(np.tanh(R.conv(a1, Q["W2"]) + Q["b2"]))
    return (R.conv(a2, Q["W3"]) + Q["b3"])[..., 0]


def _setup(shape, seed):
    rng = np.random.default_rng(seed)
    Pd = R.init_params(seed)
    Pd["W3"] *= 10.0
    Pd["b1"] = 0.1 * rng.normal(size=16)
    Pd["b2"] = 0.1 * rng.normal(size=16)
    Pd["b3"] = np.array([0.05])
    x5 = rng.normal(size=shape + (5,))
    x = np.zeros(shape + (16,))
    x[..., :5] = x5
    net = CL.Network(*shape)
    net.set_params(CL.flat_from_dict(Pd))
    return rng, Pd, x5, x, net


@pytest.mark.parametrize("shape", [(6, 5, 16), (5, 7, 21), (12, 10, 40)])
def test_forward_matches_reference(shape):
    """Grids whose k rows are a multiple of the 16-cell MFMA tile and not
    (tail cells masked), with the zero padding on every face."""
    rng, Pd, x5, x, net = _setup(shape, 1)
    y = net.forward(x5.astype(np.float32))
    e_bf = np.linalg.norm(y - forward_bf16(x, Pd)) / np.linalg.norm(forward_bf16(x, Pd))
    yref = R.forward(x, Pd)[0]
    e64 = np.linalg.norm(y - yref) / np.linalg.norm(yref)
    print("[cnn forward %s] vs bf16-stored reference %.2e, vs FP64 %.2e" % (shape, e_bf, e64), flush=True)
    assert e_bf <= 2e-3 and e64 <= 3e-2
    # the parameters round-trip
    assert np.allclose(net.get_params(), CL.flat_from_dict(Pd).astype(np.float32), rtol=0, atol=0)
    net.close()


def test_gradients_match_reference():
    """Loss and every parameter's gradient of one sample against the FP64
    reference's analytic gradient (oracle/cnn_ref.py, pinned by central
    differences in tests/test_cnn_ref.py)."""
    shape = (8, 6, 20)
    rng, Pd, x5, x, net = _setup(shape, 2)
    t = rng.normal(size=shape)
    mask = (rng.uniform(size=shape) > 0.2).astype(np.uint8)
    i = net.add_sample(x5, t, mask)
    L, g = net.loss_grad(i)
    Lr, Gr, _ = R.loss_and_grads(x, t, mask.astype(bool), Pd)
    gr = CL.flat_from_dict({k: Gr[k] for k in Gr})
    print("[cnn grad] loss gpu %.6e ref %.6e" % (L, Lr), flush=True)
    assert abs(L - Lr) <= 1e-2 * Lr
    w = 27 * 256
    blocks = {"W1": slice(0, w), "b1": slice(w, w + 16), "W2": slice(w + 16, 2 * w + 16),
              "b2": slice(2 * w + 16, 2 * w + 32), "W3": slice(2 * w + 32, 3 * w + 32),
              "b3": slice(3 * w + 32, 3 * w + 48)}
    for name, sl in blocks.items():
        a, b = g[sl], gr[sl]
        e = np.linalg.norm(a - b) / np.linalg.norm(b)
        cos = a @ b / (np.linalg.norm(a) * np.linalg.norm(b))
        print("[cnn grad] %s rel L2 %.2e cos %.6f" % (name, e, cos), flush=True)
        assert e <= 3e-2 and cos >= 0.999, name
    net.close()


def test_training_fits_a_teacher_network():
    """Adam on the device fits the outputs of a fixed teacher network of
    the same shape: the loss falls by more than 10x in 400 steps (measured:
    16x)."""
    shape = (16, 12, 32)
    rng = np.random.default_rng(3)
    teacher = R.init_params(11)
    teacher["W3"] *= 10.0
    xs, ts = [], []
    net = CL.Network(*shape)
    mask = np.ones(shape, np.uint8)
    for s in range(4):
        x5 = rng.normal(size=shape + (5,))
        x = np.zeros(shape + (16,))
        x[..., :5] = x5
        t = R.forward(x, teacher)[0]
        net.add_sample(x5, t, mask)
    net.set_params(CL.flat_from_dict(R.init_params(12)))
    losses = net.train(400, [0, 1, 2, 3], 3e-3)
    print("[cnn train] loss %.3e -> %.3e" % (losses[:4].mean(), losses[-4:].mean()), flush=True)
    assert losses[-4:].mean() < losses[:4].mean() / 10
    net.close()


def _coarse(precision="mixed"):
    prm = P.make_params(32, 16, 16, Vratio=3.0, Pe=1000.0, Inf_thick=0.15)
    c0 = P.ic_fill(prm, P.IC_MULTIMODE, 6, 0.05)
    return prm, c0


def test_zero_network_equals_plain_run():
    """A closure whose network outputs zero leaves the run unchanged."""
    prm, c0 = _coarse()
    a = P.Sim(prm, "mixed")
    a.init(c0)
    b = P.Sim(prm, "mixed")
    b.init(c0)
    net = CL.Network(prm.nx, prm.ny, prm.nz)
    net.set_params(np.zeros(CL.nparams()))
    net.set_norm(np.zeros(5), np.ones(5), 1.0)
    CL.set_closure(b, net)
    a.run(4)
    b.run(4)
    for f in ("C_OLD", "U_OLD", "W_OLD", "PHI"):
        assert np.array_equal(a.download(f), b.download(f)), f
    net.close()


def test_guard_falls_back_to_plain_weno():
    """A network whose correction is far too large: every pair of steps
    breaks the bounds or the mass check, is run again without the
    correction, and the result equals the plain run; a small correction
    passes the guard and changes the run.  The bound allows the plain
    scheme's own overshoot (WENO and Crank-Nicolson are not bounded in
    [0, 1]): 0.05."""
    prm, c0 = _coarse()
    plain = P.Sim(prm, "mixed")
    plain.init(c0)
    bad = P.Sim(prm, "mixed")
    bad.init(c0)
    net = CL.Network(prm.nx, prm.ny, prm.nz)
    Pd = R.init_params(5)
    Pd["b3"] = np.array([1.0])
    net.set_params(CL.flat_from_dict(Pd))
    net.set_norm(np.zeros(5), np.ones(5), 1e3)
    CL.set_closure(bad, net)
    fb, worst = CL.run_guarded(bad, 4, 2, 0.05, 1e-6)
    plain.run(4)
    print("[guard] fallback steps %d of 4 (worst bound excess %.2e, mass ratio %.2e)" % (fb, worst[0], worst[1]),
          flush=True)
    assert fb == 4
    for f in ("C_OLD", "U_OLD", "W_OLD", "PHI"):
        assert np.array_equal(plain.download(f), bad.download(f)), f
    # a small correction passes the guard (the parameters and the output
    # scale are read by the captured graph from device memory)
    Pd = R.init_params(6)
    Pd["W3"] *= 10.0
    net.set_params(CL.flat_from_dict(Pd))
    net.set_norm(np.zeros(5), np.ones(5), 1e-3)
    fb, worst = CL.run_guarded(bad, 2, 2, 0.05, 1e-3)
    c = plain.download("C_OLD")[3:-3, 1:-1, 1:-1]
    print("[guard] plain run c in [%.4f, %.4f]; small correction fallbacks %d (worst %.2e, %.2e)"
          % (c.min(), c.max(), fb, worst[0], worst[1]), flush=True)
    assert fb == 0 and worst[1] <= 1e-9
    plain.run(2)
    assert not np.array_equal(plain.download("C_OLD"), bad.download("C_OLD"))
    net.close()


def test_correction_enters_concentration_term():
    """The correction a simulation adds to N is out_scale * (y - mean y) of
    the network on the gathered, shifted and scaled cell inputs (c,
    cell-centred u, v, w, mu_old.cell), the mean over the corrected cells
    (no mass added), zero within 4 cells of the inlet and outlet."""
    prm, c0 = _coarse()
    a = P.Sim(prm, "mixed")
    a.init(c0)
    a.run(3)
    b = P.Sim(prm, "mixed")
    b.init(c0)
    b.run(3)
    net = CL.Network(prm.nx, prm.ny, prm.nz)
    Pd = R.init_params(8)
    Pd["W3"] *= 10.0
    net.set_params(CL.flat_from_dict(Pd))
    shift = np.array([0.5, 0.7, 0.0, 0.0, 3.0])
    scale = np.array([2.0, 1.5, 20.0, 20.0, 0.5])
    net.set_norm(shift, scale, 0.25)
    CL.set_closure(b, net)
    dt = 1e-3
    for s in (a, b):
        vb = s.op("BULK_VELO")
        s.op("CONCENTRATION", dt, 0, vb)
    st = {f: b.download(f) for f in ("C_OLD", "U_OLD", "V_OLD", "W_OLD", "MUO_CELL")}
    x = (CL.features(st["C_OLD"], st["U_OLD"], st["V_OLD"], st["W_OLD"], st["MUO_CELL"], prm) - shift) * scale
    y = net.forward(x.astype(np.float32))
    inner = slice(CL.XMARGIN, prm.nx - CL.XMARGIN)
    want = 0.25 * (y - y[inner].mean())
    want[:CL.XMARGIN] = 0.0
    want[prm.nx - CL.XMARGIN:] = 0.0
    got = (b.download("C_N") - a.download("C_N"))[3:3 + prm.nx, 1:1 + prm.ny, 1:1 + prm.nz]
    e = np.linalg.norm(got - want) / np.linalg.norm(want)
    print("[closure] N correction vs network %.2e" % e, flush=True)
    assert e <= 1e-2
    assert np.all(got[:CL.XMARGIN] == 0.0) and np.all(got[prm.nx - CL.XMARGIN:] == 0.0)
    assert abs(got[inner].sum()) <= 1e-5 * np.abs(got[inner]).sum()
    net.close()


def test_trained_closure_gains_over_no_correction():
    """The whole pipeline at a small size: box-filtered 64x32x32 plain runs
    (R ~ U[2,4], Pe ~ U[1000,3000]), targets the coarse solver's error over
    4 steps, 600 Adam steps on the GPU; on a held-out run the trained
    network's loss is below that of no correction (the network gains over
    plain coarse WENO), and a guarded coarse run with it stays finite and
    bounded."""
    fine = (64, 32, 32)
    runs = CL.draw_training_runs(2, 700)
    X, T = CL.collect(fine, runs, 120, 20, 20, print, 4)
    Xv, Tv = CL.collect(fine, CL.draw_training_runs(1, 800), 120, 40, 40, print, 4)
    fp = P.make_params(*fine, Vratio=3.0, Pe=2000.0, Inf_thick=0.02)
    cp = CL.coarse_params(fp)
    mask = CL.loss_mask(cp)
    net, norm, losses = CL.train(X, T, mask, (cp.nx, cp.ny, cp.nz), 450, 1, print)
    trained, none, gain = CL.validation_gain(net, norm, Xv, Tv, mask)
    print("[closure pipeline] held-out loss trained %.4f none %.4f gain %.3f" % (trained, none, gain), flush=True)
    assert np.all(np.isfinite(losses)) and gain > 0.2
    c0 = P.ic_fill(fp, P.IC_MULTIMODE, 2, 0.01)
    shift, scale, osc = norm
    net.set_norm(shift, scale, 0.5 * osc)
    s = P.Sim(cp, "mixed")
    s.init(CL.coarse_c0(fp, c0))
    CL.set_closure(s, net, 4)
    fb, worst = CL.run_guarded(s, 40, 10, 2e-2, 1e-4)
    c = s.download("C_OLD")[3:-3, 1:-1, 1:-1]
    print("[closure pipeline] 40 guarded steps: fallbacks %d, c in [%.4f, %.4f]" % (fb, c.min(), c.max()), flush=True)
    assert np.all(np.isfinite(c)) and c.min() > -0.02 and c.max() < 1.02
    net.close()


@pytest.mark.parametrize("every", [0, 1, 3])
def test_refresh_period_graph_equals_eager(monkeypatch, every):
    """A correction made every `every` steps (0: every substage) and kept in
    between: pairs of steps replayed from the captured graphs (one per
    pattern of corrections within the pair) equal eager steps
    (PLLC_GRAPH=0) in every bit, over 9 steps, so the patterns repeat and
    the odd last step runs eagerly; and the refresh period changes the
    result (the correction is really held)."""
    fp = P.make_params(32, 16, 16, Vratio=3.0, Pe=2000.0, Inf_thick=0.02)
    cp = CL.coarse_params(fp)
    c0 = P.ic_fill(fp, P.IC_MULTIMODE, 2, 0.01)
    Pd = R.init_params(5)
    Pd["W3"] *= 10.0
    net = CL.Network(cp.nx, cp.ny, cp.nz)
    net.set_params(CL.flat_from_dict(Pd))
    net.set_norm(np.zeros(5), np.ones(5), 0.05)
    out = {}
    for mode in ("graph", "eager"):
        if mode == "eager":
            monkeypatch.setenv("PLLC_GRAPH", "0")
        else:
            monkeypatch.delenv("PLLC_GRAPH", raising=False)
        s = P.Sim(cp, "mixed")
        s.init(CL.coarse_c0(fp, c0))
        CL.set_closure(s, net, every)
        print("[closure refresh] every %d %s" % (every, mode), flush=True)
        s.run(9)
        out[mode] = s.download("C_OLD")
        s.close()
    assert np.array_equal(out["graph"], out["eager"]), every
    monkeypatch.delenv("PLLC_GRAPH", raising=False)
    s = P.Sim(cp, "mixed")
    s.init(CL.coarse_c0(fp, c0))
    CL.set_closure(s, net, 2 if every != 2 else 1)
    s.run(9)
    assert not np.array_equal(s.download("C_OLD"), out["graph"]), every
    s.close()
    net.close()


def test_training_repeats_bit_for_bit():
    """Two trainings from the same parameters, samples and sample order end
    with equal parameters in every bit: the weight-gradient sums of a
    block's waves go in a fixed order."""
    shape = (8, 6, 20)
    rng = np.random.default_rng(5)
    x = rng.normal(size=shape + (5,)).astype(np.float32)
    t = rng.normal(size=shape).astype(np.float32)
    mask = np.ones(shape, np.uint8)
    out = []
    for rep in range(2):
        net = CL.Network(*shape)
        net.set_params(CL.flat_from_dict(R.init_params(2)))
        net.set_norm(np.zeros(5), np.ones(5), 1.0)
        i = net.add_sample(x, t, mask)
        losses = net.train(60, [i], 1e-3)
        out.append(net.get_params())
        net.close()
        print("[cnn determinism] run %d: loss %.4e -> %.4e" % (rep, losses[0], losses[-1]), flush=True)
    assert np.array_equal(out[0], out[1])
