"""GPU parity of the BF16 MFMA 3^3 convolution kernels against a float64
torch reference of the same op (periodic conv3d), on BF16-representable
inputs, plus the full LocalNet3D forward/backward against autograd."""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("needs a GPU")
    return torch.device("cuda:0")


def ref_conv(x_cf, w_t, b=None):
    """float64 periodic conv3d on the CPU. x (B, C, G, G, G), w (co, ci, 3, 3, 3)."""
    xp = F.pad(x_cf.double().cpu(), (1, 1, 1, 1, 1, 1), mode="circular")
    return F.conv3d(xp, w_t.double().cpu(), None if b is None else b.double().cpu())


def bf16_round(t):
    return t.to(torch.bfloat16).to(torch.float32)


@pytest.mark.parametrize("cin", [4, 16, 32])
def test_conv_forward_modes(dev, cin):
    from augur_amd import surrogate as S
    g = torch.Generator().manual_seed(cin)
    B, G = 2, 32
    cinp = S.pad16(cin)
    x = bf16_round(torch.randn(B, cin, G, G, G, generator=g))
    w = bf16_round(torch.randn(27, 32, cin, generator=g) * 0.1)
    b = torch.randn(32, generator=g) * 0.1
    xc = S.pack_cl(x.to(dev), cinp)
    wk = S.kernel_weights(w.to(dev), cinp)
    ref = ref_conv(x, S.from_taps(w), b)                     # (B, 32, G, G, G) float64
    # mode 0: FP32 channels-first, only the first 7 channels
    y = S.conv3_fwd(xc, wk, b.to(dev), S.MODE_OUT_CF, 7).cpu().double()
    assert (y - ref[:, :7]).abs().max() < 1e-4 * ref.abs().max()
    # mode 1: z and gelu(z) in BF16
    z = torch.empty(B, G, G, G, 32, dtype=torch.bfloat16, device=dev)
    h = S.conv3_fwd(xc, wk, b.to(dev), S.MODE_GELU, 32, z_out=z)
    refz = ref.permute(0, 2, 3, 4, 1)
    assert (z.cpu().double() - refz).abs().max() < 1e-2 * refz.abs().max()
    assert (h.cpu().double() - F.gelu(refz, approximate='tanh')).abs().max() < 1e-2 * refz.abs().max()
    if cin == 32:
        h2 = S.conv3_fwd(xc, wk, b.to(dev), S.MODE_RES_GELU, 32)
        exp = x.double().permute(0, 2, 3, 4, 1) + F.gelu(refz, approximate='tanh')
        assert (h2.cpu().double() - exp).abs().max() < 1.5e-2 * exp.abs().max()
        # narrow outputs (cout <= 16 runs the 16-channel kernel), channels-first and channels-last
        y16 = S.conv3_fwd(xc, wk, b.to(dev), S.MODE_OUT_CF, 16).cpu().double()
        assert (y16 - ref[:, :16]).abs().max() < 1e-4 * ref.abs().max()
        y17 = S.conv3_fwd(xc, wk, b.to(dev), S.MODE_OUT_CF, 17).cpu().double()
        assert (y17 - ref[:, :17]).abs().max() < 1e-4 * ref.abs().max()
        ref_nb = (ref - b.double().view(1, 32, 1, 1, 1)).permute(0, 2, 3, 4, 1)   # no bias in mode 3
        add = torch.randn(B, G, G, G, 5, generator=g)
        y5 = S.conv3_fwd(xc, wk, None, S.MODE_OUT_CL, 5, addend=add.to(dev)).cpu().double()
        exp5 = ref_nb[..., :5] + add.double()
        assert (y5 - exp5).abs().max() < 1e-4 * exp5.abs().max()
        # all 32 channels channels-last (the backward data-gradient layout), with and without addend
        add32 = torch.randn(B, G, G, G, 32, generator=g)
        y32 = S.conv3_fwd(xc, wk, None, S.MODE_OUT_CL, 32, addend=add32.to(dev)).cpu().double()
        exp32 = ref_nb + add32.double()
        assert (y32 - exp32).abs().max() < 1e-4 * exp32.abs().max()
        y32n = S.conv3_fwd(xc, wk, None, S.MODE_OUT_CL, 32).cpu().double()
        assert (y32n - ref_nb).abs().max() < 1e-4 * ref_nb.abs().max()


def test_conv_data_gradient_is_transposed_conv(dev):
    from augur_amd import surrogate as S
    g = torch.Generator().manual_seed(7)
    B, G = 1, 32
    dz = bf16_round(torch.randn(B, 32, G, G, G, generator=g))
    w = bf16_round(torch.randn(27, 3, 32, generator=g) * 0.1)     # an output layer: 32 -> 3
    x = torch.randn(B, 32, G, G, G, dtype=torch.float64, requires_grad=True)
    y = ref_conv(x, S.from_taps(w))
    y.backward(dz[:, :3].double())
    wk = S.kernel_weights(w.to(dev), 32)
    wt = S.flip_weights(wk, 32, 3, 32)
    dzc = S.pack_cl(dz[:, :3].to(dev), 32)
    dx = S.conv3_fwd(dzc, wt, None, S.MODE_OUT_CL, 32).cpu().double().permute(0, 4, 1, 2, 3)
    assert (dx - x.grad).abs().max() < 1e-4 * x.grad.abs().max()


@pytest.mark.parametrize("cin,B,G", [(4, 2, 32), (4, 1, 64), (16, 2, 32), (32, 2, 32)])
def test_conv_weight_gradient(dev, cin, B, G):
    """Weight and bias gradients against float64 autograd; cin 4 is the input
    layer's own kernel (columns = tap x channel pairs, one partial per wave)."""
    from augur_amd import surrogate as S
    g = torch.Generator().manual_seed(11 + cin + G)
    x = bf16_round(torch.randn(B, cin, G, G, G, generator=g))
    dz = bf16_round(torch.randn(B, 32, G, G, G, generator=g))
    wt = torch.zeros(32, cin, 3, 3, 3, dtype=torch.float64, requires_grad=True)
    bt = torch.zeros(32, dtype=torch.float64, requires_grad=True)
    ref_conv(x, wt, bt).backward(dz.double())
    xc = S.pack_cl(x.to(dev), cin)
    dzc = S.pack_cl(dz.to(dev), 32)
    for nwg in (1, 37, 256):
        dw, db = S.conv3_wgrad(xc, dzc, nwg=nwg)
        ref = S.to_taps(wt.grad)
        assert (dw.cpu().double() - ref).abs().max() < 1e-4 * ref.abs().max()
        assert (db.cpu().double() - bt.grad).abs().max() < 1e-4 * bt.grad.abs().max()


def test_localnet3d_forward_backward_vs_autograd(dev):
    from augur_amd import surrogate as S
    from oracle.torch_cpu import localnet3d
    B, G = 1, 32
    net = S.LocalNet3D(cin=4, cout=3, layers=5, device=dev, init_std=0.05)
    g = torch.Generator().manual_seed(3)
    x = torch.randn(B, 4, G, G, G, generator=g)
    y = net.forward(x.to(dev), save=True)
    tw = [(w.detach().cpu().double().requires_grad_(True), b.detach().cpu().double().requires_grad_(True))
          for w, b in net.torch_weights()]
    # the reference sees the same BF16-rounded weights and input
    with torch.no_grad():
        for (w, b) in tw:
            w.copy_(bf16_round(w.float()).double())
    yr = localnet3d(bf16_round(x).double(), [w for w, _ in tw], [b for _, b in tw])
    rel = ((y.cpu().double() - yr) ** 2).sum().sqrt() / (yr ** 2).sum().sqrt()
    assert rel < 3e-2, rel
    dy = torch.randn(yr.shape, generator=g, dtype=torch.float64)
    yr.backward(dy)
    grads = net.backward(dy.float().to(dev))
    for k, ((dw, db), (w, b)) in enumerate(zip(grads, tw)):
        ref = S.to_taps(w.grad)
        rel = ((dw.cpu().double() - ref) ** 2).sum().sqrt() / (ref ** 2).sum().sqrt()
        assert rel < 5e-2, (k, float(rel))
        relb = ((db.cpu().double() - b.grad) ** 2).sum().sqrt() / (b.grad ** 2).sum().sqrt()
        assert relb < 5e-2, (k, float(relb))


def test_adam_matches_torch(dev):
    from augur_amd import surrogate as S
    g = torch.Generator().manual_seed(5)
    p0 = torch.randn(1000, generator=g)
    p = p0.clone().to(dev)
    m = torch.zeros_like(p)
    v = torch.zeros_like(p)
    tp = p0.clone().double().requires_grad_(True)
    opt = torch.optim.Adam([tp], lr=3e-4)
    for step in range(1, 4):
        gr = torch.randn(1000, generator=g)
        S.adam_update(p, gr.to(dev), m, v, 3e-4, step)
        tp.grad = gr.double()
        opt.step()
    assert (p.cpu().double() - tp.detach()).abs().max() < 1e-6


# layer signatures of the surrogate's conv calls: (cin, mode, cout, with addend)
TILE_LAYERS = [(16, 1, 32, False), (32, 1, 32, False), (32, 2, 32, False), (32, 0, 3, False),
               (32, 0, 32, False), (32, 3, 32, True), (32, 3, 16, False), (16, 3, 16, True)]


@pytest.mark.parametrize("cin,mode,cout,with_add", TILE_LAYERS)
def test_every_conv_tile_matches_reference(dev, cin, mode, cout, with_add):
    """Every (kernel family, z planes per workgroup) tile the library offers for a
    layer agrees with the float64 conv; within a family every zc gives the same bits."""
    from augur_amd import autotune as A
    from augur_amd import surrogate as S
    g = torch.Generator().manual_seed(100 * cin + 10 * mode + cout)
    B, G = 2, 64
    x = bf16_round(torch.randn(B, cin, G, G, G, generator=g))
    w = bf16_round(torch.randn(27, 32, cin, generator=g) * 0.1)
    w[:, cout:] = 0
    b = torch.randn(32, generator=g) * 0.1
    xc = S.pack_cl(x.to(dev), cin)
    wk = S.kernel_weights(w.to(dev), cin)
    ref = ref_conv(x, S.from_taps(w), None if mode == 3 else b)[:, :cout]   # (B, cout, G, G, G)
    add = torch.randn(B, G, G, G, cout, generator=g) if with_add else None
    if mode == 3:
        ref = ref.permute(0, 2, 3, 4, 1) + (add.double() if with_add else 0)
    elif mode in (1, 2):
        z = ref.permute(0, 2, 3, 4, 1)
        ref = F.gelu(z, approximate='tanh') + (x.double().permute(0, 2, 3, 4, 1) if mode == 2 else 0)
    cands = A.candidates(cin, mode, cout, G)
    fams = {v for v, _ in cands}
    assert A.TILE_ZM in fams and len(cands) == 4 * len(fams)
    if cin == 32 and cout <= 16 and mode in (0, 3):
        assert A.TILE_NARROW in fams
    if cin == 32 and cout == 32 and mode in (1, 2, 3):
        assert A.TILE_SPLIT in fams
    tol = (1e-4 if mode in (0, 3) else 1.5e-2) * ref.abs().max()
    first = {}
    for tile in cands:
        y = S.conv3_fwd(xc, wk, b.to(dev), mode, cout, addend=None if add is None else add.to(dev), tile=tile)
        err = (y.cpu().double() - ref).abs().max()
        assert err < tol, (tile, float(err))
        if tile[0] in first:
            assert torch.equal(y, first[tile[0]]), tile
        else:
            first[tile[0]] = y
    # the default entry point is the library's default family at zc 16
    y0 = S.conv3_fwd(xc, wk, b.to(dev), mode, cout, addend=None if add is None else add.to(dev))
    assert any(torch.equal(y0, v) for v in first.values())


def test_conv_tile_rejects_unavailable_family(dev):
    from augur_amd import _lib
    from augur_amd import autotune as A
    from augur_amd import surrogate as S
    x = torch.zeros(1, 32, 32, 32, 16, dtype=torch.bfloat16, device=dev)
    wk = torch.zeros(27, 32, 16, dtype=torch.bfloat16, device=dev)
    b = torch.zeros(32, device=dev)
    with pytest.raises(_lib.AugurError, match="tile variant"):
        S.conv3_fwd(x, wk, b, S.MODE_GELU, 32, tile=(A.TILE_SPLIT, 16))   # split needs 32 input channels
    with pytest.raises(_lib.AugurError, match="zc"):
        S.conv3_fwd(x, wk, b, S.MODE_GELU, 32, tile=(A.TILE_ZM, 24))


def test_autotuned_localnet_matches_default_tiles(dev, tmp_path):
    """autotune() times every candidate of every conv call of a forward and
    backward pass; the tuned net's outputs and gradients agree with the
    default tiles', and the table round-trips through its JSON file."""
    from augur_amd import autotune as A
    from augur_amd.surrogate import LocalNet3D
    B, G = 2, 64
    net = LocalNet3D(cin=4, cout=3, device=dev, init_std=0.05)
    g = torch.Generator(device=dev).manual_seed(3)
    x = torch.randn(B, 4, G, G, G, device=dev, generator=g)
    dy = torch.randn(B, 3, G, G, G, device=dev, generator=g)
    y_def = net.forward(x, save=True)
    g_def = net.backward(dy)
    tuner = net.autotune(B, G, reps=2)
    rep = tuner.report()
    # 6 forward calls: conv_in (the 4-channel packed-tap layer), 4 residual layers with one shape,
    # conv_out; backward: the output layer's data gradient, the residual layers' (with addend), the
    # last of them writing dz only; two weight-gradient shapes (4 and 32 input channels)
    fwd = [k for k in rep if k.startswith("fwd:")]
    wg = [k for k in rep if k.startswith("wgrad:")]
    assert len(fwd) == 6 and len(wg) == 2, sorted(rep)
    assert any(k.startswith("fwd:cin4:") for k in fwd) and any(k.startswith("wgrad:cin4:") for k in wg)
    assert any(k.endswith(":noy") for k in fwd)
    for k, v in rep.items():
        assert v["best"] in v["ms"] and min(v["ms"].values()) == v["ms"][v["best"]], (k, v)
    y_t = net.forward(x, save=True)
    g_t = net.backward(dy)
    assert (y_t - y_def).abs().max() <= 1e-2 * y_def.abs().max() + 1e-6
    for (dw0, db0), (dw1, db1) in zip(g_def, g_t):
        assert (dw1 - dw0).abs().max() <= 2e-2 * dw0.abs().max() + 1e-6
        assert (db1 - db0).abs().max() <= 2e-2 * db0.abs().max() + 1e-6
    path = str(tmp_path / "tiles.json")
    tuner.save(path)
    t2 = A.ConvTuner(path=path)
    assert t2.table == tuner.table


def test_four_channel_training_path_matches_padded_image(dev):
    """Training on the 4-channel image (packed-tap input layer, 4-channel weight
    gradient) and on the zero-padded 16-channel image give the same loss
    gradients within BF16 rounding of the activations."""
    from augur_amd import surrogate as S
    B, G = 2, 32
    net = S.LocalNet3D(cin=4, cout=3, device=dev, init_std=0.05)
    g = torch.Generator(device=dev).manual_seed(17)
    x = torch.randn(B, 4, G, G, G, device=dev, generator=g)
    dy = torch.randn(B, 3, G, G, G, device=dev, generator=g)
    assert net.cpack == 4
    assert torch.equal(S.pack_cl(x, 4), S.pack_cl(x, 16)[..., :4])   # the 4-channel pack kernel
    y4 = net.forward_packed(S.pack_cl(x, 4), save=True)
    g4 = net.backward(dy)
    y16 = net.forward_packed(S.pack_cl(x, 16), save=True)
    g16 = net.backward(dy)
    assert (y4 - y16).abs().max() <= 1e-2 * y16.abs().max()
    for (dw4, db4), (dw16, db16) in zip(g4, g16):
        assert dw4.shape == dw16.shape and db4.shape == db16.shape
        assert (dw4 - dw16).abs().max() <= 2e-2 * dw16.abs().max() + 1e-6
        assert (db4 - db16).abs().max() <= 2e-2 * db16.abs().max() + 1e-6
    # the input layer's weight gradient itself: same saved activations' gradient, both kernels
    dz = torch.randn(B, G, G, G, 32, device=dev, generator=g).to(torch.bfloat16)
    dwa, dba = S.conv3_wgrad(S.pack_cl(x, 4), dz)
    dwb, dbb = S.conv3_wgrad(S.pack_cl(x, 16), dz)
    assert (dwa - dwb[:, :, :4]).abs().max() <= 1e-5 * dwb.abs().max()
    assert (dba - dbb).abs().max() <= 1e-5 * dbb.abs().max()


@pytest.mark.parametrize("with_add", [False, True])
def test_fused_gelu_backward_data_gradient(dev, with_add):
    """Mode 4 = mode 3 (data gradient, + addend) followed by augur_gelu_bwd on
    the previous layer's pre-activation, in one launch: the same FP32 output
    and the same BF16 dz."""
    from augur_amd import surrogate as S
    g = torch.Generator().manual_seed(41 + int(with_add))
    B, G = 2, 64
    dz_in = bf16_round(torch.randn(B, G, G, G, 32, generator=g)).to(dev).to(torch.bfloat16)
    w = bf16_round(torch.randn(27, 32, 32, generator=g) * 0.1)
    wk = S.kernel_weights(w.to(dev), 32)
    z_prev = (2.0 * torch.randn(B, G, G, G, 32, generator=g)).to(dev).to(torch.bfloat16)
    add = torch.randn(B, G, G, G, 32, generator=g).to(dev) if with_add else None
    y3 = S.conv3_fwd(dz_in, wk, None, S.MODE_OUT_CL, 32, addend=add)
    dz3 = S.gelu_bwd(y3, z_prev)
    y4, dz4 = S.conv3_fwd(dz_in, wk, None, S.MODE_OUT_CL_DGELU, 32, addend=add, z_in=z_prev)
    assert torch.equal(y4, y3)
    y5, dz5 = S.conv3_fwd(dz_in, wk, None, S.MODE_OUT_CL_DGELU, 32, addend=add, z_in=z_prev, want_y=False)
    assert y5 is None and torch.equal(dz5, dz4)
    d = (dz4.float() - dz3.float()).abs()
    assert float(d.max()) <= float((2.0 ** -7 * dz3.float().abs()).max()) + 1e-30
    assert float((dz4 != dz3).float().mean()) < 1e-3


def test_input_layer_four_channels(dev):
    """The 4-channel input layer (taps packed into the MFMA K dimension) against
    the float64 conv, and against the same layer run on 16 padded channels."""
    from augur_amd import autotune as A
    from augur_amd import surrogate as S
    g = torch.Generator().manual_seed(44)
    B, G = 2, 64
    x = bf16_round(torch.randn(B, 4, G, G, G, generator=g))
    w = bf16_round(torch.randn(27, 32, 4, generator=g) * 0.1)
    b = torch.randn(32, generator=g) * 0.1
    ref = ref_conv(x, S.from_taps(w), b).permute(0, 2, 3, 4, 1)
    x4, x16 = S.pack_cl(x.to(dev), 4), S.pack_cl(x.to(dev), 16)
    w4, w16 = S.kernel_weights(w.to(dev), 4), S.kernel_weights(w.to(dev), 16)
    z4 = torch.empty(B, G, G, G, 32, dtype=torch.bfloat16, device=dev)
    z16 = torch.empty_like(z4)
    h16 = S.conv3_fwd(x16, w16, b.to(dev), S.MODE_GELU, 32, z_out=z16)
    first = None
    for tile in A.candidates(4, S.MODE_GELU, 32, G):
        h4 = S.conv3_fwd(x4, w4, b.to(dev), S.MODE_GELU, 32, z_out=z4, tile=tile)
        assert (z4.cpu().double() - ref).abs().max() < 1e-2 * ref.abs().max()
        assert (h4.cpu().double() - F.gelu(ref, approximate='tanh')).abs().max() < 1e-2 * ref.abs().max()
        assert (h4.float() - h16.float()).abs().max() <= 2 ** -7 * h16.float().abs().max()
        if first is None:
            first = h4.clone()
        assert torch.equal(h4, first), tile
