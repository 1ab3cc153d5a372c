"""Neural-texture decoder on the GPU.

- the BF16 matrix-core GEMM against float64 products of the same
  bfloat16-rounded operands (all transpositions, epilogues, ragged sizes,
  split-K), and bfloat16 operands bit-equal to float32 ones;
- the decoder's forward and backward against the oracle's float64 restatement;
- end to end on a config-1-style scene: textures refined by the plain method,
  a decoder fitted to them, PSNR of decoded vs explicit renders, memory, and the
  gain over the same memory spent without the network.

GEMM tolerance: the kernel accumulates in float32, the reference in float64, on
identical bfloat16 operands, so they agree to float32 accumulation error
(1e-5 relative to the largest entry)."""
import pytest
import torch
import torch.nn.functional as F

import textured_gaussians_amd as tg
from oracle.neural import gemm_reference, mlp_backward, mlp_forward
from textured_gaussians_amd import neural, scenes
from textured_gaussians_amd.cuda import _ops

from gpu_util import DEV

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a, b = a.double().cpu(), b.double().cpu()
    return ((a - b).abs().max() / b.abs().max().clamp_min(1e-30)).item()


@pytest.mark.parametrize("ta", [False, True])
@pytest.mark.parametrize("tb", [False, True])
@pytest.mark.parametrize("epi", ["none", "relu", "sigmoid", "relu_mask"])
def test_gemm_matches_bf16_reference(ta, tb, epi):
    g = torch.Generator().manual_seed(1)
    M, N, K = 77, 130, 45  # ragged in every dimension
    a = torch.randn((K, M) if ta else (M, K), generator=g)
    b = torch.randn((N, K) if tb else (K, N), generator=g)
    bias = torch.randn(N, generator=g)
    mask = torch.randn(M, N, generator=g).clamp_min(0)
    code = dict(none=_ops.EPI_NONE, relu=_ops.EPI_RELU, sigmoid=_ops.EPI_SIGMOID, relu_mask=_ops.EPI_RELU_MASK)[epi]
    out = _ops.gemm_bf16(a.to(DEV), b.to(DEV), trans_a=ta, trans_b=tb, epilogue=code,
                         bias=None if epi == "relu_mask" else bias.to(DEV),
                         mask=mask.to(DEV) if epi == "relu_mask" else None)
    ref = gemm_reference(a, b, ta, tb, None if epi == "relu_mask" else bias, epi, mask)
    assert _rel(out, ref) < 1e-5


@pytest.mark.parametrize("splits", [1, 3, 16])
def test_gemm_split_k_accumulates(splits):
    g = torch.Generator().manual_seed(2)
    rows, I, O = 5000, 40, 70  # dW = X^T G with the batch as K
    X = torch.randn(rows, I, generator=g)
    G = torch.randn(rows, O, generator=g)
    C0 = torch.randn(I, O, generator=g)
    out = _ops.gemm_bf16(X.to(DEV), G.to(DEV), trans_a=True, epilogue=_ops.EPI_ATOMIC, out=C0.to(DEV).clone(),
                         k_splits=splits)
    ref = C0.double() + gemm_reference(X, G, True, False)
    assert _rel(out, ref) < 1e-5


@pytest.mark.parametrize("ta", [False, True])
@pytest.mark.parametrize("tb", [False, True])
@pytest.mark.parametrize("shape", [(77, 130, 45), (4099, 257, 96)])
def test_gemm_bfloat16_operands_equal_float_operands(ta, tb, shape):
    """An operand given as bfloat16 is staged as is; given as float32 it is rounded
    to nearest even when staged, as torch's conversion rounds: the products agree
    bit for bit.  Leading dimensions of 45, 130, 257 and 4099 take the element
    loads, 96 the 16-byte loads; the second shape's tile count (65 x 5) is not a
    multiple of 8."""
    g = torch.Generator().manual_seed(11)
    M, N, K = shape
    a = torch.randn((K, M) if ta else (M, K), generator=g).to(DEV)
    b = torch.randn((N, K) if tb else (K, N), generator=g).to(DEV)
    mask = torch.randn(M, N, generator=g).clamp_min(0).to(DEV)
    for epi, kw in ((_ops.EPI_NONE, {}), (_ops.EPI_RELU_MASK, dict(mask=mask))):
        ref = _ops.gemm_bf16(a, b, trans_a=ta, trans_b=tb, epilogue=epi, **kw)
        for a_, b_ in ((a.bfloat16(), b), (a, b.bfloat16()), (a.bfloat16(), b.bfloat16())):
            out = _ops.gemm_bf16(a_, b_, trans_a=ta, trans_b=tb, epilogue=epi, **kw)
            assert torch.equal(out, ref), (epi, a_.dtype, b_.dtype)
    full = gemm_reference(a.cpu(), b.cpu(), ta, tb)
    assert _rel(_ops.gemm_bf16(a.bfloat16(), b.bfloat16(), trans_a=ta, trans_b=tb), full) < 1e-5


def test_gemm_split_k_bfloat16_operand_and_empty_k():
    g = torch.Generator().manual_seed(12)
    X = torch.randn(9000, 128, generator=g).to(DEV)
    G = torch.randn(9000, 1024, generator=g).to(DEV)
    ref = gemm_reference(X.cpu(), G.bfloat16().float().cpu(), True, False)
    out = _ops.gemm_bf16(X, G.bfloat16(), trans_a=True, epilogue=_ops.EPI_ATOMIC,
                         out=torch.zeros(128, 1024, device=DEV), k_splits=3)
    assert _rel(out, ref) < 1e-5
    # K = 0: the product is empty and the epilogue sees the bias alone
    bias = torch.randn(70, generator=g).to(DEV)
    out = _ops.gemm_bf16(torch.empty(33, 0, device=DEV), torch.empty(0, 70, device=DEV), epilogue=_ops.EPI_RELU,
                         bias=bias)
    assert torch.equal(out, bias.clamp_min(0).expand(33, 70))


def test_sigmoid_bwd_colsum_bfloat16_output():
    g = torch.Generator().manual_seed(13)
    o = torch.rand(3000, 1024, generator=g).to(DEV)
    go = torch.randn(3000, 1024, generator=g).to(DEV)
    g32, db32 = _ops.sigmoid_bwd_colsum(go, o)
    g16, db16 = _ops.sigmoid_bwd_colsum(go, o, g_dtype=torch.bfloat16)
    assert g16.dtype == torch.bfloat16 and torch.equal(g16, g32.bfloat16())
    assert _rel(db16, db32.double()) < 1e-6  # sums of the same float32 values, atomic order aside
    assert _rel(db32, (go.double() * o * (1 - o)).sum(0)) < 1e-5


def test_colsum_and_sigmoid_bwd():
    g = torch.Generator().manual_seed(3)
    G = torch.randn(9000, 130, generator=g)
    db = torch.randn(130, generator=g)
    out = _ops.colsum_(G.to(DEV), db.to(DEV).clone())
    assert _rel(out, db.double() + G.double().sum(0)) < 1e-5
    # 4-vector path (column count a multiple of 4), rows not a multiple of the block's
    for rows in (1, 37, 9001):
        G4 = torch.randn(rows, 128, generator=g)
        db4 = torch.randn(128, generator=g)
        out = _ops.colsum_(G4.to(DEV), db4.to(DEV).clone())
        assert _rel(out, db4.double() + G4.double().sum(0)) < 1e-5, rows
    o = torch.rand(1000, generator=g)
    go = torch.randn(1000, generator=g)
    assert _rel(_ops.sigmoid_bwd(go.to(DEV), o.to(DEV)), go * o * (1 - o)) < 1e-6


def test_decoder_forward_backward_match_oracle():
    dec = neural.NeuralTextureDecoder(device=DEV, seed=3)
    g = torch.Generator().manual_seed(4)
    z = torch.randn(300, 32, generator=g).to(DEV).requires_grad_(True)
    out = neural.decode_mlp(z, dec.W1, dec.b1, dec.W2, dec.b2, dec.W3, dec.b3)
    gout = torch.randn(out.shape, generator=g).to(DEV)
    (out * gout).sum().backward()
    P = [t.detach().double().cpu() for t in (z, dec.W1, dec.b1, dec.W2, dec.b2, dec.W3, dec.b3)]
    ref_out, h1, h2 = mlp_forward(*P)
    assert _rel(out.detach(), ref_out) < 1e-4
    ref = mlp_backward(gout.double().cpu(), P[0], P[1], P[3], P[5], h1, h2, ref_out)
    mine = [z.grad, dec.W1.grad, dec.b1.grad, dec.W2.grad, dec.b2.grad, dec.W3.grad, dec.b3.grad]
    for name, m, r in zip(["dz", "dW1", "db1", "dW2", "db2", "dW3", "db3"], mine, ref):
        # float32 activations re-rounded to bfloat16 can land on the neighbouring value
        # where float64 does not: compare in the L2 norm
        e = ((m.double().cpu() - r).norm() / r.norm()).item()
        assert e < 2e-3, (name, e)


@pytest.mark.parametrize("n", [1, 63, 64, 65, 300, 4096 + 17])
def test_fused_decode_matches_oracle_and_gemm_path(n):
    dec = neural.NeuralTextureDecoder(device=DEV, seed=5)
    g = torch.Generator().manual_seed(6)
    z = (torch.randn(n, 32, generator=g) * 2).to(DEV)
    packed = _ops.decoder_pack(dec.W1.detach(), dec.W2.detach(), dec.W3.detach())
    out = _ops.decode_textures(z, packed, dec.b1.detach(), dec.b2.detach(), dec.b3.detach())
    assert out.shape == (n, 16, 16, 4)
    P = [t.detach().double().cpu() for t in (z, dec.W1, dec.b1, dec.W2, dec.b2, dec.W3, dec.b3)]
    ref, _, _ = mlp_forward(*P)
    # The kernel accumulates in float32, the oracle in float64: a hidden activation within
    # ~1e-7 of a bfloat16 rounding boundary rounds to the neighbouring value (one bfloat16
    # ulp, 2^-8 relative), which moves that latent's outputs by up to ~1e-3.  The other
    # latents agree to float32 accumulation error.
    d = (out.reshape(n, -1).double().cpu() - ref).abs()
    assert float(d.max()) < 2e-3 and float(d.mean()) < 1e-6
    assert float((d.amax(1) > 1e-5).double().mean()) < 0.02
    with torch.no_grad():
        gemm = neural.decode_mlp(z, dec.W1, dec.b1, dec.W2, dec.b2, dec.W3, dec.b3)
    dg = (out.reshape(n, -1) - gemm).abs()
    assert float(dg.max()) < 2e-3 and float((dg.amax(1) > 1e-5).float().mean()) < 0.02
    # inference through the module (no autograd) takes the fused kernel
    with torch.no_grad():
        assert torch.equal(dec(z), out)


def test_fused_decode_writes_only_its_rows():
    dec = neural.NeuralTextureDecoder(device=DEV, seed=7)
    z = torch.randn(100, 32, device=DEV)
    buf = torch.full((101, 16, 16, 4), -7.0, device=DEV)
    dec.decode(z, out=buf[:100])
    assert float(buf[100].max()) == -7.0 and float(buf[:100].min()) > 0.0


def test_decoder_memory_and_output_range():
    dec = neural.NeuralTextureDecoder(device=DEV)
    assert sum(p.numel() for p in dec.parameters()) == 32 * 128 + 128 + 128 * 128 + 128 + 128 * 1024 + 1024
    t = dec(torch.randn(10, 32, device=DEV))
    assert t.shape == (10, 16, 16, 4) and (t > 0).all() and (t < 1).all()


def _neural_pipeline(n=10_000, refine_steps=300, pre_steps=3000, fine_steps=4000, threshold=0.125):
    import math
    sc = scenes.neural_scene(n=n, size=256, n_cams=16).to(DEV)
    bg = torch.zeros(16, 3, device=DEV)
    vm_h = scenes.look_at_viewmat(scenes.sphere_cameras(1, offset=1.0)[0])[None].to(DEV)  # a view not trained on

    def render(tex, held_out=False):
        vm, K, b = (vm_h, sc.Ks[:1], bg[:1]) if held_out else (sc.viewmats, sc.Ks, bg)
        return tg.rasterization_textured_gaussians(sc.means, sc.quats, sc.scales, sc.opacities, sc.colors, tex,
                                                   vm, K, sc.width, sc.height, sh_degree=None, backgrounds=b)[0]

    with torch.no_grad():
        targets = render(sc.textures)
    # 1. the plain method's refinement from neutral textures (no colour, full alpha)
    init = torch.zeros_like(sc.textures)
    init[..., 3] = 1.0
    refined, seen = neural.refine_textures(render, init, targets, steps=refine_steps)
    w = neural.observation_weights(seen)
    with torch.no_grad():
        ref_img, ref_held = render(refined), render(refined, True)
    # 2. decoder pretrained on the refined textures; 3. fallback set; 4. render-space fine-tune
    dec, z, hist = neural.fit_decoder(refined, steps=pre_steps, weights=w)
    err = neural.texture_errors(dec, z, refined, w)
    err_q = [round(float(x), 4) for x in torch.quantile(err, torch.tensor([0.5, 0.9, 0.95, 0.99], device=DEV))]
    fb = err > threshold
    z, hist2 = neural.fit_decoder_render(dec, z, render, ref_img, steps=fine_steps, explicit=refined, fallback=fb)
    stats = neural.neural_memory(dec, z, refined, fb)
    with torch.no_grad():
        tex_n = neural.mixed_textures(dec, z, refined, fb)
        p_neural = neural.psnr(render(tex_n), ref_img)
        p_held = neural.psnr(render(tex_n, True), ref_held)
        p_refine = neural.psnr(ref_img, targets)

    # the same per-Gaussian memory without the network: 2x2 RGBA float32 textures (64 B, like a
    # 32-d bfloat16 latent), bilinearly upsampled, refined on the same renders the same way
    def up(t2):
        return F.interpolate(t2.permute(0, 3, 1, 2), size=(16, 16), mode="bilinear",
                             align_corners=False).permute(0, 2, 3, 1).contiguous()

    small0 = F.interpolate(refined.permute(0, 3, 1, 2), size=(2, 2), mode="area").permute(0, 2, 3, 1).contiguous()
    small, _ = neural.refine_textures(lambda t2: render(up(t2)), small0, ref_img, steps=fine_steps // 3)
    with torch.no_grad():
        p_base = neural.psnr(render(up(small)), ref_img)
    return dict(p_neural=p_neural, p_held_out=p_held, p_base=p_base, p_refine=p_refine, err_q=err_q,
                hist=hist + hist2, **stats)


def test_neural_textures_end_to_end():
    r = _neural_pipeline()
    print({k: v for k, v in r.items() if k != "hist"}, r["hist"][-3:])
    assert r["p_refine"] > 40  # the refinement reproduces the ground-truth renders
    assert 0.0 < r["fallback_frac"] <= 0.10
    assert r["memory_ratio"] >= 8.0
    assert r["p_neural"] >= 40.0
    assert r["p_neural"] > r["p_base"] + 5.0  # what the network gains at equal memory
