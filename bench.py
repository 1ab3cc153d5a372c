"""Benchmark: textured-2DGS forward+backward at scene scale on one MI355X.

Workload (the headline configuration): N = 1,000,000 Gaussians, 50x50 RGBA
float32 textures per Gaussian (40 GB), SH degree 3, one 1297x840 camera (fov
60 deg) at radius 5 inside the cloud, 16x16 tiles, white background, L1 loss
to a seeded random target.  A step is one full render-path iteration through
the public API: projection -> tile intersection + sort -> SH -> textured
rasterizer forward -> L1 loss -> backward through rasterizer, SH and
projection (gradients for means, quats, scales, opacities, SH coefficients and
textures; ``.grad`` is reset to None before each step, as an optimizer's
zero_grad does).  Inputs are resident in HBM before the timed region.

Prints ONE JSON line (rank 0).  ``roofline`` is the rasterizer pair: the
algorithmic bytes of forward and backward over their summed kernel time, as
launched by the product path (which zero-fills the 40 GB texture gradient along
the two rasterizer walks); the same pair with the fill done by one dense launch
beforehand is reported beside it (``raster_fill_free``), with the fill bytes as
their own figure.  ``cpu_baseline`` times the CPU oracle on a bounded tile sample
(median of 3 runs after a warm-up) and extrapolates.
"""
import argparse
import json
import math
import os
import platform
import sys
import time

import torch

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "fwd+bwd ms/iter at 1M textured 2DGS, 1297x840; % HBM roofline vs plain"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1_000_000, help="Gaussians (default: the headline 1M)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--cpu-tiles", type=int, default=40, help="tiles in the CPU oracle sample")
    ap.add_argument("--cpu-runs", type=int, default=3, help="timed CPU runs (after one warm-up); median")
    ap.add_argument("--no-neural", action="store_true", help="skip the neural-texture leg")
    ap.add_argument("--no-config5", action="store_true", help="skip the full-training-iteration leg")
    ap.add_argument("--config5-only", action="store_true", help="run only the full-training-iteration leg")
    ap.add_argument("--no-layouts", action="store_true", help="skip the mipmapped2 and packed-layout legs")
    ap.add_argument("--fwd-no-cull", action="store_true",
                    help="forward without the candidate pre-pass (every list entry visited; same outputs)")
    ap.add_argument("--lds-atomics", action="store_true",
                    help="also measure the GPU's LDS atomic-add rates (float32, float64, uint32, plain)")
    ap.add_argument("--grad-paths", action="store_true",
                    help="also time the texture-gradient variants (records + gather, hybrid) on the headline scene")
    return ap.parse_args()


def make_inputs(n, dev):
    from textured_gaussians_amd import scenes
    sc = scenes.config2(n=n, texture_device=dev).to(dev)
    params = dict(means=sc.means, quats=sc.quats, scales=sc.scales,
                  opacities_logit=torch.logit(sc.opacities.clamp(1e-6, 1 - 1e-6)), sh=sc.colors, textures=sc.textures)
    for p in params.values():
        p.requires_grad_(True)
    return sc, params


def step(tg, sc, params, textured=True, **kw):
    for p in params.values():
        p.grad = None
    op = torch.sigmoid(params["opacities_logit"])
    if textured:
        out = tg.rasterization_textured_gaussians(
            params["means"], params["quats"], params["scales"], op, params["sh"], params["textures"], sc.viewmats,
            sc.Ks, sc.width, sc.height, sh_degree=3, backgrounds=sc.backgrounds, **kw)
    else:
        out = tg.rasterization_2dgs(params["means"], params["quats"], params["scales"], op, params["sh"], sc.viewmats,
                                    sc.Ks, sc.width, sc.height, sh_degree=3, backgrounds=sc.backgrounds)
    loss = (out[0] - sc.target).abs().mean()
    loss.backward()
    return out


def timed_steps(tg, sc, params, steps, warmup, textured, **kw):
    from textured_gaussians_amd.cuda import _ops
    for _ in range(warmup):
        step(tg, sc, params, textured, **kw)
    torch.cuda.synchronize()
    timer = {}
    _ops.set_kernel_timer(timer)
    t0 = time.perf_counter()
    out = None
    for _ in range(steps):
        out = step(tg, sc, params, textured, **kw)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    _ops.set_kernel_timer(None)
    kt = {k: sum(a.elapsed_time(b) for a, b in v) / len(v) for k, v in timer.items()}
    return dt * 1e3, kt, out


def texels_read(tg, sc, params):
    """U: distinct (Gaussian, texel) pairs the forward reads, counted with a byte map the
    forward marks (measurement hook of the library), on one extra iteration."""
    from textured_gaussians_amd.cuda import _ops
    tex = params["textures"]
    marks = torch.zeros(tex.shape[:3], dtype=torch.uint8, device=tex.device)
    _ops.set_texel_marks(marks)
    try:
        step(tg, sc, params, True)
        torch.cuda.synchronize()
    finally:
        _ops.set_texel_marks(None)
    U = int(marks.count_nonzero())
    del marks
    return U


def algorithmic_bytes(out, params, sc, textured, K=3, U=None):
    """Per-launch algorithmic HBM bytes of the rasterizer forward and backward.

    I = intersections, R = 76 B per intersection record (4 id + 8 mean2d + 36
    ray transform + 4 opacity + 12 colour + 12 normal), P = pixels,
    U = distinct (Gaussian, texel) pairs read (16 B each; `texels_read`),
    U_g = texels with a nonzero gradient, V = visible (camera, Gaussian) pairs.
      forward : I*R + 16U + 4*C*tiles + P*(4K+32)
      backward: I*R + 16U + P*(4K+36) + 32U_g + 2V*(68+4K)
    """
    meta = out[-1]
    I = meta["flatten_ids"].numel()
    C = sc.viewmats.shape[0]
    P = C * sc.width * sc.height
    tiles = meta["tile_width"] * meta["tile_height"]
    V = int((meta["radii"] > 0).sum())
    Ug = 0
    if textured:
        g = params["textures"].grad
        Ug = int((g != 0).any(-1).sum())  # the gradient is already resident
    U = Ug if U is None else U
    R = 76
    bf = I * R + 16 * U + 4 * C * tiles + P * (4 * K + 32)
    bb = I * R + 16 * U + P * (4 * K + 36) + 32 * Ug + 2 * V * (68 + 4 * K)
    return dict(I=I, P=P, V=V, U=U, U_g=Ug, tiles=tiles, bytes_fwd=bf, bytes_bwd=bb)


def stream_rates(dev):
    """Measured streaming rates on this device, GB/s of bytes moved (read + written):
    a device-to-device copy, and c = a + b (two reads, one write) over 4 GiB arrays."""
    n = 1 << 30
    a = torch.empty(n, dtype=torch.float32, device=dev)
    b = torch.empty_like(a)
    c = torch.empty_like(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    out = {}
    for name, fn, nbytes in (("copy", lambda: b.copy_(a), 2 * n * 4),
                             ("add", lambda: torch.add(a, b, out=c), 3 * n * 4)):
        for _ in range(2):
            fn()
        torch.cuda.synchronize()
        e0.record()
        for _ in range(5):
            fn()
        e1.record()
        torch.cuda.synchronize()
        out[name] = round(nbytes / (e0.elapsed_time(e1) / 5 * 1e-3) / 1e9, 1)
    del a, b, c
    return out


def host_cpus():
    """CPUs this process may use: its affinity set, capped by a cgroup CPU quota."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    try:
        quota, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if quota != "max":
            n = min(n, max(1, int(int(quota) / int(period))))
    except (OSError, ValueError):
        pass
    return n


def host_info():
    try:
        model = [l for l in open("/proc/cpuinfo") if l.startswith("model name")][0].split(":")[1].strip()
    except (OSError, IndexError):
        model = platform.processor()
    try:
        kb = int([l for l in open("/proc/meminfo") if l.startswith("MemTotal")][0].split()[1])
        ram = round(kb / 1024 ** 2, 1)
    except (OSError, IndexError, ValueError):
        ram = None
    return model, ram


def cpu_baseline(sc, meta_gpu, n_tiles_sample, runs=3):
    """The CPU oracle (float32 torch on the host, every CPU this process may use) on a
    bounded sample of one iteration: projection + SH forward/backward and tile
    intersection of all Gaussians, and the rasterizer forward+backward on every k-th
    tile (on the GPU's projected inputs), extrapolated to all tiles by the per-tile
    work sum(I_tile * 256).  One warm-up, then the median of ``runs`` timed runs."""
    import numpy as np
    from oracle import isect as O_isect
    from oracle.projection import project_2dgs
    from oracle.raster import rasterize_to_pixels
    from oracle.sh import spherical_harmonics

    threads = host_cpus()
    torch.set_num_threads(threads)
    f = torch.float32
    cpu = lambda t: t.detach().cpu().to(f) if t.is_floating_point() else t.detach().cpu()  # noqa: E731
    means, quats, scales = cpu(sc.means), cpu(sc.quats), cpu(sc.scales)
    viewmats, Ks = cpu(sc.viewmats), cpu(sc.Ks)
    sh, opac = cpu(sc.colors), cpu(sc.opacities)
    W, H = sc.width, sc.height
    tw, th = math.ceil(W / 16), math.ceil(H / 16)

    def front():
        # (1) projection + SH, forward and backward, all Gaussians
        t0 = time.perf_counter()
        m = means.clone().requires_grad_(True)
        q = quats.clone().requires_grad_(True)
        s = scales.clone().requires_grad_(True)
        shc = sh.clone().requires_grad_(True)
        radii, m2, dep, rt, nrm = project_2dgs(m, q, s, viewmats, Ks, W, H)
        dirs = m[None] - torch.inverse(viewmats)[:, None, :3, 3]
        col = torch.clamp_min(spherical_harmonics(3, dirs, shc[None], radii > 0) + 0.5, 0.0)
        ((m2.sum() + rt.sum() + nrm.sum() + col.sum())).backward()
        t1 = time.perf_counter()
        # (2) tile intersection + sort + offsets (numpy), all Gaussians
        _, ids, flat = O_isect.isect_tiles(m2.detach().numpy(), radii.numpy(), dep.detach().numpy(), 16, tw, th)
        offs = O_isect.isect_offset_encode(ids, 1, tw, th)
        t2 = time.perf_counter()
        return t1 - t0, t2 - t1, col.detach(), ids, flat, offs

    t_proj, t_isect, col, ids, flat, offs = front()
    # (3) the rasterizer sample: every k-th tile, compact sub-problem of the Gaussians it references
    offs_flat = offs.reshape(-1).astype(np.int64)
    I = ids.size
    ends = np.append(offs_flat[1:], I)
    per_tile = (ends - offs_flat).clip(min=0)
    n_t = offs_flat.size
    stride = max(1, n_t // n_tiles_sample)
    sample = list(range(stride // 2, n_t, stride))[:n_tiles_sample]
    sel = np.concatenate([flat[offs_flat[t]:ends[t]] for t in sample]) if sample else np.zeros(0, np.int32)
    uniq, inv = np.unique(sel, return_inverse=True)
    ut = torch.from_numpy(uniq.astype(np.int64))
    tex0 = sc.textures[ut.to(sc.textures.device)].detach().cpu().to(f)
    m2_0 = cpu(meta_gpu["means2d"])[0, ut][None]
    rt_0 = cpu(meta_gpu["ray_transforms"])[0, ut][None]
    nrm_0 = cpu(meta_gpu["normals"])[0, ut][None]
    new_flat = torch.from_numpy(inv.astype(np.int32))
    tiles_list, o = [], 0
    for t in sample:
        tiles_list.append((0, t // tw, t % tw, o, o + int(per_tile[t])))
        o += int(per_tile[t])
    loc_offs = torch.zeros(1, th, tw, dtype=torch.int32)
    tgt, bg = cpu(sc.target), cpu(sc.backgrounds)

    def raster():
        t0 = time.perf_counter()
        leaves = [x.clone().requires_grad_(True) for x in (m2_0, rt_0, col[0, ut][None], opac[ut][None], nrm_0,
                                                            tex0)]
        outr = rasterize_to_pixels(leaves[0], leaves[1], leaves[2], leaves[3], leaves[4], W, H, 16, loc_offs,
                                   new_flat, textures=leaves[5], backgrounds=bg, uv_grad="none",
                                   distort_grad="reference", tiles=tiles_list)
        loss = (outr["render_colors"] - tgt).abs().sum() / (H * W * 3)
        loss.backward()
        return time.perf_counter() - t0

    raster()  # warm-up (the front stages warmed up above)
    work_all = float((per_tile * 256).sum())
    work_smp = float((per_tile[sample] * 256).sum())
    totals, parts = [], []
    for _ in range(runs):
        tp, ti = front()[:2]
        ts = raster()
        tr = ts * (work_all / max(work_smp, 1.0))
        totals.append(tp + ti + tr)
        parts.append((tp, ti, ts))
    k = int(np.argsort(totals)[len(totals) // 2])
    tp, ti, ts = parts[k]
    model, ram = host_info()
    return dict(value=round(totals[k] * 1e3, 1), unit="ms/iter", cores=torch.get_num_threads(), kind="port",
                runs=runs, runs_ms=[round(t * 1e3, 1) for t in totals], cpu_model=model, ram_GB=ram,
                os_cpu_count=os.cpu_count(),
                sample=(f"extrapolated, median of {runs} runs after 1 warm-up: full projection+SH fwd/bwd "
                        f"({tp:.2f}s) and tile intersection ({ti:.2f}s) of all {means.shape[0]} Gaussians, "
                        f"rasterizer fwd+bwd on {len(sample)} of {n_t} tiles ({ts:.1f}s) scaled by "
                        f"sum(I_tile*256) ({work_all / max(work_smp, 1.0):.0f}x); fp32 torch oracle, scene opacities"))


def grad_paths_leg(tg, sc, params, steps, warmup):
    """The headline iteration under each texture-gradient variant of the autograd layer
    (all give the same gradients): per-pair records summed per row by the gather kernel,
    the backward reading the forward's pair samples (the product path: the backward walk
    zero-fills the unlisted rows), the same with the backward reading the texels, the same
    with the forward walk filling the unlisted rows, float atomics into the row-filled
    gradient, the same with one dense fill, and records for the rows with at least 64 pairs
    only (atomics for the rest).  The records variants run the rasterizer backward without
    texture-gradient atomics; the gather runs after it, inside the backward op."""
    from textured_gaussians_amd.cuda import _wrapper
    variants = [("records_gather", {}),
                ("records_texels", {"PAIR_SAMPLES": False}),
                ("records_forward_fills_unlisted", {"TEXTURE_GRAD_RECORDS": True, "FWD_FILL_RESIDUES": 0xFF}),
                ("atomics_fused_fill", {"TEXTURE_GRAD_RECORDS": False}),
                ("atomics_dense_fill", {"TEXTURE_GRAD_RECORDS": False, "FUSED_FILL": False}),
                ("records_rows_ge64", {"TEXTURE_GRAD_RECORDS": False, "HEAVY_ROW_PAIRS": 64})]
    res = {}
    for name, cfg in variants:
        saved = {k: getattr(_wrapper, k) for k in cfg}
        try:
            for k, v in cfg.items():
                setattr(_wrapper, k, v)
            ms, kt, out = timed_steps(tg, sc, params, steps, warmup, True)
            del out
        finally:
            for k, v in saved.items():
                setattr(_wrapper, k, v)
        res[name] = {"ms_per_iter": round(ms, 3), "kernels_ms": {k: round(v, 3) for k, v in kt.items()}}
    return res


DECODER_FLOP = 2 * (32 * 128 + 128 * 128 + 128 * 1024)  # per latent


def neural_leg(tg, n, steps, warmup, dev):
    """The neural-texture form at scene scale: the headline scene's N Gaussians with
    textures decoded from per-Gaussian 32-d latents (N(0,1)) by the 32 -> 128 -> 128
    -> 1024 decoder (16x16 RGBA).  Times (a) the fused inference kernel alone, HIP
    events on its stream, and (b) one training step: decode (matrix-core GEMMs under
    autograd) -> textured render -> L1 loss -> backward into latents, decoder and
    Gaussians."""
    from textured_gaussians_amd import neural, scenes
    from textured_gaussians_amd.cuda import _ops
    sc = scenes.config2(n=n, tex_res=None).to(dev)
    z = torch.randn(n, 32, generator=torch.Generator().manual_seed(2)).to(dev)
    dec = neural.NeuralTextureDecoder(device=dev, seed=3)
    packed = _ops.decoder_pack(dec.W1.detach(), dec.W2.detach(), dec.W3.detach())
    bias = [dec.b1.detach(), dec.b2.detach(), dec.b3.detach()]
    tex = torch.empty(n, 16, 16, 4, device=dev)
    for _ in range(2):
        _ops.decode_textures(z, packed, *bias, out=tex)
    torch.cuda.synchronize()
    reps = 10
    st = torch.cuda.current_stream(dev)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(reps):
        _ops.decode_textures(z, packed, *bias, out=tex)
    e1.record(st)
    torch.cuda.synchronize()
    t_dec = e0.elapsed_time(e1) / reps
    dec_bytes = n * (32 * 4 + 1024 * 4) + packed.numel() * 2 + 4 * (128 + 128 + 1024)
    del tex

    zp = z.clone().requires_grad_(True)
    gauss = dict(means=sc.means.clone().requires_grad_(True), quats=sc.quats.clone().requires_grad_(True),
                 scales=sc.scales.clone().requires_grad_(True),
                 opacities_logit=torch.logit(sc.opacities.clamp(1e-6, 1 - 1e-6)).requires_grad_(True),
                 sh=sc.colors.clone().requires_grad_(True))
    train = [zp, *dec.parameters(), *gauss.values()]

    def nstep():
        for p in train:
            p.grad = None
        textures = dec(zp)
        out = tg.rasterization_textured_gaussians(
            gauss["means"], gauss["quats"], gauss["scales"], torch.sigmoid(gauss["opacities_logit"]), gauss["sh"],
            textures, sc.viewmats, sc.Ks, sc.width, sc.height, sh_degree=3, backgrounds=sc.backgrounds)
        loss = (out[0] - sc.target).abs().mean()
        loss.backward()

    for _ in range(warmup):
        nstep()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        nstep()
    torch.cuda.synchronize()
    t_step = (time.perf_counter() - t0) / steps * 1e3
    achieved = dec_bytes / (t_dec * 1e-3) / 1e9
    return {
        "workload": f"N={n} Gaussians of the headline scene, 32-d latents -> 16x16x4 f32 textures, 1297x840",
        "decode_ms": round(t_dec, 4), "decode_tflops": round(n * DECODER_FLOP / (t_dec * 1e-3) / 1e12, 1),
        "decode_roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": round(achieved / HBM_PEAK_GBS, 4), "kernel": "decode_kernel",
                            "algorithmic_bytes": dec_bytes},
        "train_step_ms": round(t_step, 3),
        "texture_bytes": {"explicit_16x16_f32": n * 4096, "latents_bf16_plus_decoder": n * 64 + 151_680 * 2},
    }


def _scene_params(sc):
    params = dict(means=sc.means, quats=sc.quats, scales=sc.scales,
                  opacities_logit=torch.logit(sc.opacities.clamp(1e-6, 1 - 1e-6)), sh=sc.colors, textures=sc.textures)
    for p in params.values():
        p.requires_grad_(True)
    return params


def filters_leg(tg, steps, warmup, dev, n=1_000_000):
    """The mipmapped2 filter beside bilinear on the headline scene with 32x32 textures
    (mipmapped2 needs power-of-two textures): whole fwd+bwd iterations and the
    rasterizer kernels (HIP events)."""
    from textured_gaussians_amd import scenes
    sc = scenes.config2(n=n, tex_res=32, texture_device=dev).to(dev)
    params = _scene_params(sc)
    res = {"workload": f"N={n}, 32x32 RGBA f32 textures, SH3, 1297x840, fwd+bwd through the API"}
    for filt in ("bilinear", "mipmapped2"):
        ms, kt, out = timed_steps(tg, sc, params, steps, warmup, True, filtering=filt)
        res[filt] = {"ms_per_iter": round(ms, 3), "raster_ms": {k: round(v, 3) for k, v in kt.items()}}
        del out
    return res


def packed_leg(tg, steps, warmup, dev, n=500_000):
    """The packed layout beside the grid layout on the batched-camera scene (8 cameras
    800x800, N=500k, 32x32 textures): whole fwd+bwd iterations."""
    from textured_gaussians_amd import scenes
    sc = scenes.config3(n=n, texture_device=dev).to(dev)
    params = _scene_params(sc)
    res = {"workload": f"N={n}, 8 cameras 800x800, 32x32 RGBA f32 textures, SH3, fwd+bwd through the API"}
    for packed in (False, True):
        ms, kt, out = timed_steps(tg, sc, params, steps, warmup, True, packed=packed)
        res["packed" if packed else "grid"] = {"ms_per_iter": round(ms, 3),
                                               "raster_ms": {k: round(v, 3) for k, v in kt.items()},
                                               "entries": int(out[-1]["radii"].numel())}
        del out
    return res


def config5_leg(tg, steps, warmup, dev, n=2_000_000):
    """Full training iteration at scale: N=2M Gaussians, 64x64 RGBA textures stored
    bfloat16 with a float32 master (262 GB of texture state in HBM), 1920x1080, SH3;
    projection + tile sort + textured raster forward/backward + L1 + 0.2 D-SSIM + fused
    selective Adam over every parameter (the textures' step runs inside the backward from
    per-pair gradient records).  Times whole iterations and, with HIP events on their
    stream, the rasterizer and texture-Adam kernels."""
    from textured_gaussians_amd import scenes, train
    from textured_gaussians_amd.cuda import _ops
    sc = scenes.config5(n=n).to(dev)
    texopt = train.scene_texture_adam(n, 64, dev)
    state = train.init_state(sc.means, sc.quats, sc.scales, sc.opacities, sc.colors, texopt)
    torch.cuda.synchronize()
    torch.cuda.reset_peak_memory_stats()
    args = (sc.viewmats, sc.Ks, sc.width, sc.height, sc.target, sc.backgrounds)
    for _ in range(warmup):
        train.train_step(state, *args)
    torch.cuda.synchronize()
    timer = {}
    _ops.set_kernel_timer(timer)
    t0 = time.perf_counter()
    for _ in range(steps):
        loss = train.train_step(state, *args)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / steps * 1e3
    _ops.set_kernel_timer(None)
    kt = {k: sum(a.elapsed_time(b) for a, b in v) / len(v) for k, v in timer.items()}
    with torch.no_grad():
        radii = tg.fully_fused_projection_2dgs(state.params["means"], state.params["quats"],
                                               torch.exp(state.params["scales"]), sc.viewmats, sc.Ks, sc.width,
                                               sc.height)[0]
    vis = int((radii > 0).sum())
    adam_bytes = vis * 64 * 64 * 4 * 16  # hi, lo, m, v: 2 bytes each, read and written
    achieved = adam_bytes / (kt["texture_adam"] * 1e-3) / 1e9
    traffic = None
    tf = os.path.join(ROOT, "profiles", "texture_adam_traffic_config5.json")
    if os.path.exists(tf) and n == 2_000_000:
        traffic = json.load(open(tf)).get("bytes_per_launch")
    return {
        "workload": f"full training iteration, N={n}, 64x64 RGBA bf16 textures + f32 master, 1920x1080, SH3, "
                    "L1+0.2 D-SSIM, selective Adam on all parameters",
        "ms_per_step": round(ms, 3), "loss": round(float(loss), 6),
        "kernels_ms": {k: round(v, 3) for k, v in kt.items()},
        "visible": vis, "state_bytes": train.state_bytes(state),
        "peak_step_GB": round(torch.cuda.max_memory_allocated() / 1e9, 1),
        "texture_adam_roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                                  "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                                  "kernel": "tex_adam_kernel", "algorithmic_bytes": adam_bytes},
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)
    import textured_gaussians_amd as tg
    from textured_gaussians_amd.cuda._backend import lib
    lib()
    dev = torch.device("cuda:0")
    torch.manual_seed(0)
    if args.fwd_no_cull:
        from textured_gaussians_amd.cuda import _ops
        _ops.FWD_CULL = False
    if args.config5_only:
        print("[bench] full training iteration leg only", file=sys.stderr, flush=True)
        print(json.dumps({"config5": config5_leg(tg, args.steps, args.warmup, dev)}), flush=True)
        return
    sc, params = make_inputs(args.n, dev)
    torch.cuda.synchronize()
    print("[bench] inputs resident; timing textured steps", file=sys.stderr, flush=True)

    ms, kt, out = timed_steps(tg, sc, params, args.steps, args.warmup, True)
    ab = algorithmic_bytes(out, params, sc, True, U=texels_read(tg, sc, params))
    # the backward's texture gradient is complete after the gather of its records (when
    # the records path runs): the gather's time counts with the backward kernel's
    t_gat = kt.get("tex_gather", 0.0)
    t_bwd = kt.get("raster_bwd", float("nan")) + t_gat
    t_fwd = kt.get("raster_fwd", float("nan"))
    B = ab["bytes_fwd"] + ab["bytes_bwd"]
    achieved_pair = B / ((t_fwd + t_bwd) * 1e-3) / 1e9
    achieved_bwd = ab["bytes_bwd"] / (t_bwd * 1e-3) / 1e9
    from textured_gaussians_amd.cuda import _wrapper
    records = _wrapper._use_records(params["textures"], torch.float32)
    # the backward reads the forward's pair samples (the lean instantiation 2) on the records path
    samples = records and _wrapper.PAIR_SAMPLES is not False
    listed_rows = int(torch.unique(out[-1]["flatten_ids"].long() % args.n).numel())
    meta_ref = {k: out[-1][k].detach() for k in ("means2d", "ray_transforms", "normals")}
    del out
    # the same iteration with the texture gradient's zero-fill as one launch before the
    # forward (the unlisted rows on the records path, whose gather writes the listed
    # ones; all of it with float atomics): the rasterizer kernels without the fill, and
    # the fill alone
    print(f"[bench] textured {ms:.2f} ms/iter; timing the dense-fill variant", file=sys.stderr, flush=True)
    _wrapper.FUSED_FILL = False
    try:
        fms, fkt, fout = timed_steps(tg, sc, params, max(3, args.steps // 2), 1, True)
    finally:
        _wrapper.FUSED_FILL = True
    del fout
    ff_fwd = fkt.get("raster_fwd", float("nan"))
    ff_bwd = fkt.get("raster_bwd", float("nan")) + fkt.get("tex_gather", 0.0)
    t_fill = fkt.get("grad_fill", float("nan"))
    row_bytes = params["textures"][0].numel() * 4
    fill_bytes = (args.n - listed_rows if records else args.n) * row_bytes
    achieved_ff = B / ((ff_fwd + ff_bwd) * 1e-3) / 1e9
    fill_gbs = fill_bytes / (t_fill * 1e-3) / 1e9
    plain = None
    print("[bench] timing plain 2DGS", file=sys.stderr, flush=True)
    if not args.no_plain:
        pms, pkt, pout = timed_steps(tg, sc, params, max(3, args.steps // 2), 1, False)
        pab = algorithmic_bytes(pout, params, sc, False)
        pt = pkt.get("raster_fwd", float("nan")) + pkt.get("raster_bwd", float("nan"))
        p_ach = (pab["bytes_fwd"] + pab["bytes_bwd"]) / (pt * 1e-3) / 1e9
        plain = dict(ms_per_iter=round(pms, 3), raster_fwd_ms=round(pkt.get("raster_fwd", 0), 3),
                     raster_bwd_ms=round(pkt.get("raster_bwd", 0), 3), raster_hbm_frac=round(p_ach / HBM_PEAK_GBS, 4),
                     algorithmic_bytes=pab["bytes_fwd"] + pab["bytes_bwd"])
        del pout
    streams = stream_rates(dev)
    streams["zero_fill"] = round(fill_gbs, 1)
    practical = max(streams.values())
    traffic = None
    tf = os.path.join(ROOT, "profiles", "raster_traffic_config2.json")
    if os.path.exists(tf):
        traffic = json.load(open(tf)).get("pair_bytes_per_launch")
    line = {
        "metric": METRIC, "value": round(ms, 3), "unit": "ms/iter", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded scene of the headline workload)",
        "config": {"workload": "textured-2DGS fwd+bwd, N=1M, 50x50 RGBA f32 textures, SH3, 1297x840, 1 camera",
                   "n_gaussians": args.n, "width": sc.width, "height": sc.height, "tile": 16,
                   "n_isects": ab["I"], "visible": ab["V"], "texels_read": ab["U"], "texels_with_gradient": ab["U_g"]},
        "mpix_per_s": round(ab["P"] / (ms * 1e-3) / 1e6, 2),
        "raster_ms": {"fwd": round(t_fwd, 3), "bwd": round(t_bwd - t_gat, 3), "gather": round(t_gat, 3)},
        "texture_gradient": (("per-pair records + row gather (float64 LDS sums)"
                              + (", backward reading the forward's pair samples" if samples else "")) if records
                             else "float atomics into the row-filled gradient"),
        "raster_hbm_frac_fwd_bwd": round(achieved_pair / HBM_PEAK_GBS, 4),
        "plain_2dgs": plain,
        # the rasterizer pair without the fused fill, and the fill as one dense launch
        "raster_fill_free": {"ms_per_iter": round(fms, 3), "fwd_ms": round(ff_fwd, 3), "bwd_ms": round(ff_bwd, 3),
                             "achieved_GBs": round(achieved_ff, 1), "frac": round(achieved_ff / HBM_PEAK_GBS, 4),
                             "frac_of_stream_peak": round(achieved_ff / practical, 4)},
        "grad_fill": {"bytes": fill_bytes, "ms": round(t_fill, 3), "GBs": round(fill_gbs, 1),
                      "note": ("zero-fill of the texture-gradient rows absent from the tile lists (the gather writes "
                               "the listed rows whole); end-to-end bytes, outside B_f+B_b" if records else
                               "dense zero-fill of the 40 GB texture gradient (end-to-end bytes, outside B_f+B_b)")},
        "stream_GBs": streams,
        "roofline": {"bound": "hbm", "achieved": round(achieved_pair, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved_pair / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "frac_of_stream_peak": round(achieved_pair / practical, 4), "stream_peak_GBs": practical,
                     "kernel": ("raster_fwd_kernel<float,3,1> + "
                                + ("raster_bwd_kernel<float,3,1,2>" if samples else
                                   "raster_bwd_kernel<float,3,1,1>" if records else "raster_bwd_kernel<float,3,1,0>")
                                + (" + tex_gather2_kernel<6,10>" if records else "")),
                     "algorithmic_bytes": B, "bytes_fwd": ab["bytes_fwd"], "bytes_bwd": ab["bytes_bwd"],
                     "time_ms": round(t_fwd + t_bwd, 3)},
        "raster_bwd_only": {"achieved": round(achieved_bwd, 1), "frac": round(achieved_bwd / HBM_PEAK_GBS, 4)},
    }
    if args.lds_atomics:
        from textured_gaussians_amd.cuda import _ops
        rates = _ops.lds_atomic_rates(dev)
        line["lds_atomics"] = {"unit": "lanes per clock per CU", "addresses": "random, 40 KB LDS buffer",
                               **{k: round(v, 3) for k, v in rates.items()}}
    if args.grad_paths:
        print("[bench] texture-gradient variants", file=sys.stderr, flush=True)
        line["grad_paths"] = grad_paths_leg(tg, sc, params, max(3, args.steps // 2), 2)
    if not args.no_cpu_baseline:
        print("[bench] CPU oracle baseline (bounded sample)", file=sys.stderr, flush=True)
        try:
            line["cpu_baseline"] = cpu_baseline(sc, meta_ref, args.cpu_tiles, args.cpu_runs)
        except MemoryError as e:  # pragma: no cover
            line["cpu_baseline"] = {"error": str(e)}
    del sc, params  # the other legs build their own scenes
    torch.cuda.empty_cache()
    if not args.no_neural:
        print("[bench] neural-texture leg", file=sys.stderr, flush=True)
        line["neural"] = neural_leg(tg, args.n, max(3, args.steps // 2), 1, dev)
        torch.cuda.empty_cache()
    if not args.no_layouts:
        print("[bench] filter and layout legs", file=sys.stderr, flush=True)
        line["filters"] = filters_leg(tg, max(3, args.steps // 2), 1, dev)
        torch.cuda.empty_cache()
        line["layouts"] = packed_leg(tg, max(3, args.steps // 2), 1, dev)
        torch.cuda.empty_cache()
    if not args.no_config5:
        print("[bench] full training iteration leg", file=sys.stderr, flush=True)
        line["config5"] = config5_leg(tg, max(3, args.steps // 2), 2, dev)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
