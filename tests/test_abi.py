"""CPU-side checks of the native boundary: the library loads, exports every
entry point the C header declares, and the host-side validation of the ops
raises the reference's errors (no GPU needed)."""
import ctypes
import os
import re

import pytest
import torch

import textured_gaussians_amd as tg
from textured_gaussians_amd.cuda import _ops
from textured_gaussians_amd.cuda._backend import LIB_PATH, lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "tgs_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(tgs_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    L = lib()
    names = _declared()
    assert len(names) >= 15
    for n in names:
        assert hasattr(L, n), n
    assert L.tgs_abi_version() == 2


def test_ctypes_prototypes_match_header_arity():
    from textured_gaussians_amd.cuda._backend import _PROTOS
    src = re.sub(r"/\*.*?\*/", "", open(os.path.join(ROOT, "include", "tgs_hip.h")).read(), flags=re.S)
    for name in _declared():
        params = re.search(r"\b" + name + r"\s*\(([^)]*)\)", src).group(1)
        n = len([a for a in params.split(",") if a.strip() and a.strip() != "void"])
        assert name in _PROTOS, name
        assert len(_PROTOS[name][1]) == n, (name, len(_PROTOS[name][1]), n)


def _ctype_kind(t):
    if t is ctypes.c_void_p or (isinstance(t, type) and issubclass(t, ctypes._Pointer)):
        return "ptr"
    if t is ctypes.c_double:
        return "f64"
    return {ctypes.sizeof(ctypes.c_int32): "i32", 8: "i64"}[ctypes.sizeof(t)]


def _c_kind(param):
    p = param.strip()
    if "*" in p or p.startswith("tgs_stream_t"):
        return "ptr"
    if p.startswith("double"):
        return "f64"
    if p.startswith(("int64_t", "size_t", "uint64_t")):
        return "i64"
    if p.startswith(("int32_t", "int ", "uint32_t", "uint8_t")):
        return "i32"
    raise AssertionError(f"unclassified parameter {p!r}")


def test_ctypes_prototypes_match_header_types():
    """Argument by argument: a pointer where the header has a pointer, a 64-bit integer
    where it has int64_t/size_t, and so on (a shifted argument list passes the arity
    check but not this one)."""
    from textured_gaussians_amd.cuda._backend import _PROTOS
    src = re.sub(r"/\*.*?\*/", "", open(os.path.join(ROOT, "include", "tgs_hip.h")).read(), flags=re.S)
    for name in _declared():
        params = [a for a in re.search(r"\b" + name + r"\s*\(([^)]*)\)", src).group(1).split(",")
                  if a.strip() and a.strip() != "void"]
        got = [_ctype_kind(t) for t in _PROTOS[name][1]]
        want = [_c_kind(p) for p in params]
        assert got == want, (name, [(i, p.strip(), g) for i, (p, g, w) in enumerate(zip(params, got, want))
                                    if g != w])


def test_host_validation_errors_without_gpu():
    L = lib()
    null = None
    # 5 channels are rejected before any device work (the reference dispatches 1..5 but
    # its python layer pads 5 -> 8, which it rejects)
    rc = L.tgs_rasterize_fwd(0, 1, 10, 5, *([null] * 6), 0, 0, 3.0, 3.0, null, null, null, 0, 16, 16, 16, 1, 1, null, null,
                             0, 0.0, 1.0, *([null] * 16), 0, 0, null, 0, null)
    assert rc != 0 and b"channels" in L.tgs_last_error()
    # any tile size whose square is at most 1024 (the reference's tile_size^2 threads)
    rc = L.tgs_rasterize_fwd(0, 1, 10, 3, *([null] * 6), 0, 0, 3.0, 3.0, null, null, null, 0, 16, 16, 33, 1, 1, null, null,
                             0, 0.0, 1.0, *([null] * 16), 0, 0, null, 0, null)
    assert rc != 0 and b"tile_size" in L.tgs_last_error()
    # the packed-record workspace is required and sized by tgs_rasterize_workspace_bytes
    assert L.tgs_rasterize_workspace_bytes(0, 2, 1000) >= 2 * 1000 * 20 * 4
    assert L.tgs_rasterize_workspace_bytes(1, 2, 1000) >= 2 * 1000 * 20 * 8
    rc = L.tgs_rasterize_fwd(0, 1, 10, 3, *([null] * 6), 0, 0, 3.0, 3.0, null, null, null, 0, 16, 16, 16, 1, 1, null, null,
                             0, 0.0, 1.0, *([null] * 16), 0, 0, null, 0, null)
    assert rc != 0 and b"workspace" in L.tgs_last_error()
    # camera and tile ids must fit in 32 bits
    rc = L.tgs_isect_tiles_count(1 << 20, 1, null, null, 16, 4096, 4096, null, null, null, 0, null)
    assert rc != 0 and b"32 bits" in L.tgs_last_error()
    rc = L.tgs_sh_fwd(0, 10, 4, 5, null, null, null, null, null)
    assert rc != 0 and b"degree" in L.tgs_last_error()
    # the views kernels: float32 with K * 3 <= 48 only
    rc = L.tgs_sh_fwd_cams(1, 2, 10, 16, 3, null, null, null, null, null)
    assert rc == 3 and b"float32" in L.tgs_last_error()
    rc = L.tgs_sh_bwd_cams(0, 2, 10, 17, 3, null, null, null, null, null, null, null)
    assert rc == 3 and b"K * 3" in L.tgs_last_error()
    # decoder pieces: operand flags OR-ed into the epilogue, argument rules
    rc = L.tgs_gemm_bf16(0, 0, 7 | 0x10, 4, 4, 4, null, 4, null, 4, null, null, 4, null, 0, 1, null)
    assert rc == 1 and b"epilogue" in L.tgs_last_error()
    rc = L.tgs_gemm_bf16(0, 0, 0x30, 4, 4, 4, null, 4, null, 4, null, null, 4, null, 0, 1, null)
    assert rc == 1 and b"null operand" in L.tgs_last_error()
    rc = L.tgs_sigmoid_bwd_colsum(10, 6, null, null, null, 1, null, null)
    assert rc == 1 and b"multiple of 4" in L.tgs_last_error()
    # the measurement hook: forward (0), backward (1) or texture gather (2), both events or neither
    rc = L.tgs_set_raster_events(3, null, null)
    assert rc == 1 and b"which" in L.tgs_last_error()
    assert L.tgs_set_raster_events(2, null, null) == 0
    # the streaming gather: float32 compute (bf16 textures too), rows whose float64 sums fit
    # twice in LDS (50x50 yes, 64x64 no); never for float64
    assert L.tgs_texture_gather_fast(0, 50, 50) == 1 and L.tgs_texture_gather_fast(0 | 0x100, 32, 32) == 1
    assert L.tgs_texture_gather_fast(0, 64, 64) == 0 and L.tgs_texture_gather_fast(1, 16, 16) == 0
    assert L.tgs_texture_gather_fast(0, 0, 16) == 0
    rc = L.tgs_set_raster_events(0, 1234, null)
    assert rc == 1 and b"both events" in L.tgs_last_error()
    assert L.tgs_set_raster_events(1, null, null) == 0


def test_ops_reject_cpu_and_noncontiguous_tensors():
    with pytest.raises(RuntimeError, match="must be a CUDA tensor"):
        _ops.isect_offset_encode(torch.zeros(4, dtype=torch.int64), 1, 2, 2)
    with pytest.raises(RuntimeError, match="must be a CUDA tensor"):
        tg.rasterization_2dgs(torch.zeros(3, 3), torch.ones(3, 4), torch.ones(3, 3), torch.ones(3), torch.ones(3, 3),
                              torch.eye(4)[None], torch.eye(3)[None], 8, 8)


def test_wrapper_channel_rules():
    with pytest.raises(ValueError, match="Unsupported number of color channels"):
        tg.rasterize_to_pixels_2dgs(torch.zeros(1, 2, 2), torch.zeros(1, 2, 3, 3), torch.zeros(1, 2, 0),
                                    torch.zeros(1, 2), torch.zeros(1, 2, 3), torch.zeros(1, 2, 2), 8, 8, 16,
                                    torch.zeros(1, 1, 1, dtype=torch.int32), torch.zeros(0, dtype=torch.int32))
    with pytest.raises(Exception, match="Unsupported filter type"):
        tg.rasterize_to_pixels_textured_gaussians(
            torch.zeros(1, 2, 2), torch.zeros(1, 2, 3, 3), torch.zeros(1, 2, 3), torch.zeros(1, 2),
            torch.zeros(2, 4, 4, 4), 3.0, 3.0, torch.zeros(1, 2, 3), torch.zeros(1, 2, 2), 8, 8, 16,
            torch.zeros(1, 1, 1, dtype=torch.int32), torch.zeros(0, dtype=torch.int32), filtering="mipmapped")


def test_product_package_never_imports_the_oracle():
    pkg = os.path.join(ROOT, "textured_gaussians_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith(".py"):
                txt = open(os.path.join(dirpath, f)).read()
                assert "import oracle" not in txt and "from oracle" not in txt, f
