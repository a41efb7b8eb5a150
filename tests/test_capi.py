"""C-ABI checks that need no GPU: the library builds for gfx950, loads, exports
every symbol include/gazpart_gpu.h declares, its ctypes mirrors match the C
struct layouts, and it refuses to run without a device (no CPU fallback)."""
import os
import re
import subprocess

import pytest

import gazpart_amd as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gazpart_gpu.h")


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(G.LIB_PATH):
        from gazpart_amd import _build
        _build.build()
    return G.load_library()


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(gp_[a-z_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported(lib):
    syms = declared_symbols()
    assert len(syms) >= 15
    assert set(syms) == set(G.EXPORTS)
    for s in syms:
        assert hasattr(lib, s), s


def test_struct_layouts_match(tmp_path):
    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "gazpart_gpu.h"\n'
        "int main(void){printf(\"%zu %zu %zu %zu %zu %zu\\n\", sizeof(gp_params),"
        " offsetof(gp_params, dpmax), offsetof(gp_params, prefilter), sizeof(gp_stats),"
        " offsetof(gp_stats, greedy_rounds), offsetof(gp_stats, nlm)); return 0;}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    import ctypes as C
    assert int(out[0]) == C.sizeof(G.Params)
    assert int(out[1]) == G.Params.dpmax.offset
    assert int(out[2]) == G.Params.prefilter.offset
    assert int(out[3]) == C.sizeof(G.Stats)
    assert int(out[4]) == G.Stats.greedy_rounds.offset
    assert int(out[5]) == G.Stats.nlm.offset


def test_no_cpu_fallback_without_gpu(lib):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    with pytest.raises(G.GpError) as e:
        G.Context(16, 6.28, 0.01, 1.0, 1e-3)
    assert e.value.code == 12


def test_invalid_arguments_rejected_before_device(lib):
    import ctypes as C
    p = G.Params()
    p.n, p.interp_scheme, p.precision, p.L, p.visc, p.dtime = 16, 7, 0, 6.28, 0.01, 1e-3
    h = C.c_void_p()
    assert lib.gp_create(C.byref(p), C.byref(h)) == 10
    assert lib.gp_create(None, C.byref(h)) == 10


def test_tile_lds_limit_checked_before_device(lib):
    """gp_params.tile: the staged block, (tile+S-1)^2 (tile+3) nodes of six
    values, must fit 64 KiB of LDS; larger tiles are refused (GP_ERR_ARG)
    before any device call, smaller ones get as far as the device."""
    import ctypes as C
    import torch
    ok = [(G.PREC_MIXED, 3, 10), (G.PREC_MIXED, 4, 10), (G.PREC_FP64, 3, 8),
          (G.PREC_FP64, 4, 7)]
    bad = [(G.PREC_MIXED, 3, 11), (G.PREC_MIXED, 4, 11), (G.PREC_FP64, 3, 9),
           (G.PREC_FP64, 4, 8)]
    for (prec, scheme, tile), refused in [(c, False) for c in ok] + [(c, True) for c in bad]:
        # closed form of the block's bytes
        S = 4 if scheme == 3 else 5
        nbytes = (tile + S - 1) ** 2 * (tile + 3) * 6 * (4 if prec == G.PREC_MIXED else 8)
        assert (nbytes > 65536) == refused, (prec, scheme, tile, nbytes)
        p = G.Params()
        p.n, p.interp_scheme, p.precision, p.L, p.visc, p.dtime = 16, scheme, prec, 6.28, 0.01, 1e-3
        p.tile = tile
        h = C.c_void_p()
        rc = lib.gp_create(C.byref(p), C.byref(h))
        if refused:
            assert rc == 10, (prec, scheme, tile)
        elif torch.cuda.is_available():
            assert rc == 0
            lib.gp_destroy(h)
        else:
            assert rc == 12
    p.tile = -1
    assert lib.gp_create(C.byref(p), C.byref(h)) == 10
