"""The random-gather probe reads the cells it says it reads: each thread's output equals the sum of
the 8 corners of its 16 hashed cells, taken from the maps in their standard layout."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def fmix(x):
    x = np.uint32(x)
    x ^= x >> np.uint32(16)
    x = np.uint32((int(x) * 0x85EBCA6B) & 0xFFFFFFFF)
    x ^= x >> np.uint32(13)
    x = np.uint32((int(x) * 0xC2B2AE35) & 0xFFFFFFFF)
    x ^= x >> np.uint32(16)
    return int(x)


@pytest.mark.parametrize("span_kind", ["all", "one_map"])
def test_gather_probe_reads_the_hashed_cells(gpu, receptor, box, span_kind):
    import torch
    from unidock_engine import synthetic as S
    sc = gpu.Scene(receptor, box)
    sc.build_grids(S.GRID_SPACING, [2, 3, 4, 5, 6, 10, 11, 15])
    n, _, _, nt = sc.grid_info()
    maps = sc.download_maps()
    cx, cy, cz = int(n[0]) - 1, int(n[1]) - 1, int(n[2]) - 1
    per_map = cx * cy * cz
    span = per_map * nt if span_kind == "all" else per_map
    n_thr, seed = 256, 77
    out = torch.zeros(n_thr, device="cuda")
    gpu.check(gpu.lib().ud2_scene_gather_probe(sc.h, n_thr * 16, 0 if span_kind == "all" else span, seed,
                                               C.c_void_p(out.data_ptr()), gpu.stream_ptr()))
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    want = np.zeros(n_thr)
    for t in range(n_thr):
        h = (t ^ seed) & 0xFFFFFFFF
        for k in range(16):
            c = (fmix((h + k * 0x9E3779B9) & 0xFFFFFFFF) * span) >> 32
            slot, r = divmod(c, per_map)
            iz, r = divmod(r, cx * cy)
            iy, ix = divmod(r, cx)
            want[t] += maps[slot, iz:iz + 2, iy:iy + 2, ix:ix + 2].astype(np.float64).sum()
    assert np.allclose(got, want, rtol=1e-5, atol=1e-3 * np.abs(want).max()), np.abs(got - want).max()


def test_gather_probe_rejects_bad_spans(gpu, receptor, box):
    import torch
    from unidock_engine import synthetic as S
    sc = gpu.Scene(receptor, box)
    out = torch.zeros(4, device="cuda")
    assert gpu.lib().ud2_scene_gather_probe(sc.h, 64, 0, 1, C.c_void_p(out.data_ptr()), None) != 0  # no maps yet
    sc.build_grids(S.GRID_SPACING, [2, 3])
    assert gpu.lib().ud2_scene_gather_probe(sc.h, 64, 10 ** 9, 1, C.c_void_p(out.data_ptr()), None) != 0
    assert gpu.lib().ud2_scene_gather_probe(sc.h, 0, 0, 1, C.c_void_p(out.data_ptr()), None) != 0
