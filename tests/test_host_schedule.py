"""Host-side logic of batching, tile autotuning and the learned
initialisation that needs no GPU."""

from __future__ import annotations

import pytest

from wholistic_registration_amd import autotune, learned, train_init
from wholistic_registration_amd.batch import max_layer


@pytest.mark.parametrize("shape,r,want", [
    ((32, 128, 128), 8, 3),       # 128 / 8 = 16 >= 10, 128 / 16 = 8 < 10
    ((64, 512, 1024), 8, 4),      # capped at 4 levels
    ((10, 20, 20), 8, 1),
    ((10, 19, 40), 8, 0),         # 19 / 2 = 9 < 10
    ((96, 1024, 2048), 16, 4),
])
def test_max_layer_keeps_a_patch_centre_on_the_coarsest_level(shape, r, want):
    L = max_layer(shape, r)
    assert L == want
    _, ny, nx = shape
    # the coarsest level holds a patch centre (calFlow3d_Wei_v1.py:434-436: range(r, n - 1, 2r + 1))
    assert len(range(r, int(nx / 2 ** L) - 1, 2 * r + 1)) >= 1
    assert len(range(r, int(ny / 2 ** L) - 1, 2 * r + 1)) >= 1


def test_tile_candidates_fit_one_workgroup():
    for r in (1, 4, 8, 16, 40, 123):
        c = autotune.candidates(r)
        assert c[0] == 1
        assert c[-1] * (2 * r + 1) <= autotune.MAX_TILE_COLS or c == [1]
        assert c == list(range(1, len(c) + 1))


def test_tile_cache_keys_separate_geometries():
    class V:
        def __init__(self, shape, dtype="torch.uint16"):
            self.shape, self.dtype = shape, dtype
    k = autotune.TileCache.key
    assert k(V((8, 64, 64)), 2, 8) == k(V((8, 64, 64)), 2, 8)
    assert k(V((8, 64, 64)), 2, 8) != k(V((8, 64, 96)), 2, 8)
    assert k(V((8, 64, 64)), 2, 8) != k(V((8, 64, 64)), 3, 8)
    assert k(V((8, 64, 64)), 2, 8, "bf16") != k(V((8, 64, 64)), 2, 8)
    assert k(V((8, 64, 64)), 2, 8, mask_ref=object()) != k(V((8, 64, 64)), 2, 8)


def test_network_flops_and_recipes():
    # 2 * 27 * cin * cout per voxel and layer, at the three resolutions
    assert learned.forward_flops(64, 64, 128, channels=(16, 32, 64)) == 69_759_664_128
    assert learned.forward_flops(64, 64, 128, channels=(16, 32, 64, 128)) == 78_366_375_936
    k = 2 ** (len(learned.CHANNELS) - 1)
    for name, rc in train_init.RECIPES.items():
        z, y, x = rc["zyx"]
        d, h, w = z, y // 2 ** learned.LEVEL, x // 2 ** learned.LEVEL
        assert d % k == 0 and h % k == 0 and w % (8 * k) == 0, name      # the network's input rule
        assert max_layer(rc["zyx"], rc["lk"]["r"]) >= rc["lk"]["layer"], name


def test_pair_batches_share_one_shape():
    import numpy as np

    class X:
        def __init__(self, shape):
            self.shape = shape
    ps = train_init.PairSet.__new__(train_init.PairSet)
    ps.mov = [X((64, 64, 128))] * 6 + [X((32, 64, 128))] * 3
    ps.groups = {}
    for i, m in enumerate(ps.mov):
        ps.groups.setdefault(tuple(m.shape), []).append(i)
    rng = np.random.default_rng(0)
    seen = set()
    for _ in range(200):
        idx = ps.sample(rng, 4)
        shapes = {ps.mov[i].shape for i in idx}
        assert len(shapes) == 1 and len(set(idx)) == len(idx)
        seen |= shapes
        assert len(idx) == (4 if shapes == {(64, 64, 128)} else 3)
    assert len(seen) == 2
