"""Level starts and first-update maxima: the kernels of the automatic path
equal the reference-shaped ones they replace, bit for bit.

* Every level below the coarsest starts from imresize(.., 'bilinear') of the
  previous level's end field (calFlow3d_Wei_v1.py:381-410).  The automatic
  path forms it from the previous level's control points (k_level_start);
  ``resample_start=True`` writes the dense end field and upsamples it with
  the resampler.  Fields, logs and schedules must be identical, including
  levels that end without an update (their end field is their start field)
  and a start from the control-point field itself.
* The |delta motion| maximum of a level's first update (calFlow3d:641-654)
  has an LDS-staged and a gather kernel; ``maxdiff_gather=True`` forces the
  gather kernel.  Update counts and logs must be identical, also on a volume
  wide enough that the automatic path picks the gather kernel itself.
"""

from __future__ import annotations

import numpy as np
import pytest

from wholistic_registration_amd import synth

pytestmark = pytest.mark.gpu


def _pair(cuda, seed, zyx, dmax=2.0, mask_fraction=0.0, uint16=False):
    import torch
    n_blobs = max(20, int(2e-3 * np.prod(zyx)))
    p = synth.make_pair_np(seed, zyx, n_blobs, 2, dmax, mask_fraction=mask_fraction)
    mov, ref = p.moving, p.template
    if uint16:
        mov = np.clip(np.rint(mov), 0, 65535).astype(np.uint16)
        ref = np.clip(np.rint(ref), 0, 65535).astype(np.uint16)
    mask = None if p.mask is None else torch.from_numpy(p.mask.astype(np.float32)).to(cuda)
    return torch.from_numpy(np.ascontiguousarray(mov)).to(cuda), \
        torch.from_numpy(np.ascontiguousarray(ref)).to(cuda), mask


def _same(a, b):
    import torch
    assert torch.equal(a.motion, b.motion)
    assert a.records == b.records and a.updates == b.updates
    for k in ("diff_error", "penalty_error", "current_error", "max_diff_motion"):
        np.testing.assert_array_equal(getattr(a, k), getattr(b, k), err_msg=k)
    assert a.final_error == b.final_error


# (seed, (Z, Y, X), layer, r, iter, mask fraction, uint16)
START_CASES = {
    "halving_512x256": (1201, (6, 256, 512), 3, 4, 8, 0.0, False),
    "odd_sizes": (1202, (5, 75, 99), 2, 3, 8, 0.0, False),
    "masked_u16": (1203, (4, 160, 224), 3, 5, 10, 0.1, True),
    "r16_layer4": (1204, (4, 320, 640), 4, 16, 20, 0.1, True),
}


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
@pytest.mark.parametrize("name", sorted(START_CASES))
def test_start_from_control_points_equals_resampled_dense_field(cuda, name, precision):
    from wholistic_registration_amd import device as D
    seed, zyx, layer, r, n_iter, mf, u16 = START_CASES[name]
    mov, ref, mask = _pair(cuda, seed, zyx, mask_fraction=mf, uint16=u16)
    prm = dict(layer=layer, iter=n_iter, r=r, zRatio=1.0, movRange=5.0, smoothPenalty=40.0, tol=1e-3,
               mask_ref=mask, precision=precision)
    a = D.get_motion(mov, ref, **prm)
    used = [D.stats().start_from_cp[L] for L in range(layer + 1)]
    b = D.get_motion(mov, ref, resample_start=True, **prm)
    assert [D.stats().start_from_cp[L] for L in range(layer + 1)] == [0] * (layer + 1)
    print(f"{name} {precision}: levels started from control points {used}, schedule {a.schedule()}")
    _same(a, b)
    if name != "odd_sizes":
        assert all(used[:layer]) and not used[layer]


def test_level_without_update_hands_on_its_start_field(cuda):
    """A forced schedule with no update on levels 2 and 1: their end field is
    their start field, which the next start reads from the kept copy."""
    from wholistic_registration_amd import device as D
    mov, ref, _ = _pair(cuda, 1205, (4, 192, 256))
    prm = dict(layer=3, iter=6, r=4, zRatio=1.0, movRange=5.0, smoothPenalty=40.0, tol=1e-3)
    sched = {3: (4, 3), 2: (1, 0), 1: (1, 0), 0: (6, 5)}
    a = D.get_motion(mov, ref, schedule=sched, **prm)
    assert [D.stats().start_from_cp[L] for L in range(4)] == [1, 1, 1, 0]
    b = D.get_motion(mov, ref, schedule=sched, resample_start=True, **prm)
    assert a.schedule() == sched
    _same(a, b)


def test_start_after_control_point_start(cuda):
    """The coarsest level starts from a given CP field (init_cp) and its end
    field feeds the next level's start."""
    import torch

    from wholistic_registration_amd import device as D
    mov, ref, _ = _pair(cuda, 1206, (4, 192, 256))
    layer, r = 2, 4
    ncx, ncy = D.cp_grid(tuple(mov.shape), r, layer)
    g = torch.Generator(device="cpu").manual_seed(7)
    cp = (0.5 * torch.randn((3, mov.shape[0], ncy, ncx), generator=g)).to(cuda)
    for sched in (None, {2: (1, 0), 1: (3, 2), 0: (3, 2)}):
        prm = dict(layer=layer, iter=5, r=r, zRatio=1.0, movRange=5.0, smoothPenalty=40.0, tol=1e-3,
                   motion_init_cp=cp, schedule=sched)
        a = D.get_motion(mov, ref, **prm)
        b = D.get_motion(mov, ref, resample_start=True, **prm)
        _same(a, b)


@pytest.mark.parametrize("zyx,layer", [((5, 96, 160), 2), ((2, 24, 6400), 1)])
def test_first_update_maximum_gather_equals_lds_form(cuda, zyx, layer):
    """(2, 24, 6400): level 0 rows too wide for the LDS-staged kernel, so the
    automatic path itself runs the gather kernel there."""
    from wholistic_registration_amd import device as D
    mov, ref, _ = _pair(cuda, 1207 + layer, zyx, dmax=1.5)
    prm = dict(layer=layer, iter=6, r=3, zRatio=1.0, movRange=5.0, smoothPenalty=40.0, tol=1e-3)
    a = D.get_motion(mov, ref, **prm)
    b = D.get_motion(mov, ref, maxdiff_gather=True, **prm)
    assert all(u >= 1 for u in a.updates)
    _same(a, b)
