"""Many frames of a time series registered against one template on one GPU.

The reference registers the frames of a batch one after another
(``wbi_registration_3d``'s frame loop, registration.py:299-346; each
``getMotion`` call rebuilds both pyramids).  Here ``FrameBatch`` keeps
``workers`` library contexts, each with its own stream pair and workspace, and
registers frames on all of them at once: the short, latency-bound launches of
the coarse levels of one frame overlap with the other frames' work.  Within a
worker, every frame after the first reuses the template's pyramid levels
(``wr_params.reuse_template``), so the template is resampled once per worker
instead of once per frame.  Each frame's result is exactly what a single
``getMotion`` call returns (the kernels are deterministic and the frames
independent).
"""

from __future__ import annotations

import threading

from . import _lib
from . import autotune as _tune
from . import device as _dev
from ._util import torch_module as _torch


class FrameBatch:
    """Registration contexts and streams for batches of frames on one device.

    ``register(frames, ref, **params)`` registers every ``(Z, Y, X)`` frame of
    ``frames`` (a sequence of device tensors, or one ``(T, Z, Y, X)`` tensor)
    onto ``ref``; ``params`` are the keyword arguments of
    ``device.get_motion`` (``layer``, ``iter``, ``r``, ``zRatio``,
    ``movRange``, ``smoothPenalty``, ``tol``, ``mask_ref``, ``precision``).
    With ``autotune=True`` every geometry runs with per-level LK tile widths
    tuned on its first frame (``autotune.TileCache``).
    """

    def __init__(self, workers=4, device=0, autotune=False):
        torch = _torch()
        if workers < 1:
            raise ValueError("workers must be at least 1")
        self.device = device
        lib = _lib.load()
        self._lib = lib
        self.contexts = []
        for _ in range(workers):
            h = lib.wr_create(device)
            if not h:
                self.close()
                raise RuntimeError(f"wr_create({device}): {_lib.last_error()}")
            self.contexts.append(h)
        with torch.cuda.device(device):
            self.streams = [torch.cuda.Stream() for _ in range(workers)]
        self._ref_seen = [None] * workers
        self.tiles = _tune.TileCache() if autotune else None

    def invalidate(self):
        """Forget the template pyramids (call after modifying the template
        tensor in place: the contexts recognise it by its address)."""
        self._ref_seen = [None] * len(self.contexts)

    def close(self):
        for h in getattr(self, "contexts", []):
            self._lib.wr_destroy(h)
        self.contexts = []

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def register(self, frames, ref, motion_out=None, warped_out=None, masks_mov=None, net=None, **params):
        """Register every frame; returns ``(motion (T, Z, Y, X, 3), results)``
        where ``results[i]`` is frame i's ``device.MotionResult``.  With
        ``warped_out`` (a ``(T, Z, Y, X)`` float32 tensor) every frame is also
        warped by its field (``correctMotion``).  With ``net`` (a
        ``learned.InitNet``) every frame starts from the network's field
        (``learned.register_learned``; pass ``learned.SCHEDULE``'s layer and
        iter), each worker on its own fork of the network."""
        torch = _torch()
        frames = list(frames)
        n = len(frames)
        if n == 0:
            raise ValueError("no frames")
        shape = tuple(frames[0].shape)
        if any(tuple(f.shape) != shape for f in frames):
            raise ValueError("frames must share one shape")
        if tuple(ref.shape) != shape:
            raise ValueError("ref must have the frames' shape")
        if masks_mov is not None and len(masks_mov) != n:
            raise ValueError("masks_mov needs one mask per frame")
        if motion_out is None:
            motion_out = torch.empty((n,) + shape + (3,), dtype=torch.float32, device=ref.device)
        ref_level, forks = None, None
        if net is not None:
            from . import learned as _learned
            if masks_mov is not None:
                raise ValueError("the learned start takes no mask_mov")
            ref_level = _dev.pyramid_level(ref, _learned.LEVEL)
            forks = [net.fork(ctx) for ctx in self.contexts]
        # inputs come from the caller's stream: let it finish before the workers start
        torch.cuda.current_stream(ref.device).synchronize()
        # template pyramids are reused within one call only: between calls the
        # caller may free the template and the allocator hand its address to a
        # new one
        self.invalidate()
        results = [None] * n
        errors = []
        w = len(self.contexts)

        def work(k):
            try:
                ctx, strm = self.contexts[k], self.streams[k]
                for i in range(k, n, w):
                    key = (ref.data_ptr(), shape, params.get("layer"), params.get("precision", "fp32"),
                           None if params.get("mask_ref") is None else params["mask_ref"].data_ptr())
                    reuse = self._ref_seen[k] == key
                    extra = dict(params)
                    if self.tiles is not None:
                        # tuning (first frame of a geometry) runs on the same template, size,
                        # layer, precision and mask: the context's template levels stay valid
                        extra["tile_patches_x"] = self.tiles.tiles(frames[i], ref, context=ctx, stream=strm,
                                                                   **params)
                    if warped_out is not None:   # correctMotion formed with the field
                        extra.update(warp_src=frames[i], warped_out=warped_out[i])
                    if forks is not None:
                        results[i], info = _learned.register_learned(
                            frames[i], ref, forks[k], ref_level=ref_level, motion_out=motion_out[i],
                            context=ctx, stream=strm, reuse_template=reuse, **extra)
                        # a fallback run rebuilt the context's template levels for its own layer
                        self._ref_seen[k] = key if "plain_error" not in info else None
                    else:
                        results[i] = _dev.get_motion(
                            frames[i], ref, motion_out=motion_out[i], context=ctx, stream=strm,
                            reuse_template=reuse, mask_mov=None if masks_mov is None else masks_mov[i], **extra)
                        self._ref_seen[k] = key
                strm.synchronize()
            except BaseException as e:          # re-raised on the calling thread
                self._ref_seen[k] = None
                errors.append(e)

        threads = [threading.Thread(target=work, args=(k,)) for k in range(min(w, n))]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        if errors:
            raise errors[0]
        return motion_out, results

    def register_batched(self, frames, ref, chunk=16, motion_out=None, warped_out=None, masks_mov=None,
                         **params):
        """Register every frame of ``frames`` (a contiguous ``(T, Z, Y, X)``
        device tensor) with the batched launch (``device.get_motion_batch``):
        chunks of ``chunk`` frames, each chunk one launch sequence in which
        every kernel of a level-iteration is one launch over the chunk's
        frames, the chunks spread over the workers' contexts.  ``masks_mov``:
        ``(T, Z, Y, X)`` float32 or None.  Returns ``(motion, results)`` as
        ``register`` does, with the same values bit for bit."""
        torch = _torch()
        if frames.dim() != 4 or not frames.is_contiguous():
            raise ValueError("frames must be a contiguous (T, Z, Y, X) tensor")
        if chunk < 1:
            raise ValueError("chunk must be at least 1")
        n = int(frames.shape[0])
        shape = tuple(frames.shape[1:])
        if tuple(ref.shape) != shape:
            raise ValueError("ref must have the frames' shape")
        if motion_out is None:
            motion_out = torch.empty((n,) + shape + (3,), dtype=torch.float32, device=ref.device)
        torch.cuda.current_stream(ref.device).synchronize()
        self.invalidate()
        chunks = [(i, min(n, i + chunk)) for i in range(0, n, chunk)]
        results = [None] * n
        errors = []
        w = len(self.contexts)

        def work(k):
            try:
                ctx, strm = self.contexts[k], self.streams[k]
                for c0, c1 in chunks[k::w]:
                    key = (ref.data_ptr(), shape, params.get("layer"), params.get("precision", "fp32"),
                           None if params.get("mask_ref") is None else params["mask_ref"].data_ptr())
                    res = _dev.get_motion_batch(
                        frames[c0:c1], ref, motion_out=motion_out[c0:c1], context=ctx, stream=strm,
                        reuse_template=self._ref_seen[k] == key,
                        mask_mov=None if masks_mov is None else masks_mov[c0:c1],
                        warp_src=None if warped_out is None else frames[c0:c1],
                        warped_out=None if warped_out is None else warped_out[c0:c1], **params)
                    self._ref_seen[k] = key
                    results[c0:c1] = res
                strm.synchronize()
            except BaseException as e:          # re-raised on the calling thread
                self._ref_seen[k] = None
                errors.append(e)

        threads = [threading.Thread(target=work, args=(k,)) for k in range(min(w, len(chunks)))]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        if errors:
            raise errors[0]
        return motion_out, results

    def register_crops(self, pairs, layer=None, warp=False, net=None, learned_min_depth=48, plain_iter=15, **params):
        """Register crops of uneven sizes: ``pairs`` is a sequence of ``(mov,
        ref)`` or ``(mov, ref, mask_ref)`` device volumes, each pair of one
        shape.  The crops are spread over the worker contexts, largest first,
        each to the worker with the fewest voxels so far.  ``layer=None``
        uses up to 4 levels, fewer where a crop's coarsest level would be
        smaller than one patch (``max_layer``).  With ``autotune=True`` each
        geometry runs with the tile widths tuned on its first crop.  Returns
        ``(results, warped)``: ``results[i]`` is crop i's
        ``device.MotionResult`` and ``warped[i]`` its ``correctMotion`` output
        (None unless ``warp``).  With ``net`` (a ``learned.InitNet``) every
        crop starts from the network's field at full resolution
        (``learned.register_learned``, ``layer`` 0 unless given; pass
        ``learned.SCHEDULE``'s iter), each worker on its own fork; crops
        shallower than ``learned_min_depth`` planes run the plain pyramid
        method instead (``max_layer`` levels, ``plain_iter`` iterations): on
        32-plane crops the full-resolution schedule stops outside the
        0.1-voxel gate from any start."""
        torch = _torch()
        pairs = [tuple(p) for p in pairs]
        n = len(pairs)
        if n == 0:
            raise ValueError("no crops")
        for p in pairs:
            if tuple(p[0].shape) != tuple(p[1].shape):
                raise ValueError("each crop's mov and ref must share one shape")
        r = params["r"]
        forks = None
        use_net = [net is not None and p[0].shape[0] >= learned_min_depth for p in pairs]
        if net is not None:
            from . import learned as _learned
            forks = [net.fork(ctx) for ctx in self.contexts]

        def crop_layer(shape, learned_start):
            if learned_start:
                return 0 if layer is None else int(layer)
            if net is not None or layer is None:          # the plain method: as deep as the crop allows
                return max_layer(shape, r)
            return int(layer)

        layers = [crop_layer(p[0].shape, u) for p, u in zip(pairs, use_net)]
        w = len(self.contexts)
        load, owner = [0] * w, [[] for _ in range(w)]
        for i in sorted(range(n), key=lambda i: -pairs[i][0].numel()):
            k = min(range(w), key=lambda k: load[k])
            owner[k].append(i)
            load[k] += pairs[i][0].numel()
        torch.cuda.current_stream(pairs[0][0].device).synchronize()
        results, warped, errors = [None] * n, [None] * n, []

        def work(k):
            try:
                ctx, strm = self.contexts[k], self.streams[k]
                with torch.cuda.stream(strm):
                    for i in owner[k]:
                        mov, ref = pairs[i][0], pairs[i][1]
                        extra = dict(params)
                        if len(pairs[i]) > 2 and pairs[i][2] is not None:
                            extra["mask_ref"] = pairs[i][2]
                        if self.tiles is not None:
                            extra["tile_patches_x"] = self.tiles.tiles(mov, ref, layer=layers[i], context=ctx,
                                                                       stream=strm, **extra)
                        if use_net[i]:
                            results[i], _ = _learned.register_learned(mov, ref, forks[k], layer=layers[i], context=ctx,
                                                                      stream=strm, **extra)
                        else:
                            if net is not None:          # the plain method for a crop too shallow
                                extra["iter"] = plain_iter
                            results[i] = _dev.get_motion(mov, ref, layer=layers[i], context=ctx, stream=strm, **extra)
                        if warp:
                            warped[i] = _dev.correct_motion(mov, results[i].motion, context=ctx, stream=strm)
                    strm.synchronize()
            except BaseException as e:          # re-raised on the calling thread
                errors.append(e)

        threads = [threading.Thread(target=work, args=(k,)) for k in range(w) if owner[k]]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        if errors:
            raise errors[0]
        return results, warped


def max_layer(shape_zyx, r, cap=4):
    """Deepest pyramid level (at most ``cap``) whose x and y extents
    int(n / 2^L) still hold one LK patch centre (n_L >= r + 2,
    calFlow3d_Wei_v1.py:434-436)."""
    _, ny, nx = shape_zyx
    L = 0
    while L < cap and int(nx / 2 ** (L + 1)) >= r + 2 and int(ny / 2 ** (L + 1)) >= r + 2:
        L += 1
    return L
