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
from . import device as _dev


def _torch():
    import torch
    return torch


class FrameBatch:
    """Registration contexts and streams for batches of frames on one device.

    ``register(frames, ref, **params)`` registers every ``(Z, Y, X)`` frame of
    ``frames`` (a sequence of device tensors, or one ``(T, Z, Y, X)`` tensor)
    onto ``ref``; ``params`` are the keyword arguments of
    ``device.get_motion`` (``layer``, ``iter``, ``r``, ``zRatio``,
    ``movRange``, ``smoothPenalty``, ``tol``, ``mask_ref``, ``precision``).
    """

    def __init__(self, workers=4, device=0):
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

    def register(self, frames, ref, motion_out=None, warped_out=None, masks_mov=None, **params):
        """Register every frame; returns ``(motion (T, Z, Y, X, 3), results)``
        where ``results[i]`` is frame i's ``device.MotionResult``.  With
        ``warped_out`` (a ``(T, Z, Y, X)`` float32 tensor) every frame is also
        warped by its field (``correctMotion``)."""
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
        # inputs come from the caller's stream: let it finish before the workers start
        torch.cuda.current_stream(ref.device).synchronize()
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
                    results[i] = _dev.get_motion(
                        frames[i], ref, motion_out=motion_out[i], context=ctx, stream=strm,
                        reuse_template=reuse, mask_mov=None if masks_mov is None else masks_mov[i], **params)
                    self._ref_seen[k] = key
                    if warped_out is not None:
                        _dev.correct_motion(frames[i], motion_out[i], out=warped_out[i], context=ctx, stream=strm)
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
