"""Helpers shared by the package's Python modules: torch is imported lazily
(the package imports without it, for the host-side tests), and every device
entry point needs the current HIP device."""

from __future__ import annotations


def torch_module():
    import torch
    return torch


def is_tensor(a):
    try:
        import torch
    except ImportError:        # pragma: no cover
        return False
    return isinstance(a, torch.Tensor)


def hip_device():
    """torch's current device; raises when no HIP device is visible (there is
    no CPU path)."""
    torch = torch_module()
    if not torch.cuda.is_available():
        raise RuntimeError("wholistic_registration_amd needs a HIP device (none is visible)")
    return torch.device("cuda", torch.cuda.current_device())
