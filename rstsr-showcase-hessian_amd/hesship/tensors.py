"""Column-major (Fortran-order) device tensor helpers for the C ABI."""
import torch


def is_fortran(t):
    return t.dim() <= 1 and t.is_contiguous() or t.permute(*reversed(range(t.dim()))).is_contiguous()


def fortran(t, dtype=torch.float64):
    """Return a column-major view/copy of t (same logical shape), float64 on t's device."""
    t = t.to(dtype)
    if is_fortran(t):
        return t
    rev = list(reversed(range(t.dim())))
    return t.permute(*rev).contiguous().permute(*rev)


def empty_f(shape, device, dtype=torch.float64):
    """Uninitialised column-major tensor of logical `shape`."""
    rev = list(reversed(range(len(shape))))
    return torch.empty(tuple(reversed(shape)), device=device, dtype=dtype).permute(*rev)


def stream_of(device):
    return torch.cuda.current_stream(device).cuda_stream
