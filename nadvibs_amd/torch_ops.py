"""PyTorch-ROCm binding of the library: H*v on device tensors, enqueued on
torch's current stream.  Device memory and streams are torch's; the
arithmetic is the library's HIP kernels (there is no torch fallback)."""
import torch

from .engine import Engine


def hv(eng: Engine, x: torch.Tensor, prec: str = "f64", out: torch.Tensor = None) -> torch.Tensor:
    """y = H x for a float64 CUDA tensor holding nstates*dimen values in the
    state-major layout."""
    if not x.is_cuda or x.dtype != torch.float64 or not x.is_contiguous() or x.numel() != eng.n:
        raise ValueError("x must be a contiguous float64 CUDA tensor of nstates*dimen values")
    y = out if out is not None else torch.empty_like(x)
    stream = torch.cuda.current_stream(x.device).cuda_stream
    eng.hv_device(x.data_ptr(), y.data_ptr(), prec, stream)
    return y
