"""A plain device copy of known size (2**27 float64 = 1 GiB read, 1 GiB written): the calibration run for
HBM byte counters (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE) of the kernels of this package."""
import torch

N = 1 << 27
src = torch.ones(N, dtype=torch.float64, device="cuda")
dst = torch.empty_like(src)
torch.cuda.synchronize()
dst.copy_(src)
torch.cuda.synchronize()
print(f"copied {N * 8} bytes")
