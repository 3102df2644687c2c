"""Byte calibration for rocprofv3 PMC counters: copies a 1 GiB device buffer three times.

Run under `rocprofv3 --pmc FETCH_SIZE` (then WRITE_SIZE); the counter value of the copy kernel
divided into 1 GiB gives the scale factor applied to the step kernels' FETCH_SIZE / WRITE_SIZE.
"""
import torch

n = (1 << 30) // 8
a = torch.ones(n, dtype=torch.float64, device="cuda")
b = torch.empty_like(a)
for _ in range(3):
    b.copy_(a)
torch.cuda.synchronize()
print("copied 1 GiB x3")
