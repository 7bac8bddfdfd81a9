"""Calibration for HBM-traffic counters: one device-to-device copy of a known
size (256 MiB read + 256 MiB written), run under rocprofv3 --pmc FETCH_SIZE /
WRITE_SIZE to obtain the counter-to-bytes factor of this GPU."""
import torch

n = 256 * 1024 * 1024 // 8
a = torch.rand(n, dtype=torch.float64, device="cuda")
b = torch.empty_like(a)
for _ in range(3):
    b.copy_(a)
torch.cuda.synchronize()
print("copied", a.numel() * 8, "bytes x3")
