"""Device copy of a known size (1 GiB), used to calibrate HBM byte counters
(FETCH_SIZE / WRITE_SIZE) against algorithmic bytes."""
import torch

N = 1 << 30
a = torch.empty(N, dtype=torch.uint8, device="cuda")
a.fill_(1)
b = torch.empty_like(a)
torch.cuda.synchronize()
b.copy_(a)
torch.cuda.synchronize()
print(f"copied {N} bytes", flush=True)
