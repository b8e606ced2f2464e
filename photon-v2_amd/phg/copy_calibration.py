"""Plain device copy of a known size (1 GiB, three times), run under
`rocprofv3 --pmc FETCH_SIZE` / `--pmc WRITE_SIZE` to calibrate those counters:
    rocprofv3 --pmc FETCH_SIZE -d out -o copy -- python3 photon-v2_amd/phg/copy_calibration.py
"""
import torch

if __name__ == "__main__":
    n = 1 << 28  # 2^28 floats = 1 GiB
    a = torch.ones(n, device="cuda", dtype=torch.float32)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print("copy done", flush=True)
