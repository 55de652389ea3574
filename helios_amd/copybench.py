"""A plain device-to-device copy of known size, for calibrating the HBM
traffic counters (FETCH_SIZE / WRITE_SIZE) of a profiler run:

    rocprofv3 --pmc FETCH_SIZE -- python3 helios_amd/copybench.py

reads 1 GiB and writes 1 GiB per copy (three copies, the first one warms up).
"""
import sys

import torch

NBYTES = 1 << 30


def main():
    n = NBYTES // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(f"copied {NBYTES} bytes x 3", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
