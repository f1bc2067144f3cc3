"""Copy kernel of known size, the reference point for HBM traffic counters.

``python -m tensor_train_amd.calibrate [MiB]`` copies a float64 device buffer
of MiB mebibytes into another (default 1024) five times with torch's
elementwise copy kernel: each dispatch reads and writes exactly MiB MiB.  The
per-dispatch FETCH_SIZE / WRITE_SIZE that rocprofv3 reports for it, divided
by MiB * 1024 KiB, is the factor by which those counters over- or
under-count HBM bytes on this device.
"""
import sys

import torch


def main():
    mib = int(sys.argv[1]) if len(sys.argv) > 1 else 1024
    n = mib * (1 << 20) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"copied {mib} MiB x 5", flush=True)


if __name__ == "__main__":
    main()
