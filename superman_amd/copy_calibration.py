"""Calibration run for HBM traffic counters: one plain device-to-device copy
kernel of a known size, so that FETCH_SIZE / WRITE_SIZE (KiB per dispatch,
as rocprofv3 --pmc reports them) can be converted into bytes.

    rocprofv3 --pmc FETCH_SIZE -- python3 -m superman_amd.copy_calibration [--mib 256]
    rocprofv3 --pmc WRITE_SIZE -- python3 -m superman_amd.copy_calibration [--mib 256]

The copy kernel reads and writes exactly `mib` MiB; the scale factor for a
counter is (mib * 2^20) / (counter KiB * 1024).
"""
from __future__ import annotations

import argparse
import sys


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--mib", type=int, default=256)
    ap.add_argument("--reps", type=int, default=3)
    args = ap.parse_args(argv)
    import torch

    n = args.mib * (1 << 20) // 8
    src = torch.arange(n, dtype=torch.float64, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(args.reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    if not torch.equal(dst[-4:], src[-4:]):
        raise RuntimeError("copy mismatch")
    print(f"copied {args.mib} MiB x {args.reps}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
