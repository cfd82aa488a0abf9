"""Plain device-to-device copy of a known size, used to calibrate the HBM
counters (FETCH_SIZE, WRITE_SIZE) that rocprofv3 reports per dispatch.

    rocprofv3 --pmc FETCH_SIZE --kernel-trace -- python3 -m radtranbook_amd.profile_calibration

Each of the 5 copies reads and writes exactly ``--mib`` MiB.
"""
import argparse

import torch


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--mib", type=int, default=512)
    args = ap.parse_args()
    n = args.mib * 1024 * 1024 // 4
    a = torch.arange(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"copied {args.mib} MiB x 5, checksum {float(b[-1])}", flush=True)


if __name__ == "__main__":
    main()
