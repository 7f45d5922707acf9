"""Copy kernel of known size for calibrating HBM byte counters.

    rocprofv3 --pmc FETCH_SIZE -- python3 -m titan_amd.calibrate
    rocprofv3 --pmc WRITE_SIZE -- python3 -m titan_amd.calibrate

Copies a 1 GiB device buffer into another one (torch's copy kernel) three
times; the counter per dispatch divided by 1 GiB (1,048,576 KiB) is the
factor that converts the counter into bytes for the same run conditions.
"""
import sys

import torch

NBYTES = 1 << 30


def main():
    a = torch.ones(NBYTES // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print("copied %d bytes x3" % NBYTES, file=sys.stderr)


if __name__ == "__main__":
    main()
