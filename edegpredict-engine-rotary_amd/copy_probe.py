"""Plain device copy of a known size: the calibration kernel for the
rocprofv3 FETCH_SIZE / WRITE_SIZE counters.

Runs `out = in * 1` over a float32 buffer (an elementwise kernel that reads
and writes every byte once, which the runtime cannot turn into a DMA copy)
`--reps` times and prints the bytes each launch must move.  The ratio of those
bytes to what the counters report for the same launch scales the counters of
the solver's kernels to HBM bytes.
"""
import argparse
import json

import torch


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--mib", type=int, default=1024, help="size of each buffer in MiB")
    ap.add_argument("--reps", type=int, default=3)
    args = ap.parse_args()
    n = args.mib * (1 << 20) // 4
    a = torch.empty(n, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(args.reps):
        torch.mul(a, 1.0, out=b)
    torch.cuda.synchronize()
    print(json.dumps({"kernel": "elementwise mul (copy)", "read_bytes": 4 * n, "write_bytes": 4 * n,
                      "launches": args.reps}), flush=True)


if __name__ == "__main__":
    main()
