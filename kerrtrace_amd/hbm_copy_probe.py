"""A plain device-to-device copy of known size, for calibrating HBM byte
counters: under `rocprofv3 --pmc FETCH_SIZE` (then `WRITE_SIZE`) the copy
kernel should read and write exactly --mib MiB; the ratio of the known size
to what the counter reports is the scale applied to the tracer's counts.

    python -m kerrtrace_amd.hbm_copy_probe --mib 1024
"""
import argparse


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--mib", type=int, default=1024)
    ap.add_argument("--reps", type=int, default=3)
    args = ap.parse_args()
    import torch
    n = args.mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(args.reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(f"copied {args.mib} MiB x {args.reps} ({args.mib * 1024} KiB read and written per copy)", flush=True)


if __name__ == "__main__":
    main()
