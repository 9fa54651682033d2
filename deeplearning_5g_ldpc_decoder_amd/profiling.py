"""Reference workloads for reading rocprofv3 counters.

``python -m deeplearning_5g_ldpc_decoder_amd.profiling copy --mib 1024``
runs an elementwise device copy of a known size (one read and one write of
every byte) a few times.  Its FETCH_SIZE / WRITE_SIZE per dispatch, divided
into the known byte counts, gives the factor that converts the same counters
measured on the decode kernel into HBM bytes per launch.
"""
import argparse
import sys


def copy_workload(mib, reps):
    import torch
    n = mib * 1024 * 1024 // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(reps):
        torch.mul(src, 1.0, out=dst)   # elementwise: reads n*4 bytes, writes n*4 bytes
    torch.cuda.synchronize()
    print(f"copy: {reps} dispatches of {n * 4} bytes read + {n * 4} bytes written",
          flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("copy")
    c.add_argument("--mib", type=int, default=1024)
    c.add_argument("--reps", type=int, default=3)
    args = ap.parse_args(argv)
    if args.cmd == "copy":
        copy_workload(args.mib, args.reps)
    return 0


if __name__ == "__main__":
    sys.exit(main())
