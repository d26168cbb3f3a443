"""A plain device copy of known size, for scaling HBM traffic counters.

rocprofv3's FETCH_SIZE / WRITE_SIZE count L2 fill and write-back traffic in units that depend on
the request size; running them over a copy of exactly `MiB` MiB read and written gives the factor
that turns a kernel's counter into bytes:

    rocprofv3 --pmc FETCH_SIZE --output-format csv -d <dir> -o copy -- \\
        python3 -m ps26119_amd.copyprobe [MiB] [repeats]

Each repeat is one torch copy kernel of float64 data (the source written once beforehand); the
measured rate of the timed repeats is printed as a cross-check against the 8 TB/s HBM peak.
"""
import sys
import time


def main(argv=None):
    import torch
    argv = sys.argv[1:] if argv is None else argv
    mib = int(argv[0]) if argv else 64
    reps = int(argv[1]) if len(argv) > 1 else 10
    if not torch.cuda.is_available():
        raise SystemExit("copyprobe needs a GPU")
    n = mib * (1 << 20) // 8
    src = torch.arange(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t) / reps
    print(f"copy {mib} MiB x {reps}: {dt * 1e6:.1f} us each, {2 * mib * (1 << 20) / dt / 1e9:.0f} GB/s (read + write)",
          flush=True)


if __name__ == "__main__":
    main()
