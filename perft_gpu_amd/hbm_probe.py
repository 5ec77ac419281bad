"""Plain-copy reference run: launches the library's 16-byte copy kernel on a
buffer of known size, so HBM byte counters (FETCH_SIZE / WRITE_SIZE) of the
perft kernels can be scaled by what they read for a known traffic.

    python3 -m perft_gpu_amd.hbm_probe [MiB] [iters]
"""
import sys

import perft_gpu_amd as pg


def main():
    mib = int(sys.argv[1]) if len(sys.argv) > 1 else 1024
    iters = int(sys.argv[2]) if len(sys.argv) > 2 else 5
    with pg.Device(0, arena_bytes=(2 * mib + 64) << 20) as d:
        ms, gbs = d.copy_bandwidth(mib << 20, iters)
        print("copy %d MiB: %.3f ms per launch, %.1f GB/s (read+write)" % (mib, ms, gbs))


if __name__ == "__main__":
    main()
