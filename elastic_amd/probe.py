"""Streaming-kernel probe of the achievable HBM bandwidth (and the
calibration run for hardware byte counters): copies, reads or writes a
buffer of the given size with the library's own kernel and prints GB/s.

    python -m elastic_amd.probe [bytes] [reps] [mode]   (mode 0 copy, 1 non-temporal copy, 2 read, 3 write)
"""
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

from elastic_amd import lib  # noqa: E402


def main(argv):
    nbytes = int(argv[1]) if len(argv) > 1 else 1 << 30
    reps = int(argv[2]) if len(argv) > 2 else 3
    mode = int(argv[3]) if len(argv) > 3 else 0
    gbs, ms = lib.bandwidth_probe(nbytes, reps, mode)
    print(f"bytes {nbytes} mode {mode}: {gbs:.1f} GB/s, {ms:.4f} ms per pass")
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
