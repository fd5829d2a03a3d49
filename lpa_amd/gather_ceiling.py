"""The random-gather ceiling of the device: how many uniformly random label
rows the memory system delivers per second (lpa_calibration_gather), beside
a plain streaming copy (lpa_calibration_copy).  A step of label propagation
gathers one label row per arc (per 8 label vectors), so its arcs per second
are bounded by this rate, whatever the bytes per row.

Usage: python3 -m lpa_amd.gather_ceiling
prints, per case, the rows per second and the row bytes per second."""
import lpa_amd as L

CASES = (
    # (footprint bytes, 32-bit words per row, gathers): the label arrays of
    # R-MAT 27 at K = 8 and K = 1, and of R-MAT 22 at K = 8
    (4 * 8 * 134217728, 8, 1 << 30),
    (4 * 134217728, 1, 1 << 30),
    (4 * 8 * 4194304, 8, 1 << 28),
)


def main():
    e = L.Engine(0)
    c = e.calibration_copy(256 << 20, 3)
    print(f"copy 256 MiB: {c:.3f} ms ({2 * 256 * 1.048576 / c:.0f} GB/s)", flush=True)
    for nbytes, words, gathers in CASES:
        ms = e.calibration_gather(nbytes, words, gathers, 3)
        print(f"gather {gathers} rows of {4 * words} B over {nbytes / 1e9:.2f} GB: {ms:.2f} ms, "
              f"{gathers / ms / 1e6:.1f} G rows/s = {gathers * 4 * words / ms / 1e6:.0f} GB/s of row bytes",
              flush=True)
    e.close()


if __name__ == "__main__":
    main()
