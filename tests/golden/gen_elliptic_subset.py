"""Copy a subset of the reference's Carlson fixture into tests/golden/.

Source: geokerr_port/validation/test_data/elliptic_reference.dat (columns:
kind, x, y, z, p, value; written by geokerr_batch.f).  Keeps the first 120
RF, RC and RD rows.  RJ rows are dropped: that file's RJ values come from a
wrong formula (dnm variant) and pin nothing.
"""
import os
import sys

SRC = "/root/reference/geokerr_port/validation/test_data/elliptic_reference.dat"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "elliptic_reference_subset.txt")


def main() -> int:
    keep = {"RF": [], "RC": [], "RD": []}
    with open(SRC) as f:
        next(f)
        for line in f:
            parts = line.split()
            kind = parts[0]
            if kind in keep and len(keep[kind]) < 120:
                x, y, z, p, val = (float(v) for v in parts[1:6])
                if kind == "RC":
                    keep[kind].append(f"RC {x!r} {y!r} {val!r}")
                else:
                    keep[kind].append(f"{kind} {x!r} {y!r} {z!r} {val!r}")
    with open(DST, "w") as f:
        f.write("# subset of geokerr_port/validation/test_data/elliptic_reference.dat\n")
        f.write("# kind args... value   (RC: x y value; RF/RD: x y z value)\n")
        for k in ("RF", "RC", "RD"):
            f.write("\n".join(keep[k]) + "\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
