"""Write tests/golden/rng_golden.txt from the reference-built RNG table.

The table comes from oracle/_ref/ref_rng_table, which `make -C oracle ref`
builds from the reference's own module (TSLB_newBCS/vars_single_3D_module.f90,
rand_noseeded at :102-165, gauss_noseeded at :78-100) compiled with amdflang
and linked to oracle/ref_rng_harness.f90.  Each row holds the seed (i, j, k, l)
and the IEEE-754 bit patterns of rand_noseeded and gauss_noseeded (REAL(4)).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def main():
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_rng_table")
    if not os.path.exists(exe):
        subprocess.check_call(["make", "-C", os.path.join(ROOT, "oracle"), "ref"])
    out = subprocess.check_output([exe], text=True)
    rows = [ln.split() for ln in out.strip().splitlines()]
    dst = os.path.join(HERE, "rng_golden.txt")
    with open(dst, "w") as fh:
        fh.write("# i j k l rand_noseeded_bits gauss_noseeded_bits  (reference-built, amdflang)\n")
        for r in rows:
            fh.write(" ".join(r) + "\n")
    print(f"wrote {len(rows)} rows to {dst}", file=sys.stderr)


if __name__ == "__main__":
    main()
