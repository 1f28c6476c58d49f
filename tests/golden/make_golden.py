"""Writes the golden vectors of tests/golden from the reference's own REXI and
SPHIdentities headers, compiled from the reference tree into oracle/_ref
(oracle/Makefile, target `ref`).  Run in a tree where /root/reference exists:

    make -C oracle ref && python tests/golden/make_golden.py

Outputs (plain text, one row per entry, full double precision):
  rexi_h<h>_M<M>_<half|full>.txt   Re alpha, Im alpha, Re beta_re, Im beta_re, Re beta_im,
                                   Im beta_im per pole
  identities_L20.txt               n, m, R, S, A, B, C, D, E, G, H of SPHIdentities
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
DUMP = os.path.join(HERE, "..", "..", "oracle", "_ref", "ref_rexi_dump")

REXI_CASES = [(0.2, 64, True), (0.2, 128, True), (0.2, 188, True), (0.15, 32, False)]


def rexi_name(h, M, half):
    return "rexi_h%s_M%d_%s.txt" % (repr(h), M, "half" if half else "full")


def main():
    if not os.path.exists(DUMP):
        sys.exit("oracle/_ref/ref_rexi_dump is not built (make -C oracle ref)")
    for h, M, half in REXI_CASES:
        out = subprocess.run([DUMP, "rexi", repr(h), str(M), str(int(half))], check=True,
                             capture_output=True, text=True).stdout
        with open(os.path.join(HERE, rexi_name(h, M, half)), "w") as f:
            f.write(out)
    out = subprocess.run([DUMP, "ident", "20"], check=True, capture_output=True, text=True).stdout
    with open(os.path.join(HERE, "identities_L20.txt"), "w") as f:
        f.write(out)


if __name__ == "__main__":
    main()
