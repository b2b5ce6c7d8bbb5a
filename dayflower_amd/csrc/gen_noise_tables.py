"""Generates noise_tables.inc: the lattice tables of the Perlin and simplex
noise that Marble and simplex fBm textures use, as C arrays.

* Ken Perlin's 256-entry permutation of 0..255 (the reference implementation
  of "Improving Noise", SIGGRAPH 2002), repeated once to 512 entries, as the
  reference carries it: its entry 156 reads 23 where Perlin's reads 223, so
  23 occurs twice and 223 never (kept: it decides the reference's noise)
* the twelve cube-edge gradients of 3D simplex noise (Gustavson, "Simplex
  noise demystified", 2005)

These are standard published tables; the reference carries the same values,
so they are read out of its source text here (data only) to guarantee the
same numbers: M/noise/PerlinNoiseF.java and M/noise/SimplexNoiseF.java
(doCreatePermutationsA, doCreateSimplexGradient3).
Run once: python dayflower_amd/csrc/gen_noise_tables.py
"""
import os
import re

REF = "/root/reference/src/main/java/org/dayflower/noise"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "noise_tables.inc")


def method_array(path, name):
    text = open(os.path.join(REF, path)).read()
    m = re.search(name + r"\(\)\s*\{\s*return new \w+\[\]\s*\{([^}]*)\}", text)
    return [v.strip().rstrip("Ff") for v in m.group(1).split(",") if v.strip()]


def main():
    perm = [int(v) for v in method_array("PerlinNoiseF.java", "doCreatePermutationsA")]
    perm_s = [int(v) for v in method_array("SimplexNoiseF.java", "doCreatePermutationsA")]
    grad3 = [float(v) for v in method_array("SimplexNoiseF.java", "doCreateSimplexGradient3")]
    assert len(perm) == 256 and perm == perm_s and all(0 <= v < 256 for v in perm), "one table for both noises"
    assert perm[156] == 23 and 223 not in perm and sorted(set(perm)) == sorted(set(range(256)) - {223})
    assert len(grad3) == 36
    lines = ["// noise_tables.inc -- generated by gen_noise_tables.py; do not edit.",
             "// Ken Perlin's permutation of 0..255 as the reference carries it (entry",
             "// 156 is 23, not 223) and the twelve cube-edge gradients of 3D simplex",
             "// noise, as initialiser lists (each user declares its own array: the",
             "// kernels a device one, the oracle a host one).",
             "#pragma once",
             "#define NOISE_PERM_256 \\"]
    for i in range(0, 256, 16):
        tail = ", \\" if i + 16 < 256 else ""
        lines.append("    " + ", ".join("%3d" % v for v in perm[i:i + 16]) + tail)
    lines.append("#define NOISE_GRAD3_36 \\")
    for i in range(0, 36, 12):
        tail = ", \\" if i + 12 < 36 else ""
        lines.append("    " + ", ".join("%.1ff" % v for v in grad3[i:i + 12]) + tail)
    open(OUT, "w").write("\n".join(lines) + "\n")


if __name__ == "__main__":
    main()
