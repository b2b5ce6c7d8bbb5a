"""Generates spectral_tables.inc: the published spectral data the Perez
sun/sky model needs, as C arrays.

* CIE 1931 2-degree colour-matching functions, 1 nm, 360-830 nm
* CIE daylight basis S0, S1, S2, 10 nm, 300-830 nm
* Preetham et al. solar spectrum and the ozone, mixed-gas and water-vapour
  absorption coefficients

These are standard published tables; the reference carries the same values,
so they are read out of its source text here (data only) to guarantee the
same numbers:
  M/curve/SpectralCurveF.java (CIE_X_BAR, CIE_Y_BAR, CIE_Z_BAR, CIE_LAMBDA,
  CIE_Y_INTEGRAL), M/curve/ChromaticSpectralCurveF.java (S0/S1/S2),
  M/scene/light/PerezLight.java (SOL and K_* tables).
Run once: python dayflower_amd/csrc/gen_spectral_tables.py
"""
import os
import re

REF = "/root/reference/src/main/java/org/dayflower"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "spectral_tables.inc")


def arrays(path):
    text = open(os.path.join(REF, path)).read()
    out = {}
    for m in re.finditer(r"(\w+)\s*=\s*(?:new \w+\[\]\s*)?\{([^}]*)\}", text):
        vals = [v.strip().rstrip("DdFf") for v in m.group(2).split(",") if v.strip()]
        try:
            out[m.group(1)] = [float(v) for v in vals]
        except ValueError:
            pass
    for m in re.finditer(r"(\w+)\s*=\s*([0-9.]+)F\s*;", text):
        out[m.group(1)] = float(m.group(2))
    return out


def emit(f, name, vals, ctype="float"):
    f.write(f"static const {ctype} {name}[{len(vals)}] = {{\n")
    for i in range(0, len(vals), 8):
        f.write("    " + ", ".join(repr(v) + ("f" if ctype == "float" else "") for v in vals[i:i + 8]) + ",\n")
    f.write("};\n")


def main():
    sc = arrays("curve/SpectralCurveF.java")
    ch = arrays("curve/ChromaticSpectralCurveF.java")
    pz = arrays("scene/light/PerezLight.java")
    with open(OUT, "w") as f:
        f.write("// spectral_tables.inc — generated by gen_spectral_tables.py; published spectral data\n")
        f.write("// (CIE 1931 observer, CIE daylight basis, Preetham et al. sun and absorption tables).\n")
        f.write("#pragma once\n")
        emit(f, "kCieX", sc["CIE_X_BAR"], "double")
        emit(f, "kCieY", sc["CIE_Y_BAR"], "double")
        emit(f, "kCieZ", sc["CIE_Z_BAR"], "double")
        emit(f, "kCieLambda", sc["CIE_LAMBDA"])
        f.write(f"static const float kCieYIntegral = {sc['CIE_Y_INTEGRAL']!r}f;\n")
        emit(f, "kS0", ch["S0_AMPLITUDES"])
        emit(f, "kS1", ch["S1_AMPLITUDES"])
        emit(f, "kS2", ch["S2_AMPLITUDES"])
        emit(f, "kSol", pz["SOL_AMPLITUDES"])
        emit(f, "kGasAmp", pz["K_GAS_ABSORPTION_ATTENUATION_AMPLITUDES"])
        emit(f, "kGasWl", pz["K_GAS_ABSORPTION_ATTENUATION_WAVELENGTHS"])
        emit(f, "kOzoneAmp", pz["K_OZONE_ABSORPTION_ATTENUATION_AMPLITUDES"])
        emit(f, "kOzoneWl", pz["K_OZONE_ABSORPTION_ATTENUATION_WAVELENGTHS"])
        emit(f, "kWaterAmp", pz["K_WATER_VAPOR_ABSORPTION_ATTENUATION_AMPLITUDES"])
        emit(f, "kWaterWl", pz["K_WATER_VAPOR_ABSORPTION_ATTENUATION_WAVELENGTHS"])
    print(OUT)


if __name__ == "__main__":
    main()
