"""Writes compiled_layout.json: the layout constants of the reference's
compiled scene (every `public static final int NAME = VALUE;` of
M/scene/compiler/Compiled*Cache.java) and the type IDs (`public static final
int ID = N;`) of the classes whose records it holds.  Run from the repository
root in a tree that has the reference at /root/reference; the JSON is the
fixture the tests read (the reference never travels)."""
import json
import os
import re

REF = "/root/reference/src/main/java/org/dayflower"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "compiled_layout.json")

CACHES = ["CompiledBoundingVolume3FCache", "CompiledCameraCache", "CompiledLightCache", "CompiledMaterialCache",
          "CompiledModifierCache", "CompiledPrimitiveCache", "CompiledShape3FCache", "CompiledTextureCache"]
ID_DIRS = {"geometry/shape": r".*3F$", "scene/material": r".*Material$", "scene/texture": r".*Texture$",
           "scene/light": r".*Light$", "scene/modifier": r".*Modifier$", "geometry/boundingvolume": r".*3F$",
           "geometry/boundingvolume/hierarchy": r".*BVHNode3F$"}
CONST = re.compile(r"public static final int ([A-Z0-9_]+)\s*=\s*(-?\d+)\s*;")


def main():
    constants, sources = {}, {}
    for c in CACHES:
        path = os.path.join(REF, "scene/compiler", c + ".java")
        for m in CONST.finditer(open(path).read()):
            constants[m.group(1)] = int(m.group(2))
            sources[m.group(1)] = f"M/scene/compiler/{c}.java"
    ids = {}
    for d, pat in ID_DIRS.items():
        for f in sorted(os.listdir(os.path.join(REF, d))):
            name = f[:-5]
            if not f.endswith(".java") or not re.match(pat, name):
                continue
            m = re.search(r"public static final int ID\s*=\s*(\d+)\s*;", open(os.path.join(REF, d, f)).read())
            if m:
                ids[name] = int(m.group(1))
    json.dump({"_source": "extracted from the reference's Java source text by tests/golden/gen_compiled_layout.py; "
                          "M/ = src/main/java/org/dayflower/",
               "constants": dict(sorted(constants.items())), "constant_sources": dict(sorted(sources.items())),
               "ids": dict(sorted(ids.items()))}, open(OUT, "w"), indent=1)
    print(f"{len(constants)} constants, {len(ids)} type IDs -> {OUT}")


if __name__ == "__main__":
    main()
