"""Write tests/golden/hu_materials.json from the reference's own data files
(read as text): the HU ranges of core/gate/data/HU_to_material.txt and the
densities of core/gate/data/GateMaterials.db.  Run in a container that has the
reference checkout; the JSON is committed."""
import json
import os
import re

REF = "/root/reference/pyvoxeldosimetry/core/gate/data"


def main():
    ranges = []
    for line in open(os.path.join(REF, "HU_to_material.txt")):
        line = line.strip()
        if not line or line.startswith("#"):
            continue
        lo, hi, name = line.split()
        ranges.append([float(lo), float(hi), name])
    dens = {}
    for line in open(os.path.join(REF, "GateMaterials.db")):
        m = re.match(r"\s*([A-Za-z0-9_-]+):\s*d\s*=\s*([0-9.]+)\s*(mg|g)/cm3", line)
        if m:
            dens[m.group(1)] = float(m.group(2)) * (1e-3 if m.group(3) == "mg" else 1.0)
    out = {"ranges": ranges, "densities": {r[2]: dens.get(r[2]) for r in ranges}}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hu_materials.json")
    json.dump(out, open(path, "w"), indent=1)
    print(path)


if __name__ == "__main__":
    main()
