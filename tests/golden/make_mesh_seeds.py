"""Writes tests/golden/mesh_seeds.json from the reference's mesh seed data
files (meshSeeds/cubedSphereSeed.dat, meshSeeds/icosTriSphereSeed.dat): the
vertex coordinates as the files' decimal strings, the edge table
(orig, dest, left face, right face) and each root face's vertex and edge
lists.  Only the data files are read; nothing of the reference is run.

    python tests/golden/make_mesh_seeds.py [/root/reference/meshSeeds]
"""
import json
import sys
from pathlib import Path

SECTIONS = ("x", "edgeO", "faceverts", "faceedges", "vertexdegree")


def parse(path: Path, nverts: int) -> dict:
    rows, key = {k: [] for k in SECTIONS}, None
    for line in path.read_text().splitlines():
        tok = line.split()
        if not tok:
            continue
        if tok[0] in SECTIONS:
            key = tok[0]
            continue
        rows[key].append(tok)
    return {"xyz": [r[:3] for r in rows["x"][:nverts]],
            "edges": [[int(v) for v in r[:4]] for r in rows["edgeO"]],
            "face_verts": [[int(v) for v in r] for r in rows["faceverts"]],
            "face_edges": [[int(v) for v in r] for r in rows["faceedges"]]}


def main() -> None:
    src = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/meshSeeds")
    out = {"source": "meshSeeds/cubedSphereSeed.dat, meshSeeds/icosTriSphereSeed.dat",
           "cubed": parse(src / "cubedSphereSeed.dat", 8),
           "icos": parse(src / "icosTriSphereSeed.dat", 12)}
    dest = Path(__file__).resolve().parent / "mesh_seeds.json"
    dest.write_text(json.dumps(out, indent=1) + "\n")
    print("wrote", dest)


if __name__ == "__main__":
    main()
