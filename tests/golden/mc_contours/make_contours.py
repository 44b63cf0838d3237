"""Writes contours.json: for each of the 256 marching-cubes cases of the
reference's TriTable (src/Utils/MC/MarchingCube_Kernel.cuh, read as text),
the outline of its polygons -- the pairs of cube edges joined by a triangle
side that belongs to exactly one triangle of the case -- and the set of
crossed edges (its EdgeTable entry).  Run here, where the reference is
mounted; the JSON is the committed fixture.

    python tests/golden/mc_contours/make_contours.py /root/reference
"""
import json
import re
import sys
from pathlib import Path


def main(ref_root: str) -> None:
    text = (Path(ref_root) / "src/Utils/MC/MarchingCube_Kernel.cuh").read_text(errors="replace")
    edge_body = text[text.index("EdgeTable[256]"):]
    edge_body = edge_body[edge_body.index("{") + 1:edge_body.index("};")]
    edge_table = [int(x, 16) for x in re.findall(r"0x[0-9a-fA-F]+", edge_body)]
    tri_body = text[text.index("TriTable[256][16]"):]
    tri_body = tri_body[:tri_body.index("};")]
    rows = re.findall(r"\{([-0-9, ]+)\}", tri_body)
    assert len(edge_table) == 256 and len(rows) == 256
    cases = []
    for c, row in enumerate(rows):
        v = [int(t) for t in row.split(",") if int(t) >= 0]
        outline = {}
        for k in range(0, len(v), 3):
            a, b, d = v[k:k + 3]
            for p, q in ((a, b), (b, d), (d, a)):
                key = (min(p, q), max(p, q))
                outline[key] = outline.get(key, 0) + 1
        pairs = sorted([list(k) for k, n in outline.items() if n == 1])
        crossed = [e for e in range(12) if edge_table[c] >> e & 1]
        cases.append({"case": c, "crossed": crossed, "outline": pairs,
                      "triangles": len(v) // 3})
    out = Path(__file__).with_name("contours.json")
    out.write_text(json.dumps(cases, separators=(",", ":")) + "\n")
    print(f"wrote {out} ({len(cases)} cases)")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
