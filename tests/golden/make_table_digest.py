#!/usr/bin/env python3
"""Writes table_digest.json: SHA-256 digests and sizes of the tessellation
table words exactly as the upstream header defines them, read from the
output of oracle/_ref/table_dump (compiled from that header by
`make -C oracle ref`)."""
import hashlib
import json
import subprocess
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
DUMP = HERE.parent.parent / "oracle" / "_ref" / "table_dump"


def parse(text):
    lines = text.split("\n")
    seg = int(lines[0].split()[1])
    v = [int(l[2:], 16) for l in lines if l.startswith("V ")]
    t = [int(l[2:], 16) for l in lines if l.startswith("T ")]
    c = [int(l[2:]) for l in lines if l.startswith("C ")]
    return seg, np.array(v, np.uint32), np.array(t, np.uint32), np.array(c, np.uint16)


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main():
    seg, v, t, c = parse(subprocess.run([str(DUMP)], capture_output=True, text=True, check=True).stdout)
    out = {"max_edge_segments": seg, "num_vertices": int(v.size), "num_triangles": int(t.size),
           "num_configs": int(c.size // 4), "vertices_sha256": digest(v), "triangles_sha256": digest(t),
           "configs_sha256": digest(c)}
    (HERE / "table_digest.json").write_text(json.dumps(out, indent=1) + "\n")
    print(out)


if __name__ == "__main__":
    sys.exit(main())
