#!/usr/bin/env python3
"""Regenerate ``csrc/tess_table_nv_data.inc`` from the upstream table header.

The tessellation patterns (vertex UVs, triangle indices and per-config ranges)
are data published under Apache-2.0 in
``src/tessellation_table_nv_raw.hpp`` of nvpro-samples/vk_tessellated_clusters
(format documented at ``README.md:30-92``).  The per-frame path must produce
bit-identical vertices and indices, so the numbers are taken verbatim; this
script only parses them and re-emits them as flat C arrays with the upstream
licence header and an attribution note.

Usage (only where the upstream tree is present):
    python gen_tess_table.py /path/to/tessellation_table_nv_raw.hpp
"""
import re
import sys
from pathlib import Path

OUT = Path(__file__).resolve().parent.parent / "csrc" / "tess_table_nv_data.inc"


def parse(text: str):
    def scalar(name):
        m = re.search(r"uint32_t\s+%s\s*=\s*(\d+)\s*;" % name, text)
        return int(m.group(1))

    def array(name, ctype):
        m = re.search(r"%s\s+%s\[\]\s*=\s*\{(.*?)\};" % (ctype, name), text, re.S)
        body = re.sub(r"//[^\n]*", "", m.group(1))
        return [int(tok, 0) for tok in re.findall(r"0x[0-9a-fA-F]+|\d+", body)]

    return {
        "max_edge_segments": scalar("max_edge_segments"),
        "max_vertices": scalar("max_vertices"),
        "max_triangles": scalar("max_triangles"),
        "max_configs": scalar("max_configs"),
        "vertices": array("vertices", "uint32_t"),
        "triangles": array("triangles", "uint32_t"),
        "configs": array("configs", "uint16_t"),
    }


def emit(d) -> str:
    assert len(d["vertices"]) == d["max_vertices"]
    assert len(d["triangles"]) == d["max_triangles"]
    assert len(d["configs"]) == 4 * d["max_configs"]
    lines = [
        "// SPDX-FileCopyrightText: Copyright (c) 2026 NVIDIA CORPORATION.",
        "// SPDX-License-Identifier: Apache-2.0",
        "//",
        "// Triangle tessellation pattern table, taken verbatim as data from",
        "// nvpro-samples/vk_tessellated_clusters src/tessellation_table_nv_raw.hpp",
        "// (Apache-2.0).  Regenerate with data/gen_tess_table.py; do not edit.",
        "//",
        "// vertices : u | v << 16, each component quantized to [0, 32768]",
        "// triangles: 3 x 8-bit counter-clockwise indices local to the config",
        "// configs  : {firstTriangle, firstVertex, numTriangles, numVertices},",
        "//            ordered x >= y >= z by increasing edge segments",
        "#define TC_TABLE_MAX_EDGE_SEGMENTS %d" % d["max_edge_segments"],
        "#define TC_TABLE_NUM_VERTICES %d" % d["max_vertices"],
        "#define TC_TABLE_NUM_TRIANGLES %d" % d["max_triangles"],
        "#define TC_TABLE_NUM_CONFIGS %d" % d["max_configs"],
    ]

    def arr(name, ctype, vals, per_line, fmt):
        out = ["static const %s %s[] = {" % (ctype, name)]
        for i in range(0, len(vals), per_line):
            out.append("  " + ", ".join(fmt % v for v in vals[i:i + per_line]) + ",")
        out.append("};")
        return out

    lines += arr("tc_table_nv_vertices", "uint32_t", d["vertices"], 8, "0x%08x")
    lines += arr("tc_table_nv_triangles", "uint32_t", d["triangles"], 8, "0x%08x")
    lines += arr("tc_table_nv_configs", "uint16_t", d["configs"], 12, "%d")
    return "\n".join(lines) + "\n"


def main():
    src = Path(sys.argv[1] if len(sys.argv) > 1 else
               "/root/reference/src/tessellation_table_nv_raw.hpp")
    OUT.write_text(emit(parse(src.read_text())))
    print("wrote", OUT)


if __name__ == "__main__":
    main()
