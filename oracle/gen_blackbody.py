"""Writes oracle/blackbody_rgb.inc: the blackbody colour table of the
reference (Const.h:7-61 == kernel.cuh:45-100), 391 RGB8 entries for
1000 K .. 40000 K in 100 K steps, as data.  Run once from the reference
checkout; the generated file is committed."""
import re
import sys

src = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/BlackTracer2/BlackTracer/Const.h"
text = open(src).read()
body = re.search(r"tempToRGBx\[1173\]\s*=\s*\{(.*?)\}", text, re.S).group(1)
vals = [int(v) for v in body.replace("\n", " ").split(",") if v.strip()]
assert len(vals) == 1173
out = ["/* Blackbody RGB8 colours, 1000 K + 100 K * index, index 0..390. */"]
for k in range(391):
    r, g, b = vals[3 * k:3 * k + 3]
    out.append(f"{r:3d}, {g:3d}, {b:3d},  /* {1000 + 100 * k:5d} K */")
open(sys.argv[2] if len(sys.argv) > 2 else "oracle/blackbody_rgb.inc", "w").write("\n".join(out) + "\n")
