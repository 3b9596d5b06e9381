"""Writes tests/golden/svg_fixtures.npz: the reference's two SVG assets (assets/tiger.svg,
assets/features.svg) as input bytes, and what the reference's SVG ingest makes of them -- nanosvg
(third_party/nanosvg.h) and SvgScene's steps (core/svg.cpp:128-206) compiled from the unchanged
sources into oracle/_ref/libpfref.so by oracle/build_ref.sh: per draw path its kind (fill or
stroke), fill rule, paint, stroke style and dash pattern, the outline before stroking, and the
outline the scene receives (for a stroke: dashed with offset 0, then stroked, as
Canvas::stroke_path does).

Run in the build container, where the reference tree is present:
    python tests/golden/gen_svg_golden.py
tests/test_svg.py and tests/test_gpu_svg.py read the fixtures.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle  # noqa: E402

REF = os.environ.get("PF_REFERENCE", "/root/reference")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "svg_fixtures.npz")
DOCS = ("tiger", "features")


def main():
    arrays = {}
    for name in DOCS:
        with open(os.path.join(REF, "assets", f"{name}.svg"), "rb") as f:
            text = f.read()
        r = oracle.ref_svg(text)
        arrays[f"{name}/svg"] = np.frombuffer(text, np.uint8)
        for k, v in r.items():
            arrays[f"{name}/{k}"] = v
        print(f"{name}: {len(r['path_i32'])} draw paths, {len(r['out_flags'])} points")
    arrays["names"] = np.array(DOCS)
    np.savez_compressed(OUT, **arrays)
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
