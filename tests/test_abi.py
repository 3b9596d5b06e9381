"""The C-ABI library loads and exports every entry point include/pfh.h declares
(no device calls: runs without a GPU)."""
import ctypes
import os
import re
import subprocess

import pfh

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "pfh.h")).read()
    return sorted(set(re.findall(r"\b(pfh_[a-z0-9_]+)\s*\(", src)))


def test_header_and_binding_agree():
    assert set(declared_functions()) == set(pfh.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(pfh.LIB_PATH)
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_create_rejects_bad_device_without_crashing():
    lib = pfh.load()
    h = ctypes.c_void_p()
    assert lib.pfh_create(-1, ctypes.byref(h)) != 0
    assert not h.value
    assert lib.pfh_rasterize(None, None, 0, 0) == -1


def test_struct_layouts_match_header(tmp_path):
    # sizes and field offsets of the ABI structs, as a C compiler lays them out from the header
    fields = {"pfh_stats": pfh.Stats, "pfh_server_stats": pfh.ServerStats, "pfh_scene_desc": pfh.SceneDesc,
              "pfh_schedule": pfh.Schedule}
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "pfh.h"', "int main(void) {"]
    for cname, py in fields.items():
        lines.append(f'    printf("{cname} %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'    printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    lines += ["    return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = dict(line.split() for line in subprocess.run([str(exe)], check=True, capture_output=True,
                                                        text=True).stdout.splitlines())
    for cname, py in fields.items():
        assert int(got[cname]) == ctypes.sizeof(py), cname
        for f, _ in py._fields_:
            assert int(got[f"{cname}.{f}"]) == getattr(py, f).offset, f"{cname}.{f}"
