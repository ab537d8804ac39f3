"""The ctypes mirrors of the C-ABI structs have the header's layout: a small C
program compiled against include/hamers_convective.h prints sizeof and
offsetof of every field, and the Python structs must agree (CPU only)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _c_layout(tmp_path, structs):
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "hamers_convective.h"', "int main(void) {"]
    for name, fields in structs.items():
        lines.append(f'printf("{name} sizeof %zu\\n", sizeof({name}));')
        for f in fields:
            lines.append(f'printf("{name} {f} %zu\\n", offsetof({name}, {f}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.check_call(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)])
    out = subprocess.check_output([str(exe)]).decode().split("\n")
    res = {}
    for line in out:
        if line.strip():
            s, f, v = line.split()
            res[(s, f)] = int(v)
    return res


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_struct_layouts_match_header(tmp_path):
    from hamers_amd import capi

    structs = {
        "hm_config": capi.hm_config,
        "hm_patch_view": capi.hm_patch_view,
        "hm_rk_patch_view": capi.hm_rk_patch_view,
        "hm_vjp_patch_view": capi.hm_vjp_patch_view,
        "hm_transport": capi.hm_transport,
    }
    c = _c_layout(tmp_path, {k: [f[0] for f in v._fields_] for k, v in structs.items()})
    for name, cls in structs.items():
        import ctypes

        assert c[(name, "sizeof")] == ctypes.sizeof(cls), name
        for f in cls._fields_:
            assert c[(name, f[0])] == getattr(cls, f[0]).offset, f"{name}.{f[0]}"
