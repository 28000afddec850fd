"""The C ABI without a GPU: the library builds, loads, exports every entry point the
header declares, and its ctypes mirror of bm_columns has the C layout."""
import ctypes
import os
import re
import subprocess

import pytest

from bearing_amd import native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "bearing_mi355x.h")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|double|const char\*)\s+(bm_\w+)\s*\(", text, re.M)))


def test_header_declares_what_the_binding_knows():
    assert declared() == sorted(native.EXPORTS)


def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", native.LIB], capture_output=True, text=True, check=True).stdout
    names = set(line.split()[-1] for line in out.splitlines() if line.strip())
    missing = [n for n in declared() if n not in names]
    assert not missing, f"not exported: {missing}"


def test_library_loads_without_a_gpu():
    lib = native.load()
    for name in declared():
        assert hasattr(lib, name)


def test_ctypes_layout_matches_the_header(tmp_path):
    fields = ["n_worlds"] + [f for f, _, _, _ in native.FIELDS]
    src = tmp_path / "layout.c"
    body = "".join(f'  printf("%s %zu\\n", "{f}", offsetof(bm_columns, {f}));\n' for f in fields)
    src.write_text('#include <stddef.h>\n#include <stdio.h>\n#include "bearing_mi355x.h"\nint main(void) {\n'
                   f'  printf("size %zu\\n", sizeof(bm_columns));\n{body}  return 0;\n}}\n')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.dirname(HEADER), str(src), "-o", str(exe)], check=True)
    got = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True, text=True).stdout.splitlines())
    assert int(got["size"]) == ctypes.sizeof(native.Columns)
    for f in fields:
        assert int(got[f]) == getattr(native.Columns, f).offset, f


def test_create_without_a_gpu_fails_loudly():
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    with pytest.raises(native.BearingError):
        native.Context(0)


@pytest.mark.parametrize("struct, mirror", [("bm_bodies", native.Bodies), ("bm_world_state", native.WorldState)])
def test_world_structs_match_the_header(tmp_path, struct, mirror):
    fields = [f for f, _ in mirror._fields_]
    src = tmp_path / "layout.c"
    body = "".join(f'  printf("%s %zu\\n", "{f}", offsetof({struct}, {f}));\n' for f in fields)
    src.write_text('#include <stddef.h>\n#include <stdio.h>\n#include "bearing_mi355x.h"\nint main(void) {\n'
                   f'  printf("sizeof %zu\\n", sizeof({struct}));\n{body}  return 0;\n}}\n')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.dirname(HEADER), str(src), "-o", str(exe)], check=True)
    got = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True, text=True).stdout.splitlines())
    assert int(got["sizeof"]) == ctypes.sizeof(mirror)
    for f in fields:
        assert int(got[f]) == getattr(mirror, f).offset, f


def test_world_create_without_a_gpu_fails_loudly():
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    with pytest.raises(native.BearingError):
        native.World(0)


def test_world_state_layout_matches_the_header(tmp_path):
    fields = [f for f, _ in native.WorldState._fields_]
    src = tmp_path / "state.c"
    body = "".join(f'  printf("%s %zu\\n", "{f}", offsetof(bm_world_state, {f}));\n' for f in fields)
    src.write_text('#include <stddef.h>\n#include <stdio.h>\n#include "bearing_mi355x.h"\nint main(void) {\n'
                   f'  printf("size %zu\\n", sizeof(bm_world_state));\n{body}  return 0;\n}}\n')
    exe = tmp_path / "state"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.dirname(HEADER), str(src), "-o", str(exe)], check=True)
    got = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True, text=True).stdout.splitlines())
    assert int(got["size"]) == ctypes.sizeof(native.WorldState)
    for f in fields:
        assert int(got[f]) == getattr(native.WorldState, f).offset, f
