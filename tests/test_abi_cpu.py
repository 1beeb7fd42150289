"""CPU-side checks of the C-ABI library: it loads and exports every symbol the headers declare
(no compute calls: there is no GPU in the build container)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "fe_gpu_amd", "lib", "libfe_gpu_hip.so")


def declared(header):
    src = open(os.path.join(ROOT, "include", header)).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"^\s*int\s+(\w+)\s*\(", src, flags=re.M)))


def test_headers_declare_the_reference_abi():
    names = declared("fe_gpu_runtime.h")
    # the 28 entry points of src/gpu/fe_gpu_runtime.h:23-131, plus the allocation-free SYRK, which
    # has no counterpart there
    assert len(names) == 29 and "fe_gpu_syrk_noalloc" in names
    for n in ["fe_gpu_runtime_init", "fe_gpu_fe_accumulate", "fe_gpu_build_multi_cluster_ids", "fe_gpu_gemm",
              "fe_gpu_cluster_meat", "fe_gpu_absmax", "fe_gpu_axpy"]:
        assert n in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_library_exports_every_declared_symbol():
    lib = C.CDLL(LIB)
    for h in ("fe_gpu_runtime.h", "fehip.h"):
        for name in declared(h):
            assert hasattr(lib, name), name


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_python_bindings_cover_headers():
    from fe_gpu_amd import _lib, _coarse_sigs
    bound = {n for n, _, _ in _lib.ABI} | {n for n, _, _ in _coarse_sigs.COARSE}
    for h in ("fe_gpu_runtime.h", "fehip.h"):
        for name in declared(h):
            assert name in bound, name
    L = _lib.load()
    assert L is not None


def test_ctypes_struct_layouts_match_headers():
    from fe_gpu_amd import _coarse_sigs as S
    from fe_gpu_amd._lib import DeviceInfo
    # fe_gpu_device_info: int, size_t, int, int, int, char[256] -> 288 bytes (header layout)
    assert C.sizeof(DeviceInfo) == 288
    assert C.sizeof(S.DemeanOpts) == 32
    assert C.sizeof(S.Opts) == 32 + 4 + 16 + 4 + 8 + 4 + 4 + 8


def test_ctypes_structs_match_the_c_compiler(tmp_path):
    """sizeof and every field offset of the fehip.h structs, as gcc lays them out."""
    import shutil
    import subprocess
    from fe_gpu_amd import _coarse_sigs as S
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    structs = {"fehip_panel_info": S.PanelInfo, "fehip_demean_opts": S.DemeanOpts,
               "fehip_demean_stats": S.DemeanStats, "fehip_opts": S.Opts, "fehip_result": S.Result,
               "fehip_request": S.Request, "fehip_server_opts": S.ServerOpts,
               "fehip_server_stats": S.ServerStats, "fehip_bin_header": S.BinHeader}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "fehip.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {(a, b): int(c) for a, b, c in (ln.split() for ln in out if ln)}
    for cname, py in structs.items():
        assert got[(cname, "sizeof")] == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert got[(cname, f)] == getattr(py, f).offset, (cname, f)


def test_product_never_imports_the_oracle():
    pkg = os.path.join(ROOT, "fe_gpu_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".cpp", ".hpp", ".h")):
                txt = open(os.path.join(dirpath, f)).read()
                assert "import oracle" not in txt and "from oracle" not in txt and "fe_oracle" not in txt, f
