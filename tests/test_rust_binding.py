"""The psr-gpu-sys crate (rust/psr-gpu-sys) against include/psr_gpu.h, on the
CPU (no Rust toolchain is needed): every C entry point is declared in the
crate with the same arity, every #[repr(C)] struct has the C struct's field
names, offsets and size (the C side measured by compiling a probe with gcc
against the header), and the call-site patches apply to the reference's
files (crates/cli/src/main.rs:462-473, crates/search/src/confirmation.rs:528-536)."""
import os
import re
import shutil
import subprocess

import pytest

from conftest import REFERENCE, ROOT

CRATE = os.path.join(ROOT, "rust", "psr-gpu-sys", "src", "lib.rs")
HEADER = os.path.join(ROOT, "include", "psr_gpu.h")
PATCHES = os.path.join(ROOT, "rust", "patches")


def rust_source():
    return open(CRATE).read()


def c_functions():
    t = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    out = {}
    for m in re.finditer(r"^\s*(?:const\s+)?\w+\s*\*?\s*(psr_\w+)\s*\(([^;]*?)\);", t, re.M | re.S):
        args = " ".join(m.group(2).split())
        out[m.group(1)] = 0 if args == "void" else args.count(",") + 1
    return out


def rust_functions():
    src = rust_source()
    block = src[src.index('extern "C" {'):src.index("pub mod gpu")]
    out = {}
    for m in re.finditer(r"pub fn (psr_\w+)\((.*?)\)\s*(?:->\s*[\w:*\s]+)?;", block, re.S):
        args = " ".join(m.group(2).split()).strip().rstrip(",")
        out[m.group(1)] = 0 if not args else args.count(":")
    return out


def rust_structs():
    src = rust_source()
    out = {}
    for m in re.finditer(r"#\[repr\(C\)\]\s*(?:#\[derive\([^)]*\)\]\s*)?pub struct (psr_\w+) \{(.*?)\n\}", src, re.S):
        fields = re.findall(r"pub (\w+): ([^,\n]+),", m.group(2))
        if fields:
            out[m.group(1)] = fields
    return out


PRIM = {"f64": (8, 8), "u64": (8, 8), "usize": (8, 8), "u32": (4, 4), "c_int": (4, 4), "f32": (4, 4)}


def rust_layout(name, structs, memo):
    """(size, align, [(field, offset)]) of a #[repr(C)] struct."""
    if name in memo:
        return memo[name]
    off, align, fields = 0, 1, []
    for f, ty in structs[name]:
        size, al = rust_type(ty.strip(), structs, memo)
        off = (off + al - 1) // al * al
        fields.append((f, off))
        off += size
        align = max(align, al)
    size = (off + align - 1) // align * align
    memo[name] = (size, align, fields)
    return memo[name]


def rust_type(ty, structs, memo):
    if ty in PRIM:
        return PRIM[ty]
    if ty.startswith("*const") or ty.startswith("*mut"):
        return (8, 8)
    m = re.match(r"\[(.+);\s*(\d+)\]$", ty)
    if m:
        s, a = rust_type(m.group(1).strip(), structs, memo)
        return (s * int(m.group(2)), a)
    s, a, _ = rust_layout(ty, structs, memo)
    return (s, a)


def test_every_entry_point_is_declared_with_its_arity():
    c, r = c_functions(), rust_functions()
    assert len(c) >= 30
    assert set(c) == set(r), (set(c) ^ set(r))
    for name in c:
        assert c[name] == r[name], (name, c[name], r[name])


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_struct_layouts_match_the_header(tmp_path):
    structs = rust_structs()
    assert {"psr_candidate", "psr_search_cfg", "psr_confirmed", "psr_batch_opts"} <= set(structs)
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "psr_gpu.h"', "int main(void) {"]
    for s, fields in structs.items():
        lines.append('printf("%s size %%zu\\n", sizeof(%s));' % (s, s))
        for f, _ in fields:
            lines.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (s, f, s, f))
    lines.append("return 0; }")
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-I", os.path.dirname(HEADER), str(src), "-o", str(exe)], check=True)
    got = dict(l.rsplit(" ", 1) for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                         check=True).stdout.splitlines())
    memo = {}
    for s in structs:
        size, _, fields = rust_layout(s, structs, memo)
        assert int(got["%s size" % s]) == size, s
        for f, off in fields:
            assert int(got["%s.%s" % (s, f)]) == off, (s, f)


def test_patches_apply_to_the_reference(tmp_path):
    files = ["Cargo.toml", "crates/cli/Cargo.toml", "crates/search/Cargo.toml", "crates/cli/src/main.rs",
             "crates/search/src/confirmation.rs"]
    for f in files:
        if not os.path.exists(os.path.join(REFERENCE, f)):
            pytest.skip("reference tree not present")
        os.makedirs(tmp_path / os.path.dirname(f), exist_ok=True)
        shutil.copy(os.path.join(REFERENCE, f), tmp_path / f)
    for p in sorted(os.listdir(PATCHES)):
        r = subprocess.run(["patch", "-p1", "-i", os.path.join(PATCHES, p)], cwd=tmp_path, capture_output=True,
                           text=True)
        assert r.returncode == 0, (p, r.stdout, r.stderr)
    main = (tmp_path / "crates/cli/src/main.rs").read_text()
    assert "obs.search(&trials, discovery, &gpu_config)" in main and ".par_iter()" not in main.split("let lists")[1][:200]
