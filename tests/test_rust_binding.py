"""The Rust binding (rustfs-erasure-codec_amd/bindings/rust/ffi.rs) against the C
header: every function include/rsgpu.h declares has an extern item with the same
number of parameters, and every struct has the same field names in the same
order.  No Rust toolchain is needed: both files are parsed as text."""
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "rsgpu.h")
RUST = os.path.join(ROOT, "rustfs-erasure-codec_amd", "bindings", "rust")


def _strip_c(text):
    return re.sub(r"/\*.*?\*/", "", text, flags=re.S)


def _count_params(args):
    args = args.strip()
    if args in ("", "void"):
        return 0
    depth, n = 0, 1
    for ch in args:
        if ch in "([":
            depth += 1
        elif ch in ")]":
            depth -= 1
        elif ch == "," and depth == 0:
            n += 1
    return n - (1 if args.rstrip().endswith(",") else 0)


def c_functions():
    text = _strip_c(open(HEADER).read())
    out = {}
    for m in re.finditer(r"\b(rsgpu_[a-z0-9_]+)\s*\(([^;{]*?)\)\s*;", text):
        name, args = m.group(1), m.group(2)
        if "(*" in args or name.endswith("_cb"):
            continue
        out[name] = _count_params(args)
    return out


def rust_functions():
    text = open(os.path.join(RUST, "ffi.rs")).read()
    block = text[text.index('extern "C" {'):]
    return {m.group(1): _count_params(m.group(2))
            for m in re.finditer(r"pub fn (rsgpu_[a-z0-9_]+)\s*\((.*?)\)\s*(?:->[^;]*)?;", block, flags=re.S)}


def c_structs():
    text = _strip_c(open(HEADER).read())
    out = {}
    for m in re.finditer(r"typedef struct\s*(\w*)\s*\{(.*?)\}\s*(\w+)\s*;", text, flags=re.S):
        fields = []
        for decl in m.group(2).split(";"):
            decl = decl.strip()
            if not decl:
                continue
            names = [re.sub(r"[\s*]+", " ", p).split()[-1] for p in decl.split(",")]
            fields += names
        out[m.group(3)] = fields
    return out


def rust_structs():
    text = open(os.path.join(RUST, "ffi.rs")).read()
    return {m.group(1): re.findall(r"pub (\w+):", m.group(2))
            for m in re.finditer(r"pub struct (rsgpu_\w+)\s*\{(.*?)\}", text, flags=re.S)}


def test_every_header_function_is_bound_with_its_arity():
    c, r = c_functions(), rust_functions()
    assert len(c) > 70
    assert sorted(set(c) - set(r)) == []
    assert sorted(set(r) - set(c)) == []
    assert {n: (c[n], r[n]) for n in c if c[n] != r[n]} == {}


def test_struct_layouts_match():
    c, r = c_structs(), rust_structs()
    for name, fields in c.items():
        assert name in r, name
        assert r[name] == fields, name


def test_binding_sources_are_present():
    for f in ("ffi.rs", "mod.rs", "stream_io.rs", "build.rs"):
        assert os.path.exists(os.path.join(RUST, f)), f
    mod = open(os.path.join(RUST, "mod.rs")).read()
    assert "rsgpu_server_release" in mod and "pub mod stream_io;" in mod
