"""The ihs-sys Rust crate (botheadstudios_amd/rust/ihs-sys) against include/ihs_sph.h, without a Rust toolchain:
every entry point declared with the same parameter and return types, every struct with the same fields in the same
order and types, every constant with the same value, and the crate's compile-time size checks equal to what the C
compiler lays out for the header."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ihs_sph.h")
CRATE = os.path.join(ROOT, "botheadstudios_amd", "rust", "ihs-sys")

SCALARS = {"int": "c_int", "int32_t": "i32", "int64_t": "i64", "uint32_t": "u32", "uint64_t": "u64", "float": "f32",
           "double": "f64", "char": "c_char", "void": "c_void"}


def _header():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return re.sub(r"//[^\n]*", "", text)


def _rust():
    return open(os.path.join(CRATE, "src", "lib.rs")).read()


def _c_type(decl: str) -> str:
    """A C parameter or field type as the Rust spelling it must have."""
    decl = decl.strip()
    const = decl.startswith("const ")
    base = decl[6:] if const else decl
    stars = base.count("*")
    base = base.replace("*", "").strip()
    t = SCALARS.get(base, base)
    for k in range(stars):
        t = ("*const " if const and k == 0 else "*mut ") + t
    return t


def _c_functions():
    out = {}
    for ret, name, params in re.findall(r"\b(void|int|int64_t|const char\s*\*)\s+(ihs_\w+)\s*\(([^)]*)\)\s*;", _header()):
        ps = [] if params.strip() in ("", "void") else [p.strip() for p in params.split(",")]
        types = []
        for p in ps:
            m = re.match(r"(.*?)(\w+)$", p.replace("* ", "*").replace(" *", "*"))
            types.append(_c_type(m.group(1)))
        r = None if ret == "void" else _c_type(" ".join(ret.split()))
        out[name] = (types, r)
    return out


def _rust_functions():
    out = {}
    block = _rust()[_rust().index('extern "C" {'):]
    for name, params, ret in re.findall(r"pub fn (ihs_\w+)\(([^)]*)\)\s*(?:->\s*([^;]+))?;", block, flags=re.S):
        ps = [p.strip() for p in params.replace("\n", " ").split(",") if p.strip()]
        types = [re.sub(r"\s+", " ", p.split(":", 1)[1].strip()) for p in ps]
        out[name] = (types, ret.strip() if ret else None)
    return out


def _c_structs():
    out = {}
    for body, name in re.findall(r"typedef struct\s*\{(.*?)\}\s*(ihs_\w+)\s*;", _header(), flags=re.S):
        fields = []
        for line in body.split(";"):
            line = " ".join(line.split())
            if not line:
                continue
            m = re.match(r"((?:const )?\w+)\s+(.*)$", line)
            ctype = _c_type(m.group(1))
            for d in m.group(2).split(","):
                d = d.strip()
                a = re.match(r"(\w+)\[(\d+)\]$", d)
                fields.append((a.group(1), f"[{ctype}; {a.group(2)}]") if a else (d, ctype))
        out[name] = fields
    return out


def _rust_structs():
    out = {}
    for name, body in re.findall(r"pub struct (ihs_\w+) \{(.*?)\n\}", _rust(), flags=re.S):
        fields = re.findall(r"pub (\w+): ([^,\n]+),", body)
        out[name] = [(f, t.strip()) for f, t in fields]
    return out


def _c_constants():
    text = _header()
    out = {k: int(v) for k, v in re.findall(r"#define (IHS_\w+) \(?(-?\d+)\)?", text)}
    for body in re.findall(r"enum\s*\{(.*?)\}", text, flags=re.S):
        for k, v in re.findall(r"(IHS_\w+)\s*=\s*(-?\d+)", body):
            out[k] = int(v)
    return out


def test_every_entry_point_is_declared_with_the_header_types():
    c, r = _c_functions(), _rust_functions()
    assert len(c) >= 44
    assert set(c) == set(r), (sorted(set(c) - set(r)), sorted(set(r) - set(c)))
    for name, (ctypes_, cret) in c.items():
        rtypes, rret = r[name]
        assert rtypes == ctypes_, (name, rtypes, ctypes_)
        assert rret == cret, (name, rret, cret)


def test_every_struct_has_the_header_fields_in_order():
    c, r = _c_structs(), _rust_structs()
    assert set(c) == set(r) - {"ihs_handle"}, (set(c), set(r))
    for name, fields in c.items():
        assert r[name] == fields, (name, r[name], fields)


def test_every_constant_has_the_header_value():
    c = _c_constants()
    rust = {k: int(v) for k, v in re.findall(r"pub const (IHS_\w+): \w+ = (-?\d+);", _rust())}
    assert set(c) == set(rust), (sorted(set(c) - set(rust)), sorted(set(rust) - set(c)))
    for k, v in c.items():
        assert rust[k] == v, (k, rust[k], v)


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs a C compiler")
def test_struct_sizes_match_the_c_layout(tmp_path):
    asserted = {k: int(v) for k, v in re.findall(r"size_of::<(ihs_\w+)>\(\) == (\d+)", _rust())}
    assert set(asserted) == set(_c_structs())
    src = tmp_path / "sizes.c"
    src.write_text('#include <stdio.h>\n#include "ihs_sph.h"\nint main(void) {\n' +
                   "".join(f'  printf("{k} %zu\\n", sizeof({k}));\n' for k in sorted(asserted)) + "  return 0;\n}\n")
    exe = tmp_path / "sizes"
    subprocess.run(["gcc", "-I", os.path.dirname(HEADER), str(src), "-o", str(exe)], check=True)
    got = dict(line.split() for line in subprocess.run([str(exe)], check=True, capture_output=True,
                                                        text=True).stdout.splitlines())
    assert {k: int(v) for k, v in got.items()} == asserted
