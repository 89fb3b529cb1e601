"""The Rust binding a plugin would link: the `-sys` crate
(global-illumination_amd/rust/vgi-sys/src/lib.rs, generated from
include/vgi.h) and the `GpuTracer` adapter crate
(global-illumination_amd/rust/bevy_voxel_gi_amd/src/lib.rs) against the C
header, by parsing all three (no Rust toolchain is needed; none runs here).

- the generated binding equals what the generator writes from vgi.h now;
- every function vgi.h declares is in the extern block with the same
  parameter count, pointer kinds (`*const` for `const T *`, `*mut` else) and
  return presence, and nothing else is;
- every struct has the header's fields in order, with the header's sizes
  (32-byte tasks, 96-byte results, 48-byte lights: gpu.rs:136-177,
  voxel_trace.wgsl:5-20);
- the adapter calls every extern function, each call with the declared
  number of arguments.
"""
import importlib.util
import os
import re

from conftest import ROOT

RUST = os.path.join(ROOT, "global-illumination_amd", "rust")
SYS = os.path.join(RUST, "vgi-sys", "src", "lib.rs")
ADAPTER = os.path.join(RUST, "bevy_voxel_gi_amd", "src", "lib.rs")
HEADER = os.path.join(ROOT, "include", "vgi.h")

C_SIZES = {"int32_t": 4, "uint32_t": 4, "float": 4, "int": 4, "int64_t": 8, "uint64_t": 8, "double": 8, "uint16_t": 2}
RUST_SIZES = {"i32": 4, "u32": 4, "f32": 4, "c_int": 4, "i64": 8, "u64": 8, "f64": 8, "u16": 2}


def header_text():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return re.sub(r"#ifdef __cplusplus.*?#endif", "", text, flags=re.S)


def header_functions():
    """name -> (list of 'const'|'mut'|'value' per parameter, has return)."""
    out = {}
    for m in re.finditer(r"([A-Za-z_][\w ]*?[\s*]+)(vgi_\w+)\s*\(([^)]*)\)\s*;", header_text()):
        ret, name, args = m.group(1).strip(), m.group(2), m.group(3).strip()
        kinds = []
        if args not in ("", "void"):
            for p in args.split(","):
                p = " ".join(p.split())
                if "*" in p or "[" in p:
                    kinds.append("const" if p.startswith("const ") and p.count("*") == 1 else "mut")
                else:
                    kinds.append("value")
        out[name] = (kinds, ret != "void")
    return out


def header_structs():
    out = {}
    for m in re.finditer(r"typedef\s+struct\s+(\w+)\s*\{(.*?)\}\s*\1\s*;", header_text(), flags=re.S):
        fields = []
        for decl in m.group(2).split(";"):
            decl = " ".join(decl.split())
            if decl:
                fm = re.match(r"^(\w+)\s+(\w+)(?:\[(\d+)\])?$", decl)
                fields.append((fm.group(2), C_SIZES[fm.group(1)] * int(fm.group(3) or 1)))
        out[m.group(1)] = fields
    return out


def rust_functions():
    text = open(SYS).read()
    block = re.search(r'extern "C" \{(.*)\n\}', text, flags=re.S).group(1)
    out = {}
    for m in re.finditer(r"pub fn (vgi_\w+)\(([^)]*)\)(\s*->\s*[^;]+)?;", block):
        kinds = []
        for p in [p for p in m.group(2).split(",") if p.strip()]:
            t = p.split(":", 1)[1].strip()
            kinds.append("const" if t.startswith("*const") else "mut" if t.startswith("*mut") else "value")
        out[m.group(1)] = (kinds, m.group(3) is not None)
    return out


def rust_structs():
    out = {}
    for m in re.finditer(r"pub struct (vgi_\w+) \{(.*?)\n\}", open(SYS).read(), flags=re.S):
        fields = []
        for fm in re.finditer(r"pub (\w+): (\[(\w+); (\d+)\]|\w+),", m.group(2)):
            if fm.group(3):
                fields.append((fm.group(1), RUST_SIZES[fm.group(3)] * int(fm.group(4))))
            else:
                fields.append((fm.group(1), RUST_SIZES[fm.group(2)]))
        if fields:
            out[m.group(1)] = fields
    return out


def test_binding_is_generated_from_the_header():
    spec = importlib.util.spec_from_file_location("gen_sys", os.path.join(RUST, "gen_sys.py"))
    gen = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gen)
    assert open(SYS).read() == gen.generate(), "run global-illumination_amd/rust/gen_sys.py"


def test_extern_block_matches_the_header():
    want, got = header_functions(), rust_functions()
    assert len(want) > 60
    assert sorted(got) == sorted(want), sorted(set(got) ^ set(want))
    for name in want:
        assert got[name] == want[name], (name, got[name], want[name])


def test_structs_match_the_header():
    want, got = header_structs(), rust_structs()
    assert sorted(got) == sorted(want)
    for name in want:
        assert got[name] == want[name], name
    size = {n: sum(s for _, s in f) for n, f in got.items()}
    assert size["vgi_face_task"] == 32 and size["vgi_face_result"] == 96 and size["vgi_light"] == 48


def call_arities(text, name):
    """Argument counts of every `sys::name(...)` call in adapter source."""
    out = []
    for m in re.finditer(r"sys::%s\(" % name, text):
        depth, i, args, cur = 1, m.end(), 0, ""
        while depth:
            c = text[i]
            if c in "([{":
                depth += 1
            elif c in ")]}":
                depth -= 1
            elif c == "," and depth == 1:
                args += 1
            if depth:
                cur += c
            i += 1
        out.append(args + (1 if cur.strip().rstrip(",").strip() else 0) - (1 if cur.strip().endswith(",") else 0))
    return out


def test_adapter_calls_every_entry_point():
    text = open(ADAPTER).read()
    declared = rust_functions()
    for name, (kinds, _) in declared.items():
        arities = call_arities(text, name)
        assert arities, "the adapter never calls %s" % name
        assert all(a == len(kinds) for a in arities), (name, arities, len(kinds))
    called = set(re.findall(r"sys::(vgi_\w+)\(", text))
    assert called <= set(declared), called - set(declared)
