"""Write the Rust `-sys` binding (vgi-sys/src/lib.rs) from include/vgi.h.

The binding is generated so that it cannot drift from the C header: the CPU
tests regenerate it and compare with the committed file. Run from anywhere:

    python global-illumination_amd/rust/gen_sys.py          # rewrite lib.rs
    python global-illumination_amd/rust/gen_sys.py --check  # exit 1 on drift

Only the constructs vgi.h uses are understood: `#define NAME <u32 literal>`,
opaque and field-list `typedef struct`s, and function prototypes whose
parameters are scalars, pointers, pointer-to-pointers or sized arrays.
"""
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
HEADER = os.path.normpath(os.path.join(HERE, "..", "..", "include", "vgi.h"))
OUTPUT = os.path.join(HERE, "vgi-sys", "src", "lib.rs")

SCALARS = {
    "int": "c_int",
    "char": "c_char",
    "float": "f32",
    "double": "f64",
    "int32_t": "i32",
    "uint32_t": "u32",
    "int64_t": "i64",
    "uint64_t": "u64",
    "uint16_t": "u16",
    "size_t": "usize",
}


def strip_cplusplus(text):
    return re.sub(r"#ifdef __cplusplus.*?#endif", "", text, flags=re.S)


def comment_text(block):
    """The lines of a /* ... */ comment without its frame."""
    body = block[2:-2]
    lines = []
    for ln in body.split("\n"):
        ln = re.sub(r"^\s*\*\s?", "", ln.rstrip())
        lines.append(ln.strip())
    while lines and not lines[0]:
        lines.pop(0)
    while lines and not lines[-1]:
        lines.pop()
    if lines:  # section banners: "---- title ----"
        lines[0] = re.sub(r"^-+\s*", "", lines[0])
        lines[-1] = re.sub(r"\s*-+$", "", lines[-1])
    return lines


def rust_type(ctype, structs):
    """Map a C parameter or return type (without the name) to Rust."""
    ctype = ctype.strip()
    stars = ctype.count("*")
    const = ctype.startswith("const ")
    base = ctype.replace("const ", "").replace("*", "").strip()
    if base == "void" and stars == 0:
        return None
    if base in SCALARS:
        rt = SCALARS[base]
    elif base in structs:
        rt = base
    elif base == "void":
        rt = "c_void"
    else:
        raise ValueError("unknown C type %r" % ctype)
    for level in range(stars):
        # Only the innermost pointee carries the const of `const T *`.
        rt = ("*const " if (const and level == 0) else "*mut ") + rt
    return rt


def parse_param(param, structs):
    param = " ".join(param.split())
    m = re.match(r"^(.*?)([A-Za-z_][A-Za-z0-9_]*)\s*\[(\d+)\]$", param)
    if m:  # a sized array parameter decays to a pointer
        ctype, name = m.group(1).strip() + " *", m.group(2)
    else:
        m = re.match(r"^(.*?[\s*])([A-Za-z_][A-Za-z0-9_]*)$", param)
        if not m:
            raise ValueError("cannot parse parameter %r" % param)
        ctype, name = m.group(1), m.group(2)
    if name in ("in", "type", "ref", "loop", "move", "fn", "use", "mod", "match", "box"):
        name = name + "_"
    return name, rust_type(ctype, structs)


def parse(text):
    text = strip_cplusplus(text)
    items = []  # (kind, data, docs)
    structs = set()
    pos = 0
    pending_doc = None
    token = re.compile(
        r"(?P<comment>/\*.*?\*/)"
        r"|(?P<define>#define[ \t]+(?P<dname>VGI_[A-Z0-9_]+)[ \t]+(?P<dvalue>[0-9][0-9a-fA-FxuU]*)[^\n]*)"
        r"|(?P<pre>#[^\n]*)"
        r"|(?P<opaque>typedef\s+struct\s+(?P<oname>\w+)\s+(?P=oname)\s*;)"
        r"|(?P<struct>typedef\s+struct\s+(?P<sname>\w+)\s*\{(?P<sbody>.*?)\}\s*(?P=sname)\s*;)"
        r"|(?P<func>(?P<fret>[A-Za-z_][A-Za-z0-9_ ]*?[\s*]+)(?P<fname>vgi_\w+)\s*\((?P<fargs>[^)]*)\)\s*;)"
        r"|(?P<blank>\n\s*\n)"
        r"|(?P<other>\S)",
        re.S,
    )
    for m in token.finditer(text):
        kind = m.lastgroup
        if kind in ("dname", "dvalue", "oname", "sname", "sbody", "fret", "fname", "fargs"):
            kind = next(k for k in ("comment", "define", "pre", "opaque", "struct", "func", "blank", "other") if m.group(k))
        if kind == "comment":
            pending_doc = comment_text(m.group("comment"))
            continue
        if kind == "blank":
            pending_doc = None
            continue
        if kind == "define":
            value = m.group("dvalue").rstrip("uU")
            tail = re.search(r"/\*(.*?)\*/", m.group("define"))
            items.append(("const", (m.group("dname"), int(value, 0)), [tail.group(1).strip()] if tail else []))
            continue  # a section comment above a block of defines documents the function after them
        elif kind == "opaque":
            structs.add(m.group("oname"))
            items.append(("opaque", m.group("oname"), pending_doc or []))
        elif kind == "struct":
            name = m.group("sname")
            structs.add(name)
            fields = []
            body = re.sub(r"/\*.*?\*/", "", m.group("sbody"), flags=re.S)
            for decl in body.split(";"):
                decl = " ".join(decl.split())
                if not decl:
                    continue
                fm = re.match(r"^(\w+)\s+(\w+)(?:\[(\d+)\])?$", decl)
                if not fm:
                    raise ValueError("cannot parse field %r of %s" % (decl, name))
                base = SCALARS[fm.group(1)]
                fields.append((fm.group(2), "[%s; %s]" % (base, fm.group(3)) if fm.group(3) else base))
            items.append(("struct", (name, fields), pending_doc or []))
        elif kind == "func":
            args = m.group("fargs").strip()
            params = [] if args in ("", "void") else [parse_param(p, structs) for p in args.split(",")]
            items.append(("fn", (m.group("fname"), params, rust_type(m.group("fret"), structs)), pending_doc or []))
        pending_doc = None
    return items


def render(items):
    out = [
        "// Generated by global-illumination_amd/rust/gen_sys.py from include/vgi.h; do not edit.",
        "//! Raw bindings of `libvgi.so`, the MI355X backend of bevy_voxel_gi's GPU",
        "//! transport seam (`GpuTracer`, crates/bevy_voxel_gi/src/gpu.rs:309-1236).",
        "//! Every item mirrors include/vgi.h one to one; the safe adapter lives in",
        "//! the `bevy_voxel_gi_amd` crate.",
        "#![allow(non_camel_case_types)]",
        "",
        "use std::os::raw::{c_char, c_int};",
        "",
    ]

    def docs(lines, indent=""):
        return [indent + ("/// " + ln if ln else "///") for ln in lines]

    funcs = []
    for kind, data, doc in items:
        if kind == "const":
            name, value = data
            out += docs(doc)
            out.append("pub const %s: u32 = %s;" % (name, "0x%x" % value if value > 0xFFFF else str(value)))
        elif kind == "opaque":
            out += docs(doc)
            out.append("#[repr(C)]")
            out.append("pub struct %s {" % data)
            out.append("    _private: [u8; 0],")
            out.append("}")
        elif kind == "struct":
            name, fields = data
            out += docs(doc)
            out.append("#[repr(C)]")
            out.append("#[derive(Clone, Copy, Debug, Default, PartialEq)]")
            out.append("pub struct %s {" % name)
            for fname, ftype in fields:
                out.append("    pub %s: %s," % (fname, ftype))
            out.append("}")
        elif kind == "fn":
            funcs.append((data, doc))
            continue
        out.append("")
    out.append('#[link(name = "vgi")]')
    out.append('extern "C" {')
    for k, ((name, params, ret), doc) in enumerate(funcs):
        if k:
            out.append("")
        out += docs(doc, "    ")
        args = ", ".join("%s: %s" % (n, t) for n, t in params)
        out.append("    pub fn %s(%s)%s;" % (name, args, "" if ret is None else " -> " + ret))
    out.append("}")
    out.append("")
    text = "\n".join(out)
    # Only the raw types actually used are imported.
    for t in ("c_char", "c_int"):
        if text.count(t) == 1:
            text = text.replace("use std::os::raw::{c_char, c_int};", "use std::os::raw::{%s};" % ("c_int" if t == "c_char" else "c_char"))
    return text


def generate():
    return render(parse(open(HEADER).read()))


def main():
    text = generate()
    if "--check" in sys.argv[1:]:
        current = open(OUTPUT).read() if os.path.exists(OUTPUT) else ""
        if current != text:
            print("%s is out of date with %s" % (OUTPUT, HEADER))
            return 1
        return 0
    with open(OUTPUT, "w") as f:
        f.write(text)
    return 0


if __name__ == "__main__":
    sys.exit(main())
