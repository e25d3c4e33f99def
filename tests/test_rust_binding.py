"""The Rust -sys binding (petalo_amd/rust/petalo-hip-sys/src/lib.rs) against
include/petalo_hip.h, without a Rust toolchain (none is installed here):

  * every struct of the header has a #[repr(C)] struct of the same name with
    the same fields in the same order, at the offsets and with the size the C
    compiler gives the header's struct (gcc offsetof/sizeof, printed by a
    generated C program); the Rust layout is computed with the C rules that
    #[repr(C)] follows;
  * every #define and enum constant has a Rust constant of the same value;
  * every function has an extern "C" declaration with the same name, return
    type and parameter types (C arrays in parameters are pointers).
A binding that would make the library read past a caller's struct (the
24-byte options struct of an older snippet against the header's 28 bytes)
fails here.
"""
import os
import re
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "petalo_hip.h")
RUST = os.path.join(ROOT, "petalo_amd", "rust", "petalo-hip-sys", "src", "lib.rs")

C_TO_RUST = {"float": "f32", "double": "f64", "int32_t": "i32", "uint32_t": "u32", "uint64_t": "u64",
             "int64_t": "i64", "size_t": "usize", "int": "c_int", "char": "c_char", "void": "c_void"}
RUST_SIZE = {"f32": 4, "f64": 8, "i32": 4, "u32": 4, "u64": 8, "i64": 8, "usize": 8, "c_int": 4, "c_char": 1,
             "u8": 1}


def _preprocessed():
    return subprocess.check_output(["gcc", "-E", "-P", "-x", "c", HEADER], text=True)


def header_structs():
    """{name: [(field, c_type, dims)]} of the header's typedef structs."""
    text = _preprocessed()
    out = {}
    for body, name in re.findall(r"typedef\s+struct\s*\{([^{}]*?)\}\s*(\w+)\s*;", text, re.S):
        if not name.startswith("petalo_"):
            continue
        fields = []
        for decl in body.split(";"):
            decl = " ".join(decl.split())
            if not decl:
                continue
            m = re.match(r"((?:const\s+)?[\w]+)\s+(.*)", decl)
            ctype, names = m.group(1), m.group(2)
            for nm in names.split(","):
                nm = nm.strip()
                dims = [int(d) for d in re.findall(r"\[(\d+)\]", nm)]
                fields.append((re.sub(r"\[.*", "", nm), ctype, dims))
        out[name] = fields
    return out


def c_layout(structs):
    """{name: (size, [(field, offset)])} from the C compiler."""
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void) {"]
    for name, fields in structs.items():
        lines.append(f'  printf("S {name} %zu\\n", sizeof({name}));')
        for f, _, _ in fields:
            lines.append(f'  printf("F {name} {f} %zu\\n", offsetof({name}, {f}));')
    lines.append("  return 0; }")
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "l.c"), os.path.join(d, "l")
        open(src, "w").write("\n".join(lines))
        subprocess.check_call(["gcc", "-o", exe, src])
        txt = subprocess.check_output([exe], text=True)
    lay = {}
    for ln in txt.splitlines():
        k = ln.split()
        if k[0] == "S":
            lay[k[1]] = [int(k[2]), []]
        else:
            lay[k[1]][1].append((k[2], int(k[3])))
    return lay


def rust_structs():
    """{name: [(field, rust_type)]} of the #[repr(C)] structs in lib.rs."""
    src = open(RUST).read()
    out = {}
    for m in re.finditer(r"#\[repr\(C\)\]\s*(?:#\[derive\([^)]*\)\]\s*)?pub struct (\w+)\s*\{(.*?)\n\}", src, re.S):
        fields = re.findall(r"(?:pub\s+)?(\w+)\s*:\s*([^,\n]+),", m.group(2))
        out[m.group(1)] = [(f, t.strip()) for f, t in fields]
    return out


def rust_layout(name, rs, memo):
    """(size, align, offsets) of a #[repr(C)] struct by the C layout rules."""
    if name in memo:
        return memo[name]

    def size_align(t):
        t = t.strip()
        m = re.match(r"\[(.+);\s*(\d+)\]$", t)
        if m:
            s, a = size_align(m.group(1))
            return s * int(m.group(2)), a
        if t in RUST_SIZE:
            return RUST_SIZE[t], RUST_SIZE[t]
        if t.startswith("*"):
            return 8, 8
        s, a, _ = rust_layout(t, rs, memo)
        return s, a

    off, align, offs = 0, 1, []
    for f, t in rs[name]:
        s, a = size_align(t)
        off = (off + a - 1) // a * a
        offs.append((f, off))
        off += s
        align = max(align, a)
    size = (off + align - 1) // align * align
    memo[name] = (size, align, offs)
    return memo[name]


def test_structs_match_header():
    hs = header_structs()
    assert len(hs) >= 12
    lay = c_layout(hs)
    rs = rust_structs()
    memo = {}
    for name, fields in hs.items():
        assert name in rs, f"{name} has no #[repr(C)] struct in the binding"
        assert [f for f, _ in rs[name]] == [f for f, _, _ in fields], f"{name}: fields differ"
        size, _, offs = rust_layout(name, rs, memo)
        csize, coffs = lay[name]
        assert offs == coffs, f"{name}: offsets {offs} vs {coffs}"
        assert size == csize, f"{name}: {size} bytes in Rust, {csize} in C"
        for (f, ct, dims), (_, rt) in zip(fields, rs[name]):
            base = C_TO_RUST.get(ct, ct)
            want = base
            for d in reversed(dims):
                want = f"[{want}; {d}]"
            assert rt.replace(" ", "") == want.replace(" ", ""), f"{name}.{f}: {rt} vs {want}"


def header_constants():
    raw = open(HEADER).read()
    consts = {k: int(v) for k, v in re.findall(r"#define\s+(PETALO_\w+)\s+(-?\d+)", raw)}
    for body in re.findall(r"enum\s*\{(.*?)\}", _preprocessed(), re.S):
        for k, v in re.findall(r"(PETALO_\w+)\s*=\s*(\d+)", body):
            consts[k] = int(v)
    return consts


def test_constants_match_header():
    hc = header_constants()
    assert len(hc) >= 25
    rc = {k: int(v) for k, v in re.findall(r"pub const (PETALO_\w+)\s*:\s*\w+\s*=\s*(-?\d+);", open(RUST).read())}
    assert hc == rc


def _c_param_type(decl):
    """C parameter (or return) declaration -> the Rust type of the binding."""
    decl = " ".join(decl.replace("*", " * ").split())
    decl = re.sub(r"\s*\[\d*\]", " [] ", decl).strip()
    toks = decl.split()
    # drop the parameter name (the last identifier after the type)
    if len(toks) > 1 and re.match(r"^[A-Za-z_]\w*$", toks[-1]) and toks[-1] not in C_TO_RUST and toks[-1] != "const" \
            and not toks[-1].endswith("_t"):
        toks = toks[:-1]
    elif len(toks) > 2 and toks[-1] == "[]":
        toks = toks[:-2] + ["*"]
    if "[]" in toks:
        i = toks.index("[]")
        toks = toks[:i - 1] + ["*"] + toks[i + 1:]
    consts, ptrs, base = [], [], None
    q = False
    for t in toks:
        if t == "const":
            q = True
        elif t == "*":
            ptrs.append(q)
            q = False
        else:
            base = t
            consts.append(q)
            q = False
    base_const = consts[0] if consts else False
    rust = C_TO_RUST.get(base, base)
    # pointer k points to an object qualified by the const written just
    # before its '*' (for the first pointer, also a const before the base type)
    for k, q in enumerate(ptrs):
        rust = f"*{'const' if (q or (k == 0 and base_const)) else 'mut'} {rust}"
    return rust


def header_functions():
    text = " ".join(_preprocessed().split())
    out = {}
    for ret, name, params in re.findall(r"((?:const\s+)?\w+\s*\*?)\s*(petalo_\w+)\s*\(([^)]*)\)\s*;", text):
        ps = [] if params.strip() in ("void", "") else [_c_param_type(p) for p in params.split(",")]
        out[name] = (_c_param_type(ret), ps)
    return out


def rust_functions():
    src = open(RUST).read()
    block = src[src.index('extern "C" {'):]
    out = {}
    for name, params, ret in re.findall(r"pub fn (\w+)\s*\((.*?)\)\s*(?:->\s*([^;]+))?;", block, re.S):
        ps = [" ".join(p.split(":", 1)[1].split()) for p in params.split(",") if p.strip()]
        out[name] = (" ".join((ret or "void").split()), ps)
    return out


def test_functions_match_header():
    hf = header_functions()
    rf = rust_functions()
    assert len(hf) >= 45
    assert sorted(hf) == sorted(rf), set(hf) ^ set(rf)
    for name, (ret, ps) in hf.items():
        rret, rps = rf[name]
        assert rret == ret, f"{name}: returns {rret}, header {ret}"
        assert rps == ps, f"{name}: parameters {rps} vs header {ps}"


def test_header_struct_sizes_are_the_documented_ones():
    lay = c_layout(header_structs())
    assert lay["petalo_lor_t"][0] == 32
    assert lay["petalo_mlem_options_t"][0] == 4 * len(header_structs()["petalo_mlem_options_t"]) + 8
