"""The Rust drop-in crate (cfgable-matmul_amd/rust/cfgable-matmul-hip) cannot be compiled in
this environment (no Rust toolchain), so its raw declarations are checked against the C header
here: every `extern "C"` function of src/ffi.rs is declared in include/cfgmm.h with the same
number of parameters and corresponding types, and the `cfgmm_rho_problem` struct has the
header's fields in the header's order and types.  The crate's public call forms are checked
to be the reference's (names and parameter lists of src/impl_*.rs)."""
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CRATE = os.path.join(ROOT, "cfgable-matmul_amd", "rust", "cfgable-matmul-hip")
HEADER = os.path.join(ROOT, "include", "cfgmm.h")

C_TO_RUST = {
    "double *": "*mut f64", "const double *": "*const f64", "size_t": "usize", "int": "c_int",
    "const uint8_t *": "*const u8", "void *": "*mut c_void", "void **": "*mut *mut c_void",
    "const void *": "*const c_void", "const cfgmm_rho_problem *": "*const cfgmm_rho_problem",
    "const char *": "*const c_char",
}


def _norm_c_type(decl: str) -> str:
    """'const double *rhos' -> 'const double *' (parameter name dropped)."""
    decl = " ".join(decl.replace("*", " * ").split())
    toks = decl.split(" ")
    if toks and re.fullmatch(r"\w+", toks[-1]) and len(toks) > 1:
        toks = toks[:-1]
    t = " ".join(toks)
    return t.replace(" *", " *").replace("* *", "**").replace(" * ", " *").strip()


def header_functions():
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    out = {}
    for m in re.finditer(r"(?:int|const char \*)\s*(cfgmm_\w+)\s*\(([^)]*)\)\s*;", text):
        params = [p.strip() for p in m.group(2).split(",")]
        if params == ["void"]:
            params = []
        out[m.group(1)] = [_norm_c_type(p) for p in params]
    return out


def rust_functions():
    text = open(os.path.join(CRATE, "src", "ffi.rs")).read()
    text = re.sub(r"//[^\n]*", "", text)
    block = text[text.index('extern "C" {'):]
    out = {}
    for m in re.finditer(r"pub fn (cfgmm_\w+)\s*\(([^)]*)\)\s*->", block):
        params = [p.strip() for p in m.group(2).split(",") if p.strip()]
        out[m.group(1)] = [" ".join(p.split(":", 1)[1].split()) for p in params]
    return out


def test_every_rust_declaration_matches_the_header():
    hdr, rs = header_functions(), rust_functions()
    assert len(rs) >= 20
    for name, rtypes in rs.items():
        assert name in hdr, name
        ctypes_ = hdr[name]
        assert len(ctypes_) == len(rtypes), name
        for c, r in zip(ctypes_, rtypes):
            assert C_TO_RUST.get(c) == r, (name, c, r)


def test_rho_problem_struct_layout_matches_the_header():
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    body = re.search(r"typedef struct cfgmm_rho_problem\s*\{(.*?)\}\s*cfgmm_rho_problem;", text, re.S).group(1)
    cfields = []
    for decl in body.split(";"):
        decl = " ".join(decl.split())
        if not decl:
            continue
        base, names = re.match(r"((?:const )?\w+ ?\*?)\s*(.*)", decl).groups()
        star = "*" in base or names.startswith("*")
        for nm in names.split(","):
            nm = nm.strip()
            is_ptr = star or nm.startswith("*")
            cfields.append((nm.lstrip("*").strip(), _norm_c_type(base.replace("*", "").strip() + (" *" if is_ptr else ""))))
    rtext = open(os.path.join(CRATE, "src", "ffi.rs")).read()
    rbody = re.search(r"pub struct cfgmm_rho_problem \{(.*?)\}", rtext, re.S).group(1)
    rfields = [(m.group(1), " ".join(m.group(2).split())) for m in re.finditer(r"pub (\w+):\s*([^,]+),", rbody)]
    assert [f for f, _ in cfields] == [f for f, _ in rfields]
    for (f, c), (_, r) in zip(cfields, rfields):
        want = {"double *": "*mut f64", "const double *": "*const f64", "size_t": "usize",
                "const uint8_t *": "*const u8"}[c]
        assert want == r, (f, c, r)


def test_public_call_forms_are_the_reference_ones():
    """Module paths, function names and parameter names of the reference's public calls
    (src/impl_matmul.rs:155-248, src/impl_matmul_non0tab.rs:210-326, src/impl_aodmao2rho.rs:171-261)."""
    want = {
        "impl_matmul.rs": {
            "matmul_loop_parallel_mnk_pack_a": "c a b m n k lda ldb ldc transa transb",
            "matmul_loop_macro_mb": "c a b m n k lda ldb ldc transa transb",
            "matmul_anyway_full": "c a b m n k lda ldb ldc transa transb"},
        "impl_matmul_non0tab.rs": {
            "matmul_loop_parallel_mnk_pack_a_non0tab": "c a b m n k lda ldb ldc transa transb tab ldtab",
            "matmul_anyway_full_non0tab": "c a b m n k lda ldb ldc transa transb tab ldtab"},
        "impl_aodmao2rho.rs": {
            "aodmao2rho_loop_parallel": "rhos dms ao_lhs ao_rhs strides shape nset tab ldtab",
            "aodmao2rho_anyway": "rhos dms ao_lhs ao_rhs strides shape nset tab ldtab"},
    }
    for fname, fns in want.items():
        text = open(os.path.join(CRATE, "src", fname)).read()
        for fn, params in fns.items():
            m = re.search(r"pub fn " + fn + r"(?:<const BLKSIZE: usize>)?\s*\((.*?)\)\s*\{", text, re.S)
            assert m, (fname, fn)
            names = [p.split(":")[0].strip() for p in m.group(1).split(",") if p.strip()]
            assert names == params.split(), (fname, fn, names)
    lib = open(os.path.join(CRATE, "src", "lib.rs")).read()
    for mod in ("impl_aodmao2rho", "impl_matmul", "impl_matmul_non0tab", "structs", "prelude"):
        assert f"pub mod {mod};" in lib
    structs = open(os.path.join(CRATE, "src", "structs.rs")).read()
    assert re.search(r"pub struct MatmulLoops<\s*T,\s*const MC: usize,\s*const KC: usize,\s*const NC: usize,\s*"
                     r"const MR: usize,\s*const NR_LANE: usize,\s*const LANE: usize,\s*const MB: usize = 0,?\s*>",
                     structs)
    cargo = open(os.path.join(CRATE, "Cargo.toml")).read()
    assert re.search(r'\[lib\]\s*name = "cfgable_matmul"', cargo)
