"""The Rust crate's raw bindings (mosaic-art-rust_amd/rust/src/hip_ffi.rs)
against the C ABI (include/mosaic_hip.h), checked on the CPU without cargo:
the same constants with the same values, the same structures with their
fields in the same order and of the same types (so #[repr(C)] lays them out
as C does), and the same functions with the same parameter and return
types.  Both files are parsed here; neither is generated from the other at
test time."""
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mosaic_hip.h")
CRATE = os.path.join(ROOT, "mosaic-art-rust_amd", "rust")
FFI = os.path.join(CRATE, "src", "hip_ffi.rs")

C_SCALARS = {"int32_t": "i32", "int64_t": "i64", "uint32_t": "u32", "uint64_t": "u64",
             "uint8_t": "u8", "float": "f32", "double": "f64", "size_t": "usize", "int": "c_int",
             "char": "c_char", "void": "c_void"}


def _c_type(decl):
    """Canonical type of a C declaration fragment without its name, e.g.
    'const float *' -> ('ptr', 'const', 'f32')."""
    decl = decl.strip()
    stars = decl.count("*")
    words = decl.replace("*", " ").split()
    const = "const" in words[:1]
    words = [w for w in words if w != "const"]
    t = C_SCALARS.get(words[0], words[0])
    for i in range(stars):
        t = ("ptr", "const" if (const and i == 0) else "mut", t)
    return t


def parse_header():
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    consts = {}
    for m in re.finditer(r"#define\s+(MH_[A-Z0-9_]+)\s+(.+)", text):
        v = m.group(2).strip()
        if v.startswith("("):
            v = v[1:-1]
        if v.endswith("ull"):
            v = v[:-3]
        consts[m.group(1)] = int(eval(v.replace("ull", "")))  # e.g. 64 << 20
    structs = {}
    for m in re.finditer(r"typedef struct\s*\{(.*?)\}\s*(\w+)\s*;", text, flags=re.S):
        fields = []
        for line in m.group(1).split(";"):
            line = line.strip()
            if not line:
                continue
            # "int32_t a, b" / "const uint8_t *p" / "int32_t x[2], y[2]"
            first = re.match(r"((?:const\s+)?\w+)\s*(.*)", line)
            base, rest = first.group(1), first.group(2)
            for part in rest.split(","):
                part = part.strip()
                pm = re.match(r"(\**)\s*(\w+)\s*(?:\[(\w+)\])?", part)
                stars, name, dim = pm.group(1), pm.group(2), pm.group(3)
                t = _c_type(base + " " + stars)
                if dim:
                    t = ("array", t, consts[dim] if dim in consts else int(dim))
                fields.append((name, t))
        structs[m.group(2)] = fields
    opaque = set(re.findall(r"typedef struct (\w+) \1;", text))
    funcs = {}
    for m in re.finditer(r"^([a-z][\w \*]*?)\b(mh_\w+)\s*\(([^;]*?)\)\s*;", text, flags=re.M | re.S):
        ret = _c_type(m.group(1))
        params = []
        if m.group(3).strip() != "void":
            for p in m.group(3).split(","):
                p = " ".join(p.split())
                pm = re.match(r"(.*?)(\w+)\s*(?:\[(\w+)\])?$", p)
                t = _c_type(pm.group(1))
                if pm.group(3):  # an array parameter is a pointer
                    t = ("ptr", "const" if pm.group(1).strip().startswith("const") else "mut", t)
                params.append(t)
        funcs[m.group(2)] = (ret, params)
    return consts, structs, opaque, funcs


def _r_type(t):
    t = t.strip()
    m = re.match(r"\*(const|mut)\s+(.*)", t)
    if m:
        return ("ptr", m.group(1), _r_type(m.group(2)))
    m = re.match(r"\[(.*);\s*(\d+)\]", t)
    if m:
        return ("array", _r_type(m.group(1)), int(m.group(2)))
    return t


def parse_rust():
    text = re.sub(r"//[^\n]*", "", open(FFI).read())
    consts = {m.group(1): int(m.group(2).replace("_", ""))
              for m in re.finditer(r"pub const (MH_\w+):\s*\w+\s*=\s*(-?[\d_]+)\s*;", text)}
    structs, opaque = {}, set()
    for m in re.finditer(r"pub struct (\w+)\s*\{(.*?)\}", text, flags=re.S):
        body = m.group(2).strip()
        if body.startswith("_"):
            opaque.add(m.group(1))
            continue
        fields = []
        for f in re.finditer(r"pub (\w+)\s*:\s*([^,]+?)\s*(?:,|$)", body):
            fields.append((f.group(1), _r_type(f.group(2))))
        structs[m.group(1)] = fields
    funcs = {}
    for m in re.finditer(r"pub fn (mh_\w+)\s*\((.*?)\)\s*(?:->\s*([^;]+))?;", text, flags=re.S):
        params = []
        for p in m.group(2).split(","):
            p = p.strip()
            if p:
                params.append(_r_type(p.split(":", 1)[1]))
        ret = _r_type(m.group(3)) if m.group(3) else "c_void"
        funcs[m.group(1)] = (ret, params)
    return consts, structs, opaque, funcs


def test_crate_files_exist():
    for f in ("Cargo.toml", "build.rs", os.path.join("src", "lib.rs"), os.path.join("src", "hip_ffi.rs")):
        assert os.path.exists(os.path.join(CRATE, f)), f
    build = open(os.path.join(CRATE, "build.rs")).read()
    assert "rustc-link-lib=dylib=mosaic_hip" in build


def test_constants_match():
    hc, _, _, _ = parse_header()
    rc, _, _, _ = parse_rust()
    assert rc == hc


def test_structures_match_field_for_field():
    _, hs, ho, _ = parse_header()
    _, rs, ro, _ = parse_rust()
    assert ro == ho
    assert sorted(rs) == sorted(hs)
    for name, fields in hs.items():
        assert rs[name] == fields, name


def test_functions_match_signature_for_signature():
    _, _, _, hf = parse_header()
    _, _, _, rf = parse_rust()
    assert sorted(rf) == sorted(hf)
    for name, sig in hf.items():
        assert rf[name] == sig, name


def test_parsers_see_the_whole_header(mh):
    # every exported symbol of the binding is a parsed function
    _, _, _, hf = parse_header()
    assert sorted(hf) == sorted(mh.EXPORTS)


def test_safe_wrappers_name_the_reference_api():
    src = open(os.path.join(CRATE, "src", "lib.rs")).read()
    for name in ("pub struct SimilarityDatabase", "pub fn build_similarities", "pub fn get_similarity",
                 "pub struct AdjacencyPenaltyCalculator", "pub fn calculate_penalty",
                 "pub fn calculate_total_cost", "pub fn calculate_swap_delta",
                 "pub struct MosaicOptimizer", "pub fn optimize", "pub fn optimize_greedy",
                 "pub struct OptimizationConfig", "pub struct OptimizationResult",
                 "pub struct ColorAdjustment", "pub fn adjust_pixel", "pub fn apply_to_image",
                 "pub fn calculate_optimal_adjustment", "pub fn calculate_average_lab",
                 "pub fn generate_mosaic"):
        assert name in src, name
