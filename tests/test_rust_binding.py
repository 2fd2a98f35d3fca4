"""The Rust binding (bindings/rust) against the C header it binds, without a Rust toolchain:
every function the header declares has exactly one `extern "C"` item in src/ffi.rs with the
same name, the same number of arguments and, argument by argument, the same kind (scalar
type, or pointer with the same constness and pointee); the #[repr(C)] structs list the
header structs' fields in order with matching types; the safe wrappers in src/lib.rs call
only declared items."""
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "hyperprec_hip.h")
FFI = os.path.join(ROOT, "bindings", "rust", "src", "ffi.rs")
LIBRS = os.path.join(ROOT, "bindings", "rust", "src", "lib.rs")

C_SCALAR = {"size_t": "usize", "int": "c_int", "double": "f64", "uint64_t": "u64", "float": "f32"}
C_POINTEE = {"double": "f64", "char": "c_char", "uint8_t": "u8", "size_t": "usize", "int": "c_int", "int64_t": "i64",
             "void": "c_void", "hp_solution": "HpSolution", "hp_refine_stats": "HpRefineStats", "hp_batch": "HpBatch",
             "hp_adjoint": "HpAdjoint", "hp_adjoint_info": "HpAdjointInfo", "hp_gemm_plan": "HpGemmPlan"}


def c_type_to_rust(t):
    """'const double*' -> '*const f64', 'double* const*' -> '*const *mut f64', 'size_t' -> 'usize'."""
    t = re.sub(r"\s+", " ", t.strip()).replace(" *", "*")
    if "*" not in t:
        return C_SCALAR[t]
    # peel pointer levels from the right: each '*' optionally followed by ' const' (pointer constness)
    parts = re.findall(r"\*( const)?", t)
    base = t[:t.index("*")].strip()
    base_const = base.startswith("const ")
    base = base.replace("const ", "").strip()
    rust = C_POINTEE[base]
    # innermost pointer's pointee constness comes from the base; outer pointers' from the
    # 'const' after the previous '*'
    consts = [base_const] + [bool(p) for p in parts[:-1]]
    for c in consts:
        rust = ("*const " if c else "*mut ") + rust
    return rust


def header_decls():
    src = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    out = {}
    for ret, name, args in re.findall(r"([A-Za-z_][A-Za-z0-9_ \*]*?)\b(hp_[a-z0-9_]+)\s*\(([^)]*)\)\s*;", src):
        args = " ".join(args.split())
        params = []
        if args and args != "void":
            for a in args.split(","):
                m = re.match(r"(.*?)([A-Za-z_][A-Za-z0-9_]*)$", a.strip())
                params.append(c_type_to_rust(m.group(1)))
        ret = " ".join(ret.split())
        out[name] = (params, None if ret == "void" else c_type_to_rust(ret))
    return out


def rust_decls():
    src = open(FFI).read()
    block = src[src.index('extern "C" {'):]
    out = {}
    for name, args, ret in re.findall(r"pub fn (hp_[a-z0-9_]+)\(([^)]*)\)(?:\s*->\s*([^;]+))?;", block):
        params = [re.sub(r"\s+", " ", a.split(":", 1)[1].strip()) for a in args.split(",") if a.strip()]
        assert name not in out, f"{name} declared twice"
        out[name] = (params, ret.strip() if ret else None)
    return out


def test_every_header_function_is_bound_with_matching_arguments():
    h, r = header_decls(), rust_decls()
    assert len(h) > 60
    assert set(h) == set(r), (sorted(set(h) - set(r)), sorted(set(r) - set(h)))
    for name, (params, ret) in h.items():
        rp, rr = r[name]
        assert len(params) == len(rp), (name, params, rp)
        assert params == rp, (name, params, rp)
        assert ret == rr, (name, ret, rr)


def header_structs():
    src = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    out = {}
    for body, name in re.findall(r"typedef struct \{(.*?)\}\s*(hp_[a-z_]+);", src, flags=re.S):
        fields = []
        for line in body.split(";"):
            line = line.strip()
            if not line:
                continue
            m = re.match(r"(.+?)\s+([a-z_0-9]+)(\[(\d+)\])?$", line)
            t = C_SCALAR[m.group(1).strip()]
            fields.append((m.group(2), f"[{t}; {m.group(4)}]" if m.group(3) else t))
        out[name] = fields
    return out


def test_repr_c_structs_match_the_header():
    rs = open(FFI).read()
    names = {"hp_solution": "HpSolution", "hp_refine_stats": "HpRefineStats", "hp_gemm_plan": "HpGemmPlan",
             "hp_adjoint_info": "HpAdjointInfo"}
    hs = header_structs()
    assert set(hs) == set(names), sorted(hs)
    for cname, rname in names.items():
        body = re.search(r"#\[repr\(C\)\][^{]*pub struct " + rname + r" \{(.*?)\n\}", rs, flags=re.S).group(1)
        fields = [(m.group(1), re.sub(r"\s+", " ", m.group(2).strip()))
                  for m in re.finditer(r"pub ([a-z_0-9]+): ([^,]+),", body)]
        assert fields == hs[cname], (cname, fields, hs[cname])


def test_wrappers_call_only_declared_functions():
    used = set(re.findall(r"ffi::(hp_[a-z0-9_]+)\(", open(LIBRS).read()))
    used |= set(re.findall(r"ffi::(hp_[a-z0-9_]+)[,)]", open(LIBRS).read()))
    assert used and used <= set(rust_decls()), sorted(used - set(rust_decls()))
    for f in ("gemm", "gemm_atb", "mf_gemm", "mf_gemm_atb", "cholesky", "cholesky_blocked", "forward_solve",
              "backward_solve", "solve_cholesky", "matvec", "cond_estimate", "dot", "gemv", "jacobi_eigen",
              "solve_spd", "solve_spd_f128", "mf_dot"):
        assert re.search(r"pub fn " + f + r"\b", open(LIBRS).read()), f
