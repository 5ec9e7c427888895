"""The Julia module julia/HybridSmoothersMI355X.jl against the C ABI (no Julia needed).

Every `ccall((:hs_*, libhsmooth), RetT, (ArgT...), ...)` in the module is
parsed and compared with the prototype of the same function in
include/hsmooth.h: return type, argument count and each argument's C type
(pointee type and pointer depth; handles are opaque pointers).  The Julia
structs passed by reference (PCGStats, BatchConfig, BatchResult) are compared
field by field with hs_pcg_stats, hs_batch_config and hs_batch_result.  The
module binds the methods the reference's API needs (the builder functor and its
Symmetric method, ldiv! for host and device vectors, `\\`, the field
properties, the Chebyshev smoothers).
"""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "hsmooth.h")
JULIA = os.path.join(ROOT, "julia", "HybridSmoothersMI355X.jl")

# Julia ccall type -> (C base type, pointer depth); handles (hs_*_t) are pointers
_JL = {
    "Cint": ("int", 0), "Int64": ("int64_t", 0), "UInt64": ("uint64_t", 0),
    "Cdouble": ("double", 0), "Float64": ("double", 0), "Cstring": ("char", 1),
    "Ptr{Cvoid}": ("void", 1), "Ptr{Int64}": ("int64_t", 1), "Ref{Int64}": ("int64_t", 1),
    "Ptr{Float64}": ("double", 1), "Ptr{Cdouble}": ("double", 1), "Ref{Float64}": ("double", 1),
    "Ref{Ptr{Cvoid}}": ("void", 2), "Ref{PCGStats}": ("hs_pcg_stats", 1),
    "Ref{BatchConfig}": ("hs_batch_config", 1), "Ptr{BatchResult}": ("hs_batch_result", 1),
}
_STRUCTS = {"PCGStats": "hs_pcg_stats", "BatchConfig": "hs_batch_config",
            "BatchResult": "hs_batch_result"}
_JL_FIELD = {"Int64": "int64_t", "Cint": "int", "Cdouble": "double", "Float64": "double"}


def _strip_c_comments(txt):
    txt = re.sub(r"/\*.*?\*/", " ", txt, flags=re.S)
    return re.sub(r"//[^\n]*", " ", txt)


def _c_type(decl, handles):
    """'const int64_t *ptr' -> ('int64_t', 1); a handle type counts one pointer level."""
    decl = decl.replace("const", " ").replace("*", " * ")
    toks = decl.split()
    depth = toks.count("*")
    words = [t for t in toks if t != "*"]
    if len(words) > 1:  # drop the parameter name
        words = words[:-1]
    base = " ".join(words)
    if base in handles:
        return "void", depth + 1
    return base, depth


def _header():
    txt = _strip_c_comments(open(HEADER).read())
    handles = set(re.findall(r"typedef\s+struct\s+\w+\s*\*\s*(hs_\w+_t)\s*;", txt))
    protos = {}
    for ret, name, args in re.findall(r"([\w\s\*]+?)\b(hs_\w+)\s*\(([^)]*)\)\s*;", txt):
        ret = ret.strip().split("\n")[-1]
        arglist = [] if args.strip() in ("", "void") else [a.strip() for a in args.split(",")]
        protos[name] = (_c_type(ret + " r", handles), [_c_type(a, handles) for a in arglist])
    structs = {}
    for body, name in re.findall(r"typedef\s+struct\s*\{(.*?)\}\s*(hs_\w+)\s*;", txt, flags=re.S):
        fields = []
        for line in body.split(";"):
            line = line.strip()
            if not line:
                continue
            m = re.match(r"([\w]+)\s+(.*)", line)
            typ, names = m.group(1), m.group(2)
            fields += [(n.strip(), typ) for n in names.split(",")]
        structs[name] = fields
    return protos, structs


def _split_top(s):
    """Split at commas outside braces/parentheses."""
    out, depth, cur = [], 0, ""
    for ch in s:
        if ch in "({[":
            depth += 1
        elif ch in ")}]":
            depth -= 1
        if ch == "," and depth == 0:
            out.append(cur.strip())
            cur = ""
        else:
            cur += ch
    if cur.strip():
        out.append(cur.strip())
    return out


def _balanced(s, i):
    """s[i] == '('; returns the index after its matching ')'."""
    depth = 0
    for j in range(i, len(s)):
        if s[j] == "(":
            depth += 1
        elif s[j] == ")":
            depth -= 1
            if depth == 0:
                return j + 1
    raise ValueError("unbalanced")


def _julia_ccalls():
    src = open(JULIA).read()
    calls = []
    for m in re.finditer(r"ccall\(\(:(hs_\w+),\s*libhsmooth\),", src):
        i = m.end()
        # return type: up to the next top-level comma
        depth, j = 0, i
        while True:
            ch = src[j]
            if ch in "({[":
                depth += 1
            elif ch in ")}]":
                depth -= 1
            elif ch == "," and depth == 0:
                break
            j += 1
        ret = src[i:j].strip()
        k = src.index("(", j)
        e = _balanced(src, k)
        args = _split_top(src[k + 1:e - 1])
        calls.append((m.group(1), ret, args))
    return calls


def _julia_structs():
    src = open(JULIA).read()
    out = {}
    for name, body in re.findall(r"\nstruct (\w+)\n(.*?)\nend", src, flags=re.S):
        fields = []
        for line in body.split("\n"):
            line = line.split("#")[0].strip()
            m = re.match(r"(\w+)::(\w+)$", line)
            if m:
                fields.append((m.group(1), m.group(2)))
        out[name] = fields
    return out


def _problems(name, ret, args, protos):
    if name not in protos:
        return [f"{name} is not declared in hsmooth.h"]
    cret, cargs = protos[name]
    out = []
    if _JL.get(ret) != cret:
        out.append((name, "return", ret, cret))
    if len(args) != len(cargs):
        return out + [(name, "argument count", len(args), len(cargs))]
    for k, (ja, ca) in enumerate(zip(args, cargs)):
        if _JL.get(ja) != ca:
            out.append((name, "argument", k, ja, ca))
    return out


def test_every_ccall_matches_the_header():
    protos, _ = _header()
    calls = _julia_ccalls()
    assert len(calls) >= 35
    bad = [p for c in calls for p in _problems(*c, protos)]
    assert not bad, bad


def test_checker_catches_mistakes():
    """A dropped argument, a wrong pointee type and a wrong return type are reported."""
    protos, _ = _header()
    good = ("hs_l1gs_apply", "Cint",
            ["Ptr{Cvoid}", "Ptr{Float64}", "Ptr{Float64}", "Cint", "Ptr{Cvoid}", "Cint"])
    assert not _problems(*good, protos)
    assert _problems(good[0], good[1], good[2][:-1], protos)
    assert _problems(good[0], good[1], ["Ptr{Cvoid}", "Ptr{Int64}"] + good[2][2:], protos)
    assert _problems(good[0], "Int64", good[2], protos)
    assert _problems("hs_no_such_call", "Cint", [], protos)


def test_binds_the_reference_api():
    protos, _ = _header()
    bound = {name for name, _, _ in _julia_ccalls()}
    needed = {
        "hs_init", "hs_last_error", "hs_last_error_aux",
        "hs_l1gs_build", "hs_l1gs_apply", "hs_l1gs_info", "hs_l1gs_get", "hs_l1gs_destroy",
        "hs_l1gs_save", "hs_l1gs_load", "hs_l1gs_smooth", "hs_l1gs_smooth_noalloc",
        "hs_csr_create", "hs_csr_spmv", "hs_csr_info", "hs_csr_get", "hs_csr_destroy",
        "hs_cheb_build", "hs_cheb_smooth", "hs_cheb_apply", "hs_cheb_info", "hs_cheb_get",
        "hs_cheb_matrix", "hs_cheb_destroy", "hs_pcg_solve_true",
        "hs_batch_create", "hs_batch_submit", "hs_batch_step", "hs_batch_collect",
        "hs_batch_info", "hs_batch_destroy",
    }
    assert needed <= set(protos)
    assert needed <= bound, sorted(needed - bound)
    src = open(JULIA).read()
    # the reference's method signatures (l1_gauss_seidel.jl:287-313, 367-407;
    # chebyshev.jl:141, 158, 258, 278) for the ROCBackend
    for pat in [
        r"function \(builder::L1GSPrecBuilder\{<:ROCBackend, <:GPUConfig\}\)\(\s*A::Union\{SparseMatrixCSC",
        r"function \(builder::L1GSPrecBuilder\{<:ROCBackend, <:GPUConfig\}\)\(\s*A::Symmetric",
        r"isSymA::Bool = false,\s*η = 1\.5,\s*sweep::HS\.AbstractSweep = SymmetricSweep\(\),",
        r"function LinearSolve\.ldiv!\(y::Vector\{Float64\}, P::MI355XL1GSPreconditioner",
        r"function LinearSolve\.ldiv!\(y::ROCVector\{Float64\}, P::MI355XL1GSPreconditioner",
        r"function \(\\\)\(P::MI355XL1GSPreconditioner, x::AbstractVector\)",
        r"HS\._build_smoother\(config::ChebyshevFourth\{<:ROCBackend\}",
        r"HS\._build_smoother\(config::ChebyshevFirst\{<:ROCBackend\}",
        r"function HS\.smooth!\(x::Vector\{Float64\}, s::MI355XChebyshevSmoother",
        r"function HS\.smooth!\(x::ROCVector\{Float64\}, s::MI355XChebyshevSmoother",
        r"throw\(SingularException\(",
    ]:
        assert re.search(pat, src), pat
    # the fields the reference's tests read (test_preconditioners.jl:36-39, 121-126;
    # test_chebyshev.jl:38-47, 104-112)
    for prop in [":sweep", ":op", ":lop", ":uop", ":D_DL1", ":SLbuffer", ":SUbuffer",
                 ":l1inv", ":ρ", ":a", ":r", ":z", ":tmp", ":A"]:
        assert re.search(r"=== " + re.escape(prop) + r"\b", src), prop
    assert "partitioning::BlockPartitioning{Int64, ROCBackend}" in src


@pytest.mark.parametrize("jl,c", sorted(_STRUCTS.items()))
def test_struct_layouts_match(jl, c):
    _, cstructs = _header()
    jstructs = _julia_structs()
    assert jl in jstructs and c in cstructs
    jf, cf = jstructs[jl], cstructs[c]
    assert [n for n, _ in jf] == [n for n, _ in cf], (jl, jf, cf)
    assert [_JL_FIELD[t] for _, t in jf] == [t for _, t in cf], (jl, jf, cf)


def test_parser_sees_known_prototypes():
    """The header parser itself: a few prototypes spelled out by hand."""
    protos, structs = _header()
    assert protos["hs_l1gs_apply"] == (("int", 0), [("void", 1), ("double", 1), ("double", 1),
                                                    ("int", 0), ("void", 1), ("int", 0)])
    assert protos["hs_last_error"] == (("char", 1), [])
    assert protos["hs_cheb_matrix"] == (("void", 1), [("void", 1)])
    assert protos["hs_l1gs_build"][1][-1] == ("void", 2)
    assert structs["hs_batch_result"][4:6] == [("rnorm0", "double"), ("rnorm", "double")]
