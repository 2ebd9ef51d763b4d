"""The Julia binding (krakenn.jl_amd/julia/KrakenMI355X.jl) against the C header.

Julia is not installed here, so the module is checked as text: every `ccall` names a symbol
the header declares, with the header's argument count, matching argument and return types;
the Julia mirrors of kmi_face / kmi_desc / kmi_batch_desc / kmi_nn_desc have the header's
fields in order; the module adds Kraken methods only with a backend-typed first positional
argument (Julia does not dispatch on keywords, so a method with Kraken's own positional
signature would replace the reference's); and every `KrakenMI355X.<name>` the integration
notes call exists in the module.
"""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
JL = os.path.join(ROOT, "krakenn.jl_amd", "julia", "KrakenMI355X.jl")
HDR = os.path.join(ROOT, "include", "kraken_mi355x.h")


def _strip_c_comments(s):
    return re.sub(r"/\*.*?\*/", "", s, flags=re.S)


def _strip_jl_comments(s):
    s = re.sub(r'"""(.*?)"""', lambda m: '""', s, flags=re.S)  # docstrings
    return "\n".join(re.sub(r"#.*$", "", ln) if '"' not in ln else ln for ln in s.splitlines())


def c_kind(t):
    """Canonical kind of a C type: ptr, i32, i64, u64, f64, f32, size, void."""
    t = " ".join(t.replace("const ", "").split())
    if "*" in t:
        return "ptr"
    return {"int": "i32", "int32_t": "i32", "int64_t": "i64", "uint64_t": "u64", "double": "f64",
            "float": "f32", "size_t": "size", "void": "void"}[t]


def jl_kind(t):
    t = t.strip()
    if t.startswith(("Ptr{", "Ref{")) or t == "Cstring":
        return "ptr"
    return {"Int32": "i32", "Cint": "i32", "Int64": "i64", "UInt64": "u64", "Float64": "f64",
            "Cdouble": "f64", "Float32": "f32", "Cfloat": "f32", "Csize_t": "size", "Cvoid": "void"}[t]


def header_functions():
    txt = _strip_c_comments(open(HDR).read())
    out = {}
    for m in re.finditer(r"([A-Za-z_][\w\s\*]*?)\b(kmi_\w+)\s*\(([^)]*)\)\s*;", txt):
        ret, name, params = m.group(1), m.group(2), m.group(3).strip()
        ret = ret.split("}")[-1].strip()
        args = [] if params in ("", "void") else [p.strip() for p in params.split(",")]
        kinds = [c_kind(re.sub(r"\b\w+$", "", a).strip() if not a.endswith("*") else a) for a in args]
        out[name] = (c_kind(ret), kinds)
    return out


def header_structs():
    txt = _strip_c_comments(open(HDR).read())
    out = {}
    for m in re.finditer(r"typedef struct\s*\{(.*?)\}\s*(\w+)\s*;", txt, flags=re.S):
        fields = []
        for decl in m.group(1).split(";"):
            decl = " ".join(decl.split())
            if not decl:
                continue
            typ, names = decl.split(" ", 1)
            for nm in names.split(","):
                nm = nm.strip()
                arr = re.match(r"(\w+)\[(\d+)\]", nm)
                fields.append((arr.group(1), typ, int(arr.group(2))) if arr else (nm, typ, 1))
        out[m.group(2)] = fields
    return out


def _split_top(s):
    """Split on top-level commas (outside (), {}, [])."""
    parts, depth, cur = [], 0, ""
    for ch in s:
        if ch in "({[":
            depth += 1
        elif ch in ")}]":
            depth -= 1
        if ch == "," and depth == 0:
            parts.append(cur)
            cur = ""
        else:
            cur += ch
    if cur.strip():
        parts.append(cur)
    return [p.strip() for p in parts]


def _balanced(s, i):
    """Text inside the parentheses opening at s[i]."""
    depth = 0
    for j in range(i, len(s)):
        depth += {"(": 1, ")": -1}.get(s[j], 0)
        if depth == 0:
            return s[i + 1:j]
    raise ValueError("unbalanced")


def julia_ccalls():
    src = _strip_jl_comments(open(JL).read())
    calls = []
    for m in re.finditer(r"\bccall\(", src):
        body = _balanced(src, m.end() - 1)
        parts = _split_top(body)
        sym = re.match(r"\(:(\w+)\s*,\s*libkmi\)", parts[0])
        assert sym, parts[0]
        types = _split_top(_balanced(parts[2], 0)) if parts[2].startswith("(") else []
        calls.append((sym.group(1), parts[1], types, parts[3:]))
    return calls


def test_every_ccall_matches_the_header():
    hdr = header_functions()
    calls = julia_ccalls()
    assert len(calls) >= 40
    for name, ret, types, args in calls:
        assert name in hdr, f"{name} is not declared in the header"
        cret, cargs = hdr[name]
        assert jl_kind(ret) == cret, (name, ret, cret)
        assert len(types) == len(cargs), (name, types, cargs)
        assert len(args) == len(types), (name, "argument values vs types", args, types)
        for jt, ck in zip(types, cargs):
            assert jl_kind(jt) == ck, (name, jt, ck)


def test_binding_covers_the_lattice_batch_and_network_entry_points():
    used = {c[0] for c in julia_ccalls()}
    for need in ("kmi_ctx_create", "kmi_ctx_destroy", "kmi_lattice_create", "kmi_lattice_destroy", "kmi_step",
                 "kmi_step_residual", "kmi_upload", "kmi_download", "kmi_moments", "kmi_reduce",
                 "kmi_init_equilibrium", "kmi_init_temperature", "kmi_set_mask", "kmi_set_face",
                 "kmi_set_face_profile", "kmi_batch_create", "kmi_batch_destroy", "kmi_batch_run",
                 "kmi_nn_create", "kmi_nn_destroy", "kmi_nn_forward", "kmi_nn_train_epochs"):
        assert need in used, need


JL_OF_C = {"int32_t": "Int32", "int64_t": "Int64", "double": "Float64", "uint64_t": "UInt64",
           "kmi_face": "KmiFace"}


@pytest.mark.parametrize("cname,jname", [("kmi_face", "KmiFace"), ("kmi_desc", "KmiDesc"),
                                         ("kmi_batch_desc", "KmiBatchDesc"), ("kmi_nn_desc", "KmiNNDesc")])
def test_struct_layouts_match_the_header(cname, jname):
    src = _strip_jl_comments(open(JL).read())
    m = re.search(r"\bstruct\s+%s\b(.*?)\bend\b" % jname, src, flags=re.S)
    assert m, jname
    jfields = [tuple(x.strip() for x in ln.split("::")) for ln in m.group(1).splitlines() if "::" in ln]
    cfields = header_structs()[cname]
    assert [f[0] for f in jfields] == [f[0] for f in cfields], (jname, jfields, cfields)
    for (jn, jt), (cn, ct, n) in zip(jfields, cfields):
        want = JL_OF_C[ct] if n == 1 else "NTuple{%d,%s}" % (n, JL_OF_C[ct])
        assert jt == want, (jname, jn, jt, want)


def test_no_kraken_method_is_overwritten():
    src = _strip_jl_comments(open(JL).read())
    # Kraken names may only be extended qualified, with MI355XBackend first
    assert not re.search(r"^\s*import\s+Kraken\s*:", src, flags=re.M), "import Kraken: ... allows overwriting"
    ext = re.findall(r"^function\s+Kraken\.(\w+)\(([^;)]*)", src, flags=re.M)
    assert {n for n, _ in ext} >= {"run_cavity_2d", "run_cavity_3d", "run_rayleigh_benard_2d",
                                    "run_natural_convection_3d", "run_simulation"}
    for name, params in ext:
        first = params.split(",")[0].strip()
        assert first.endswith("::MI355XBackend"), (name, first)
    # and no unqualified definitions of Kraken's driver names
    for name in ("run_cavity_2d", "run_cavity_3d", "run_simulation", "run_rayleigh_benard_2d",
                 "run_natural_convection_3d"):
        assert not re.search(r"^function\s+%s\(" % name, src, flags=re.M), name


def test_integration_notes_call_only_defined_names():
    src = _strip_jl_comments(open(JL).read())
    defined = set(re.findall(r"^(?:function\s+|mutable struct\s+|struct\s+)(?:Kraken\.)?(\w+)", src, flags=re.M))
    defined |= set(re.findall(r"^(\w+)\(.*\)\s*=", src, flags=re.M))
    notes = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    called = set(re.findall(r"KrakenMI355X\.(\w+)", notes)) - {"jl"}  # not the file name
    assert called, "the integration notes name no module function"
    missing = sorted(called - defined)
    assert not missing, missing
