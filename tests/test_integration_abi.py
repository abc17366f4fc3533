"""The Julia binding written out in INTEGRATION.md against include/xcore_hip.h, on the CPU: every
`ccall(sym(:xc_...), ...)` names an entry point the header declares, with the same number of arguments of the
same kinds (by-value type, or pointer to the same struct / scalar), and every `struct Xc...` mirror lists the C
struct's fields in the same order with the same types."""
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "xcore_hip.h")
DOC = os.path.join(ROOT, "INTEGRATION.md")

JL2C = {"XcMat": "xc_mat", "XcBMat": "xc_bmat", "XcFluid": "xc_fluid", "XcPhase": "xc_phase", "XcBnchm": "xc_bnchm",
        "XcNoise": "xc_noise", "XcGrid": "xc_grid", "XcPar": "xc_par", "XcScales": "xc_scales",
        "XcFmOpts": "xc_fm_opts", "XcSolveStats": "xc_solve_stats", "XcStepOpts": "xc_step_opts",
        "XcStepState": "xc_step_state", "XcStepStats": "xc_step_stats"}
JL_SCALAR = {"Cint": "int", "Cuint": "unsigned", "Cdouble": "double", "Csize_t": "size_t", "Int64": "int64_t",
             "UInt64": "uint64_t", "UInt32": "uint32_t", "Float64": "double", "UInt8": "uint8_t", "Cvoid": "void"}


def _header():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    src = re.sub(r"//[^\n]*", "", src)
    protos = {}
    for m in re.finditer(r"(?:^|\n)\s*(int|void\*|const char\*)\s+(xc_\w+)\s*\(([^)]*)\)\s*;", src):
        protos[m.group(2)] = (m.group(1), [a.strip() for a in m.group(3).split(",") if a.strip()])
    structs = {}
    for m in re.finditer(r"typedef\s+struct\s*\{(.*?)\}\s*(\w+)\s*;", src, flags=re.S):
        fields = []
        for decl in m.group(1).split(";"):
            decl = decl.strip()
            if not decl:
                continue
            toks = decl.split(None, 1)
            ctype, names = toks[0], toks[1]
            if ctype == "const":
                ctype, names = names.split(None, 1)
            for n in names.split(","):
                n = n.strip()
                star = n.startswith("*")
                n = n.lstrip("*")
                arr = re.match(r"(\w+)\[(\d+)\]", n)
                if arr:
                    fields.append((arr.group(1), ctype, int(arr.group(2))))
                else:
                    fields.append((n, ctype + ("*" if star else ""), 1))
        structs[m.group(2)] = fields
    return protos, structs


def _c_kind(param):
    p = re.sub(r"\bconst\b", "", param).strip()
    arr = re.match(r"(.*?)\s*(\w+)\s*\[[^\]]*\]$", p)
    if arr:
        return "ptr:" + arr.group(1).strip()
    base = re.match(r"([\w ]+?)\s*(\*+)\s*\w*$", p)
    if base:
        t = base.group(1).strip()
        return "ptr:ptr" if len(base.group(2)) > 1 else "ptr:" + ("void" if t == "xc_ctx" else t)
    t = p.rsplit(None, 1)[0] if len(p.split()) > 1 else p
    return "val:" + t.strip()


def _jl_kind(t):
    t = t.strip()
    m = re.match(r"(?:Ptr|Ref)\{(.*)\}$", t)
    if m:
        inner = m.group(1).strip()
        if re.match(r"Ptr\{", inner):
            return "ptr:ptr"
        return "ptr:" + JL2C.get(inner, JL_SCALAR.get(inner, inner))
    return "val:" + JL2C.get(t, JL_SCALAR.get(t, t))


def _balanced(s, i):
    """s[i] == '(' -> index just past the matching ')'."""
    depth = 0
    for j in range(i, len(s)):
        if s[j] in "({":
            depth += 1
        elif s[j] in ")}":
            depth -= 1
            if depth == 0:
                return j + 1
    raise ValueError("unbalanced")


def _split_top(s):
    out, depth, cur = [], 0, ""
    for ch in s:
        if ch in "({":
            depth += 1
        elif ch in ")}":
            depth -= 1
        if ch == "," and depth == 0:
            out.append(cur)
            cur = ""
        else:
            cur += ch
    if cur.strip():
        out.append(cur)
    return [x.strip() for x in out]


def _julia():
    doc = open(DOC).read()
    return "\n".join(re.findall(r"```julia\n(.*?)```", doc, flags=re.S))


def _ccalls(code):
    calls = []
    for m in re.finditer(r"ccall\(sym\(:(xc_\w+)\),\s*(\w+),\s*", code):
        i = m.end()
        assert code[i] == "(", m.group(1)
        j = _balanced(code, i)
        calls.append((m.group(1), m.group(2), _split_top(code[i + 1:j - 1])))
    return calls


def _jl_structs(code):
    out = {}
    for m in re.finditer(r"\bstruct\s+(Xc\w+)\b(.*?)\bend\b", code, flags=re.S):
        fields = []
        for f in re.finditer(r"(\w+)::(NTuple\{\d+,\s*\w+\}|Ptr\{\w+\}|\w+)", m.group(2)):
            name, t = f.group(1), f.group(2)
            nt = re.match(r"NTuple\{(\d+),\s*(\w+)\}", t)
            if nt:
                ct = JL2C.get(nt.group(2), JL_SCALAR.get(nt.group(2), nt.group(2)))
                fields.append((name, ct, int(nt.group(1))))
            else:
                pm = re.match(r"Ptr\{(\w+)\}", t)
                ct = (JL_SCALAR[pm.group(1)] + "*") if pm else JL2C.get(t, JL_SCALAR.get(t, t))
                fields.append((name, ct, 1))
        out[m.group(1)] = fields
    return out


def _flat(fields):
    return [t for _, t, n in fields for _ in range(n)]


def test_every_ccall_matches_a_prototype():
    protos, _ = _header()
    calls = _ccalls(_julia())
    assert len(calls) >= 25
    names = set()
    for name, ret, args in calls:
        assert name in protos, f"{name} is not declared in include/xcore_hip.h"
        cret, cparams = protos[name]
        assert {"Cint": "int", "Cstring": "const char*"}[ret] == cret, (name, ret, cret)
        assert len(args) == len(cparams), (name, len(args), len(cparams))
        for k, (a, p) in enumerate(zip(args, cparams)):
            assert _jl_kind(a) == _c_kind(p), (name, k, a, p, _jl_kind(a), _c_kind(p))
        names.add(name)
    # the functions the shim binds on the reference's hot path
    for n in ("xc_advect", "xc_advect_with_flux", "xc_advect_centered", "xc_diffus", "xc_diffus_with_flux",
              "xc_diffus_centered", "xc_fill_ghosts", "xc_embed_interior", "xc_phsevo", "xc_update",
              "xc_update_dt", "xc_store_previous", "xc_update_phase_velocities", "xc_noise_step", "xc_store_noise",
              "xc_fluidmech", "xc_run_step", "xc_history_stats", "xc_initialize", "xc_malloc", "xc_free",
              "xc_create", "xc_last_error"):
        assert n in names, n


def test_struct_mirrors_match_the_header():
    _, structs = _header()
    jl = _jl_structs(_julia())
    for jname, cname in JL2C.items():
        assert jname in jl, f"no Julia mirror of {cname}"
        assert cname in structs, cname
        cf, jf = structs[cname], jl[jname]
        assert _flat(jf) == _flat(cf), (jname, _flat(jf), _flat(cf))
        if not any(n > 1 and t in ("xc_mat", "xc_bmat") for _, t, n in jf):
            assert [n for n, _, _ in jf] == [n for n, _, _ in cf], jname   # named fields: same names in order


def test_parser_catches_a_mismatch():
    # the checker itself: a swapped argument and a missing field are reported
    protos, structs = _header()
    bad = _ccalls("ccall(sym(:xc_advect), Cint, (Ptr{Cvoid}, Cint, XcMat, XcMat, XcMat, Cdouble, XcMat), x)")
    name, _, args = bad[0]
    assert [_jl_kind(a) for a in args] != [_c_kind(p) for p in protos[name][1]]
    jl = _jl_structs("struct XcGrid; Nz::Int64; h::Cdouble; D::Cdouble; L::Cdouble; end")
    assert _flat(jl["XcGrid"]) != _flat(structs["xc_grid"])
