"""The Julia seam (itensornetworks.jl_amd/julia/ITensorNetworksMI355X.jl) against
the C ABI (include/itnbp.h), without a Julia runtime: every ccall names a
declared symbol of libitnbp with the declared return type, the declared number
of arguments, C-compatible argument types, and passes that many arguments;
the mirrored option/result structs match the header's field order and types;
and the required cache interface of the reference
(src/caches/abstractbeliefpropagationcache.jl:44-69) is implemented."""
import os
import re

from helpers import HEADER, ROOT

SHIM = os.path.join(ROOT, "itensornetworks.jl_amd", "julia", "ITensorNetworksMI355X.jl")

# C parameter type -> Julia ccall types that pass it correctly
COMPAT = {
    "int": {"Cint"},
    "int64_t": {"Int64"},
    "double": {"Cdouble", "Float64"},
    "const char*": {"Cstring"},
    "itnbp_ctx*": {"Ptr{Cvoid}"},
    "itnbp_net*": {"Ptr{Cvoid}"},
    "itnbp_lnet*": {"Ptr{Cvoid}"},
    "itnbp_ctx**": {"Ref{Ptr{Cvoid}}", "Ptr{Ptr{Cvoid}}"},
    "itnbp_net**": {"Ref{Ptr{Cvoid}}", "Ptr{Ptr{Cvoid}}"},
    "itnbp_lnet**": {"Ref{Ptr{Cvoid}}", "Ptr{Ptr{Cvoid}}"},
    "const int*": {"Ptr{Cint}", "Ref{Cint}"},
    "int*": {"Ptr{Cint}", "Ref{Cint}"},
    "const double*": {"Ptr{Float64}", "Ref{Float64}", "Ptr{Cdouble}", "Ref{Cdouble}"},
    "double*": {"Ptr{Float64}", "Ref{Float64}", "Ptr{Cdouble}", "Ref{Cdouble}"},
    "const void*": {"Ptr{Cvoid}", "Ptr{Float64}"},
    "void*": {"Ptr{Cvoid}", "Ptr{Float64}"},
    "const itnbp_update_opts*": {"Ref{UpdateOpts}", "Ptr{UpdateOpts}"},
    "itnbp_update_result*": {"Ref{UpdateResult}", "Ptr{UpdateResult}"},
}
RET = {"int": "Cint", "int64_t": "Int64", "const char*": "Cstring"}
FIELD = {"int": "Cint", "double": "Cdouble", "const int*": "Ptr{Cint}"}


def _norm_ctype(p):
    """C type of a declaration "type name" (the name is the last identifier)."""
    p = re.sub(r"/\*.*?\*/", "", p).strip()
    m = re.match(r"^(.*?)\b(\w+)\s*$", p)
    t = m.group(1).strip()
    t = re.sub(r"\s*\*", "*", t)
    return re.sub(r"\s+", " ", t)


def header_prototypes():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    protos = {}
    for m in re.finditer(r"^\s*(int|int64_t|const char\*)\s+(itnbp_\w+)\s*\(([^)]*)\);", txt, re.M):
        params = [_norm_ctype(p) for p in m.group(3).split(",") if p.strip() and p.strip() != "void"]
        protos[m.group(2)] = (m.group(1), params)
    return protos


def header_struct(name):
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    m = re.search(r"typedef struct \{([^{}]*)\}\s*" + name + r";", txt)
    fields = []
    for line in m.group(1).split(";"):
        line = line.strip()
        if line:
            ctype = _norm_ctype(line)
            fields.append((line.split()[-1].lstrip("*"), ctype))
    return fields


def _split_top(s):
    out, depth, cur = [], 0, ""
    for ch in s:
        if ch in "([{":
            depth += 1
        elif ch in ")]}":
            depth -= 1
        if ch == "," and depth == 0:
            out.append(cur.strip())
            cur = ""
        else:
            cur += ch
    if cur.strip():
        out.append(cur.strip())
    return out


def shim_ccalls():
    txt = open(SHIM).read()
    txt = "\n".join(line.split("#")[0] for line in txt.splitlines())
    calls = []
    for m in re.finditer(r"ccall\(", txt):
        i, depth = m.end(), 1
        while depth:
            depth += {"(": 1, ")": -1}.get(txt[i], 0)
            i += 1
        parts = _split_top(txt[m.end():i - 1])
        target, ret, argtypes, args = parts[0], parts[1], parts[2], parts[3:]
        sym = re.match(r"\(:(\w+),\s*(\w+)\)", target)
        types = _split_top(argtypes.strip()[1:-1])
        calls.append((sym.group(1), sym.group(2), ret, types, args))
    return calls


def test_every_ccall_matches_the_header():
    protos = header_prototypes()
    calls = shim_ccalls()
    assert len(calls) >= 20
    for name, lib, ret, types, args in calls:
        assert lib == "libitnbp", name
        assert name in protos, f"{name} is not declared in include/itnbp.h"
        cret, params = protos[name]
        assert RET[cret] == ret, (name, ret, cret)
        assert len(types) == len(params), (name, types, params)
        for jt, ct in zip(types, params):
            assert jt in COMPAT[ct], f"{name}: Julia {jt} for C {ct}"
        assert len(args) == len(types), (name, args)


def test_mirrored_structs_match_the_header():
    txt = open(SHIM).read()
    for name in ("UpdateOpts", "UpdateResult"):
        cname = {"UpdateOpts": "itnbp_update_opts", "UpdateResult": "itnbp_update_result"}[name]
        m = re.search(r"struct " + name + r"\n(.*?)\nend", txt, re.S)
        jfields = [tuple(f.strip().split("::")) for f in m.group(1).strip().splitlines()]
        cfields = header_struct(cname)
        assert [f for f, _ in jfields] == [f for f, _ in cfields], name
        for (jf, jt), (cf, ct) in zip(jfields, cfields):
            assert FIELD[ct] == jt, (name, jf, jt, ct)


def test_precision_codes_match_the_header():
    txt = open(SHIM).read()
    hdr = open(HEADER).read()
    codes = dict(re.findall(r"ITNBP_PREC_(\w+)\s*=\s*(\d+)", hdr))
    jl = dict(re.findall(r'"(\w+)" => Cint\((\d+)\)', txt))
    assert jl == {k.lower(): v for k, v in codes.items()}


def test_required_cache_interface_is_implemented():
    """abstractbeliefpropagationcache.jl:44-69 declares these not_implemented()."""
    txt = open(SHIM).read()
    for fn in ("ITensorNetworks.partitioned_tensornetwork", "ITensorNetworks.messages", "Base.copy",
               "ITensorNetworks.default_update_alg", "ITensorNetworks.default_message_update_alg",
               "ITensorNetworks.default_bp_maxiter", "ITensorNetworks.default_bp_edge_sequence",
               "ITensorNetworks.environment", "ITensorNetworks.partitions", "PartitionedGraphs.quotientedges",
               "PartitionedGraphs.partitioned_vertices"):
        assert re.search(r"(^|\n)(function )?" + re.escape(fn) + r"\((::)?\w*::?MI355XCache", txt) or \
            re.search(re.escape(fn) + r"\(bpc::MI355XCache", txt), fn
    assert txt.count("function ITensorNetworks.region_scalar(bpc::MI355XCache") == 2
    for fn in ("ITensorNetworks.update(alg::Algorithm\"bp\", bpc::MI355XCache)",
               "ITensorNetworks.set_default_kwargs(alg::Algorithm\"bp\", bpc::MI355XCache)",
               "ITensorNetworks.message(bpc::MI355XCache", "ITensorNetworks.set_message!(bpc::MI355XCache",
               "ITensorNetworks.set_messages!(bpc::MI355XCache", "ITensorNetworks.logscalar(bpc::MI355XCache)",
               "ITensorNetworks.rescale(bpc::MI355XCache", "ITensors.apply(os::Vector{ITensor}",
               "ChainRulesCore.rrule(::typeof(bp_energy)"):
        assert fn in txt, fn
