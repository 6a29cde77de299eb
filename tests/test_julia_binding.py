"""The Julia binding (octopus.jl_amd/julia/OctopusHIP) checked without Julia,
which this image does not have: every ccall names an entry point that
include/octohip.h declares, with the same number of arguments and
compatible argument and return types; the Julia mirror structs have the
header's fields in the same order with the same types; the package
declares its extensions and they exist (ChainRulesCore rrules of
build_edges_diff and apply_zsort, AMDGPU arrays and streams)."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
JL = os.path.join(ROOT, "octopus.jl_amd", "julia", "OctopusHIP")
HDR = os.path.join(ROOT, "include", "octohip.h")


def _jl_sources():
    out = {}
    for d, _, files in os.walk(JL):
        for f in files:
            if f.endswith(".jl"):
                out[os.path.relpath(os.path.join(d, f), JL)] = open(os.path.join(d, f)).read()
    return out


def _c_type(t):
    """(base, pointer depth) of a C parameter or return type."""
    t = re.sub(r"\bconst\b", "", t).strip()
    depth = t.count("*")
    base = t.replace("*", "").strip()
    return base, depth


def header_prototypes():
    src = re.sub(r"/\*.*?\*/", "", open(HDR).read(), flags=re.S)
    protos = {}
    for ret, name, params in re.findall(r"(int|const char\s*\*)\s+(oct_\w+)\s*\(([^)]*)\)\s*;", src, re.S):
        args = []
        for p in params.split(","):
            p = " ".join(p.split())
            if not p or p == "void":
                continue
            m = re.match(r"(.*?)(\w+)$", p)
            args.append(_c_type(m.group(1)))
        protos[name] = (_c_type(ret), args)
    return protos


def header_structs():
    src = re.sub(r"/\*.*?\*/", "", open(HDR).read(), flags=re.S)
    out = {}
    for name, body in re.findall(r"typedef struct (oct_\w+) \{(.*?)\}", src, re.S):
        fields = []
        for line in body.split(";"):
            line = " ".join(line.split())
            if not line:
                continue
            m = re.match(r"(.*?)(\w+)(\[(\d+)\])?$", line)
            t, fname, _, n = m.groups()
            fields.append((fname, _c_type(t), int(n) if n else None))
        out[name] = fields
    return out


_JL_SCALAR = {"Cint": ("int", 0), "Int64": ("int64_t", 0), "Cfloat": ("float", 0), "Float32": ("float", 0),
              "Int32": ("int32_t", 0), "UInt32": ("uint32_t", 0), "UInt64": ("uint64_t", 0),
              "Cdouble": ("double", 0), "Float64": ("double", 0), "Cstring": ("char", 1)}
_JL_STRUCT = {"OctEdges": "oct_edges", "OctTree": "oct_tree", "OctView": "oct_view",
              "OctNeighbors": "oct_neighbors", "OctBuildInfo": "oct_build_info"}


def _jl_type(t):
    t = t.strip()
    if t in _JL_SCALAR:
        return _JL_SCALAR[t]
    m = re.match(r"(?:Ptr|Ref)\{(.*)\}$", t)
    if m:
        inner = m.group(1)
        if inner == "Cvoid":
            return ("void", 1)
        if inner in _JL_STRUCT:
            return (_JL_STRUCT[inner], 1)
        b, d = _jl_type(inner)
        return (b, d + 1)
    raise AssertionError(f"unmapped Julia type {t}")


def _compatible(jl, c):
    if jl == c:
        return True
    # Ptr{Cvoid}: the opaque context or an untyped buffer
    if jl == ("void", 1) and c[1] == 1 and c[0] in ("void", "oct_ctx"):
        return True
    if jl == ("void", 2) and c == ("oct_ctx", 2):
        return True
    return False


def _split_top(s):
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


def julia_ccalls():
    calls = []
    for f, src in _jl_sources().items():
        for m in re.finditer(r"ccall\(\(:(oct_\w+),\s*libocto\),", src):
            i = m.end()
            # the return type, then the parenthesized argument type tuple
            depth, j = 0, i
            while True:
                ch = src[j]
                if ch in "({[":
                    if ch == "(" and depth == 0:
                        break
                    depth += 1
                elif ch in ")}]":
                    depth -= 1
                j += 1
            ret = src[i:j].strip().rstrip(",").strip()
            depth, k = 0, j
            while True:
                if src[k] == "(":
                    depth += 1
                elif src[k] == ")":
                    depth -= 1
                    if depth == 0:
                        break
                k += 1
            args = _split_top(src[j + 1:k])
            calls.append((f, m.group(1), ret, args))
    return calls


def test_every_ccall_matches_the_header():
    protos = header_prototypes()
    calls = julia_ccalls()
    assert len(calls) >= 20
    for f, name, ret, args in calls:
        assert name in protos, f"{f}: {name} is not declared in include/octohip.h"
        cret, cargs = protos[name]
        assert _compatible(_jl_type(ret), cret), f"{f}: {name} returns {ret}, header {cret}"
        assert len(args) == len(cargs), f"{f}: {name} takes {len(cargs)} arguments, ccall passes {len(args)}"
        for k, (a, c) in enumerate(zip(args, cargs)):
            assert _compatible(_jl_type(a), c), f"{f}: {name} argument {k}: Julia {a}, header {c}"


def test_binding_covers_the_drop_in_entry_points():
    names = {c[1] for c in julia_ccalls()}
    for n in ("oct_create_typed", "oct_destroy", "oct_set_stream", "oct_synchronize", "oct_set_search_radius64",
              "oct_set_refit_mode", "oct_add_point_set", "oct_update_point_set", "oct_resize_point_set",
              "oct_set_active_search", "oct_run", "oct_get_tree", "oct_build_edges", "oct_device_view",
              "oct_build_edges_backward", "oct_build_edges_backward64", "oct_materialize_neighbors", "oct_get_neighbors", "oct_apply_zsort",
              "oct_apply_zsort_adjoint", "oct_trim", "oct_last_error"):
        assert n in names, n


@pytest.mark.parametrize("jl,c", sorted(_JL_STRUCT.items()))
def test_mirror_structs_match_the_header(jl, c):
    src = _jl_sources()[os.path.join("src", "OctopusHIP.jl")]
    m = re.search(rf"struct {jl}\n(.*?)\nend", src, re.S)
    if m is None:
        pytest.skip(f"{jl} is not mirrored in Julia")
    jf = []
    for line in m.group(1).splitlines():
        line = line.split("#")[0].strip()
        if line:
            n, t = line.split("::")
            jf.append((n.strip(), t.strip()))
    cf = header_structs()[c]
    assert [n for n, _ in jf] == [n for n, _, _ in cf]
    for (n, t), (_, ct, arr) in zip(jf, cf):
        tm = re.match(r"NTuple\{(\d+),\s*(\w+)\}", t)
        if tm:
            assert arr == int(tm.group(1)) and _compatible(_jl_type(tm.group(2)), ct), n
        else:
            assert arr is None and _compatible(_jl_type(t), ct), (n, t, ct)


def test_package_declares_its_extensions():
    proj = open(os.path.join(JL, "Project.toml")).read()
    for ext in ("OctopusHIPAMDGPUExt", "OctopusHIPChainRulesCoreExt"):
        assert re.search(rf"^{ext} = ", proj, re.M), ext
        assert os.path.exists(os.path.join(JL, "ext", ext + ".jl")), ext
    crc = _jl_sources()[os.path.join("ext", "OctopusHIPChainRulesCoreExt.jl")]
    assert "ChainRulesCore.rrule(::typeof(OctopusHIP.build_edges_diff)" in crc
    assert "build_edges_diff_pullback!(" in crc
    assert "ChainRulesCore.rrule(::typeof(apply_zsort)" in crc
    amd = _jl_sources()[os.path.join("ext", "OctopusHIPAMDGPUExt.jl")]
    assert "OctopusHIP._current_stream()" in amd and "OctopusHIP._device_wrap(" in amd


def test_the_checker_catches_mismatches():
    """A wrong scalar, a wrong pointee or a wrong depth is reported."""
    assert not _compatible(_jl_type("Cfloat"), ("int", 0))
    assert not _compatible(_jl_type("Ptr{Int32}"), ("float", 1))
    assert not _compatible(_jl_type("Ptr{Cvoid}"), ("oct_ctx", 2))
    assert not _compatible(_jl_type("Int64"), ("int", 0))
    assert _compatible(_jl_type("Ref{OctEdges}"), ("oct_edges", 1))
