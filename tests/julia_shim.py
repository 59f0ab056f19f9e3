"""Read julia/VeronaHIP.jl as text (Julia is not installed here).

Two things are taken from the shim so that the tests check the file a
maintainer would load, not a Python transcription of it:

* every ``ccall((:vrn_..., lib), Ret, (ArgTypes...), ...)``: name, return
  type and argument types, compared with include/verona_hip.h;
* the body of ``while t < Tmax`` in ``HARM_HLL``: the per-kernel call
  sequence of one time step (``for d in 1:3`` loops expanded), which the GPU
  test executes through the same C entry points.
"""
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SHIM = os.path.join(ROOT, "julia", "VeronaHIP.jl")
HDR = os.path.join(ROOT, "include", "verona_hip.h")


def _text():
    with open(SHIM) as f:
        return f.read()


def _split_top(s):
    """Split a comma list at depth 0 of (), {} and []."""
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
    """Index just past the parenthesis group opening at s[i] == '('."""
    depth = 0
    for j in range(i, len(s)):
        if s[j] == "(":
            depth += 1
        elif s[j] == ")":
            depth -= 1
            if depth == 0:
                return j + 1
    raise ValueError("unbalanced parentheses in the shim")


def ccalls():
    """[(symbol, return type, [argument types])] of every ccall in the shim."""
    s = _text()
    out = []
    for m in re.finditer(r"ccall\(", s):
        end = _balanced(s, m.end() - 1)
        parts = _split_top(s[m.end():end - 1])
        sym = re.match(r"\(:(\w+),\s*lib\)", parts[0])
        assert sym, parts[0]
        ret = parts[1]
        args = parts[2].strip()
        assert args.startswith("(") and args.endswith(")"), args
        argl = _split_top(args[1:-1])
        out.append((sym.group(1), ret, [a for a in argl if a]))
    return out


def header_prototypes():
    """{name: (return C type, [argument C types])} from include/verona_hip.h."""
    txt = open(HDR).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    protos = {}
    for m in re.finditer(r"^\s*((?:const\s+)?[a-z_0-9]+\s*\*?)\s*(vrn_[a-z_0-9]+)\s*\(([^)]*)\)\s*;", txt, flags=re.M):
        ret, name, args = m.group(1), m.group(2), m.group(3)
        argl = [] if args.strip() == "void" else [" ".join(a.split()) for a in args.split(",")]
        protos[name] = (" ".join(ret.split()), argl)
    return protos


def step_params_fields():
    """[(C type, field name)] of vrn_step_params in the header, arrays as 'int[6]'."""
    txt = re.sub(r"/\*.*?\*/", "", open(HDR).read(), flags=re.S)
    body = re.search(r"typedef struct \{(.*?)\} vrn_step_params;", txt, flags=re.S).group(1)
    fields = []
    for decl in body.split(";"):
        decl = " ".join(decl.split())
        if not decl:
            continue
        ctype, names = decl.split(" ", 1)
        for nm in names.split(","):
            nm = nm.strip()
            arr = re.match(r"(\w+)\[(\d+)\]", nm)
            fields.append((f"{ctype}[{arr.group(2)}]", arr.group(1)) if arr else (ctype, nm))
    return fields


def julia_struct_fields(name="StepParams"):
    """[(Julia type, field name)] of `struct <name>` in the shim."""
    body = re.search(rf"struct {name}\n(.*?)\nend", _text(), flags=re.S).group(1)
    return [(t.strip(), n.strip()) for n, t in (ln.strip().split("::") for ln in body.splitlines() if ln.strip())]


def step_sequence():
    """The calls of one iteration of HARM_HLL's `while t < Tmax` loop, in order:
    a list of (function name, [argument expressions]), `for d in 1:3` expanded
    with d substituted; the optional host callback line is left out."""
    s = _text()
    fn = s[s.index("function HARM_HLL("):]
    fn = fn[:fn.index("\nend\n")]
    body = fn[fn.index("while t < Tmax"):]
    body = body[body.index("\n") + 1:body.index("t += dt")]
    lines = [ln.split("#", 1)[0].strip() for ln in body.splitlines()]
    lines = [ln for ln in lines if ln]
    seq, k = [], 0
    while k < len(lines):
        ln = lines[k]
        loop = re.match(r"for (\w+) in (\d+):(\d+)$", ln)
        if loop:
            var, lo, hi = loop.group(1), int(loop.group(2)), int(loop.group(3))
            inner = []
            k += 1
            while lines[k] != "end":
                inner.append(lines[k])
                k += 1
            for v in range(lo, hi + 1):
                for c in inner:
                    seq.append(_call(re.sub(rf"\b{var}\b", str(v), c)))
        elif ln.startswith("fun_bound === nothing ||"):
            pass
        else:
            seq.append(_call(ln))
        k += 1
    return seq


def _call(ln):
    m = re.match(r"(\w+)\((.*)\)$", ln)
    assert m, f"not a call: {ln}"
    return m.group(1), _split_top(m.group(2))
