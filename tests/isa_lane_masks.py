"""Static check of gfx950 assembly built from this package's HIP sources.

A lane mask written by a VALU compare (v_cmp_*_e64 s[a:b], ...) holds 0 for
every lane that was inactive when the compare ran.  When such a mask is
computed inside a loop whose lanes leave at different iterations, and read
after that loop (by a uniform branch "s_and_b64 vcc, exec, s[a:b]" or a
v_cndmask select), lanes that left the loop early see 0 whatever the
condition.  The compiler produced exactly this for a branch on a kernel
argument that three sequential loops tested: it computed the mask in the
first loop and reused it in the other two, so waves whose active lanes had
all left the first loop early took the wrong branch.

find_hazards(asm) returns (kernel, register, def_loop, use_loop, line) for
each such read; it works on the assembly printed by
`hipcc --cuda-device-only -S`, whose block comments give the loop nesting.
"""
import re

_LABEL = re.compile(r"^(\.LBB\d+_\d+|; %bb\.\d+):")
_FUNC = re.compile(r"^(_Z\S+):")
_VCMP_DEF = re.compile(r"^\s*v_cmp\w*_e64 (s\[\d+:\d+\]),")
_USES = (re.compile(r"^\s*s_(?:and|andn2)_b64 vcc, exec, (s\[\d+:\d+\])"),
         re.compile(r"^\s*v_cndmask_b32_e64 [^;]*, (s\[\d+:\d+\])\s*(?:;|$)"))
_SDST = re.compile(r"^\s*([sv]_\w+) (s\[\d+:\d+\])\s*,")
_SDST2 = re.compile(r"^\s*v_\w+(?:_co|div_scale)\w* v[\[\d:\]]+, (s\[\d+:\d+\]),")


def _blocks(lines):
    """Per line: the innermost loop header (e.g. 'BB4_10') or None; and the
    parent of every loop header."""
    parent = {}
    loop_of_line = []
    cur = None
    i = 0
    n = len(lines)
    while i < n:
        line = lines[i]
        m = _LABEL.match(line)
        if m:
            notes = [line]
            j = i + 1
            while j < n and lines[j].lstrip().startswith(";") and not _LABEL.match(lines[j]):
                notes.append(lines[j])
                j += 1
            text = " ".join(notes)
            name = m.group(1)[2:] if m.group(1).startswith(".L") else None
            if "Loop Header" in text and name:
                cur = name
                depth = int(re.search(r"Loop Header: Depth=(\d+)", text).group(1))
                for p, d in re.findall(r"Parent Loop (BB\d+_\d+) Depth=(\d+)", text):
                    if int(d) == depth - 1:
                        parent[name] = p
            else:
                h = re.search(r"in Loop: Header=(BB\d+_\d+)", text)
                cur = h.group(1) if h else None
        loop_of_line.append(cur)
        i += 1
    return loop_of_line, parent


def _contains(parent, outer, inner):
    while inner is not None:
        if inner == outer:
            return True
        inner = parent.get(inner)
    return False


def find_hazards(asm):
    found = []
    funcs = []
    name, body = None, []
    for line in asm.split("\n"):
        m = _FUNC.match(line)
        if m:
            name, body = m.group(1), []
            continue
        if name is None:
            continue
        body.append(line)
        if "s_endpgm" in line:
            funcs.append((name, body))
            name = None
    for fname, body in funcs:
        loops, parent = _blocks(body)
        defs = {}
        for k, line in enumerate(body):
            code = line.split(";")[0]
            if not code.strip():
                continue
            for use in _USES:
                u = use.match(code)
                if u and u.group(1) in defs:
                    dloop = defs[u.group(1)]
                    if dloop is not None and not _contains(parent, dloop, loops[k]):
                        found.append((fname, u.group(1), dloop, loops[k], k))
            d = _VCMP_DEF.match(code)
            if d:
                defs[d.group(1)] = loops[k]
                continue
            for pat in (_SDST, _SDST2):
                w = pat.match(code)
                if w:
                    reg = w.group(2) if pat is _SDST else w.group(1)
                    if not code.lstrip().startswith(("s_and_b64 vcc", "s_andn2_b64 vcc")):
                        defs.pop(reg, None)
    return found
