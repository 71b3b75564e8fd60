"""FORBILD ``.pha`` phantom files -> arrays of FORBILD_Shape structs.

Same conversion as ``lib/process_pha_file.py:89-388`` (``convert_ph_file_to_shapes``
and ``convert_ph_str_to_shapes``); this module is its own implementation:

1. A C-preprocessor subset: ``/* */`` and ``//`` comments, backslash line
   continuations, ``#include "file"`` (relative to the including file),
   object-like and function-like ``#define``, ``#undef``, ``#ifdef`` /
   ``#ifndef`` / ``#else`` / ``#endif``, with rescanning macro expansion.
2. The format's normalisation: ``achse`` -> ``axis``, ``dichte`` -> ``rho``;
   ``cos``/``sin``/``tan`` of an argument written with digits and ``+-*/.``
   only take degrees (any other argument, radians); ``<number>mm`` drops its
   unit; ``Detector{...}`` blocks are removed.
3. Every ``{ ... }`` block is one shape: ``[Type: ...]`` names the type and
   holds the geometry; ``key=value`` pairs anywhere in the block set scalar
   fields (``formel``/``formula`` are ignored); ``name(x,y,z)`` sets the vectors
   ``axis``, ``a_x``, ``a_y``, ``a_z``; a whitespace-free token ``lhs<rhs`` or
   ``lhs>rhs`` inside the brackets is a clip plane (``x``, ``y``, ``z`` or
   ``r(a,b,c)`` (normalised) against a value; ``<`` is mode 1, ``>`` mode 2;
   all ``<`` planes first; at most three).  An inequality written with
   spaces around its operator is not a clip plane, as in the reference
   (``process_pha_file.py:256-258`` reads clip planes as whitespace-free
   tokens).
4. ``Ellipsoid_free`` and ``Ellipt_Cyl`` with one of their three axes left out
   get it as the cross product of the other two, in right-handed order.

Values are evaluated by a whitelisting arithmetic evaluator (numbers,
``+ - * / **``, parentheses, ``sqrt exp log abs sin cos tan asin acos atan``,
``pi``); nothing in a file is executed.
"""

from __future__ import annotations

import ast
import math
import pathlib
import re

import numpy as np

from .. import _abi
from .forbild import TYPE_IDS

# ---------------------------------------------------------------- preprocessor
_IDENT = re.compile(r"[A-Za-z_][A-Za-z_0-9]*")


def _strip_comments(text: str) -> str:
    out, i, n = [], 0, len(text)
    while i < n:
        if text.startswith("/*", i):
            j = text.find("*/", i + 2)
            j = n if j < 0 else j + 2
            out.append(" " + "\n" * text.count("\n", i, j))
            i = j
        elif text.startswith("//", i):
            j = text.find("\n", i)
            i = n if j < 0 else j
        elif text[i] == '"':
            j = text.find('"', i + 1)
            j = n if j < 0 else j + 1
            out.append(text[i:j])
            i = j
        else:
            out.append(text[i])
            i += 1
    return "".join(out)


def _split_args(s: str, start: int) -> tuple[list[str], int]:
    """Arguments of a macro call whose '(' is at s[start]; returns (args, index after ')')."""
    depth, cur, args, i = 0, [], [], start
    while i < len(s):
        c = s[i]
        if c == "(":
            depth += 1
            if depth > 1:
                cur.append(c)
        elif c == ")":
            depth -= 1
            if depth == 0:
                args.append("".join(cur).strip())
                return args, i + 1
            cur.append(c)
        elif c == "," and depth == 1:
            args.append("".join(cur).strip())
            cur = []
        else:
            cur.append(c)
        i += 1
    msg = "unbalanced parentheses in a macro call"
    raise ValueError(msg)


class _Macros:
    def __init__(self) -> None:
        self.obj: dict[str, str] = {}
        self.fn: dict[str, tuple[list[str], str]] = {}

    def define(self, rest: str) -> None:
        m = re.match(r"([A-Za-z_][A-Za-z_0-9]*)(\(([^)]*)\))?\s*(.*)", rest, re.S)
        if not m:
            msg = f"bad #define: {rest!r}"
            raise ValueError(msg)
        name, params, body = m.group(1), m.group(3), m.group(4).strip()
        self.obj.pop(name, None)
        self.fn.pop(name, None)
        if m.group(2) is not None:
            self.fn[name] = ([p.strip() for p in params.split(",") if p.strip()], body)
        else:
            self.obj[name] = body

    def undef(self, name: str) -> None:
        self.obj.pop(name, None)
        self.fn.pop(name, None)

    def defined(self, name: str) -> bool:
        return name in self.obj or name in self.fn

    def expand(self, s: str, hide: frozenset = frozenset()) -> str:
        out, i = [], 0
        while i < len(s):
            m = _IDENT.match(s, i)
            if not m or (i > 0 and (s[i - 1].isalnum() or s[i - 1] in "_.")):
                if s[i] == '"':
                    j = s.find('"', i + 1)
                    j = len(s) if j < 0 else j + 1
                    out.append(s[i:j])
                    i = j
                    continue
                out.append(s[i])
                i += 1
                continue
            name = m.group(0)
            end = m.end()
            if name in hide:
                out.append(name)
                i = end
                continue
            if name in self.obj:
                out.append(self.expand(self.obj[name], hide | {name}))
                i = end
                continue
            if name in self.fn:
                j = end
                while j < len(s) and s[j] in " \t\n":
                    j += 1
                if j < len(s) and s[j] == "(":
                    args, after = _split_args(s, j)
                    params, body = self.fn[name]
                    if params == [] and args == [""]:
                        args = []
                    if len(args) != len(params):
                        msg = f"macro {name} takes {len(params)} arguments, got {len(args)}"
                        raise ValueError(msg)
                    args = [self.expand(a, hide) for a in args]
                    sub = dict(zip(params, args, strict=True))
                    body2 = _IDENT.sub(lambda mm: sub.get(mm.group(0), mm.group(0)), body)
                    out.append(self.expand(body2, hide | {name}))
                    i = after
                    continue
            out.append(name)
            i = end
        return "".join(out)


def _preprocess(text: str, base: pathlib.Path | None, macros: _Macros, depth: int = 0) -> str:
    if depth > 32:
        msg = "#include nesting too deep"
        raise ValueError(msg)
    text = _strip_comments(text.replace("\r\n", "\n")).replace("\\\n", " ")
    out: list[str] = []
    active = [True]
    pending: list[str] = []

    def flush() -> None:
        if pending:
            out.append(macros.expand("\n".join(pending)))
            pending.clear()

    for line in text.split("\n"):
        st = line.strip()
        if st.startswith("#"):
            d = st[1:].strip()
            word = d.split(None, 1)
            key = word[0] if word else ""
            rest = word[1].strip() if len(word) > 1 else ""
            if key in ("ifdef", "ifndef"):
                flush()
                ok = macros.defined(rest.split()[0]) if rest else False
                active.append(active[-1] and (ok if key == "ifdef" else not ok))
            elif key == "else":
                flush()
                parent = active[-2] if len(active) > 1 else True
                active[-1] = parent and not active[-1]
            elif key == "endif":
                flush()
                if len(active) > 1:
                    active.pop()
            elif not active[-1]:
                continue
            elif key == "define":
                flush()
                macros.define(rest)
            elif key == "undef":
                flush()
                macros.undef(rest.split()[0])
            elif key == "include":
                flush()
                m = re.match(r'"([^"]+)"|<([^>]+)>', rest)
                if not m or base is None:
                    msg = f"cannot resolve #include {rest!r}"
                    raise ValueError(msg)
                path = base / (m.group(1) or m.group(2))
                out.append(_preprocess(path.read_text(encoding="ISO-8859-1"), path.parent, macros, depth + 1))
            # other directives (#pragma, #line, ...) carry nothing for a phantom
            continue
        if active[-1]:
            pending.append(line)
    flush()
    return "\n".join(out)


# ------------------------------------------------------------------ evaluator
_FUNCS = {"sqrt": math.sqrt, "exp": math.exp, "log": math.log, "abs": abs, "sin": math.sin, "cos": math.cos,
          "tan": math.tan, "asin": math.asin, "acos": math.acos, "atan": math.atan, "fabs": abs}
_CONSTS = {"pi": math.pi, "PI": math.pi}


def _eval_node(n: ast.AST) -> float:
    if isinstance(n, ast.Expression):
        return _eval_node(n.body)
    if isinstance(n, ast.Constant) and isinstance(n.value, (int, float)):
        return float(n.value)
    if isinstance(n, ast.UnaryOp) and isinstance(n.op, (ast.UAdd, ast.USub)):
        v = _eval_node(n.operand)
        return v if isinstance(n.op, ast.UAdd) else -v
    if isinstance(n, ast.BinOp):
        a, b = _eval_node(n.left), _eval_node(n.right)
        if isinstance(n.op, ast.Add):
            return a + b
        if isinstance(n.op, ast.Sub):
            return a - b
        if isinstance(n.op, ast.Mult):
            return a * b
        if isinstance(n.op, ast.Div):
            return a / b
        if isinstance(n.op, ast.Pow):
            return a**b
    if isinstance(n, ast.Call) and isinstance(n.func, ast.Name) and n.func.id in _FUNCS and not n.keywords:
        return float(_FUNCS[n.func.id](*[_eval_node(a) for a in n.args]))
    if isinstance(n, ast.Name) and n.id in _CONSTS:
        return _CONSTS[n.id]
    msg = f"unsupported expression element {ast.dump(n)}"
    raise ValueError(msg)


def evaluate(expr: str) -> float:
    """Value of an arithmetic expression of the format (trig in radians)."""
    expr = expr.strip()
    try:
        return float(expr)
    except ValueError:
        pass
    try:
        tree = ast.parse(expr, mode="eval")
    except SyntaxError as e:
        msg = f"cannot parse expression {expr!r}"
        raise ValueError(msg) from e
    return _eval_node(tree)


# -------------------------------------------------------------- normalisation
def _normalise(s: str) -> str:
    s = s.replace("\n", " ").replace("\t", " ")
    s = s.replace("achse", "axis").replace("dichte", "rho")
    # degrees for trig of a plain arithmetic argument
    s = re.sub(r"(?<![A-Za-z_])(cos|sin|tan)\(([\-+*/0-9. ]*)\)",
               lambda m: f"{m.group(1)}(3.1415926535898/180*({m.group(2)}))", s)
    s = re.sub(r"([0-9.])mm", r"\1", s)
    s = re.sub(r"Detector\{[^{}]*\}", "", s)
    s = re.sub(r"(axis|a_x|a_y|a_z) \(", r"\1(", s)
    s = re.sub(r"\s*=\s*", "=", s)
    s = re.sub(r"-\s+([0-9.])", r"-\1", s)
    for c in "{}[]":
        s = s.replace(c, f" {c} ")
    return s


def _squeeze_parens(s: str) -> str:
    """Remove the spaces inside every (balanced) parenthesised group."""
    out, depth = [], 0
    for c in s:
        if c == "(":
            depth += 1
        elif c == ")":
            depth = max(0, depth - 1)
        if c == " " and depth > 0:
            continue
        out.append(c)
    return "".join(out)


_VEC_START = re.compile(r"(?<![A-Za-z_0-9])(axis|a_x|a_y|a_z)\(")


def _extract_vectors(body: str) -> tuple[dict[str, str], str]:
    """``name(args)`` vector settings (balanced parentheses) and the text without them."""
    found, out, i = {}, [], 0
    while True:
        m = _VEC_START.search(body, i)
        if not m:
            out.append(body[i:])
            break
        out.append(body[i:m.start()])
        depth, j = 0, m.end() - 1
        while j < len(body):
            if body[j] == "(":
                depth += 1
            elif body[j] == ")":
                depth -= 1
                if depth == 0:
                    break
            j += 1
        if depth != 0:
            msg = f"unbalanced parentheses after {m.group(1)}"
            raise ValueError(msg)
        found[m.group(1)] = body[m.end():j]
        out.append(" ")
        i = j + 1
    return found, "".join(out)


def _vec3(args: str) -> np.ndarray:
    parts, depth, cur = [], 0, []
    for c in args:
        if c == "(":
            depth += 1
        elif c == ")":
            depth -= 1
        if c == "," and depth == 0:
            parts.append("".join(cur))
            cur = []
        else:
            cur.append(c)
    parts.append("".join(cur))
    if len(parts) != 3:
        msg = f"expected three components, got {args!r}"
        raise ValueError(msg)
    return np.array([evaluate(p) for p in parts])


_SCALARS = {"x", "y", "z", "dx", "dy", "dz", "r", "r1", "r2", "l", "rho"}
_CLIP_TOKEN = re.compile(r"[^<>:=]+[<>][^<>:=]+")


def _shape_from_block(block: str) -> np.ndarray:
    block = _squeeze_parens(block)
    m = re.search(r"\[([^\]]*)\]", block)
    if not m:
        msg = f"no [Type: ...] section in {block.strip()[:80]!r}"
        raise ValueError(msg)
    inner = m.group(1)
    rest = block[: m.start()] + " " + block[m.end():]
    name_m = re.match(r"\s*([A-Za-z_0-9]+)\s*:", inner)
    if not name_m or name_m.group(1) not in TYPE_IDS:
        msg = f"unknown shape type in [{inner.strip()[:60]}]"
        raise ValueError(msg)
    tname = name_m.group(1)
    body = inner[name_m.end():]
    s = np.zeros((), dtype=_abi.forbild_shape)
    s["type"] = TYPE_IDS[tname]
    # vectors first (their commas would otherwise split tokens)
    vecs, body_novec = _extract_vectors(body)
    for k, args in vecs.items():
        s[k] = tuple(_vec3(args))
    clips = []
    for tok in body_novec.split() + rest.split():
        if "=" in tok:
            k, v = tok.split("=", 1)
            if k in ("formel", "formula") or not k:
                continue
            if k not in _SCALARS:
                msg = f"unknown field {k!r} in a {tname}"
                raise ValueError(msg)
            s[k] = evaluate(v)
    for tok in body_novec.split():
        # lhs<rhs or lhs>rhs written without spaces; anything else (a spaced
        # inequality included) is not a clip plane
        if not _CLIP_TOKEN.fullmatch(tok):
            continue
        op = "<" if "<" in tok else ">"
        lhs, rhs = tok.split(op, 1)
        if lhs in ("x", "y", "z"):
            ax = np.eye(3)["xyz".index(lhs)]
        elif lhs.startswith("r(") and lhs.endswith(")"):
            ax = _vec3(lhs[2:-1])
            ax = ax / np.linalg.norm(ax)
        else:
            msg = f"cannot read the clip condition {tok!r}"
            raise ValueError(msg)
        clips.append((op, ax, evaluate(rhs)))
    clips = [c for c in clips if c[0] == "<"] + [c for c in clips if c[0] == ">"]
    if len(clips) > 3:
        msg = "more than three clip planes"
        raise ValueError(msg)
    for n, (op, ax, d) in enumerate(clips):
        c = s[f"cas{n + 1}"]
        c["axis"] = tuple(ax)
        c["dotprod"] = d
        c["mode"] = 1 if op == "<" else 2
        s[f"cas{n + 1}"] = c
    # complete a right-handed frame from two given axes
    frames = {"Ellipsoid_free": ("a_x", "a_y", "a_z"), "Ellipt_Cyl": ("a_x", "a_y", "axis")}
    if tname in frames:
        names = frames[tname]
        vecs = [np.array(tuple(s[k].tolist()), dtype=np.float64) for k in names]
        zero = [i for i, v in enumerate(vecs) if not np.any(v)]
        if len(zero) == 1:
            i = zero[0]
            s[names[i]] = tuple(np.cross(vecs[(i + 1) % 3], vecs[(i + 2) % 3]))
    return s


def convert_ph_str_to_shapes(s: str, skip_preprocessing: bool = False, *, base_dir=None) -> np.ndarray:
    """Shapes of a ``.pha`` text; ``base_dir`` resolves its ``#include``s."""
    if not skip_preprocessing:
        s = _preprocess(s, pathlib.Path(base_dir) if base_dir is not None else None, _Macros())
    s = _normalise(s)
    blocks = re.findall(r"\{([^{}]*)\}", s)
    return np.array([_shape_from_block(b) for b in blocks], dtype=_abi.forbild_shape).reshape(-1)


def convert_ph_file_to_shapes(fn, skip_preprocessing: bool = False, append_files=None) -> np.ndarray:
    """Shapes of a ``.pha`` file, followed by those of ``append_files``."""
    paths = [pathlib.Path(fn)] + [pathlib.Path(p) for p in (append_files or [])]
    parts = []
    for p in paths:
        text = p.read_text(encoding="ISO-8859-1")
        parts.append(text if skip_preprocessing else _preprocess(text, p.parent, _Macros()))
    return convert_ph_str_to_shapes("\n".join(parts), skip_preprocessing=True)
