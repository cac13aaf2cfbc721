"""A small Quil front end for the pure-state path.

The reference parses Quil with cl-quil (not vendored; call sites
src/classical-memory-mixin.lisp:39,144).  This module reads the subset that
the pure-state path executes: DECLARE, gate applications with parameter
expressions and the DAGGER / CONTROLLED / FORKED modifiers, DEFGATE (matrix
and permutation forms), MEASURE, RESET, HALT, NOP, PRAGMA (ignored), LABEL and
the jumps JUMP / JUMP-WHEN / JUMP-UNLESS (targets resolved to instruction
indices when the program is read, as label patching does), and the classical
instructions MOVE, EXCHANGE, CONVERT, NEG, NOT, ADD, SUB, MUL, DIV, AND, IOR,
XOR, EQ, GT, GE, LT, LE (src/transition-classical-instructions.lisp).
Parameter expressions may name classical memory (``theta[0]``) and are
evaluated when the instruction executes, as ``force-parameter`` does
(src/transition.lisp:142-153).
"""
from __future__ import annotations

import math
import re
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional, Tuple

import numpy as np

from . import gates as G


class QuilParseError(ValueError):
    pass


# ---------------------------------------------------------------------------
# Expressions


_TOKEN = re.compile(r"\s*(?:(\d+\.\d*(?:[eE][-+]?\d+)?|\.\d+(?:[eE][-+]?\d+)?|\d+(?:[eE][-+]?\d+)?)"
                    r"|(%?[A-Za-z_][A-Za-z0-9_\-]*)|(\*\*|[-+*/^()\[\],]))")

_FUNCS = {
    "sin": np.sin, "cos": np.cos, "sqrt": np.sqrt, "exp": np.exp,
    "cis": lambda x: np.cos(x) + 1j * np.sin(x),
}


def _tokenize(s: str) -> List[str]:
    out, pos = [], 0
    s = s.strip()
    while pos < len(s):
        m = _TOKEN.match(s, pos)
        if not m or m.end() == pos:
            raise QuilParseError(f"bad expression near {s[pos:]!r}")
        out.append(m.group(1) or m.group(2) or m.group(3))
        pos = m.end()
    return out


Expr = Callable[[dict, dict], complex]


class _ExprParser:
    def __init__(self, toks: List[str]):
        self.t = toks
        self.i = 0

    def peek(self):
        return self.t[self.i] if self.i < len(self.t) else None

    def take(self, want=None):
        tok = self.peek()
        if tok is None or (want is not None and tok != want):
            raise QuilParseError(f"expected {want!r}, got {tok!r}")
        self.i += 1
        return tok

    def parse(self) -> Expr:
        e = self.sum()
        if self.peek() is not None:
            raise QuilParseError(f"trailing tokens {self.t[self.i:]}")
        return e

    def sum(self) -> Expr:
        e = self.prod()
        while self.peek() in ("+", "-"):
            op = self.take()
            r = self.prod()
            e = (lambda a, b: lambda env, mem: a(env, mem) + b(env, mem))(e, r) if op == "+" else \
                (lambda a, b: lambda env, mem: a(env, mem) - b(env, mem))(e, r)
        return e

    def prod(self) -> Expr:
        e = self.unary()
        while self.peek() in ("*", "/"):
            op = self.take()
            r = self.unary()
            e = (lambda a, b: lambda env, mem: a(env, mem) * b(env, mem))(e, r) if op == "*" else \
                (lambda a, b: lambda env, mem: a(env, mem) / b(env, mem))(e, r)
        return e

    def unary(self) -> Expr:
        if self.peek() == "-":
            self.take()
            e = self.unary()
            return lambda env, mem: -e(env, mem)
        if self.peek() == "+":
            self.take()
            return self.unary()
        return self.power()

    def power(self) -> Expr:
        base = self.atom()
        if self.peek() in ("^", "**"):
            self.take()
            ex = self.unary()
            return lambda env, mem: base(env, mem) ** ex(env, mem)
        return base

    def atom(self) -> Expr:
        tok = self.take()
        if tok == "(":
            e = self.sum()
            self.take(")")
            return e
        if tok[0].isdigit() or tok[0] == ".":
            v = float(tok) if any(c in tok for c in ".eE") else int(tok)
            return lambda env, mem, v=v: v
        if tok == "pi":
            return lambda env, mem: math.pi
        if tok == "i":
            return lambda env, mem: 1j
        if tok.startswith("%"):
            name = tok[1:]
            return lambda env, mem: env[name]
        if tok in _FUNCS and self.peek() == "(":
            self.take("(")
            arg = self.sum()
            self.take(")")
            fn = _FUNCS[tok]
            return lambda env, mem: fn(arg(env, mem))
        # memory reference name[idx] or name (= name[0])
        idx = 0
        if self.peek() == "[":
            self.take("[")
            idx = int(self.take())
            self.take("]")
        name = tok
        return lambda env, mem: mem[name][idx]


def parse_expression(s: str) -> Expr:
    return _ExprParser(_tokenize(s)).parse()


def _split_top(s: str, sep: str = ",") -> List[str]:
    parts, depth, cur = [], 0, []
    for ch in s:
        if ch in "([":
            depth += 1
        elif ch in ")]":
            depth -= 1
        if ch == sep and depth == 0:
            parts.append("".join(cur))
            cur = []
        else:
            cur.append(ch)
    if cur or parts:
        parts.append("".join(cur))
    return [p.strip() for p in parts]


# ---------------------------------------------------------------------------
# Instructions


@dataclass
class GateApplication:
    name: str
    params: List[Expr]
    qubits: List[int]
    modifiers: List[str] = field(default_factory=list)
    text: str = ""


@dataclass
class Measure:
    qubit: int
    target: Optional[Tuple[str, int]] = None


@dataclass
class Reset:
    qubit: Optional[int] = None


@dataclass
class Halt:
    pass


@dataclass
class Nop:
    pass


@dataclass
class Label:
    name: str


@dataclass
class Jump:
    """JUMP / JUMP-WHEN / JUMP-UNLESS: target is the label's instruction index
    after parsing; cond is (name, index) of the bit for the conditional forms."""
    label: str
    when: Optional[bool] = None
    cond: Optional[Tuple[str, int]] = None
    target: int = -1


@dataclass
class Classical:
    """A classical memory instruction: op, destination (name, index), and
    operands that are (name, index) memory references or numbers."""
    op: str
    dest: Tuple[str, int]
    args: List = field(default_factory=list)


@dataclass
class Declare:
    name: str
    type: str
    length: int


@dataclass
class GateDefinition:
    name: str
    params: List[str]
    kind: str  # "matrix" | "permutation"
    rows: List[List[Expr]] = field(default_factory=list)
    permutation: List[int] = field(default_factory=list)

    def matrix(self, values) -> np.ndarray:
        if self.kind == "permutation":
            d = len(self.permutation)
            U = np.zeros((d, d), dtype=np.complex128)
            for j, p in enumerate(self.permutation):
                U[p, j] = 1.0
            return U
        env = dict(zip(self.params, values))
        return np.array([[complex(e(env, {})) for e in row] for row in self.rows], dtype=np.complex128)


@dataclass
class ParsedProgram:
    declarations: Dict[str, Declare]
    gate_definitions: Dict[str, GateDefinition]
    executable_code: list

    def qubits_needed(self) -> int:
        m = -1
        for ins in self.executable_code:
            if isinstance(ins, GateApplication):
                m = max([m] + list(ins.qubits))
            elif isinstance(ins, Measure):
                m = max(m, ins.qubit)
            elif isinstance(ins, Reset) and ins.qubit is not None:
                m = max(m, ins.qubit)
        return m + 1


_MODIFIERS = ("DAGGER", "CONTROLLED", "FORKED")


def _parse_gate_line(line: str) -> GateApplication:
    mods = []
    rest = line
    while True:
        head = rest.split(None, 1)
        if head and head[0] in _MODIFIERS:
            mods.append(head[0])
            rest = head[1] if len(head) > 1 else ""
        else:
            break
    m = re.match(r"([A-Za-z_][A-Za-z0-9_\-]*)\s*(\((.*)\))?\s*(.*)$", rest)
    if not m:
        raise QuilParseError(f"cannot parse gate application {line!r}")
    name, params, qs = m.group(1), m.group(3), m.group(4)
    pexprs = [parse_expression(p) for p in _split_top(params)] if params is not None and params.strip() else []
    try:
        qubits = [int(q) for q in qs.split()]
    except ValueError:
        raise QuilParseError(f"bad qubit list in {line!r}")
    if not qubits:
        raise QuilParseError(f"gate {name} applied to no qubits")
    return GateApplication(name, pexprs, qubits, mods, line)


def parse_quil(text: str) -> ParsedProgram:
    """Parse a Quil program (the subset described in the module docstring)."""
    decls: Dict[str, Declare] = {}
    defs: Dict[str, GateDefinition] = {}
    code: list = []
    raw = []
    for ln in text.replace("\r", "").split("\n"):
        ln = ln.split("#", 1)[0].rstrip()
        if not ln.strip():
            continue
        indented = ln[:1] in (" ", "\t")
        for part in (ln.split(";") if not indented else [ln]):
            if part.strip():
                raw.append((indented, part))
    i = 0
    while i < len(raw):
        indented, line = raw[i]
        s = line.strip()
        i += 1
        word = s.split(None, 1)[0]
        if word == "DECLARE":
            m = re.match(r"DECLARE\s+(\w+)\s+(BIT|REAL|OCTET|INTEGER)(\[(\d+)\])?", s)
            if not m:
                raise QuilParseError(f"bad DECLARE: {s!r}")
            decls[m.group(1)] = Declare(m.group(1), m.group(2), int(m.group(4) or 1))
        elif word == "DEFGATE":
            m = re.match(r"DEFGATE\s+([A-Za-z_][\w\-]*)\s*(\(([^)]*)\))?\s*(AS\s+(\w+))?\s*:\s*$", s)
            if not m:
                raise QuilParseError(f"bad DEFGATE: {s!r}")
            name = m.group(1)
            params = [p.strip().lstrip("%") for p in m.group(3).split(",")] if m.group(3) else []
            kind = (m.group(5) or "MATRIX").lower()
            body = []
            while i < len(raw) and raw[i][0]:
                body.append(raw[i][1].strip())
                i += 1
            if kind == "permutation":
                perm = [int(x) for x in _split_top(" ".join(body))]
                defs[name] = GateDefinition(name, params, "permutation", permutation=perm)
            elif kind == "matrix":
                rows = [[parse_expression(e) for e in _split_top(r)] for r in body]
                if not rows or any(len(r) != len(rows) for r in rows):
                    raise QuilParseError(f"DEFGATE {name}: matrix must be square")
                defs[name] = GateDefinition(name, params, "matrix", rows=rows)
            else:
                raise QuilParseError(f"DEFGATE {name}: unsupported form AS {kind.upper()}")
        elif word == "MEASURE":
            parts = s.split()
            q = int(parts[1])
            tgt = None
            if len(parts) > 2:
                m = re.match(r"(\w+)(\[(\d+)\])?$", parts[2])
                if not m:
                    raise QuilParseError(f"bad MEASURE target {parts[2]!r}")
                tgt = (m.group(1), int(m.group(3) or 0))
            code.append(Measure(q, tgt))
        elif word == "RESET":
            parts = s.split()
            code.append(Reset(int(parts[1]) if len(parts) > 1 else None))
        elif word == "HALT":
            code.append(Halt())
        elif word in ("NOP", "PRAGMA", "WAIT"):
            if word == "NOP":
                code.append(Nop())
        elif word == "LABEL":
            code.append(Label(_label_name(s.split(None, 1)[1] if len(s.split()) > 1 else "")))
        elif word in ("JUMP", "JUMP-WHEN", "JUMP-UNLESS"):
            parts = s.split()
            if word == "JUMP":
                if len(parts) != 2:
                    raise QuilParseError(f"bad JUMP: {s!r}")
                code.append(Jump(_label_name(parts[1])))
            else:
                if len(parts) != 3:
                    raise QuilParseError(f"bad {word}: {s!r}")
                code.append(Jump(_label_name(parts[1]), word == "JUMP-WHEN", _mref(parts[2])))
        elif word in _CLASSICAL:
            nargs = _CLASSICAL[word]
            # a two-operand instruction's source may be a constant expression
            # with spaces in it, e.g. MOVE alpha[0] (pi / 2.0)
            parts = s.split(None, nargs)
            if len(parts) != nargs + 1 or (nargs == 3 and len(parts[3].split()) != 1):
                raise QuilParseError(f"{word} takes {nargs} operand(s): {s!r}")
            code.append(Classical(word, _mref(parts[1]), [_operand(x) for x in parts[2:]]))
        else:
            code.append(_parse_gate_line(s))
    # label patching: jumps name instruction indices
    where = {ins.name: k for k, ins in enumerate(code) if isinstance(ins, Label)}
    for ins in code:
        if isinstance(ins, Jump):
            if ins.label not in where:
                raise QuilParseError(f"jump to undefined label @{ins.label}")
            ins.target = where[ins.label]
    return ParsedProgram(decls, defs, code)


# operand counts of the classical instructions (destination included)
_CLASSICAL = {"MOVE": 2, "EXCHANGE": 2, "CONVERT": 2, "NEG": 1, "NOT": 1, "ADD": 2, "SUB": 2, "MUL": 2, "DIV": 2, "AND": 2, "IOR": 2,
              "XOR": 2, "EQ": 3, "GT": 3, "GE": 3, "LT": 3, "LE": 3}


def _label_name(tok: str) -> str:
    tok = tok.strip()
    if not tok.startswith("@") or len(tok) < 2:
        raise QuilParseError(f"bad label {tok!r}")
    return tok[1:]


def _mref(tok: str) -> Tuple[str, int]:
    m = re.match(r"([A-Za-z_]\w*)(\[(\d+)\])?$", tok)
    if not m:
        raise QuilParseError(f"bad memory reference {tok!r}")
    return (m.group(1), int(m.group(3) or 0))


def _operand(tok: str):
    """A classical operand: an integer or real literal, a memory reference,
    or a constant expression (pi / 2.0), evaluated now."""
    tok = tok.strip()
    try:
        return int(tok)
    except ValueError:
        pass
    try:
        return float(tok)
    except ValueError:
        pass
    if re.match(r"[A-Za-z_]\w*(\[\d+\])?$", tok) and tok.lower() != "pi":
        return _mref(tok)
    try:
        v = complex(parse_expression(tok)({}, {}))
    except (QuilParseError, KeyError, TypeError, ValueError, ZeroDivisionError):
        raise QuilParseError(f"bad classical operand {tok!r}")
    if v.imag != 0:
        raise QuilParseError(f"classical operand {tok!r} is not real")
    return v.real


def execute_classical(ins: Classical, memory: dict, types: Optional[Dict[str, str]] = None) -> None:
    """One classical instruction on the classical memory
    (src/transition-classical-instructions.lisp).  ``types`` maps a memory
    name to its declared type (BIT, OCTET, INTEGER, REAL); NOT and CONVERT
    depend on it (NOT of a BIT is 1 - b, of an OCTET the low 8 bits of the
    complement, of an INTEGER the 64-bit complement, lines 32-51; CONVERT to
    INTEGER rounds half to even, to BIT tests for non-zero, lines 149-191).
    INTEGER arithmetic wraps to signed 64 bits like ensure-s64; INTEGER DIV
    floors."""
    def get(x):
        if isinstance(x, tuple):
            if x[0] not in memory:
                raise QuilParseError(f"undeclared memory {x[0]}")
            return memory[x[0]][x[1]]
        return x
    name, idx = ins.dest
    if name not in memory:
        raise QuilParseError(f"undeclared memory {name}")
    cell = memory[name]
    if not (0 <= idx < cell.size):
        raise QuilParseError(f"memory reference {name}[{idx}] out of range")
    ty = (types or {}).get(name, "REAL" if cell.dtype.kind == "f" else "INTEGER")
    a = [get(x) for x in ins.args]
    op = ins.op

    def wrap(v):
        if ty == "REAL":
            return float(v)
        v = int(v) & 0xFFFFFFFFFFFFFFFF
        return v - (1 << 64) if v >> 63 else v

    if op == "MOVE":
        cell[idx] = wrap(a[0])
    elif op == "EXCHANGE":
        other = a[0]
        oname, oidx = ins.args[0]
        memory[oname][oidx] = cell[idx]
        cell[idx] = other
    elif op == "CONVERT":
        v = a[0]
        if ty == "REAL":
            cell[idx] = float(v)
        elif ty == "BIT":
            cell[idx] = int(v != 0)
        else:
            cell[idx] = wrap(round(v) if isinstance(v, (float, np.floating)) else int(v))
    elif op == "NEG":
        cell[idx] = wrap(-cell[idx])
    elif op == "NOT":
        if ty == "BIT":
            cell[idx] = 1 - int(cell[idx])
        elif ty == "OCTET":
            cell[idx] = ~int(cell[idx]) & 0xFF
        else:
            cell[idx] = wrap(~int(cell[idx]))
    elif op in ("ADD", "SUB", "MUL", "DIV"):
        x, y = (float(cell[idx]), float(a[0])) if ty == "REAL" else (int(cell[idx]), int(a[0]))
        if op == "ADD":
            r = x + y
        elif op == "SUB":
            r = x - y
        elif op == "MUL":
            r = x * y
        elif ty == "REAL":
            r = x / y
        else:
            if y == 0:
                raise QuilParseError("integer division by zero")
            r = x // y  # floor division (line 245)
        cell[idx] = wrap(r)
    elif op in ("AND", "IOR", "XOR"):
        x, y = int(cell[idx]), int(a[0])
        cell[idx] = wrap(x & y if op == "AND" else (x | y if op == "IOR" else x ^ y))
    else:  # comparisons: dest <- left OP right
        left, right = a
        cell[idx] = int({"EQ": left == right, "GT": left > right, "GE": left >= right,
                         "LT": left < right, "LE": left <= right}[op])


def resolve_gate_matrix(ins: GateApplication, defs: Dict[str, GateDefinition], memory: dict) -> np.ndarray:
    """Gate + modifiers -> matrix (Quil convention), parameters forced
    against classical memory."""
    values = [complex(p({}, memory)) for p in ins.params]
    values = [v.real if v.imag == 0 else v for v in values]
    if ins.name in defs:
        U = defs[ins.name].matrix(values)
    else:
        try:
            U = G.gate_matrix(ins.name, values)
        except KeyError:
            raise QuilParseError(f"Unknown gate {ins.name}")
    for mod in reversed(ins.modifiers):
        if mod == "DAGGER":
            U = G.dagger(U)
        elif mod == "CONTROLLED":
            U = G.controlled(U)
        elif mod == "FORKED":
            raise QuilParseError("FORKED needs per-branch parameters; not supported in this front end")
    return U
