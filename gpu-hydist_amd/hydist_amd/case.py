"""Case parameters in the reference's ``constant.cuh`` format.

The reference compiles its case parameters into the kernels: a case is a copy
of ``constant_<case>.cuh`` renamed to ``constant.cuh`` (readme.md:12-15), whose
``__constant__ T name = expr;`` lines (constant.cuh:31-118, ``PI`` from
constant.cuh:3) the CUDA compiler evaluates.  This module reads the same text
and evaluates every expression with C's typing rules in IEEE double (Python
floats), left to right as written -- so a derived literal such as
``dTchia2dX = 0.12500 / (2 * 1.25)`` (constant_songluy_980x1048.cuh:42) keeps
the value the header spells out rather than one recomputed from dX and dT.

``dT`` is single-sourced from the header: the host time step the reference
takes from ``Coeff.py`` (Engine.py:14, Coeff.py:14) is read from here.
"""
from __future__ import annotations

import ctypes
import re
from typing import Dict, Union

Number = Union[int, float]

# hd_params field order (include/hydist.h); all doubles after segment_limit.
PARAM_FIELDS = (
    "dX", "dY", "dT", "dXbp", "dX2", "dYbp", "dY2", "dTchia2dX", "dTchia2dY",
    "QuyDoiTime", "QuyDoiPi", "HaiChiadT", "kenhhepng", "kenhhepd", "NANGDAY", "H_TINH",
    "Wind", "huonggio", "Zbandau", "heso", "mu_mn", "NDnen", "NDbphai", "NDbtrai",
    "NDbtren", "NDbduoi", "Tod", "Toe", "hstoe", "ghtoe", "Mbochat", "ro", "ros", "dm",
    "d90", "muy", "Dorong", "Sx", "sohn", "soha", "sohn1", "CORIOLIS_FORCE", "g", "Ks",
    "PI",
)
DERIVED_FIELDS = ("Windx", "Windy", "Ufr", "Dxr", "wss")


class HdParams(ctypes.Structure):
    """Mirror of ``hd_params`` (include/hydist.h)."""

    _fields_ = ([("segment_limit", ctypes.c_int32), ("pad0", ctypes.c_int32)]
                + [(n, ctypes.c_double) for n in PARAM_FIELDS + DERIVED_FIELDS])


class HdStepOpts(ctypes.Structure):
    """Mirror of ``hd_step_opts`` (include/hydist.h)."""

    _fields_ = [
        ("kenhhepd", ctypes.c_int32), ("kenhhepng", ctypes.c_int32),
        ("boundary_mode", ctypes.c_int32), ("ketdinh", ctypes.c_int32),
        ("qb_mode", ctypes.c_int32), ("bed_mode", ctypes.c_int32),
        ("sediment_start", ctypes.c_double), ("bed_change_start", ctypes.c_double),
    ]


BND_AUTO, BND_SERIES, BND_ANALYTIC = 0, 1, 2
QB_REFERENCE, QB_FIXED = 0, 1
BED_PERIODIC, BED_EVERY_STEP = 0, 1


class CaseHeaderError(ValueError):
    pass


_TOKEN = re.compile(r"\s*(?:(\d+\.\d*(?:[eE][+-]?\d+)?|\.\d+(?:[eE][+-]?\d+)?|\d+[eE][+-]?\d+)"
                    r"|(\d+)|([A-Za-z_]\w*)|(.))")


def _strip_comments(text: str) -> str:
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    return re.sub(r"//[^\n]*", "", text)


class _Expr:
    """Recursive-descent evaluator of a C arithmetic expression: int and
    double literals, names, unary +/-, * / + -, parentheses.  int op int is
    int (C division truncates toward zero); any double operand promotes."""

    def __init__(self, text: str, env: Dict[str, Number]):
        self.toks = []
        pos = 0
        text = text.strip()
        while pos < len(text):
            m = _TOKEN.match(text, pos)
            if not m or m.end() == pos:
                raise CaseHeaderError(f"cannot tokenise {text!r}")
            pos = m.end()
            flt, integer, name, op = m.groups()
            if flt is not None:
                self.toks.append(("num", float(flt)))
            elif integer is not None:
                self.toks.append(("num", int(integer)))
            elif name is not None:
                if name not in env:
                    raise CaseHeaderError(f"unknown name {name!r} in {text!r}")
                self.toks.append(("num", env[name]))
            elif op is not None and op.strip():
                self.toks.append(("op", op))
        self.i = 0

    def _peek(self):
        return self.toks[self.i] if self.i < len(self.toks) else (None, None)

    def _take(self):
        t = self._peek()
        self.i += 1
        return t

    @staticmethod
    def _bin(a: Number, op: str, b: Number) -> Number:
        if isinstance(a, int) and isinstance(b, int):
            if op == "+":
                return a + b
            if op == "-":
                return a - b
            if op == "*":
                return a * b
            if b == 0:
                raise CaseHeaderError("integer division by zero")
            q = abs(a) // abs(b)
            return q if (a >= 0) == (b >= 0) else -q
        a, b = float(a), float(b)
        if op == "+":
            return a + b
        if op == "-":
            return a - b
        if op == "*":
            return a * b
        return a / b

    def parse(self) -> Number:
        v = self._sum()
        if self.i != len(self.toks):
            raise CaseHeaderError("trailing tokens in expression")
        return v

    def _sum(self) -> Number:
        v = self._prod()
        while self._peek() in (("op", "+"), ("op", "-")):
            op = self._take()[1]
            v = self._bin(v, op, self._prod())
        return v

    def _prod(self) -> Number:
        v = self._unary()
        while self._peek() in (("op", "*"), ("op", "/")):
            op = self._take()[1]
            v = self._bin(v, op, self._unary())
        return v

    def _unary(self) -> Number:
        kind, val = self._peek()
        if kind == "op" and val in "+-":
            self._take()
            x = self._unary()
            return -x if val == "-" else x
        if kind == "op" and val == "(":
            self._take()
            x = self._sum()
            if self._take() != ("op", ")"):
                raise CaseHeaderError("missing )")
            return x
        if kind == "num":
            self._take()
            return val
        raise CaseHeaderError(f"unexpected token {val!r}")


_CONST = re.compile(r"__constant__\s+(int|DOUBLE|double|float)\s+(\w+)\s*=\s*([^;]+);")
_DEFINE = re.compile(r"^[ \t]*#[ \t]*define[ \t]+(\w+)[ \t]+(.+?)[ \t]*$", re.M)


def parse_case_header(text: str) -> Dict[str, Number]:
    """Evaluate the ``__constant__`` definitions of a constant.cuh-format
    header.  Raises CaseHeaderError for a header whose DOUBLE is not double
    (e.g. a single-precision case such as constant_kenhhep.cuh:4)."""
    text = _strip_comments(text)
    env: Dict[str, Number] = {}
    for name, body in _DEFINE.findall(text):
        if name == "DOUBLE":
            if body.strip() != "double":
                raise CaseHeaderError(f"DOUBLE is {body.strip()!r}; only FP64 cases are supported")
            continue
        try:
            env[name] = _Expr(body, env).parse()
        except CaseHeaderError:
            pass  # non-numeric macro (include guards, etc.)
    out: Dict[str, Number] = {}
    for ctype, name, expr in _CONST.findall(text):
        if ctype == "float":
            raise CaseHeaderError(f"{name} is float; only FP64 cases are supported")
        val = _Expr(expr, env).parse()
        if ctype == "int":
            val = int(val)
        elif ctype in ("DOUBLE", "double"):
            val = float(val)
        env[name] = val
        out[name] = val
    if "PI" in env:
        out["PI"] = float(env["PI"])
    return out


def params_from_dict(d: Dict[str, Number]) -> HdParams:
    missing = [k for k in ("segment_limit",) + PARAM_FIELDS if k not in d]
    if missing:
        raise CaseHeaderError(f"case header lacks {missing}")
    p = HdParams()
    p.segment_limit = int(d["segment_limit"])
    for k in PARAM_FIELDS:
        setattr(p, k, float(d[k]))
    return p


def params_to_dict(p: HdParams) -> Dict[str, Number]:
    d: Dict[str, Number] = {"segment_limit": p.segment_limit}
    for k in PARAM_FIELDS + DERIVED_FIELDS:
        d[k] = getattr(p, k)
    return d


def load_case(path: str) -> HdParams:
    with open(path) as f:
        return params_from_dict(parse_case_header(f.read()))


def _fmt(x: float) -> str:
    r = repr(float(x))
    return r


def make_case_header(dX: float, dY: float, dT: float, segment_limit: int = 20,
                     **overrides: Number) -> str:
    """Write a case header in the constant.cuh format for a synthetic case.
    Derived spacing literals are spelled as expressions of dX/dY/dT, in the
    style of the reference's headers (e.g. ``dTchia2dX = 2.0 / (2 * 10.0)``)."""
    base: Dict[str, Number] = dict(
        kenhhepng=0, kenhhepd=0, NANGDAY=0.0, H_TINH=0.02, Wind=0.0, huonggio=0.0,
        Zbandau=0.0, heso=0.94, mu_mn=0.2, NDnen=0.00, NDbphai=0.5, NDbtrai=0.5,
        NDbtren=0.5, NDbduoi=0.5, Tod=0.06, Toe=0.15, hstoe=0, ghtoe=3, Mbochat=0.00001,
        ro=1000, ros=2000, dm=0.00001, d90=0.002, muy=1.01e-06, Dorong=0.5, Sx=2,
        sohn=8, soha=3, sohn1=3, CORIOLIS_FORCE=0.0, g=9.81)
    base.update(overrides)
    ks = base.pop("Ks", None)
    lines = ["#ifndef CONSTANT_CUH__", "#define CONSTANT_CUH__",
             "#define PI 3.14159265359", "#define DOUBLE double", "",
             f"__constant__ int segment_limit = {int(segment_limit)};",
             f"__constant__ DOUBLE dX = {_fmt(dX)};",
             f"__constant__ DOUBLE dY = {_fmt(dY)};",
             f"__constant__ DOUBLE dT = {_fmt(dT)};",
             f"__constant__ DOUBLE dXbp = {_fmt(dX)} * {_fmt(dX)};",
             f"__constant__ DOUBLE dX2 = 2 * {_fmt(dX)};",
             f"__constant__ DOUBLE dYbp = {_fmt(dY)} * {_fmt(dY)};",
             f"__constant__ DOUBLE dY2 = 2 * {_fmt(dY)};",
             f"__constant__ DOUBLE dTchia2dX = {_fmt(dT)} / (2 * {_fmt(dX)});",
             f"__constant__ DOUBLE dTchia2dY = {_fmt(dT)} / (2 * {_fmt(dY)});",
             "__constant__ DOUBLE QuyDoiTime = 1.0 / 3600;",
             "__constant__ DOUBLE QuyDoiPi = 1.0 / PI;",
             f"__constant__ DOUBLE HaiChiadT = 2.0 / {_fmt(dT)};"]
    for k, v in base.items():
        if k in ("segment_limit", "dX", "dY", "dT"):
            continue
        lines.append(f"__constant__ DOUBLE {k} = {_fmt(v) if isinstance(v, float) else v};")
    if ks is None:
        lines.append(f"__constant__ DOUBLE Ks = 2.5 * {_fmt(base['dm'])};")
    else:
        lines.append(f"__constant__ DOUBLE Ks = {_fmt(ks)};")
    lines += ["#endif", ""]
    return "\n".join(lines)
