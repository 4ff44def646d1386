"""Hash-consed real-valued expression graph for DFT codelets.

A codelet is built over complex values (pairs of real nodes).  Nodes are
structurally shared (common subexpressions merge automatically) and simple
algebraic identities are folded while building, so the emitted code has no
multiplications by 0/1/-1 and no double negations.  The same graph can be
evaluated numerically (numpy, any float dtype) for the generator tests and
printed as straight-line HIP C++.
"""
from __future__ import annotations

from dataclasses import dataclass
from fractions import Fraction

import mpmath


@dataclass(frozen=True)
class Const:
    """An exact-ish real constant: the double value and its float32 rounding."""
    f64: float
    f32: float

    def __neg__(self) -> "Const":
        return Const(-self.f64, -self.f32)


def _round_f32(value: mpmath.mpf) -> float:
    with mpmath.workprec(24):
        return float(+value)


_COS_CACHE: dict[tuple[int, int], tuple[Const, Const]] = {}


def cos_sin(m: int, big_m: int) -> tuple[Const, Const]:
    """(cos, sin) of 2*pi*m/M, correctly rounded to double and to float."""
    key = (m % big_m, big_m)
    if key in _COS_CACHE:
        return _COS_CACHE[key]
    frac = Fraction(m % big_m, big_m)
    with mpmath.workdps(50):
        angle = 2 * mpmath.pi * mpmath.mpf(frac.numerator) / frac.denominator
        c = mpmath.cos(angle)
        s = mpmath.sin(angle)
        # exact zeros and ones for the quarter turns
        out = []
        for v in (c, s):
            if abs(v) < mpmath.mpf(10) ** -40:
                out.append(Const(0.0, 0.0))
            else:
                out.append(Const(float(v), _round_f32(v)))
    _COS_CACHE[key] = (out[0], out[1])
    return _COS_CACHE[key]


class Node:
    __slots__ = ("kind", "args", "const", "name", "index")

    def __init__(self, kind, args=(), const=None, name=None):
        self.kind = kind        # "in", "const", "add", "sub", "mul", "neg"
        self.args = args
        self.const = const
        self.name = name
        self.index = -1

    def key(self):
        if self.kind == "in":
            return ("in", self.name)
        if self.kind == "const":
            return ("const", self.const.f64)
        if self.kind in ("add", "mul"):
            a, b = (id(x) for x in self.args)
            return (self.kind, min(a, b), max(a, b))
        return (self.kind,) + tuple(id(x) for x in self.args)


class Graph:
    """Builder with structural sharing and light constant folding."""

    def __init__(self):
        self.table: dict[tuple, Node] = {}
        self.order: list[Node] = []
        self.zero = self.const(Const(0.0, 0.0))
        self.one = self.const(Const(1.0, 1.0))

    def _intern(self, node: Node) -> Node:
        key = node.key()
        found = self.table.get(key)
        if found is not None:
            return found
        node.index = len(self.order)
        self.table[key] = node
        self.order.append(node)
        return node

    def inp(self, name: str) -> Node:
        return self._intern(Node("in", name=name))

    def const(self, c: Const) -> Node:
        return self._intern(Node("const", const=c))

    def is_const(self, a: Node, value: float) -> bool:
        return a.kind == "const" and a.const.f64 == value

    def neg(self, a: Node) -> Node:
        if a.kind == "neg":
            return a.args[0]
        if a.kind == "const":
            return self.const(-a.const)
        return self._intern(Node("neg", (a,)))

    def add(self, a: Node, b: Node) -> Node:
        if self.is_const(a, 0.0):
            return b
        if self.is_const(b, 0.0):
            return a
        if b.kind == "neg":
            return self.sub(a, b.args[0])
        if a.kind == "neg":
            return self.sub(b, a.args[0])
        return self._intern(Node("add", (a, b)))

    def sub(self, a: Node, b: Node) -> Node:
        if self.is_const(b, 0.0):
            return a
        if self.is_const(a, 0.0):
            return self.neg(b)
        if b.kind == "neg":
            return self.add(a, b.args[0])
        return self._intern(Node("sub", (a, b)))

    def mul(self, c: Const, a: Node) -> Node:
        """Constant times node."""
        if c.f64 == 0.0:
            return self.zero
        if c.f64 == 1.0:
            return a
        if c.f64 == -1.0:
            return self.neg(a)
        if c.f64 < 0.0:
            return self.neg(self._intern(Node("mul", (self.const(-c), a))))
        return self._intern(Node("mul", (self.const(c), a)))


class Cx:
    """A complex value as two graph nodes."""
    __slots__ = ("re", "im")

    def __init__(self, re: Node, im: Node):
        self.re = re
        self.im = im


def cadd(g: Graph, a: Cx, b: Cx) -> Cx:
    return Cx(g.add(a.re, b.re), g.add(a.im, b.im))


def csub(g: Graph, a: Cx, b: Cx) -> Cx:
    return Cx(g.sub(a.re, b.re), g.sub(a.im, b.im))


def cmul_omega(g: Graph, a: Cx, m: int, big_m: int, sign: int) -> Cx:
    """a * exp(sign*2*pi*i*m/M) with the quarter and eighth turns special-cased."""
    m %= big_m
    if m == 0:
        return a
    frac = Fraction(m, big_m)
    if sign > 0:
        frac = (1 - frac) % 1
    # now multiply by exp(-2*pi*i*frac)
    if frac == Fraction(1, 2):
        return Cx(g.neg(a.re), g.neg(a.im))
    if frac == Fraction(1, 4):      # * -i
        return Cx(a.im, g.neg(a.re))
    if frac == Fraction(3, 4):      # * +i
        return Cx(g.neg(a.im), a.re)
    c, s = cos_sin(frac.numerator, frac.denominator)
    # exp(-i*theta) = c - i*s
    if abs(c.f64) == abs(s.f64):
        # (a.re + i a.im)(c - i s) with |c| = |s| = h:
        # re = c*a.re + s*a.im, im = c*a.im - s*a.re
        h = Const(abs(c.f64), abs(c.f32))
        sc = 1 if c.f64 > 0 else -1
        ss = 1 if s.f64 > 0 else -1
        t_re = g.add(a.re, a.im) if sc == ss else g.sub(a.re, a.im)
        t_im = g.sub(a.im, a.re) if sc == ss else g.add(a.im, a.re)
        re = g.mul(h, t_re) if sc > 0 else g.neg(g.mul(h, t_re))
        im = g.mul(h, t_im) if sc > 0 else g.neg(g.mul(h, t_im))
        return Cx(re, im)
    re = g.add(g.mul(c, a.re), g.mul(s, a.im))
    im = g.sub(g.mul(c, a.im), g.mul(s, a.re))
    return Cx(re, im)


def evaluate(g: Graph, outputs: list[Node], inputs: dict, dtype):
    """Evaluate `outputs` given numpy arrays for the input names."""
    import numpy as np
    values: dict[int, object] = {}
    for node in g.order:
        if node.kind == "in":
            values[node.index] = np.asarray(inputs[node.name], dtype=dtype)
        elif node.kind == "const":
            v = node.const.f32 if dtype == np.float32 else node.const.f64
            values[node.index] = dtype(v)
        elif node.kind == "add":
            values[node.index] = values[node.args[0].index] + values[node.args[1].index]
        elif node.kind == "sub":
            values[node.index] = values[node.args[0].index] - values[node.args[1].index]
        elif node.kind == "mul":
            values[node.index] = values[node.args[0].index] * values[node.args[1].index]
        elif node.kind == "neg":
            values[node.index] = -values[node.args[0].index]
    return [values[o.index] for o in outputs]


def live_nodes(outputs: list[Node]) -> list[Node]:
    """Nodes reachable from outputs, in creation (topological) order."""
    seen: dict[int, Node] = {}
    stack = list(outputs)
    while stack:
        n = stack.pop()
        if n.index in seen:
            continue
        seen[n.index] = n
        stack.extend(n.args)
    return [seen[i] for i in sorted(seen)]


def op_count(outputs: list[Node]) -> dict[str, int]:
    counts = {"add": 0, "mul": 0, "neg": 0}
    for n in live_nodes(outputs):
        if n.kind in ("add", "sub"):
            counts["add"] += 1
        elif n.kind == "mul":
            counts["mul"] += 1
        elif n.kind == "neg":
            counts["neg"] += 1
    return counts
