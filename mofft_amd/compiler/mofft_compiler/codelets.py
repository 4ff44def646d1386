"""Radix-r DFT codelets: exact rewrites of the r x r DFT-matrix contraction.

Forward transform, X[k] = sum_j x[j] * w_r^(j*k) with w_r = exp(-2*pi*i/r);
the kernels obtain the backward transform by exchanging real and imaginary
parts on load and store (B(x) = swap(F(swap(x)))).

Three algebraic forms, all equal to the direct contraction in exact
arithmetic (the reference's rewrite families, rewrites.py:15-95):

* ``direct``   - the dense matrix; every coefficient explicit.
* ``paired``   - the like-terms rewrite: rows j and r-j share a cosine and a
  negated sine, so x[j] +/- x[r-j] are formed once (used for prime r).
* ``factored`` - r = p*q Cooley-Tukey: q DFTs of size p over stride q,
  constant twiddles w_r^(n2*k1), p DFTs of size q.  The reference's even-odd
  rewrite is the p = 2 case.
"""
from __future__ import annotations

from .graph import Cx, Graph, cadd, cmul_omega, cos_sin, csub

RADICES = (2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 32, 64)

# Factorisation used by the factored form: r -> (p, q), p = inner DFT size.
SPLITS = {4: (2, 2), 6: (2, 3), 8: (2, 4), 9: (3, 3), 10: (2, 5),
          12: (4, 3), 14: (2, 7), 15: (3, 5), 16: (4, 4), 32: (4, 8),
          64: (8, 8)}


def _is_prime(r: int) -> bool:
    return r > 1 and all(r % d for d in range(2, int(r ** 0.5) + 1))


def dft_direct(g: Graph, x: list[Cx]) -> list[Cx]:
    r = len(x)
    out = []
    for k in range(r):
        acc = x[0]
        for j in range(1, r):
            acc = cadd(g, acc, cmul_omega(g, x[j], j * k, r, -1))
        out.append(acc)
    return out


def dft_paired(g: Graph, x: list[Cx]) -> list[Cx]:
    """Like-terms form for odd r."""
    r = len(x)
    h = (r - 1) // 2
    s = [cadd(g, x[j], x[r - j]) for j in range(1, h + 1)]
    d = [csub(g, x[j], x[r - j]) for j in range(1, h + 1)]
    out: list[Cx | None] = [None] * r
    acc = x[0]
    for v in s:
        acc = cadd(g, acc, v)
    out[0] = acc
    for k in range(1, h + 1):
        a_re, a_im = x[0].re, x[0].im
        b_re, b_im = g.zero, g.zero
        for j in range(1, h + 1):
            c, sn = cos_sin(j * k, r)
            a_re = g.add(a_re, g.mul(c, s[j - 1].re))
            a_im = g.add(a_im, g.mul(c, s[j - 1].im))
            b_re = g.add(b_re, g.mul(sn, d[j - 1].re))
            b_im = g.add(b_im, g.mul(sn, d[j - 1].im))
        # X_k = A - i B, X_{r-k} = A + i B (forward sign)
        out[k] = Cx(g.add(a_re, b_im), g.sub(a_im, b_re))
        out[r - k] = Cx(g.sub(a_re, b_im), g.add(a_im, b_re))
    return out  # type: ignore[return-value]


def dft_factored(g: Graph, x: list[Cx], p: int, q: int) -> list[Cx]:
    r = p * q
    assert len(x) == r
    inner = []
    for n2 in range(q):
        inner.append(dft(g, [x[q * n1 + n2] for n1 in range(p)]))
    out: list[Cx | None] = [None] * r
    for k1 in range(p):
        col = [cmul_omega(g, inner[n2][k1], n2 * k1, r, -1) for n2 in range(q)]
        res = dft(g, col)
        for k2 in range(q):
            out[k1 + p * k2] = res[k2]
    return out  # type: ignore[return-value]


def dft(g: Graph, x: list[Cx], form: str = "auto") -> list[Cx]:
    r = len(x)
    if r == 1:
        return list(x)
    if r == 2:
        return [cadd(g, x[0], x[1]), csub(g, x[0], x[1])]
    if form == "direct":
        return dft_direct(g, x)
    if _is_prime(r):
        return dft_paired(g, x)
    p, q = SPLITS.get(r, (None, None))
    if p is None:
        for p in range(2, r):
            if r % p == 0:
                q = r // p
                break
    return dft_factored(g, x, p, q)


def build(r: int, form: str = "auto"):
    """Graph, input names and output nodes for a forward radix-r codelet."""
    g = Graph()
    x = [Cx(g.inp(f"r{j}"), g.inp(f"i{j}")) for j in range(r)]
    y = dft(g, x, form)
    return g, y
