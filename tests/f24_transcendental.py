"""Test cases for the F24 `&`, `|` and `^` operators (num/src/num.rs:243-248):
`a & b` = atan2(a, b), `a | b` = log_a(b) = ln(b) / ln(a) (Rust's
`b.log(a)`), `a ^ b` = powf(a, b), each computed in f32 and rounded to f24
(num.rs:81-90).

Nothing in the reference's tests exercises them, so they are pinned by
mathematics instead:

* identities whose result is exact (`x ^ 1 = x`, `x ^ 0 = 1`, `2 ^ k`,
  `log_b(b ^ k) = k`, `log_b(1) = +-0`, `atan2(+-0, x > 0) = +-0`,
  `atan2(y, 0) = +-pi/2`, `atan2(+-0, -1) = +-pi`, `atan2(1, 1) = pi/4`),
  where the f32 result is exact or lies far from an f24 rounding boundary,
  so any implementation within a few f32 ulps rounds to the same f24 value;
* seeded random operands whose exact result (computed in double precision)
  lies at least `MARGIN` f32 ulps away from an f24 rounding boundary: there
  too, every implementation within that many ulps gives the same f24 value.
  Random operands closer to a boundary are left out, since the last f32 bit
  of a libm routine (glibc's in the reference and in the oracle, the device
  library's on the GPU) decides them.

NaN results are compared by NaN-ness only: the reference's readback prints
every NaN the same way.
"""
import math
import struct

import numpy as np

F24 = 2
OPS = {"&": 12, "|": 13, "^": 14}
MARGIN = {"&": 4, "^": 4, "|": 8}


def f32(x):
    return float(np.float32(x))


def f32_bits(x):
    return struct.unpack("<I", struct.pack("<f", np.float32(x)))[0]


def f24_raw(x):
    """Round an f32 to f24 (round to nearest even on the low 8 bits)."""
    x = np.float32(x)
    if math.isnan(x):
        return None
    bits = f32_bits(x)
    hi, lo = bits >> 8, bits & 0xFF
    if lo > 0x80 or (lo == 0x80 and (hi & 1)):
        hi += 1
    return ((hi << 8) | (F24 << 3) | 3) & 0xFFFFFFFF


def f24_value(x):
    """The f32 value of an f24 number (low 8 bits cleared)."""
    raw = f24_raw(x)
    return struct.unpack("<f", struct.pack("<I", raw & ~0xFF & 0xFFFFFFFF))[0]


def exact(op, a, b):
    """The operator's result in double precision (Python's math library)."""
    try:
        if op == "&":
            return math.atan2(a, b)
        if op == "^":
            return math.pow(a, b)
        if op == "|":
            la, lb = math.log(a), math.log(b)
            return lb / la
    except (ValueError, ZeroDivisionError, OverflowError):
        return float("nan")


def near_boundary(op, v):
    """True when `v` rounded to f32 lies within MARGIN ulps of an f24 tie."""
    if math.isnan(v) or math.isinf(v) or abs(v) > 3e38 or (v != 0 and abs(v) < 1e-37):
        return True
    lo = f32_bits(v) & 0xFF
    return abs(lo - 0x80) <= MARGIN[op]


def identity_cases():
    """(op, a, b, expected f32 value) with results fixed by mathematics."""
    pi = math.pi
    cases = []
    for x in [0.75, 1.5, -3.25, 100.0, 12345.0, 0.001, -0.5]:
        cases.append(("^", x, 1.0, f24_value(x)))
        cases.append(("^", x, 0.0, 1.0))
    for k in range(-24, 25, 3):
        cases.append(("^", 2.0, float(k), 2.0 ** k))
    cases += [("^", -2.0, 3.0, -8.0), ("^", 4.0, 0.5, 2.0), ("^", 3.0, 2.0, 9.0), ("^", 10.0, 3.0, 1000.0),
              ("^", 1.0, 7.5, 1.0)]
    for base, k in [(2.0, 5), (2.0, -7), (2.0, 20), (10.0, 3), (3.0, 4), (0.5, 3), (7.0, 2)]:
        cases.append(("|", base, base ** k, float(k)))
    for base in [2.0, 10.0, 0.25]:
        cases.append(("|", base, base, 1.0))
        cases.append(("|", base, 1.0, 0.0 if base > 1 else -0.0))
    for x in [1.0, 2.5, 1e6]:
        cases.append(("&", 0.0, x, 0.0))
        cases.append(("&", -0.0, x, -0.0))
    for y in [1.0, 3.5, 1e-3]:
        cases.append(("&", y, 0.0, f32(pi / 2)))
        cases.append(("&", -y, 0.0, f32(-pi / 2)))
    cases += [("&", 0.0, -1.0, f32(pi)), ("&", -0.0, -1.0, f32(-pi)), ("&", 1.0, 1.0, f32(pi / 4)),
              ("&", -2.0, -2.0, f32(-3 * pi / 4)), ("&", 1.0, float("inf"), 0.0)]
    return cases


def random_cases(seed, n):
    """(op, a, b, exact double result) away from f24 rounding boundaries;
    operands are f24 values."""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < n:
        op = ["&", "|", "^"][len(out) % 3]
        if op == "&":
            a, b = (rng.standard_normal(2) * 10.0 ** rng.integers(-3, 4, 2)).tolist()
        elif op == "|":
            a = float(rng.uniform(0.01, 100.0))
            b = float(rng.uniform(1e-3, 1e6))
            if abs(a - 1.0) < 1e-2:
                continue
        else:
            a = float(rng.uniform(0.05, 50.0))
            b = float(rng.uniform(-8.0, 8.0))
        a, b = f24_value(a), f24_value(b)
        v = exact(op, a, b)
        if near_boundary(op, v):
            continue
        out.append((op, a, b, v))
    return out


def nan_cases():
    """Operands whose result is NaN (log of a negative, pow of a negative
    base with a non-integer exponent)."""
    return [("|", 2.0, -3.0), ("|", -2.0, 8.0), ("^", -2.5, 0.5)]


def num_literal(x):
    """An f24 operand as .hvm text: the shortest positional decimal that
    parses back to the same f32 (the grammar has no exponent form)."""
    if math.isinf(x):
        return "+inf" if x > 0 else "-inf"
    s = np.format_float_positional(np.float32(x), unique=True, trim="-")
    if "." not in s:
        s += ".0"
    return s if s.startswith("-") else "+" + s


def book(cases):
    """A book whose main is the tuple of every case's result:
    (r0 (r1 (... *))) with `& a ~ $([op] $(b r_i))` per case."""
    root = "*"
    for i in reversed(range(len(cases))):
        root = f"(r{i} {root})"
    lines = [f"@main = {root}"]
    for i, (op, a, b) in enumerate(cases):
        lines.append(f"& {num_literal(a)} ~ $([{op}] $({num_literal(b)} r{i}))")
    return "\n".join(lines) + "\n"


def tuple_items(nf):
    """The items of a normal form (x0 (x1 (... *)))."""
    items = []
    s = nf.strip()
    while s != "*":
        assert s.startswith("(") and s.endswith(")"), s
        inner = s[1:-1]
        head, _, rest = inner.partition(" ")
        items.append(head)
        s = rest
    return items
