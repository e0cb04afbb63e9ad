"""Pins the oracle's 24-bit numerics (num/src/num.rs) against closed forms
written independently here with numpy: u24 arithmetic mod 2**24, i24 two's
complement with truncating division, f24 = IEEE f32 op then round-to-nearest-
even of the low 8 mantissa bits (num.rs:81-90), saturating casts with NaN->0.
The F24 `&`, `|`, `^` operators (atan2 / log / powf) are pinned by the
identities and boundary-free random cases of f24_transcendental.py: nothing
in the reference's tests fixes their last bit, mathematics fixes these."""
import math
import struct

import numpy as np
import pytest

U24, I24, F24, SYM = 0, 1, 2, 3
OPS = {"+": 4, "-": 5, ":-": 6, "*": 7, "/": 8, ":/": 9, "%": 10, ":%": 11, "&": 12, "|": 13, "^": 14,
       "<<": 15, ":<<": 16, ">>": 17, ":>>": 18, "=": 19, "!": 20, "<": 21, ">": 22}
M = (1 << 24) - 1


def num(tag, payload):
    return ((payload & M) << 8) | (tag << 3) | 3


def u24(v):
    return num(U24, v)


def i24(v):
    return num(I24, v & M)


def sext(v):
    v &= M
    return v - (1 << 24) if v & (1 << 23) else v


def f32_bits(x):
    return struct.unpack("<I", struct.pack("<f", np.float32(x)))[0]


def f24_raw(x):
    """Independent restatement of round-to-nearest-even to 24 bits."""
    x = np.float32(x)
    bits = f32_bits(x)
    if math.isnan(x):
        return None  # NaN: only NaN-ness is compared
    hi, lo = bits >> 8, bits & 0xFF
    if lo > 0x80 or (lo == 0x80 and (hi & 1)):
        hi += 1
    return (hi << 8) | (F24 << 3) | 3


def f24_val(raw):
    return struct.unpack("<f", struct.pack("<I", raw & ~0xFF & 0xFFFFFFFF))[0]


def u24_closed(op, a, b):
    f = {
        "+": lambda: a + b, "-": lambda: a - b, ":-": lambda: b - a, "*": lambda: a * b,
        "/": lambda: a // b, ":/": lambda: b // a, "%": lambda: a % b, ":%": lambda: b % a,
        "&": lambda: a & b, "|": lambda: a | b, "^": lambda: a ^ b,
        "<<": lambda: a << (b & 31), ":<<": lambda: b << (a & 31), ">>": lambda: a >> (b & 31),
        ":>>": lambda: b >> (a & 31),
        "=": lambda: int(a == b), "!": lambda: int(a != b), "<": lambda: int(a < b), ">": lambda: int(a > b),
    }[op]
    return u24(f() & 0xFFFFFFFF)


def tdiv(x, y):
    q = abs(x) // abs(y)
    return q if (x >= 0) == (y >= 0) else -q


def i24_closed(op, a, b):
    cmp = {"=": a == b, "!": a != b, "<": a < b, ">": a > b}
    if op in cmp:
        return u24(int(cmp[op]))
    f = {
        "+": lambda: a + b, "-": lambda: a - b, ":-": lambda: b - a, "*": lambda: a * b,
        "/": lambda: tdiv(a, b), ":/": lambda: tdiv(b, a), "%": lambda: a - b * tdiv(a, b),
        ":%": lambda: b - a * tdiv(b, a), "&": lambda: a & b, "|": lambda: a | b, "^": lambda: a ^ b,
        "<<": lambda: a << (b & 31), ":<<": lambda: b << (a & 31), ">>": lambda: a >> (b & 31),
        ":>>": lambda: b >> (a & 31),
    }[op]
    return i24(f())


EDGE_U = [0, 1, 2, 3, 7, 8, 23, 24, 31, 32, 33, 255, 4096, (1 << 23) - 1, 1 << 23, M - 1, M]


@pytest.mark.parametrize("op", list(OPS))
def test_u24_binary_closed_form(oracle, op):
    rng = np.random.default_rng(OPS[op])
    pairs = [(a, b) for a in EDGE_U for b in EDGE_U] + [tuple(map(int, rng.integers(0, 1 << 24, 2))) for _ in range(800)]
    for a, b in pairs:
        if op in ("/", "%") and b == 0 or op in (":/", ":%") and a == 0:
            continue  # division by zero aborts the reference
        got = oracle.operate_binary(u24(a), OPS[op], u24(b))
        assert got == u24_closed(op, a, b), (op, a, b)


@pytest.mark.parametrize("op", list(OPS))
def test_i24_binary_closed_form(oracle, op):
    rng = np.random.default_rng(100 + OPS[op])
    vals = [0, 1, -1, 2, -2, 23, 24, 31, 32, -(1 << 23), (1 << 23) - 1, 12345, -12345]
    pairs = [(a, b) for a in vals for b in vals] + [tuple(int(x) for x in rng.integers(-(1 << 23), 1 << 23, 2))
                                                   for _ in range(800)]
    for a, b in pairs:
        if op in ("/", "%") and b == 0 or op in (":/", ":%") and a == 0:
            continue
        got = oracle.operate_binary(i24(a), OPS[op], i24(b))
        assert got == i24_closed(op, a, b), (op, a, b)


@pytest.mark.parametrize("op", ["+", "-", ":-", "*", "/", ":/", "%", ":%", "=", "!", "<", ">"])
def test_f24_binary_closed_form(oracle, op):
    rng = np.random.default_rng(200 + OPS[op])
    specials = [0.0, -0.0, 1.0, -1.5, 2.5, 1e-30, 3e38, float("inf"), float("-inf"), float("nan")]
    xs = specials + [float(np.float32(v)) for v in rng.standard_normal(120) * 10.0 ** rng.integers(-8, 9, 120)]
    rng.shuffle(xs)
    for a, b in zip(xs, xs[::-1] + xs[:3]):
        ra, rb = oracle.new_f24(a), oracle.new_f24(b)
        fa, fb = np.float32(f24_val(ra)), np.float32(f24_val(rb))
        with np.errstate(all="ignore"):
            if op in ("=", "!", "<", ">"):
                res = {"=": fa == fb, "!": fa != fb, "<": fa < fb, ">": fa > fb}[op]
                exp = u24(int(bool(res)))
            else:
                v = {"+": fa + fb, "-": fa - fb, ":-": fb - fa, "*": fa * fb, "/": fa / fb, ":/": fb / fa,
                     "%": np.fmod(fa, fb), ":%": np.fmod(fb, fa)}[op]
                exp = f24_raw(v)
        got = oracle.operate_binary(ra, OPS[op], rb)
        if exp is None:
            assert math.isnan(f24_val(got)) and (got >> 3) & 0x1F == F24, (op, a, b)
        else:
            assert got == exp, (op, a, b, hex(got), hex(exp))


def test_f24_rounding_closed_form(oracle):
    rng = np.random.default_rng(7)
    for bits in list(rng.integers(0, 1 << 32, 5000, dtype=np.uint64)) + [0x3F800080, 0x3F800180, 0x7F7FFFFF]:
        x = struct.unpack("<f", struct.pack("<I", int(bits)))[0]
        exp = f24_raw(x)
        got = oracle.new_f24(x)
        if exp is None:
            assert math.isnan(f24_val(got))
        else:
            assert got == exp, hex(int(bits))


def test_casts_closed_form(oracle):
    # u24 <-> i24 reinterpret payload bits; f24 -> int saturates, NaN -> 0 (num.rs:133-161)
    for v in [0, 1, 1234, (1 << 23) - 1, 1 << 23, M]:
        assert oracle.operate_unary(U24, i24(v)) == u24(v)
        assert oracle.operate_unary(I24, u24(v)) == i24(sext(v))
        assert oracle.operate_unary(F24, u24(v)) == f24_raw(float(np.float32(v)))
    for x, eu, ei in [(2.8, 2, 2), (-12.5, 0, -12), (16777216.0, M, (1 << 23) - 1), (float("inf"), M, (1 << 23) - 1),
                      (float("-inf"), 0, -(1 << 23)), (float("nan"), 0, 0), (-8388610.0, 0, -(1 << 23))]:
        r = oracle.new_f24(x)
        assert oracle.operate_unary(U24, r) == u24(eu), x
        assert oracle.operate_unary(I24, r) == i24(ei), x
    # partial application keeps the payload (num.rs:159) and applying a Sym is the unary op (num.rs:169)
    assert oracle.operate_unary(OPS["*"], u24(2)) == num(OPS["*"], 2)
    assert oracle.operate_binary(u24(3), SYM, num(OPS["*"], 4)) == u24(12)


def rust_f32_debug(x):
    """Rust `{:?}` for a finite f32, from numpy's shortest round-trip digits:
    exponent form when 0 < |x| < 1e-4 or |x| >= 1e16, else positional with at
    least one fractional digit (core::fmt float Debug)."""
    x = np.float32(x)
    a = abs(float(x))
    if (a != 0 and a < 1e-4) or a >= 1e16:
        m = np.format_float_scientific(x, unique=True, trim="-")
        mant, e = m.split("e")
        return f"{mant}e{int(e)}"
    s = np.format_float_positional(x, unique=True, trim="-")
    return s if "." in s else s + ".0"


@pytest.mark.parametrize("text,shown", [
    ("0", "0"), ("16777215", "16777215"), ("+5", "+5"), ("-5", "-5"), ("0x10", "16"), ("0b101", "5"),
    ("+inf", "+inf"), ("-inf", "-inf"), ("+NaN", "+NaN"), ("[+]", "[+]"), ("[u24]", "[u24]"), ("[*2]", "[*2]"),
    ("[:-7]", "[:-7]"), ("1e20", None), ("@", None),
])
def test_parse_and_display(oracle, text, shown):
    """Numbers print as num.rs:287-306 Display."""
    if shown is None:
        with pytest.raises(Exception):
            oracle.parse_num(text)
        return
    assert oracle.num_display(oracle.parse_num(text)) == shown


def test_f24_display_matches_rust_debug(oracle):
    rng = np.random.default_rng(11)
    vals = [0.1, 1.1, 1.02, 2.15, -2.15, 0.15, 1234.0, 123456.0, 16775982.0, 1e-5, 3.5e-5, 1e16, 2.5e20,
            -7.25e-9, 0.5, 100.0, 1e-4, 99999.99]
    vals += [float(v) for v in rng.standard_normal(300) * 10.0 ** rng.integers(-12, 20, 300)]
    for v in vals:
        raw = oracle.new_f24(v)
        assert oracle.num_display(raw) == rust_f32_debug(f24_val(raw)), v


def test_f24_transcendental_identities(oracle):
    """`x ^ 1 = x`, `x ^ 0 = 1`, `2 ^ k`, `log_b(b ^ k) = k`,
    `atan2(+-0, x) = +-0`, `atan2(y, 0) = +-pi/2`, ... rounded to f24."""
    import f24_transcendental as ft
    for op, a, b, want in ft.identity_cases():
        got = oracle.operate_binary(oracle.new_f24(a), ft.OPS[op], oracle.new_f24(b))
        assert got == ft.f24_raw(want), (op, a, b, hex(got), hex(ft.f24_raw(want)))


def test_f24_transcendental_random(oracle):
    """Random f24 operands whose exact result is away from an f24 rounding
    boundary round to the exact result's f24 value."""
    import f24_transcendental as ft
    for op, a, b, v in ft.random_cases(seed=5, n=600):
        got = oracle.operate_binary(oracle.new_f24(a), ft.OPS[op], oracle.new_f24(b))
        assert got == ft.f24_raw(v), (op, a, b, v, hex(got))


def test_f24_transcendental_nan(oracle):
    import f24_transcendental as ft
    for op, a, b in ft.nan_cases():
        got = oracle.operate_binary(oracle.new_f24(a), ft.OPS[op], oracle.new_f24(b))
        assert (got >> 3) & 0x1F == F24 and math.isnan(f24_val(got)), (op, a, b, hex(got))


def test_f24_transcendental_book(oracle):
    """The book the GPU test reduces: its operands parse to the intended f24
    values and the oracle's normal form lists the pinned results in order."""
    import f24_transcendental as ft
    cases = ft.identity_cases()[:40]
    nf = oracle.run(ft.book([(op, a, b) for op, a, b, _ in cases]), heap_words=1 << 20)["nf"]
    items = ft.tuple_items(nf)
    assert items == [oracle.num_display(ft.f24_raw(w)) for _, _, _, w in cases]
