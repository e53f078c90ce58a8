"""Numeric comparison of the reference's text outputs, numdiff-style.

Fields are compared as numbers (so -0 == 0 and NaN == NaN); a field printed
with d significant digits matches a value when they agree to half a unit in
the last printed digit (plus an optional relative slack).
"""
from __future__ import annotations

import math
import re

_NUM = re.compile(r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?|NaN|inf|-inf")


def tokens(text: str):
    """Split into (kind, token) with kind 'num' or 'txt'."""
    out = []
    for line in text.splitlines():
        for tok in line.split():
            if _NUM.fullmatch(tok):
                out.append(("num", tok))
            else:
                # split trailing ':'-like punctuation from numbers glued to text
                out.append(("txt", tok))
    return out


def half_ulp_printed(tok: str) -> float:
    """Half a unit in the last printed digit of a decimal token."""
    if tok in ("NaN", "inf", "-inf"):
        return 0.0
    mant, _, exp = tok.lower().partition("e")
    e = int(exp) if exp else 0
    if "." in mant:
        dec = len(mant.split(".")[1])
    else:
        dec = 0
    return 0.5 * 10.0 ** (e - dec)


def num_equal(a: str, b: str, rel: float = 0.0, abs_: float = 0.0) -> bool:
    if a == b:
        return True
    fa, fb = float(a), float(b)
    if math.isnan(fa) or math.isnan(fb):
        return math.isnan(fa) and math.isnan(fb)
    return abs(fa - fb) <= max(abs_, rel * max(abs(fa), abs(fb)))


def compare_texts(expected: str, actual: str, rel: float = 0.0, abs_: float = 0.0, digits_slack: float = 0.0):
    """Return a list of mismatches (empty = equal under numdiff semantics).

    digits_slack allows |a-b| <= digits_slack * (unit of the last printed digit).
    """
    te, ta = tokens(expected), tokens(actual)
    errs = []
    if len(te) != len(ta):
        errs.append(f"token count {len(te)} != {len(ta)}")
    for i, ((ke, ve), (ka, va)) in enumerate(zip(te, ta)):
        if ke != ka:
            errs.append(f"#{i}: kind {ke}:{ve!r} vs {ka}:{va!r}")
        elif ke == "txt":
            if ve != va:
                errs.append(f"#{i}: text {ve!r} vs {va!r}")
        else:
            slack = 2 * half_ulp_printed(ve) * digits_slack
            if not num_equal(ve, va, rel, max(abs_, slack)):
                errs.append(f"#{i}: {ve} vs {va}")
    return errs


def printed_match(value: float, tok: str, rel: float = 1e-12) -> bool:
    """Does an unrounded value print as tok (within half a unit in the last
    digit, plus rel slack)?"""
    if tok == "NaN":
        return math.isnan(value)
    t = float(tok)
    return abs(value - t) <= half_ulp_printed(tok) * (1 + 1e-9) + rel * abs(t)
