"""JSON output in the reference's serde_json layout (crates/cli/src/main.rs:119-122).

serde_json's pretty printer indents by two spaces, keeps struct fields in
declaration order and formats floats with ryu: the shortest round-trip digits,
fixed notation for decimal exponents in (-5, 16] (integral values keep ".0"),
scientific otherwise ("1e-6", "1.5e16").  f32 values (profile bins) use the
shortest f32 digits.  Non-finite floats are written as null.
"""
import math

import numpy as np


class F32(float):
    """A float that came from an f32 (formatted with f32 shortest digits)."""


def _digits(s):
    """(sign, digits, exponent) with value = digits * 10**exponent, from a repr string."""
    sign = ""
    if s[0] == "-":
        sign, s = "-", s[1:]
    mant, _, exp = s.partition("e")
    e = int(exp) if exp else 0
    ip, _, fp = mant.partition(".")
    d = (ip + fp).lstrip("0")
    e -= len(fp)
    if not d:
        return sign, "0", 0
    stripped = d.rstrip("0")
    e += len(d) - len(stripped)
    return sign, stripped, e


def format_float(x, f32=False):
    if not math.isfinite(x):
        return "null"
    if x == 0.0:
        return "-0.0" if math.copysign(1.0, x) < 0 else "0.0"
    r = repr(np.float32(x)) if f32 else repr(float(x))
    if f32:  # numpy reprs look like 'np.float32(0.1)' or '0.1'
        r = r[r.find("(") + 1:r.rfind(")")] if "(" in r else r
    sign, d, e = _digits(r)
    k = len(d)
    kk = k + e
    if 0 <= e and kk <= 16:
        return sign + d + "0" * e + ".0"
    if 0 < kk <= 16:
        return sign + d[:kk] + "." + d[kk:]
    if -5 < kk <= 0:
        return sign + "0." + "0" * (-kk) + d
    if k == 1:
        return sign + d + "e" + str(kk - 1)
    return sign + d[0] + "." + d[1:] + "e" + str(kk - 1)


def _scalar(v):
    if v is None:
        return "null"
    if isinstance(v, (bool, np.bool_)):
        return "true" if v else "false"
    if isinstance(v, F32) or isinstance(v, np.float32):
        return format_float(float(v), f32=True)
    if isinstance(v, (float, np.floating)):
        return format_float(float(v))
    if isinstance(v, (int, np.integer)):
        return str(int(v))
    if isinstance(v, str):
        out = ['"']
        for ch in v:
            if ch == '"':
                out.append('\\"')
            elif ch == "\\":
                out.append("\\\\")
            elif ch == "\n":
                out.append("\\n")
            elif ch == "\t":
                out.append("\\t")
            elif ord(ch) < 0x20:
                out.append("\\u%04x" % ord(ch))
            else:
                out.append(ch)
        out.append('"')
        return "".join(out)
    raise TypeError("unsupported JSON value %r" % type(v))


def dumps(value, indent=0):
    """serde_json::to_string_pretty of nested dicts (field order kept), lists, tuples and scalars."""
    pad = "  " * (indent + 1)
    if isinstance(value, dict):
        if not value:
            return "{}"
        items = [pad + _scalar(str(k)) + ": " + dumps(v, indent + 1) for k, v in value.items()]
        return "{\n" + ",\n".join(items) + "\n" + "  " * indent + "}"
    if isinstance(value, np.ndarray):
        value = [F32(x) if value.dtype == np.float32 else x.item() for x in value]
    if isinstance(value, (list, tuple)):
        if not value:
            return "[]"
        return "[\n" + ",\n".join(pad + dumps(v, indent + 1) for v in value) + "\n" + "  " * indent + "]"
    return _scalar(value)


def write(path, value):
    with open(path, "w") as f:
        f.write(dumps(value))
