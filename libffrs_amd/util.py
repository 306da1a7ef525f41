"""Base-64 integers of the `.ffrs` parity format (ffrs/util.py:21-48): most significant digit first,
digits 0-9 a-z A-Z - _, and zero written as the empty string."""
DIGITS = "0123456789abcdefghijklmnopqrstuvwxyzABCDEFGHIJKLMNOPQRSTUVWXYZ-_"
_VALUE = {c: i for i, c in enumerate(DIGITS)}


def b64hex(value: int) -> str:
    if value < 0:
        raise ValueError("negative value")
    out = []
    while value:
        out.append(DIGITS[value & 63])
        value >>= 6
    return "".join(reversed(out))


def b64hex_dec(text: str) -> int:
    value = 0
    for c in text:
        value = (value << 6) | _VALUE[c]
    return value
