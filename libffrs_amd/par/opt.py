"""CIRC16 geometry for a file (ffrs/par/opt.py:29-164), found by a direct search of the
reference's constraint set instead of its generic solver.

Dimensions as the reference's solver states them: the inner code in bytes (inner_message and
inner_ecc are twice the symbol counts; ffrs/util.py:108-117 halves them), the outer code in rows.
With the default ecc ratio 1/1 the outer code has outer_ecc = outer_message, a power of two from
the command line's default set. For the largest outer_ecc that admits one, the choice is: outer_interleave = interleave *
inner_message within the reference's padding window above file_size / outer_message
(opt.py:131-137), the largest inner_message, then the smallest inner_ecc with
127 <= inner_message / inner_ecc <= 512, then the smallest message, with
file_size <= message <= 2 * file_size (file_size at least 8192)."""
import math


def _pow2(lo, hi):
    v = 1
    while v <= hi:
        if v >= lo:
            yield v
        v <<= 1


# the command line's defaults (ffrs/par/common_args.py:197-239): inner ecc 4 or 8 bytes, outer
# ecc 16 .. 256 rows
INNER_ECC_BYTES = (4, 8)
OUTER_ECC = (16, 32, 64, 128, 256)


def choose_geometry(file_size, ecc_ratio=(1, 1), inner_ecc=INNER_ECC_BYTES, outer_ecc=OUTER_ECC):
    """(inner_block, inner_ecc, outer_block, outer_ecc, interleave) in CIRC16's symbol units."""
    num, den = ecc_ratio
    fs = max(int(file_size), 8192)
    max_padding = 0.05 if fs <= 16000 else 0.01
    for oe in sorted(outer_ecc, reverse=True):
        if (oe * den) % num:
            continue
        om = oe * den // num  # outer_message * num == outer_ecc * den
        if not 2 <= om <= 65536 or om + oe > 65536:
            continue
        start = math.ceil(fs / om)
        window = range(start, start + min(100, math.ceil(fs * max_padding / om)))
        best = None
        for oi in window:
            message = om * oi
            if not fs <= message <= 2 * fs:
                continue
            for im in range(2, 4097, 2):  # inner_message in bytes
                if oi % im:
                    continue
                for ie in inner_ecc:  # bytes
                    if im % ie or not 127 <= im // ie <= 512 or im + ie > 65536:
                        continue
                    I = oi // im
                    if (im // 2) * I > 0xFFFF:  # the outer code's interleave
                        continue
                    key = (-im, ie, message)
                    if best is None or key < best[0]:
                        best = (key, im, ie, I)
        if best is not None:
            _, im, ie, I = best
            return ((im + ie) // 2, ie // 2, om + oe, oe, I)
    raise ValueError("could not choose a CIRC16 geometry for %d bytes" % file_size)


def choose_codec(file_size, ecc_ratio=(1, 1), **simd):
    from .. import CIRC16
    ib, ie, ob, oe, I = choose_geometry(file_size, ecc_ratio)
    return CIRC16(ib, ie, ob, oe, I, **simd)
