"""Smallest 5-smooth FFT length (blond/toolbox/next_regular.py:9-59)."""
from __future__ import annotations


def next_regular(target: int) -> int:
    """Next regular (2^a 3^b 5^c) number >= target, with the reference's
    convention: targets <= 6 and powers of two are returned unchanged; other
    targets give 2 * (smallest regular number >= ceil(target / 2)), so the
    result always carries a factor 2."""
    target = int(target)
    if target <= 6:
        return target
    if not (target & (target - 1)):
        return target
    half = -(-target // 2)
    best = None
    p5 = 1
    while True:
        p35 = p5
        while True:
            # smallest power of two p2 with p2 * p35 >= half
            q = -(-half // p35)
            cand = (1 << (q - 1).bit_length()) * p35
            if best is None or cand < best:
                best = cand
            if p35 >= half:
                break
            p35 *= 3
        if p5 >= half:
            break
        p5 *= 5
    return 2 * best
