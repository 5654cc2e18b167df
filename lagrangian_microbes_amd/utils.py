"""Small helpers with the names and results of the reference's utils.py.

``factor`` and ``most_symmetric_integer_factorization`` fix the particle
lattice shape (utils.py:7-42, used by particle_advecter.py:28);
``pretty_time`` / ``pretty_filesize`` format log lines (utils.py:57-84).
"""
import math
from datetime import timedelta


def factor(n):
    """Sorted list of the positive divisors of n (utils.py:7-22)."""
    n = int(n)
    small = [d for d in range(1, math.isqrt(n) + 1) if n % d == 0]
    return sorted(set(small) | {n // d for d in small})


def most_symmetric_integer_factorization(N):
    """(a, N // a) with a the divisor of N closest to sqrt(N) (utils.py:25-42).

    For divisors a <= sqrt(N) <= N/a, N/a + a - 2 sqrt(N) >= 0, so the
    closest divisor is never above sqrt(N): it is the largest divisor not
    exceeding isqrt(N).
    """
    N = int(N)
    a = max(d for d in range(1, math.isqrt(N) + 1) if N % d == 0)
    return a, N // a


def closest_hour(ndt):
    """numpy datetime64 -> datetime rounded to the nearest hour (utils.py:45-54)."""
    import numpy as np
    from datetime import datetime
    secs = int((np.datetime64(ndt, "s") - np.datetime64("1970-01-01T00:00:00", "s")) // np.timedelta64(1, "s"))
    dt = datetime(1970, 1, 1) + timedelta(seconds=secs)
    base = dt.replace(minute=0, second=0, microsecond=0)
    return base + timedelta(hours=1) if dt.minute >= 30 else base


def pretty_time(t):
    """Seconds as a short human string (utils.py:57-67)."""
    for limit, scale, unit in ((1e-6, 1e9, "ns"), (1e-3, 1e6, "μs"), (1.0, 1e3, "ms"), (60.0, 1.0, "s")):
        if t < limit:
            return "{:.3g} {}".format(t * scale, unit)
    return "{:.3g} mins".format(t / 60)


# IEC binary prefixes, index k standing for 1024**k
_BINARY_PREFIXES = ("", "Ki", "Mi", "Gi", "Ti", "Pi", "Ei", "Zi", "Yi")


def pretty_filesize(num, suffix="B"):
    """Bytes as a short binary-prefixed string, one decimal: the prefix is
    the largest power of 1024 not above |num| (capped at Yi), the value num
    over that power.  The reference logs pickle sizes in this format
    (utils.py:70-84, particle_advecter.py:249)."""
    mag = abs(float(num))
    top = len(_BINARY_PREFIXES) - 1
    k = min(top, int(math.log(mag, 1024))) if mag >= 1024 else 0
    # log() may round to either side of an exact power
    if k < top and mag >= 1024.0 ** (k + 1):
        k += 1
    elif k > 0 and mag < 1024.0 ** k:
        k -= 1
    return "{:.1f} {}{}".format(num / 1024.0 ** k, _BINARY_PREFIXES[k], suffix)
