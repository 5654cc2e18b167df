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


def pretty_filesize(num, suffix="B"):
    """Bytes as a short binary-prefixed string (utils.py:70-84)."""
    for unit in ("", "Ki", "Mi", "Gi", "Ti", "Pi", "Ei", "Zi"):
        if abs(num) < 1024.0:
            return "%3.1f %s%s" % (num, unit, suffix)
        num /= 1024.0
    return "%.1f %s%s" % (num, "Yi", suffix)
