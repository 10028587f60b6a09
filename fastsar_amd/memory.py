"""Memory warnings and environment settings (fastsar/memory.py:18-55).

MemoryWarning is issued whenever a former falls back to a slower configuration for lack of device memory (fewer
first-level children per group, fewer pulses per chunk). Environment variables are parsed strictly: a malformed
value raises ValueError naming the variable (fastsar/memory.py:35-54).
"""
import os
import warnings


class MemoryWarning(UserWarning):
    """A former runs a slower configuration for lack of memory. (fastsar/memory.py:18-22)"""


def warn(msg):
    warnings.warn(msg, MemoryWarning, stacklevel=3)


def env_number(name, default, kind=int):
    """The value of an environment variable as `kind`, default when unset or empty; ValueError when malformed or
    negative. (fastsar/memory.py:35-45)"""
    v = os.environ.get(name, '').strip()
    if not v:
        return default
    try:
        x = kind(v)
    except ValueError:
        raise ValueError(f'{name} must be a number ({kind.__name__}), got {v!r}') from None
    if x < 0:
        raise ValueError(f'{name} must not be negative, got {v!r}')
    return x


def env_choice(name, default, choices):
    """The value of an environment variable among `choices`. (fastsar/memory.py:48-54)"""
    v = os.environ.get(name, '').strip() or default
    if v not in choices:
        raise ValueError(f'{name} must be one of {sorted(choices)}, got {v!r}')
    return v


def gb(n):
    return f'{n / 1e9:.1f} GB'
