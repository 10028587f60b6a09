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


# ------------------------------------------------------------------------------------------------ hip sizing model
# Device bytes of the hip former, from the plan's shapes and how each level runs (0 two passes, 1 one pass on VALU,
# 2 one pass on the matrix cores: fsh_ffbp_level_forms). The counterpart of the reference's full_speed, cuda_group
# and cuda_rest (fastsar/memory.py:149-297) for this backend's own buffers; tests check it against the bytes the
# library actually holds.

GROUP = 8                  # first-level children per group at full speed
GROUP_BUDGET = 4 << 30     # groups grow past GROUP, up to every first-level child, while their buffers stay within this


def padded_row(Ko):
    """Row stride (samples) of a two-pass level's range-filtered rows on the device: Ko rounded up to 32 from 128
    samples up, so that each row starts on a 128-byte line."""
    return (Ko + 31) // 32 * 32 if Ko >= 128 else Ko


def hip_group_bytes(plan, g, forms):
    """Bytes of the buffers of a group of g first-level children: two ping-pong buffers of the largest level output
    (and, for a two-pass level, its range-filtered intermediate) in complex64, the device-computed rotations of the
    later levels (two float32 arrays), and the final geometry (ux, uy [tiles, Pf] and three values per tile, float64)."""
    maxbuf = maxph = 0
    parents = 1
    for l, lv in enumerate(plan['levels']):
        outs = parents * (g if l == 0 else lv['C'])
        one = forms[l] != 0
        if not one:        # the range-filtered intermediate, its rows padded as the device pads them
            maxbuf = max(maxbuf, outs * lv['P'] * padded_row(lv['Ko']))
        if lv['Dp'] > 1 or one:
            maxbuf = max(maxbuf, outs * lv['Po'] * lv['Ko'])
        if l >= 2:
            maxph = max(maxph, parents * lv['C'] * lv['P'])
        parents = outs
    return 2 * 8 * maxbuf + 2 * 4 * maxph + parents * (2 * plan['final']['P'] + 3) * 8


def hip_default_group(plan, forms):
    """GROUP first-level children (fewer when there are fewer), doubled while the buffers stay within GROUP_BUDGET."""
    G = plan['levels'][0]['C']
    g = min(GROUP, G)
    budget = max(hip_group_bytes(plan, g, forms), GROUP_BUDGET)
    while g < G and hip_group_bytes(plan, min(G, 2 * g), forms) <= budget:
        g = min(G, 2 * g)
    return g


def hip_constants(plan, weights=False, window=True):
    """Bytes of the plan constants the former uploads once: taps, the host levels' rotations (float32 [B, C, P]
    twice), the later levels' track and centres (float64), the windows, the final track, tile centres and corners,
    and the aperture weights (three float32 [tiles, Pf]). The polyphase tables of the one-pass levels (a few hundred
    bytes) are left out."""
    from .plan import HOST_LEVELS
    n = 4
    for l, lv in enumerate(plan['levels']):
        n += 4 * ((lv['fir_k']['L'] if lv['Dk'] > 1 else 0) + (lv['fir_p']['L'] if lv['Dp'] > 1 else 0))
        B = lv['ref'].shape[0]
        if l < HOST_LEVELS:
            n += 2 * 4 * B * lv['C'] * lv['P']
        else:
            n += 24 * lv['P'] + 8 * lv['P'] + 24 * lv['C'] + 24 * B
    lv0, fin = plan['levels'][0], plan['final']
    tiles = fin['cen'].shape[0]
    if window:
        n += 4 * (lv0['P'] + lv0['K'])
    n += 32 * fin['P'] + 24 * tiles + 8 * tiles
    if weights:
        n += 3 * 4 * tiles * fin['P']
    return n


def hip_full_speed(plan, forms, host_input=True, weights=False, window=True, group=None):
    """Device bytes of the hip former at full speed: (total, parts). The windowed history [P, K] complex64, the
    uploaded one beside it when the history comes from the host, the image [nx, ny] complex64, the group buffers at
    the default group (or `group`), and the plan constants."""
    lv0 = plan['levels'][0]
    hist = 8 * lv0['P'] * lv0['K']
    g = hip_default_group(plan, forms) if group is None else group
    parts = dict(history=hist * (2 if host_input else 1), image=8 * plan['nx'] * plan['ny'],
                 groups=hip_group_bytes(plan, g, forms), constants=hip_constants(plan, weights, window) + 4)
    return float(sum(parts.values())), parts


def hip_group(plan, forms, free, host_input=True, weights=False, window=True):
    """First-level children per group the former runs with `free` device bytes: the default, halved while the whole
    former would not fit (at least 1)."""
    g = hip_default_group(plan, forms)
    while g > 1 and hip_full_speed(plan, forms, host_input, weights, window, group=g)[0] > free:
        g //= 2
    return g
