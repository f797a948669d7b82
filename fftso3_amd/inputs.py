"""Seeded synthetic inputs (host side).

Random numbers come from SplitMix64 (state_i = seed + i * 0x9E3779B97F4A7C15, i = 1, 2, ...)
mapped to uniforms u = (z >> 11) * 2^-53, and normals by Box-Muller on consecutive pairs
(n = sqrt(-2 ln(1 - u1)) cos(2 pi u2)), one normal per pair.  Coefficients are drawn function by
function, then l, m, n ascending, real part before imaginary part, and stored in the packed
layout off(l) + (m+l) + (n+l)(2l+1) (S^2: l^2 + m + l).
"""
import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed, count, start=0):
    i = np.arange(start + 1, start + count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + i * GOLDEN
        z = (z ^ (z >> np.uint64(30))) * M1
        z = (z ^ (z >> np.uint64(27))) * M2
    return z ^ (z >> np.uint64(31))


def uniform(seed, count, start=0):
    return (splitmix64(seed, count, start) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def normal(seed, count, start=0):
    u = uniform(seed, 2 * count, 2 * start)
    u1, u2 = u[0::2], u[1::2]
    return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)


def so3_ncoef(B):
    return B * (4 * B * B - 1) // 3


def so3_draw_permutation(B):
    """perm[i] = packed index of the i-th draw ((l, m, n) ascending, n fastest)."""
    perm = np.empty(so3_ncoef(B), dtype=np.int64)
    i = 0
    for l in range(B):
        w = 2 * l + 1
        off = l * (4 * l * l - 1) // 3
        m = np.repeat(np.arange(-l, l + 1), w)
        n = np.tile(np.arange(-l, l + 1), w)
        perm[i:i + w * w] = off + (m + l) + (n + l) * w
        i += w * w
    return perm


def so3_degree_of_packed(B):
    """l of every packed SO(3) index."""
    return np.concatenate([np.full((2 * l + 1) ** 2, l) for l in range(B)])


def s2_degree_of_packed(B):
    return np.concatenate([np.full(2 * l + 1, l) for l in range(B)])


def so3_coefficients(B, n_fn, seed, complex_=True, scale=None, start=0):
    """Packed coefficients [n_fn, ncoef] (complex128 or float64).

    scale(l) multiplies every coefficient of degree l; draws begin at draw index `start`."""
    nc = so3_ncoef(B)
    per = 2 * nc if complex_ else nc
    x = normal(seed, n_fn * per, start).reshape(n_fn, nc, 2 if complex_ else 1)
    vals = x[..., 0] + 1j * x[..., 1] if complex_ else x[..., 0]
    out = np.empty_like(vals)
    out[:, so3_draw_permutation(B)] = vals
    if scale is not None:
        out = out * scale(so3_degree_of_packed(B))[None, :]
    return out


def s2_coefficients(B, n_fn, seed, complex_=True, scale=None, start=0):
    nc = B * B
    per = 2 * nc if complex_ else nc
    x = normal(seed, n_fn * per, start).reshape(n_fn, nc, 2 if complex_ else 1)
    vals = x[..., 0] + 1j * x[..., 1] if complex_ else x[..., 0]  # draw order == packed order
    if scale is not None:
        vals = vals * scale(s2_degree_of_packed(B))[None, :]
    return vals


# ---- the five workloads ---------------------------------------------------------------------

def workload_complex_b16():
    """Complex SO(3), l_max=15, 1 function, Re/Im i.i.d. N(0,1), seed 0."""
    return so3_coefficients(16, 1, 0)


def workload_real_b64():
    """Real SO(3), l_max=63, 1 function, N(0,1)/(1+l)^1.5, seed 1."""
    return so3_coefficients(64, 1, 1, complex_=False, scale=lambda l: (1.0 + l) ** -1.5)


def workload_batch_b64(n_fn=256):
    """Complex SO(3), l_max=63, 256 functions, Re/Im N(0, 1/(1+l)^2), seed 2."""
    return so3_coefficients(64, n_fn, 2, scale=lambda l: 1.0 / (1.0 + l))


UNEVEN_BANDLIMITS = (8, 16, 24, 32, 48, 64)


def workload_uneven(n_fn=512):
    """512 real SO(3) functions, B uniform in {8,16,24,32,48,64} (seed 3: the first n_fn uniforms
    pick B), coefficients N(0, 1/(1+l)^2) drawn from the same stream after them.
    Returns (list of B, list of float64 coefficient arrays)."""
    u = uniform(3, n_fn)
    Bs = [UNEVEN_BANDLIMITS[min(5, int(6 * x))] for x in u]
    coefs, start = [], n_fn
    for B in Bs:
        c = so3_coefficients(B, 1, 3, complex_=False, scale=lambda l: 1.0 / (1.0 + l), start=start)[0]
        start += so3_ncoef(B)
        coefs.append(c)
    return Bs, coefs


def workload_b256(n_fn=8):
    """Complex SO(3), l_max=255, 8 functions, Re/Im N(0, 1/(1+l)^2), seed 4."""
    return so3_coefficients(256, n_fn, 4, scale=lambda l: 1.0 / (1.0 + l))


def workload_s2(n_fn=4096):
    """Complex S^2, l_max=127, 4096 fields, Re/Im N(0, 1/(1+l)^2), seed 5."""
    return s2_coefficients(128, n_fn, 5, scale=lambda l: 1.0 / (1.0 + l))
