"""Writes wigner_d.json and legendre.json: high-precision golden values that pin the oracle's
Wigner small-d and normalised associated Legendre functions.

Wigner d is evaluated from Wigner's explicit sum (Varshalovich et al., eq. 4.3.1(2)):
  d^l_{mn}(b) = sum_s (-1)^(m-n+s) sqrt((l+m)!(l-m)!(l+n)!(l-n)!) /
                ((l+n-s)! s! (m-n+s)! (l-m-s)!) cos(b/2)^(2l+n-m-2s) sin(b/2)^(m-n+2s)
with mpmath at 250 significant digits (the alternating sum cancels ~10^150 at l = 255), and
nP^l_m(x) = sqrt((2l+1)/(4 pi) (l-m)!/(l+m)!) P_l^m(x) with P_l^m = (-1)^m (1-x^2)^(m/2) d^m/dx^m P_l
(Condon-Shortley phase) and P_l(x) = 2^-l sum_k (-1)^k C(l,k) C(2l-2k,l) x^(l-2k), the normalisation
of src/fdcl_FFTS2.cpp:48-100.  Run: python tests/golden/make_wigner_golden.py
"""
import json
import os
import random

import mpmath as mp

mp.mp.dps = 250
HERE = os.path.dirname(os.path.abspath(__file__))


def wigner_d(l, m, n, beta):
    b = mp.mpf(beta)
    c, s = mp.cos(b / 2), mp.sin(b / 2)
    tot = mp.mpf(0)
    pref = mp.sqrt(mp.factorial(l + m) * mp.factorial(l - m) * mp.factorial(l + n) * mp.factorial(l - n))
    for k in range(max(0, n - m), min(l + n, l - m) + 1):
        den = mp.factorial(l + n - k) * mp.factorial(k) * mp.factorial(m - n + k) * mp.factorial(l - m - k)
        term = pref / den * c ** (2 * l + n - m - 2 * k) * s ** (m - n + 2 * k)
        tot += -term if (m - n + k) % 2 else term
    return tot


def legendre_norm(l, m, x):
    x = mp.mpf(x)
    acc = mp.mpf(0)
    for k in range(0, (l - m) // 2 + 1):
        e = l - 2 * k
        term = mp.binomial(l, k) * mp.binomial(2 * l - 2 * k, l) * mp.factorial(e) / mp.factorial(e - m) * x ** (e - m)
        acc += -term if k % 2 else term
    plm = acc / mp.mpf(2) ** l * (1 - x * x) ** (mp.mpf(m) / 2)
    if m % 2:
        plm = -plm
    return mp.sqrt((2 * l + 1) / (4 * mp.pi) * mp.factorial(l - m) / mp.factorial(l + m)) * plm


def main():
    rnd = random.Random(20240601)
    samples = []
    # grid betas of B = 16, 64, 256 and a few extreme angles
    betas = []
    for B in (16, 64, 256):
        for k in (0, 1, B // 2, 2 * B - 2, 2 * B - 1):
            betas.append(float(mp.mpf(2 * k + 1) * mp.pi / (4 * B)))
    betas += [1e-3, 0.3, 1.0, 1.5707963267948966, 2.5, 3.14]
    for beta in betas:
        for _ in range(6):
            l = rnd.choice([rnd.randint(0, 8), rnd.randint(0, 63), rnd.randint(64, 255)])
            m = rnd.randint(-l, l)
            n = rnd.randint(-l, l)
            samples.append([l, m, n, beta, float(wigner_d(l, m, n, beta))])
    # the exact corners that underflow in the reference seeds at large l
    for beta in (float(mp.mpf(511) * mp.pi / 1024), float(mp.mpf(1) * mp.pi / 1024)):
        for (l, m, n) in [(255, 255, -255), (255, -255, 255), (200, 150, -140), (255, 0, 0), (128, 128, 128)]:
            samples.append([l, m, n, beta, float(wigner_d(l, m, n, beta))])
    # large degrees in the oscillatory region, where a lost term or a wrong sign changes the value
    # at O(1): 48 samples with 128 <= l <= 255 and |d| > 1e-3 (rejection sampling, own stream so
    # the samples above stay as they were)
    big = random.Random(20261204)
    nbig = 0
    while nbig < 48:
        l = big.randint(128, 255)
        m = big.randint(-l, l)
        n = big.randint(-l, l)
        beta = big.uniform(0.05, 3.09)
        v = wigner_d(l, m, n, beta)
        if abs(v) > mp.mpf("1e-3"):
            samples.append([l, m, n, beta, float(v)])
            nbig += 1
    with open(os.path.join(HERE, "wigner_d.json"), "w") as f:
        json.dump({"source": "Wigner explicit sum, mpmath dps=250", "fields": ["l", "m", "n", "beta", "d"],
                   "samples": samples}, f, indent=0)

    leg = []
    for x in (-0.999, -0.5, 0.0, 0.123, 0.7, 0.9999):
        for _ in range(10):
            l = rnd.choice([rnd.randint(0, 10), rnd.randint(0, 127)])
            m = rnd.randint(0, l)
            leg.append([l, m, x, float(legendre_norm(l, m, x))])
    with open(os.path.join(HERE, "legendre.json"), "w") as f:
        json.dump({"source": "sqrt((2l+1)/4pi (l-m)!/(l+m)!) * (-1)^m (1-x^2)^(m/2) d^m/dx^m P_l(x), "
                             "P_l from its explicit polynomial, mpmath dps=250",
                   "fields": ["l", "m", "x", "nP"], "samples": leg}, f, indent=0)


if __name__ == "__main__":
    main()
