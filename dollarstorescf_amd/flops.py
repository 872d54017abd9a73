"""Algorithmic FLOP model of the Fock build (what roofline numbers divide).

Per surviving primitive quartet of class (la lb|lc ld), L = la+lb+lc+ld,
nc(l) = (l+1)(l+2)/2:

    F_prim = 40 (setup) + (27 + 3L) (Boys) + 10 N_vrr
             + 2 sum_{e=la}^{la+lb} sum_{f=lc}^{lc+ld} nc(e) nc(f)   (contraction)
    N_vrr  = sum_{e<=la+lb, f<=lc+ld} nc(e) nc(f) (L-e-f+1) - (L+1)

(the Head-Gordon-Pople VRR count at <= 5 FMA per entry).  Per surviving
shell quartet: F_hrr = 2 (bra-HRR entries x ket components + ket-HRR entries
x nc(la) nc(lb)) and F_digest = 14 n_ao(a) n_ao(b) n_ao(c) n_ao(d): per AO
integral J takes 2 FMA (the bra sum and the ket sum) and K 4 FMA (one per
exchange image), 12 FLOPs, and the permutational weight is applied to the
integral for the J and for the K terms, 2 more.  No hardware counter enters
these numbers.
"""
from __future__ import annotations


def nc(l: int) -> int:
    return (l + 1) * (l + 2) // 2


def n_vrr(la, lb, lc, ld) -> int:
    L = la + lb + lc + ld
    s = sum(nc(e) * nc(f) * (L - e - f + 1) for e in range(la + lb + 1) for f in range(lc + ld + 1))
    return s - (L + 1)


def f_prim(la, lb, lc, ld) -> int:
    L = la + lb + lc + ld
    contr = sum(nc(e) * nc(f) for e in range(la, la + lb + 1) for f in range(lc, lc + ld + 1))
    return 40 + (27 + 3 * L) + 10 * n_vrr(la, lb, lc, ld) + 2 * contr


def _hrr_entries(l1, l2) -> int:
    return sum(nc(i) * nc(j) for j in range(1, l2 + 1) for i in range(l1, l1 + l2 - j + 1))


def f_quartet(la, lb, lc, ld, spherical=True) -> int:
    nao = (lambda l: 2 * l + 1) if spherical else nc
    ket_comp = sum(nc(f) for f in range(lc, lc + ld + 1))
    hrr = 2 * (_hrr_entries(la, lb) * ket_comp + _hrr_entries(lc, ld) * nc(la) * nc(lb))
    return hrr + 14 * nao(la) * nao(lb) * nao(lc) * nao(ld)


def class_flops(cls, n_quartets, n_prim_quartets, spherical=True) -> float:
    return float(f_prim(*cls)) * n_prim_quartets + float(f_quartet(*cls, spherical=spherical)) * n_quartets


def fock_build_flops(classes, quartets, prim_quartets, spherical=True) -> float:
    return sum(class_flops(c, q, p, spherical) for c, q, p in zip(classes, quartets, prim_quartets))
