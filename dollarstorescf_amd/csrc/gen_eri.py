"""Generator for the per-class ERI device code (gfx950 HIP).

For every shell-quartet class (la lb|lc ld) (la>=lb, lc>=ld, bra class >= ket
class) it emits straight-line code for

  * the Head-Gordon-Pople form of the Obara-Saika vertical recursion
    [e0|f0]^(m) (the four-centre OS scheme of GaussianIntegral.cpp:334-559,
    reorganised: VRR on the A and C centres, then horizontal transfer),
  * contraction over primitive pairs (the loops of GaussianIntegral.cpp:573-576),
  * the horizontal recursions (a,b+1| = (a+1,b| + AB (a,b| and the ket analogue,
  * the Cartesian -> reference-AO transform (the Cartesian-term loops of
    GaussianIntegral.cpp:598-601 collapsed into one linear map per l, whose
    coefficients the host uploads into c_ao_xform),
  * J/K digestion (ExactCoulomb.cpp:103-123 restated for direct SCF).

Only the recursion *structure* is decided here; every numeric coefficient
comes from the host at run time.  Output: one .inc file included by
eri_kernels.hip.
"""
from __future__ import annotations

import sys
from fractions import Fraction

MAXL = 2


def cart(l):
    """Cartesian components of angular momentum l, x-major order."""
    out = []
    for a in range(l, -1, -1):
        for b in range(l - a, -1, -1):
            out.append((a, b, l - a - b))
    return out


def ncart(l):
    return (l + 1) * (l + 2) // 2


# ---------------------------------------------------------------------------
# Sparsity of the reference's real solid harmonics (which Cartesian monomials
# carry a nonzero weight for AO (l,m)).  Exact rational arithmetic on the
# polynomial Re/Im (x+iy)^|m| * P_l^|m|(z, r^2), the construction of
# GTO.cpp:15-88 without its normalization.
def _poly_mul_add(acc, poly, coef, dx=0, dy=0, dz=0):
    for (a, b, c), v in poly.items():
        k = (a + dx, b + dy, c + dz)
        acc[k] = acc.get(k, Fraction(0)) + v * coef


def ao_pattern(l, m):
    am = abs(m)
    base = {}
    i, x = 0, (am if m <= 0 else am - 1)
    while x >= 0:
        # binomial(am, x) with alternating sign
        from math import comb
        base[(x, am - x, 0)] = Fraction((-1) ** i * comb(am, x))
        x -= 2
        i += 1
    if l == am:
        poly = base
    else:
        prev = dict(base)
        cur = {}
        _poly_mul_add(cur, base, Fraction(2 * am + 1), dz=1)
        for ii in range(am + 1, l):
            sf = Fraction(ii - am + 1)
            nxt = {}
            _poly_mul_add(nxt, cur, Fraction(2 * ii + 1) / sf, dz=1)
            f = -Fraction(ii + am) / sf
            _poly_mul_add(nxt, prev, f, dx=2)
            _poly_mul_add(nxt, prev, f, dy=2)
            _poly_mul_add(nxt, prev, f, dz=2)
            prev, cur = cur, nxt
        poly = cur
    return [c for c in cart(l) if poly.get(c, 0) != 0]


def ao_index_pattern(l, spherical=True):
    """For each AO of an l-shell, the Cartesian component indices it uses."""
    cs = cart(l)
    if not spherical:
        return [[i] for i in range(len(cs))]
    return [[cs.index(c) for c in ao_pattern(l, m)] for m in range(-l, l + 1)]


def nao(l, spherical=True):
    return 2 * l + 1 if spherical else ncart(l)


# ---------------------------------------------------------------------------
def vname(e, f, m):
    return "v%d%d%d_%d%d%d_%d" % (e + f + (m,))


def sub1(t, i):
    return t[:i] + (t[i] - 1,) + t[i + 1:]


def pick_dir(t):
    best = -1
    for i in range(3):
        if t[i] > 0 and (best < 0 or t[i] > t[best]):
            best = i
    return best


XYZ = "xyz"


def gen_vrr(la, lb, lc, ld):
    """Statements computing [e|f]^(0) for e in shells la..la+lb and
    f in shells lc..lc+ld, from Fm[] (prefactor-scaled Boys values)."""
    E, Fk = la + lb, lc + ld
    targets = [(e, f, 0) for le in range(la, E + 1) for e in cart(le)
               for lf in range(lc, Fk + 1) for f in cart(lf)]
    order, seen = [], set()
    stmts = {}

    def need(key):
        if key in seen:
            return
        seen.add(key)
        e, f, m = key
        if sum(f) > 0:
            i = pick_dir(f)
            f1 = sub1(f, i)
            terms = [("QC%s" % XYZ[i], (e, f1, m)), ("WQ%s" % XYZ[i], (e, f1, m + 1))]
            if f1[i] > 0:
                f2 = sub1(f1, i)
                terms.append(("%d.0*o2q" % f1[i], (e, f2, m)))
                terms.append(("-%d.0*o2q*rq" % f1[i], (e, f2, m + 1)))
            if e[i] > 0:
                terms.append(("%d.0*o2pq" % e[i], (sub1(e, i), f1, m + 1)))
        elif sum(e) > 0:
            i = pick_dir(e)
            e1 = sub1(e, i)
            terms = [("PA%s" % XYZ[i], (e1, f, m)), ("WP%s" % XYZ[i], (e1, f, m + 1))]
            if e1[i] > 0:
                e2 = sub1(e1, i)
                terms.append(("%d.0*o2p" % e1[i], (e2, f, m)))
                terms.append(("-%d.0*o2p*rp" % e1[i], (e2, f, m + 1)))
        else:
            stmts[key] = None
            order.append(key)
            return
        for _, k in terms:
            need(k)
        stmts[key] = terms
        order.append(key)

    for t in targets:
        need(t)
    lines = []
    for key in order:
        e, f, m = key
        if stmts[key] is None:
            lines.append("const R %s = Fm[%d];" % (vname(e, f, m), m))
            continue
        expr = " + ".join("(%s)*%s" % (c, vname(*k)) for c, k in stmts[key])
        lines.append("const R %s = %s;" % (vname(e, f, m), expr))
    return lines, targets


def stmts_of(la, lb, lc, ld):
    """VRR statements of gen_vrr as {key: [(coefficient, source key), ...]}
    (None for the Boys base values)."""
    E, Fk = la + lb, lc + ld
    targets = [(e, f, 0) for le in range(la, E + 1) for e in cart(le)
               for lf in range(lc, Fk + 1) for f in cart(lf)]
    stmts = {}

    def need(key):
        if key in stmts:
            return
        e, f, m = key
        if sum(f) > 0:
            i = pick_dir(f)
            f1 = sub1(f, i)
            terms = [("QC%s" % XYZ[i], (e, f1, m)), ("WQ%s" % XYZ[i], (e, f1, m + 1))]
            if f1[i] > 0:
                f2 = sub1(f1, i)
                terms.append(("%d.0*o2q" % f1[i], (e, f2, m)))
                terms.append(("-%d.0*o2q*rq" % f1[i], (e, f2, m + 1)))
            if e[i] > 0:
                terms.append(("%d.0*o2pq" % e[i], (sub1(e, i), f1, m + 1)))
        elif sum(e) > 0:
            i = pick_dir(e)
            e1 = sub1(e, i)
            terms = [("PA%s" % XYZ[i], (e1, f, m)), ("WP%s" % XYZ[i], (e1, f, m + 1))]
            if e1[i] > 0:
                e2 = sub1(e1, i)
                terms.append(("%d.0*o2p" % e1[i], (e2, f, m)))
                terms.append(("-%d.0*o2p*rp" % e1[i], (e2, f, m + 1)))
        else:
            stmts[key] = None
            return
        stmts[key] = terms
        for _, k in terms:
            need(k)

    for t in targets:
        need(t)
    return stmts


def sources_of(la, lb, lc, ld):
    """VRR entries that some other entry reads."""
    out = set()
    for t in stmts_of(la, lb, lc, ld).values():
        if t:
            out.update(k for _, k in t)
    return out


def _vrr_key(line):
    """The (e, f, m) key a gen_vrr line defines."""
    name = line.split("=")[0].split()[-1]  # v<e>_<f>_<m>
    a, b, m = name[1:].split("_")
    return (tuple(int(x) for x in a), tuple(int(x) for x in b), int(m))


def acc_name(e, f):
    return "a%d%d%d_%d%d%d" % (e + f)


def gen_hrr(la, lb, lc, ld):
    """Statements from contracted accumulators a<e>_<f> to Cartesian c_<a>_<b>_<c>_<d>."""
    E, Fk = la + lb, lc + ld
    lines = []
    # bra HRR for each f in lc..Fk
    fs = [f for lf in range(lc, Fk + 1) for f in cart(lf)]
    bra_order, seen, bra_stmt = [], set(), {}

    def need_b(a, b):
        if (a, b) in seen:
            return
        seen.add((a, b))
        if sum(b) == 0:
            bra_stmt[(a, b)] = None
        else:
            i = pick_dir(b)
            b1 = sub1(b, i)
            a1 = a[:i] + (a[i] + 1,) + a[i + 1:]
            need_b(a1, b1)
            need_b(a, b1)
            bra_stmt[(a, b)] = (a1, b1, i)
        bra_order.append((a, b))

    for a in cart(la):
        for b in cart(lb):
            need_b(a, b)

    def bn(a, b, f):
        return "h%d%d%d%d%d%d_%d%d%d" % (a + b + f)

    for f in fs:
        for (a, b) in bra_order:
            st = bra_stmt[(a, b)]
            if st is None:
                lines.append("const double %s = %s;" % (bn(a, b, f), acc_name(a, f)))
            else:
                a1, b1, i = st
                lines.append("const double %s = %s + AB%s*%s;" % (bn(a, b, f), bn(a1, b1, f), XYZ[i], bn(a, b1, f)))
    # ket HRR
    k_order, kseen, k_stmt = [], set(), {}

    def need_k(c, d):
        if (c, d) in kseen:
            return
        kseen.add((c, d))
        if sum(d) == 0:
            k_stmt[(c, d)] = None
        else:
            i = pick_dir(d)
            d1 = sub1(d, i)
            c1 = c[:i] + (c[i] + 1,) + c[i + 1:]
            need_k(c1, d1)
            need_k(c, d1)
            k_stmt[(c, d)] = (c1, d1, i)
        k_order.append((c, d))

    for c in cart(lc):
        for d in cart(ld):
            need_k(c, d)

    def kn(a, b, c, d):
        return "c%d%d%d%d%d%d_%d%d%d%d%d%d" % (a + b + c + d)

    for a in cart(la):
        for b in cart(lb):
            for (c, d) in k_order:
                st = k_stmt[(c, d)]
                if st is None:
                    lines.append("const double %s = %s;" % (kn(a, b, c, d), bn(a, b, c)))
                else:
                    c1, d1, i = st
                    lines.append("const double %s = %s + CD%s*%s;" % (kn(a, b, c, d), kn(a, b, c1, d1), XYZ[i], kn(a, b, c, d1)))
    return lines, kn


def gen_xform(la, lb, lc, ld, kn, cartd=False):
    """Cartesian -> AO, out[((ia*NB+ib)*NC+ic)*ND+id].  cartd: shells with
    l >= 2 are Cartesian functions (their block of c_ao_xform is diagonal)."""
    pats = [ao_index_pattern(l, not (cartd and l >= 2)) for l in (la, lb, lc, ld)]
    cs = [cart(l) for l in (la, lb, lc, ld)]
    nb_, nc_, nd_ = (len(p) for p in pats[1:])
    lines = []
    # one index at a time: first d, then c, then b, then a
    # stage 1: t1[a][b][c][md]
    def t1(a, b, c, md):
        return "t1_%d_%d_%d_%d" % (a, b, c, md)

    def t2(a, b, mc, md):
        return "t2_%d_%d_%d_%d" % (a, b, mc, md)

    def t3(a, mb, mc, md):
        return "t3_%d_%d_%d_%d" % (a, mb, mc, md)

    def coef(l, m, ci):
        off = xform_offset_cart(l) if cartd and l >= 2 else xform_offset(l)
        return "c_ao_xform[%d]" % (off + m * ncart(l) + ci)

    for a in range(len(cs[0])):
        for b in range(len(cs[1])):
            for c in range(len(cs[2])):
                for md, comps in enumerate(pats[3]):
                    expr = " + ".join("%s*%s" % (coef(ld, md, ci), kn(cs[0][a], cs[1][b], cs[2][c], cs[3][ci])) for ci in comps)
                    lines.append("const double %s = %s;" % (t1(a, b, c, md), expr))
    for a in range(len(cs[0])):
        for b in range(len(cs[1])):
            for mc, comps in enumerate(pats[2]):
                for md in range(nd_):
                    expr = " + ".join("%s*%s" % (coef(lc, mc, ci), t1(a, b, ci, md)) for ci in comps)
                    lines.append("const double %s = %s;" % (t2(a, b, mc, md), expr))
    for a in range(len(cs[0])):
        for mb, comps in enumerate(pats[1]):
            for mc in range(nc_):
                for md in range(nd_):
                    expr = " + ".join("%s*%s" % (coef(lb, mb, ci), t2(a, ci, mc, md)) for ci in comps)
                    lines.append("const double %s = %s;" % (t3(a, mb, mc, md), expr))
    for ma, comps in enumerate(pats[0]):
        for mb in range(nb_):
            for mc in range(nc_):
                for md in range(nd_):
                    expr = " + ".join("%s*%s" % (coef(la, ma, ci), t3(ci, mb, mc, md)) for ci in comps)
                    lines.append("out[%d] = %s;" % (((ma * nb_ + mb) * nc_ + mc) * nd_ + md, expr))
    return lines


def xform_offset(l):
    off = 0
    for k in range(l):
        off += (2 * k + 1) * ncart(k)
    return off


def xform_offset_cart(l, maxl=MAXL):
    """Offset of the ncart(l) x ncart(l) Cartesian block of l >= 2, after
    the spherical blocks of l = 0..maxl."""
    return xform_offset(maxl + 1) + sum(ncart(k) ** 2 for k in range(2, l))


def xform_size(maxl=MAXL):
    return xform_offset_cart(maxl + 1, maxl)


def has_d(c):
    return max(c) >= 2


def classes(maxl=MAXL):
    pairs = [(a, b) for a in range(maxl + 1) for b in range(a + 1)]
    out = []
    for i, X in enumerate(pairs):
        for Y in pairs[:i + 1]:
            out.append(X + Y)
    return pairs, out


NO_BRA_UNROLL = {"ssss", "psss", "pppp"}


def cls_name(c):
    return "".join("spdf"[l] for l in c)


def gen_class(c, cartd=False):
    la, lb, lc, ld = c
    L = sum(c)
    vrr, targets = gen_vrr(la, lb, lc, ld)
    hrr, kn = gen_hrr(la, lb, lc, ld)
    xf = gen_xform(la, lb, lc, ld, kn, cartd)
    na = [nao(l, not (cartd and l >= 2)) for l in c]
    NAO = na[0] * na[1] * na[2] * na[3]
    name = cls_name(c) + ("_cart" if cartd else "")
    accs = sorted({(e, f) for (e, f, _) in targets})
    o = []
    o.append("// ---- class (%s) : L=%d, %d VRR terms, %d AO integrals" % (name, L, len(vrr), NAO))
    o.append("template <> struct ClassInfo<%d,%d,%d,%d,%d> { static constexpr int L = %d, NAO = %d, NA = %d, NB = %d, NC = %d, ND = %d; };"
             % (la, lb, lc, ld, 1 if cartd else 0, L, NAO, *na))
    o.append("// A group of gb*gk lanes (aligned, power of two) may share one quartet:")
    o.append("// lane (rb, rk) takes bra primitives rb, rb+gb, .. and ket primitives rk,")
    o.append("// rk+gk, ..; the [e0|f0] sums are then added across the group, so every")
    o.append("// lane of it ends with the whole quartet.  gb = gk = 1: one lane.")
    o.append("// BP: the bra primitive record (PrimPair, or PrimPairF staged for FP32 lists).")
    o.append("template <int BM, typename R, typename BP = PrimPair>")
    o.append("__device__ __forceinline__ void shell_quartet_%s(const ShellPair& bra, const ShellPair& ket,"
             " const BP* __restrict__ bra_pp, const PrimPair* __restrict__ ket_pp,"
             " const double* __restrict__ btab, double* __restrict__ out,"
             " int gb = 1, int rb = 0, int gk = 1, int rk = 0) {" % name)
    for (e, f) in accs:
        o.append("  double %s = 0.0;" % acc_name(e, f))
    o.append("  for (int ik = rk; ik < ket.nprim; ik += gk) {")
    o.append("    const PrimPair Q = ldg(ket_pp + ik);")
    # two bra primitives per trip overlap their Boys-table reads, at the cost
    # of registers; measured per class (8 waters/6-31G and C20H42/def2-SVP
    # builds): (ss|ss), (ps|ss) and (pp|pp) run faster with one per trip
    o.append("#pragma unroll %d" % (1 if name in NO_BRA_UNROLL else 2))
    o.append("    for (int ib = rb; ib < bra.nprim; ib += gb) {")
    o.append("      const BP P = bra_pp[ib];")
    o.append("      R Fm[%d];" % (L + 1))
    o.append("      PrimQuartet<R> g; prim_setup<BM, R, %d>(P, Q, bra, ket, btab, g, Fm);" % L)
    o.append("      const R PAx = g.PA[0], PAy = g.PA[1], PAz = g.PA[2];")
    o.append("      const R WPx = g.WP[0], WPy = g.WP[1], WPz = g.WP[2];")
    o.append("      const R QCx = g.QC[0], QCy = g.QC[1], QCz = g.QC[2];")
    o.append("      const R WQx = g.WQ[0], WQy = g.WQ[1], WQz = g.WQ[2];")
    o.append("      const R o2p = g.o2p, o2q = g.o2q, o2pq = g.o2pq, rp = g.rp, rq = g.rq;")
    o.append("      (void)PAx; (void)PAy; (void)PAz; (void)WPx; (void)WPy; (void)WPz; (void)QCx; (void)QCy; (void)QCz;")
    o.append("      (void)WQx; (void)WQy; (void)WQz; (void)o2p; (void)o2q; (void)o2pq; (void)rp; (void)rq;")
    # A target that no other VRR entry reads goes straight into its FP64
    # accumulator as a chain of FMAs (one instruction per term, no separate
    # sum and add); in the FP32 lists it is summed in R and converted.
    fused = {k: t for k, t in stmts_of(la, lb, lc, ld).items() if k not in sources_of(la, lb, lc, ld) and t}
    for ln in vrr:
        key = _vrr_key(ln)
        if key in fused:
            continue
        o.append("      " + ln)
    for (e, f) in accs:
        key = (e, f, 0)
        if key in fused:
            terms = fused[key]
            chain = acc_name(e, f)
            for c, k in reversed(terms):
                chain = "fma((double)(%s), (double)%s, %s)" % (c, vname(*k), chain)
            expr = " + ".join("(%s)*%s" % (c, vname(*k)) for c, k in terms)
            o.append("      if constexpr (sizeof(R) == 8) %s = %s;" % (acc_name(e, f), chain))
            o.append("      else %s += (double)(%s);" % (acc_name(e, f), expr))
        else:
            o.append("      %s += (double)%s;" % (acc_name(e, f), vname(e, f, 0)))
    o.append("    }")
    o.append("  }")
    o.append("  for (int off = 1; off < gb * gk; off <<= 1) {")
    for (e, f) in accs:
        o.append("    %s += __shfl_xor(%s, off);" % (acc_name(e, f), acc_name(e, f)))
    o.append("  }")
    o.append("  const double ABx = bra.AB[0], ABy = bra.AB[1], ABz = bra.AB[2];")
    o.append("  const double CDx = ket.AB[0], CDy = ket.AB[1], CDz = ket.AB[2];")
    o.append("  (void)ABx; (void)ABy; (void)ABz; (void)CDx; (void)CDy; (void)CDz;")
    for ln in hrr:
        o.append("  " + ln)
    for ln in xf:
        o.append("  " + ln)
    o.append("}")
    return "\n".join(o)


# ---------------------------------------------------------------------------
# Wave-cooperative (LDS) form for the classes whose accumulators do not fit in
# one thread's registers: the same VRR/HRR graph, emitted as level-ordered
# tables that the 64 lanes of a wavefront walk together.
REG_MAX_ACC = 300
REG_MAX_HRR = 2000


def is_wave(c):
    la, lb, lc, ld = c
    _, targets = gen_vrr(la, lb, lc, ld)
    acc = len({(e, f) for (e, f, _) in targets})
    hrr, _ = gen_hrr(la, lb, lc, ld)
    return acc > REG_MAX_ACC or len(hrr) > REG_MAX_HRR


def vrr_structure(la, lb, lc, ld, targets=None):
    """Structured VRR graph with the same recursion choices as gen_vrr, for
    the [e0|f0] targets of class (la lb|lc ld), or for the given (e, f, 0)
    targets."""
    E, Fk = la + lb, lc + ld
    if targets is None:
        targets = [(e, f, 0) for le in range(la, E + 1) for e in cart(le)
                   for lf in range(lc, Fk + 1) for f in cart(lf)]
    info, level = {}, {}

    def need(key):
        if key in info:
            return
        e, f, m = key
        if sum(f) > 0:
            i = pick_dir(f)
            f1 = sub1(f, i)
            src = [(e, f1, m), (e, f1, m + 1), None, None, None]
            n1 = f1[i]
            if n1 > 0:
                f2 = sub1(f1, i)
                src[2], src[3] = (e, f2, m), (e, f2, m + 1)
            n2 = e[i]
            if n2 > 0:
                src[4] = (sub1(e, i), f1, m + 1)
            info[key] = (2, i, n1, n2, src)
        elif sum(e) > 0:
            i = pick_dir(e)
            e1 = sub1(e, i)
            src = [(e1, f, m), (e1, f, m + 1), None, None, None]
            n1 = e1[i]
            if n1 > 0:
                e2 = sub1(e1, i)
                src[2], src[3] = (e2, f, m), (e2, f, m + 1)
            info[key] = (1, i, n1, 0, src)
        else:
            info[key] = (0, 0, m, 0, [None] * 5)
            level[key] = 0
            return
        for k in info[key][4]:
            if k is not None:
                need(k)
        level[key] = 1 + max(level[k] for k in info[key][4] if k is not None)

    for t in targets:
        need(t)
    keys = sorted(info, key=lambda k: (level[k], k))
    return keys, info, level, targets


def hrr_structure(la, lb, lc, ld, n_targets, tidx):
    """Level-ordered HRR entries over Y = [targets | bra HRR | ket HRR]."""
    Fk = lc + ld
    fs = [f for lf in range(lc, Fk + 1) for f in cart(lf)]
    yidx, entries = {}, []
    for key, t in tidx.items():
        yidx[("b",) + key] = t  # (e, f) of target -> ("b", a=e, b=0, f)
    # bra stage: (a, b, f)
    bra = {}

    def need_b(a, b, f):
        k = ("b", a, b, f)
        if k in bra or sum(b) == 0:
            return
        i = pick_dir(b)
        b1 = sub1(b, i)
        a1 = a[:i] + (a[i] + 1,) + a[i + 1:]
        need_b(a1, b1, f)
        need_b(a, b1, f)
        bra[k] = (("b", a1, b1, f), ("b", a, b1, f), i, sum(b))

    for f in fs:
        for a in cart(la):
            for b in cart(lb):
                need_b(a, b, f)
    ket = {}

    def need_k(a, b, c, d):
        k = ("k", a, b, c, d)
        if k in ket or sum(d) == 0:
            return
        i = pick_dir(d)
        d1 = sub1(d, i)
        c1 = c[:i] + (c[i] + 1,) + c[i + 1:]
        need_k(a, b, c1, d1)
        need_k(a, b, c, d1)
        ket[k] = (("k", a, b, c1, d1), ("k", a, b, c, d1), i, sum(d))

    for a in cart(la):
        for b in cart(lb):
            for c in cart(lc):
                for d in cart(ld):
                    need_k(a, b, c, d)

    def tgt_key(a, f):
        return ("b", a, tuple([0, 0, 0]) if False else (0, 0, 0), f)

    # map bra-stage keys with b = 0 to targets; ket-stage keys with d = 0 to bra stage
    def resolve(k):
        if k[0] == "b" and sum(k[2]) == 0:
            return yidx[("b", k[1], k[3])]
        if k[0] == "k" and sum(k[4]) == 0:
            bk = ("b", k[1], k[2], k[3])
            return resolve(bk) if sum(k[2]) == 0 else yidx[bk]
        return yidx[k]

    nxt = n_targets
    levels = []
    for stage, table in ((0, bra), (1, ket)):
        maxlev = max([v[3] for v in table.values()], default=0)
        for lev in range(1, maxlev + 1):
            start = len(entries)
            for k in sorted(kk for kk, v in table.items() if v[3] == lev):
                s1, s2, i, _ = table[k]
                entries.append((resolve(s1), resolve(s2), i, stage))
                yidx[k] = nxt
                nxt += 1
            levels.append((start, len(entries)))
    cart_y = []
    for a in cart(la):
        for b in cart(lb):
            for c in cart(lc):
                for d in cart(ld):
                    cart_y.append(resolve(("k", a, b, c, d)))
    return entries, levels, cart_y, nxt


def gen_wave_class(c, cartd=False):
    """cartd: the Cartesian-d form of the class, which shares the spherical
    form's VRR/HRR tables (generated before it) and differs in its AO count."""
    la, lb, lc, ld = c
    L = sum(c)
    name = cls_name(c)
    keys, info, level, targets = vrr_structure(la, lb, lc, ld)
    xi = {k: n for n, k in enumerate(keys)}
    nbase = sum(1 for k in keys if level[k] == 0)
    zero = len(keys)
    nlev = max(level.values())
    vrr_rows, lev_off = [], [0]
    for lev in range(1, nlev + 1):
        for k in keys:
            if level[k] != lev:
                continue
            kind, d, n1, n2, src = info[k]
            s = [xi[t] if t is not None else zero for t in src]
            assert xi[k] == nbase + len(vrr_rows)
            vrr_rows.append((s[0] | (s[1] << 16), s[2] | (s[3] << 16), s[4] | (kind << 16) | (d << 18) | (n1 << 20) | (n2 << 24)))
        lev_off.append(len(vrr_rows))
    base = [-1] * (L + 1)
    for k in keys:
        if level[k] == 0:
            base[k[2]] = xi[k]
    tlist = sorted({(e, f) for (e, f, _) in targets})
    tidx = {(e, f): n for n, (e, f) in enumerate(tlist)}
    tgt = [xi[(e, f, 0)] for (e, f) in tlist]
    hent, hlev, cart_y, ny = hrr_structure(la, lb, lc, ld, len(tlist), tidx)
    na = [nao(l, not (cartd and l >= 2)) for l in c]
    NAO = na[0] * na[1] * na[2] * na[3]
    nx = zero + 1
    # the staged AO transform writes its 1st and 3rd stages after the HRR
    # outputs (the 2nd and 4th go to the start of the table)
    st1 = ncart(la) * ncart(lb) * ncart(lc) * na[3]
    st3 = ncart(la) * na[1] * na[2] * na[3]
    lds = max(nx, ny + max(st1, st3), ny + NAO)
    o = []
    o.append("// ---- class (%s%s) : wave-cooperative, %d VRR entries in %d levels, %d HRR entries, LDS %d doubles"
             % (name, ", Cartesian d" if cartd else "", len(keys), nlev, len(hent), lds))
    o.append("template <> struct ClassInfo<%d,%d,%d,%d,%d> { static constexpr int L = %d, NAO = %d, NA = %d, NB = %d, NC = %d, ND = %d; };"
             % (la, lb, lc, ld, 1 if cartd else 0, L, NAO, *na))
    def arr(t, nm, vals):
        return "__device__ const %s %s[%d] = {%s};" % (t, nm, max(len(vals), 1), ",".join(vals) if vals else "0")
    if not cartd:
        o.append(arr("int", "wv_%s_vrr" % name, ["%d" % v for r in vrr_rows for v in r]))
        o.append(arr("int", "wv_%s_vlev" % name, ["%d" % v for v in lev_off]))
        o.append(arr("short", "wv_%s_base" % name, ["%d" % v for v in base]))
        o.append(arr("short", "wv_%s_tgt" % name, ["%d" % v for v in tgt]))
        o.append(arr("int", "wv_%s_hrr" % name, ["%d" % (s1 | (s2 << 16)) for (s1, s2, i, st) in hent] ))
        o.append(arr("char", "wv_%s_hdir" % name, ["%d" % (i | (st << 2)) for (s1, s2, i, st) in hent]))
        o.append(arr("int", "wv_%s_hlev" % name, ["%d" % v for lv in hlev for v in lv]))
        o.append(arr("short", "wv_%s_cart" % name, ["%d" % v for v in cart_y]))
    o.append("template <> struct WaveTab<%d,%d,%d,%d,%d> {" % (*c, 1 if cartd else 0))
    o.append("  static constexpr int L = %d, NBASE = %d, NVE = %d, ZERO = %d, NLEV = %d, NT = %d, NHE = %d, NHLEV = %d,"
             " NY = %d, NCART = %d, NAO = %d, LDS = %d;" % (L, nbase, len(vrr_rows), zero, nlev, len(tlist), len(hent),
                                                           len(hlev), ny, len(cart_y), NAO, lds))
    # level bounds as constants, so that a kernel can unroll its walk over
    # the levels and keep each lane's table entries in registers
    o.append("  static constexpr int VLEV[%d] = {%s};" % (len(lev_off), ",".join("%d" % v for v in lev_off)))
    o.append("  static constexpr short BASE[%d] = {%s};" % (len(base), ",".join("%d" % v for v in base)))
    o.append("  static constexpr int HLEV[%d] = {%s};" % (max(2 * len(hlev), 1),
                                                         ",".join("%d" % v for lv in hlev for v in lv) or "0"))
    for fld, nm in (("vrr", "int"), ("vlev", "int"), ("base", "short"), ("tgt", "short"), ("hrr", "int"),
                    ("hdir", "char"), ("hlev", "int"), ("cart", "short")):
        o.append("  static __device__ __forceinline__ const %s* %s() { return wv_%s_%s; }" % (nm, fld, name, fld))
    o.append("};")
    return "\n".join(o)


def gen_ao_patterns(maxl):
    cnt, idx = [], []
    for l in range(maxl + 1):
        for pat in ao_index_pattern(l, True):
            cnt.append(len(pat))
            idx.extend(pat + [0] * (3 - len(pat)))
    for l in range(2, maxl + 1):
        for pat in ao_index_pattern(l, False):
            cnt.append(len(pat))
            idx.extend(pat + [0] * (3 - len(pat)))
    cart_base = [(maxl + 1) ** 2 + sum(ncart(k) for k in range(2, l)) for l in range(maxl + 1)]
    xo = [xform_offset(l) for l in range(maxl + 1)]
    xc = [xform_offset_cart(l, maxl) if l >= 2 else xform_offset(l) for l in range(maxl + 1)]
    return ["// Cartesian components used by AO m of an l-shell: the real solid harmonics",
            "// (l = 0..maxl, m = -l..l), then the Cartesian functions of l = 2..maxl.",
            "__device__ const char ao_pat_n[%d] = {%s};" % (len(cnt), ",".join(map(str, cnt))),
            "__device__ const char ao_pat_c[%d] = {%s};" % (len(idx), ",".join(map(str, idx))),
            "// Pattern row and c_ao_xform block of AO m of an l-shell; C = 1: shells",
            "// with l >= 2 are Cartesian (x-major components, diagonal transform).",
            "__device__ __forceinline__ constexpr int ao_pat_row(int l, int m, int C) {",
            "  constexpr int cb[%d] = {%s};" % (maxl + 1, ",".join(map(str, cart_base))),
            "  return C && l >= 2 ? cb[l] + m : l * l + m;",
            "}",
            "__device__ __forceinline__ constexpr int xform_off(int l, int C = 0) {",
            "  constexpr int xs[%d] = {%s}, xc[%d] = {%s};" % (maxl + 1, ",".join(map(str, xo)), maxl + 1,
                                                              ",".join(map(str, xc))),
            "  return C ? xc[l] : xs[l];",
            "}"]


def gen_dispatch(cls_list):
    o = ["#define DSSCF_FOR_EACH_CLASS(X) \\"]
    for c in cls_list:
        o.append("  X(%d, %d, %d, %d, %s, %d) \\" % (*c, cls_name(c), 1 if is_wave(c) else 0))
    o.append("")
    o.append("// the classes with a d shell, which also have a Cartesian-d form")
    o.append("#define DSSCF_FOR_EACH_D_CLASS(X) \\")
    for c in cls_list:
        if has_d(c):
            o.append("  X(%d, %d, %d, %d, %s, %d) \\" % (*c, cls_name(c), 1 if is_wave(c) else 0))
    o.append("")
    return "\n".join(o)


def main(out_path, maxl=MAXL):
    pairs, cls_list = classes(maxl)
    text = ["// Generated by gen_eri.py -- do not edit.", "#pragma once", ""]
    text.append("#define DSSCF_MAXL %d" % maxl)
    text.append("#define DSSCF_XFORM_SIZE %d" % xform_size(maxl))
    text.append("#define DSSCF_NCLASS %d" % len(cls_list))
    text.append(gen_dispatch(cls_list))
    text.append("// Cartesian -> AO coefficients: per l = 0..maxl a (2l+1) x ncart(l) row-major block")
    text.append("// (the reference's real solid harmonics), then per l = 2..maxl an ncart(l) x ncart(l)")
    text.append("// block (Cartesian functions, diagonal).")
    text.append("static __constant__ double c_ao_xform[DSSCF_XFORM_SIZE];  // one copy per translation unit")
    text.append("")
    text.extend(gen_ao_patterns(maxl))
    text.append("")
    for c in cls_list:
        for cartd in ((False, True) if has_d(c) else (False,)):
            text.append(gen_wave_class(c, cartd) if is_wave(c) else gen_class(c, cartd))
            text.append("")
    open(out_path, "w").write("\n".join(text))


if __name__ == "__main__":
    main(sys.argv[1], int(sys.argv[2]) if len(sys.argv) > 2 else MAXL)
