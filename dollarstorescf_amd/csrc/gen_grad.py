"""Generator for the ERI nuclear-gradient device code (gfx950 HIP).

For every s/p shell-quartet class (la lb|lc ld), l <= 1, it emits
shell_grad_<class>(): the derivatives of one contracted shell quartet with
respect to the centres A, B and C, contracted with a Cartesian pair-density
block Gc, in three passes (one per centre, so that only one centre's
accumulators are live at a time):

  d/dA_x (ab|cd) = 2 alpha (a+1x b|cd) - a_x (a-1x b|cd)      (alpha: A's primitive exponent)

and likewise for B (2 beta, b+-1x) and C (2 gamma, c+-1x); D follows from
translational invariance in the kernel.  Per primitive quartet the
Head-Gordon-Pople vertical recursion gives [e0|f0] for the leaves that the
raised and lowered horizontal transfers need; the raised set accumulates
with weight 2 x exponent, the lowered one with weight 1.  This is the
derivative that DollarStoreSCF's electronRepulsionGradient /
OSrepulsionGradient (GaussianIntegral.cpp:1052-1459) evaluates per Cartesian
term and axis, reorganised into shell-quartet blocks.

Output: one .inc file included by grad_kernels.hip.
"""
from __future__ import annotations

import sys

from gen_eri import XYZ, cart, ncart, pick_dir, sub1


def add1(t, i):
    return t[:i] + (t[i] + 1,) + t[i + 1:]


def tname(t):
    return "%d%d%d" % t


def vrr_lines(targets):
    """[e0|f0]^(0) for (e, f) in targets from Fm[] (gen_eri.gen_vrr's recursion)."""
    order, seen, stmts = [], set(), {}

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

    for (e, f) in sorted(targets):
        need((e, f, 0))
    out = []
    for key in order:
        e, f, m = key
        name = "v%s_%s_%d" % (tname(e), tname(f), m)
        if stmts[key] is None:
            out.append("const double %s = Fm[%d];" % (name, m))
        else:
            expr = " + ".join("(%s)*v%s_%s_%d" % (c, tname(k[0]), tname(k[1]), k[2]) for c, k in stmts[key])
            out.append("const double %s = %s;" % (name, expr))
    return out


def hrr_leaves(pairs):
    """Leaves (e, with b = 0) and statement order of the transfer
    (a, b+1_i) = (a+1_i, b) + AB_i (a, b) for the given (a, b) targets."""
    order, seen, stmt = [], set(), {}

    def need(a, b):
        if (a, b) in seen:
            return
        seen.add((a, b))
        if sum(b) == 0:
            stmt[(a, b)] = None
        else:
            i = pick_dir(b)
            b1 = sub1(b, i)
            a1 = add1(a, i)
            need(a1, b1)
            need(a, b1)
            stmt[(a, b)] = (a1, b1, i)
        order.append((a, b))

    for a, b in pairs:
        need(a, b)
    leaves = sorted({a for (a, b) in order if stmt[(a, b)] is None})
    return order, stmt, leaves


class Side:
    """One accumulator set: its prefix, weight, bra (a, b) and ket (c, d) targets."""

    def __init__(self, prefix, weight, bra, ket):
        self.prefix, self.weight = prefix, weight
        self.bra, self.ket = sorted(set(bra)), sorted(set(ket))
        self.b_order, self.b_stmt, self.E = hrr_leaves(self.bra)
        self.k_order, self.k_stmt, self.F = hrr_leaves(self.ket)

    def acc(self, e, f):
        return "%s_%s_%s" % (self.prefix, tname(e), tname(f))

    def hrr(self, ab_name, cd_name):
        """Statements giving <prefix>c_<a>_<b>_<c>_<d> for every bra x ket target."""
        lines = []
        P = self.prefix

        def bn(a, b, f):
            return "%sh_%s_%s_%s" % (P, tname(a), tname(b), tname(f))

        for f in self.F:
            for (a, b) in self.b_order:
                st = self.b_stmt[(a, b)]
                if st is None:
                    lines.append("const double %s = %s;" % (bn(a, b, f), self.acc(a, f)))
                else:
                    a1, b1, i = st
                    lines.append("const double %s = %s + %s%s*%s;" % (bn(a, b, f), bn(a1, b1, f), ab_name, XYZ[i],
                                                                        bn(a, b1, f)))

        def kn(a, b, c, d):
            return "%sc_%s_%s_%s_%s" % (P, tname(a), tname(b), tname(c), tname(d))

        for (a, b) in self.bra:
            for (c, d) in self.k_order:
                st = self.k_stmt[(c, d)]
                if st is None:
                    lines.append("const double %s = %s;" % (kn(a, b, c, d), bn(a, b, c)))
                else:
                    c1, d1, i = st
                    lines.append("const double %s = %s + %s%s*%s;" % (kn(a, b, c, d), kn(a, b, c1, d1), cd_name,
                                                                        XYZ[i], kn(a, b, c, d1)))
        return lines, kn


def gen_class(c):
    la, lb, lc, ld = c
    A, B, C, D = cart(la), cart(lb), cart(lc), cart(ld)
    allbra = [(a, b) for a in A for b in B]
    allket = [(x, y) for x in C for y in D]
    name = "".join("spdf"[l] for l in c)
    L = la + lb + lc + ld + 1
    o = []
    o.append("// ---- gradient of class (%s)" % name)
    o.append("template <int BM>")
    o.append("__device__ __forceinline__ void shell_grad_%s(const ShellPair& bra, const ShellPair& ket,"
             " const PrimPair* __restrict__ bra_pp, const PrimPair* __restrict__ ket_pp,"
             " const double* __restrict__ btab, const double* __restrict__ Gc, double* __restrict__ g9) {" % name)
    o.append("  const double ABx = bra.AB[0], ABy = bra.AB[1], ABz = bra.AB[2];")
    o.append("  const double CDx = ket.AB[0], CDy = ket.AB[1], CDz = ket.AB[2];")
    o.append("  (void)ABx; (void)ABy; (void)ABz; (void)CDx; (void)CDy; (void)CDz;")

    def gidx(a, b, x, y):
        return ((A.index(a) * len(B) + B.index(b)) * len(C) + C.index(x)) * len(D) + D.index(y)

    passes = []
    # centre A: the bra's first shell
    up = [(add1(a, i), b) for a in A for b in B for i in range(3)]
    dn = [(sub1(a, i), b) for a in A for b in B for i in range(3) if a[i] > 0]
    passes.append(("A", "2.0 * P.alpha", Side("ra", None, up, allket), Side("la", None, dn, allket) if dn else None))
    up = [(a, add1(b, i)) for a in A for b in B for i in range(3)]
    dn = [(a, sub1(b, i)) for a in A for b in B for i in range(3) if b[i] > 0]
    passes.append(("B", "2.0 * (P.p - P.alpha)", Side("rb", None, up, allket),
                   Side("lb", None, dn, allket) if dn else None))
    up = [(add1(x, i), y) for x in C for y in D for i in range(3)]
    dn = [(sub1(x, i), y) for x in C for y in D for i in range(3) if x[i] > 0]
    passes.append(("C", "2.0 * Q.alpha", Side("rc", None, allbra, up), Side("lc", None, allbra, dn) if dn else None))

    for k, (cen, weight, rs, ls) in enumerate(passes):
        sides = [s for s in (rs, ls) if s is not None]
        o.append("  {  // centre %s" % cen)
        targets = set()
        for s in sides:
            for e in s.E:
                for f in s.F:
                    targets.add((e, f))
                    o.append("    double %s = 0.0;" % s.acc(e, f))
        o.append("    for (int ik = 0; ik < ket.nprim; ++ik) {")
        o.append("      const PrimPair Q = ldg(ket_pp + ik);")
        o.append("      for (int ib = 0; ib < bra.nprim; ++ib) {")
        o.append("        const PrimPair P = bra_pp[ib];  // the block's LDS copy (accurate mode) or global")
        o.append("        double Fm[%d];" % (L + 1))
        o.append("        PrimQuartet<double> g; prim_setup<BM, double, %d>(P, Q, bra, ket, btab, g, Fm);" % L)
        o.append("        const double PAx = g.PA[0], PAy = g.PA[1], PAz = g.PA[2];")
        o.append("        const double WPx = g.WP[0], WPy = g.WP[1], WPz = g.WP[2];")
        o.append("        const double QCx = g.QC[0], QCy = g.QC[1], QCz = g.QC[2];")
        o.append("        const double WQx = g.WQ[0], WQy = g.WQ[1], WQz = g.WQ[2];")
        o.append("        const double o2p = g.o2p, o2q = g.o2q, o2pq = g.o2pq, rp = g.rp, rq = g.rq;")
        o.append("        (void)PAx; (void)PAy; (void)PAz; (void)WPx; (void)WPy; (void)WPz; (void)QCx; (void)QCy;")
        o.append("        (void)QCz; (void)WQx; (void)WQy; (void)WQz; (void)o2p; (void)o2q; (void)o2pq; (void)rp; (void)rq;")
        for ln in vrr_lines(targets):
            o.append("        " + ln)
        o.append("        const double w = %s;" % weight)
        for e in rs.E:
            for f in rs.F:
                o.append("        %s += w * v%s_%s_0;" % (rs.acc(e, f), tname(e), tname(f)))
        if ls is not None:
            for e in ls.E:
                for f in ls.F:
                    o.append("        %s += v%s_%s_0;" % (ls.acc(e, f), tname(e), tname(f)))
        o.append("      }")
        o.append("    }")
        rl, rkn = rs.hrr("AB", "CD")
        o += ["    " + ln for ln in rl]
        if ls is not None:
            ll, lkn = ls.hrr("AB", "CD")
            o += ["    " + ln for ln in ll]
        for i in range(3):
            terms = []
            for a in A:
                for b in B:
                    for x in C:
                        for y in D:
                            gi = "Gc[%d]" % gidx(a, b, x, y)
                            if cen == "A":
                                up_t = rkn(add1(a, i), b, x, y)
                                dn_t = lkn(sub1(a, i), b, x, y) if a[i] > 0 else None
                                n_i = a[i]
                            elif cen == "B":
                                up_t = rkn(a, add1(b, i), x, y)
                                dn_t = lkn(a, sub1(b, i), x, y) if b[i] > 0 else None
                                n_i = b[i]
                            else:
                                up_t = rkn(a, b, add1(x, i), y)
                                dn_t = lkn(a, b, sub1(x, i), y) if x[i] > 0 else None
                                n_i = x[i]
                            t = up_t if dn_t is None else "(%s - %d.0 * %s)" % (up_t, n_i, dn_t)
                            terms.append("%s * %s" % (gi, t))
            o.append("    g9[%d] = %s;" % (3 * k + i, " + ".join(terms)))
        o.append("  }")
    o.append("}")
    return "\n".join(o)


def gen_density_xform(c):
    """Cartesian pair-density block from the AO block (transpose of the AO
    transform): Gc[((ca*NCB+cb)*NCC+cc)*NCD+cd] = sum_m xf..."""
    la, lb, lc, ld = c
    name = "".join("spdf"[l] for l in c)
    nc = [ncart(l) for l in c]
    na = [2 * l + 1 for l in c]

    def off(l):
        return sum((2 * k + 1) * ncart(k) for k in range(l))

    o = ["// AO-block -> Cartesian-block pair density of class (%s) (transpose of the AO transform)" % name]
    o.append("__device__ __forceinline__ void grad_density_%s(const double* __restrict__ xf,"
             " const double* __restrict__ Gs, double* __restrict__ Gc) {" % name)
    o.append("  for (int i = 0; i < %d; ++i) Gc[i] = 0.0;" % (nc[0] * nc[1] * nc[2] * nc[3]))
    o.append("  for (int ma = 0; ma < %d; ++ma) for (int mb = 0; mb < %d; ++mb)" % (na[0], na[1]))
    o.append("  for (int mc = 0; mc < %d; ++mc) for (int md = 0; md < %d; ++md) {" % (na[2], na[3]))
    o.append("    const double g = Gs[((ma * %d + mb) * %d + mc) * %d + md];" % (na[1], na[2], na[3]))
    o.append("    for (int ca = 0; ca < %d; ++ca) { const double wa = g * xf[%d + ma * %d + ca]; if (wa == 0.0) continue;"
             % (nc[0], off(la), nc[0]))
    o.append("    for (int cb = 0; cb < %d; ++cb) { const double wb = wa * xf[%d + mb * %d + cb]; if (wb == 0.0) continue;"
             % (nc[1], off(lb), nc[1]))
    o.append("    for (int cc = 0; cc < %d; ++cc) { const double wc = wb * xf[%d + mc * %d + cc]; if (wc == 0.0) continue;"
             % (nc[2], off(lc), nc[2]))
    o.append("    for (int cd = 0; cd < %d; ++cd) Gc[((ca * %d + cb) * %d + cc) * %d + cd] += wc * xf[%d + md * %d + cd];"
             % (nc[3], nc[1], nc[2], nc[3], off(ld), nc[3]))
    o.append("    } } }")
    o.append("  }")
    o.append("}")
    return "\n".join(o)


GRAD_CLASSES = [(0, 0, 0, 0), (1, 0, 0, 0), (1, 0, 1, 0), (1, 1, 0, 0), (1, 1, 1, 0), (1, 1, 1, 1)]


# ------------------------------------------------------------------------
# Wave-cooperative gradient classes (every class with a d shell).
#
# The derivative integrals of centre X are HRR transfers of [e0|f0] sums,
# so each class needs, per primitive quartet, the VRR targets of four
# accumulator kinds:
#   kind 0 (weight 1)      the lowered sets (a-1x b|, (a b-1x|, |c-1x d)
#   kind 1 (weight 2 alpha) the A-raised set (a+1x b|
#   kind 2 (weight 2 beta)  the B-raised set (a b+1x|
#   kind 3 (weight 2 gamma) the C-raised set |c+1x d)
# Each derivative set is then one HRR over its kind's accumulators to the
# Cartesian block of its shifted class, contracted with the Cartesian pair
# density Gc into the nine derivatives of A, B and C (D follows from
# translational invariance):
#   dA_x: sum Gc[a b c d] ((a+1x b|cd)_2alpha - a_x (a-1x b|cd)),  likewise B, C.
# The tables below drive one generic kernel (k_gradw, grad_kernels.hip).

WAVE_LDS_BUDGET = 65536  # bytes of LDS tables per block up to which G > 1 quartets share a wavefront


def lrange(lo, hi):
    return [t for l in range(max(lo, 0), hi + 1) for t in cart(l)]


def grad_sets(c):
    """(name, kind, shifted class, centre, +1 raised / -1 lowered)."""
    la, lb, lc, ld = c
    s = [("A+", 1, (la + 1, lb, lc, ld), 0, 1)]
    if la > 0:
        s.append(("A-", 0, (la - 1, lb, lc, ld), 0, -1))
    s.append(("B+", 2, (la, lb + 1, lc, ld), 1, 1))
    if lb > 0:
        s.append(("B-", 0, (la, lb - 1, lc, ld), 1, -1))
    s.append(("C+", 3, (la, lb, lc + 1, ld), 2, 1))
    if lc > 0:
        s.append(("C-", 0, (la, lb, lc - 1, ld), 2, -1))
    return s


def grad_wave_class(c):
    """Tables of one wave-cooperative gradient class (shared by its
    spherical and Cartesian-d forms: everything here is Cartesian)."""
    import gen_eri
    la, lb, lc, ld = c
    name = "".join("spdf"[l] for l in c)
    L1 = la + lb + lc + ld + 1
    sets = grad_sets(c)
    kinds = [set() for _ in range(4)]
    for _, k, (a, b, cc, d), _, _ in sets:
        kinds[k] |= {(e, f) for e in lrange(a, a + b) for f in lrange(cc, cc + d)}
    klist = [sorted(k) for k in kinds]
    alltg = sorted({(e, f, 0) for k in klist for (e, f) in k})
    keys, info, level, _ = gen_eri.vrr_structure(la, lb, lc, ld, targets=alltg)
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
            s_ = [xi[t] if t is not None else zero for t in src]
            assert xi[k] == nbase + len(vrr_rows)
            vrr_rows.append((s_[0] | (s_[1] << 16), s_[2] | (s_[3] << 16),
                             s_[4] | (kind << 16) | (d << 18) | (n1 << 20) | (n2 << 24)))
        lev_off.append(len(vrr_rows))
    base = [-1] * (L1 + 1)
    for k in keys:
        if level[k] == 0:
            base[k[2]] = xi[k]
    nx = zero + 1
    nc4 = ncart(la) * ncart(lb) * ncart(lc) * ncart(ld)

    def gidx(a, b, x, y):
        return ((cart(la).index(a) * ncart(lb) + cart(lb).index(b)) * ncart(lc) + cart(lc).index(x)) * ncart(ld) + \
            cart(ld).index(y)

    # HRR and contraction of each set
    hrr_rows, hrr_dir, hlev, set_rows = [], [], [], []
    ce = []  # (set, comp, y, g, coef)
    for si, (nm, k, (a_, b_, c_, d_), cen, sign) in enumerate(sets):
        tidx = {ef: n for n, ef in enumerate(klist[k])}
        ent, levs, cart_y, ny = gen_eri.hrr_structure(a_, b_, c_, d_, len(klist[k]), tidx)
        h0, l0 = len(hrr_rows), len(hlev)
        for (s1, s2, i, st) in ent:
            hrr_rows.append(s1 | (s2 << 16))
            hrr_dir.append(i | (st << 2))
        for (e0, e1) in levs:
            hlev.append((h0 + e0, h0 + e1))
        set_rows.append((k, len(klist[k]), l0, len(hlev)))
        outs = [(A, B, C, D) for A in cart(a_) for B in cart(b_) for C in cart(c_) for D in cart(d_)]
        for (A, B, C, D), y in zip(outs, cart_y):
            shifted = (A, B, C)[cen]
            for x in range(3):
                if sign > 0:
                    if shifted[x] == 0:
                        continue
                    orig = sub1(shifted, x)
                    coef = 1
                else:
                    orig = add1(shifted, x)
                    coef = -orig[x]
                q = [A, B, C]
                q[cen] = orig
                ce.append((si, 3 * cen + x, y, gidx(q[0], q[1], q[2], D), coef))
    ce.sort(key=lambda t: (t[0], t[1]))
    cstart = []
    for si in range(len(sets)):
        for comp in range(10):
            cstart.append(next((n for n, t in enumerate(ce) if (t[0], t[1]) >= (si, comp)), len(ce)))
    nacc = [len(k) for k in klist]
    # LDS per quartet group: the Cartesian pair density, then a region used
    # in turn for the density staging (two Cartesian blocks), the VRR table
    # and each set's leaves + HRR entries
    G = 8
    while True:
        LG = 64 // G
        pad = [(n + LG - 1) // LG * LG for n in nacc]
        ysz = 0
        for si, (k, n, l0, l1) in enumerate(set_rows):
            nent = sum(hlev[j][1] - hlev[j][0] for j in range(l0, l1))
            ysz = max(ysz, pad[k] + nent)
        region = max(nx, ysz, 2 * nc4)
        lds = nc4 + region
        if G == 1 or 8 * G * lds <= WAVE_LDS_BUDGET:
            break
        G //= 2
    ks = [0]
    for k in range(4):
        ks.append(ks[-1] + pad[k] // LG)
    tgt = []
    for k in range(4):
        lst = [xi[(e, f, 0)] for (e, f) in klist[k]]
        tgt += lst + [zero] * (pad[k] - len(lst))
    o = []
    o.append("// ---- gradient class (%s): wave-cooperative, %d VRR entries (L = %d), accumulators %s, "
             "%d HRR entries, %d contraction terms, G = %d, LDS %d doubles per quartet"
             % (name, len(keys), L1, "/".join(str(n) for n in nacc), len(hrr_rows), len(ce), G, lds))

    def arr(t, nm, vals):
        return "__device__ const %s %s[%d] = {%s};" % (t, nm, max(len(vals), 1), ",".join(map(str, vals)) or "0")
    o.append(arr("int", "gw_%s_vrr" % name, [v for r in vrr_rows for v in r]))
    o.append(arr("int", "gw_%s_vlev" % name, lev_off))
    o.append(arr("short", "gw_%s_base" % name, base))
    o.append(arr("short", "gw_%s_tgt" % name, tgt))
    o.append(arr("int", "gw_%s_hrr" % name, hrr_rows))
    o.append(arr("char", "gw_%s_hdir" % name, hrr_dir))
    o.append(arr("int", "gw_%s_hlev" % name, [v for lv in hlev for v in lv]))
    o.append(arr("int", "gw_%s_set" % name, [v for r in set_rows for v in r]))
    o.append(arr("int", "gw_%s_cstart" % name, cstart))
    o.append(arr("short", "gw_%s_cey" % name, [t[2] for t in ce]))
    o.append(arr("short", "gw_%s_ceg" % name, [t[3] for t in ce]))
    o.append(arr("signed char", "gw_%s_cec" % name, [t[4] for t in ce]))
    o.append("template <> struct GradWaveTab<%d,%d,%d,%d> {" % c)
    o.append("  static constexpr int L1 = %d, G = %d, LG = %d, NBASE = %d, NVE = %d, ZERO = %d, NLEV = %d, NX = %d,"
             % (L1, G, LG, nbase, len(vrr_rows), zero, nlev, nx))
    o.append("      NSET = %d, NC4 = %d, REGION = %d, LDS = %d, SLOTS = %d, KS1 = %d, KS2 = %d, KS3 = %d;"
             % (len(sets), nc4, region, lds, ks[4], ks[1], ks[2], ks[3]))
    o.append("  static constexpr int ks(int k) { return k == 0 ? 0 : k == 1 ? KS1 : k == 2 ? KS2 : k == 3 ? KS3 : SLOTS; }")
    for fld, t in (("vrr", "int"), ("vlev", "int"), ("base", "short"), ("tgt", "short"), ("hrr", "int"),
                   ("hdir", "char"), ("hlev", "int"), ("set", "int"), ("cstart", "int"), ("cey", "short"),
                   ("ceg", "short"), ("cec", "signed char")):
        o.append("  static __device__ __forceinline__ const %s* %s() { return gw_%s_%s; }" % (t, fld, name, fld))
    o.append("};")
    return "\n".join(o)


# s/p classes that would also run wave-cooperative: none.  Measured on the
# 8-water/6-31G gradient, the wave form was slower for every s/p class it was
# tried on, (pp|pp) included although its register form spills: (ps|ps) 350
# -> 1640 us, (pp|ss) 202 -> 772, (pp|ps) 462 -> 1035, (pp|pp) 1709 -> 2423.
GRAD_WAVE_SP = []


def grad_wave_classes(maxl=2):
    import gen_eri
    _, cls_list = gen_eri.classes(maxl)
    return [c for c in cls_list if max(c) >= 2 or c in GRAD_WAVE_SP]


def main():
    import gen_eri
    out = sys.argv[1]
    text = ["// Generated by gen_grad.py -- do not edit.", "#pragma once"]
    for c in GRAD_CLASSES:
        text.append(gen_density_xform(c))
        text.append(gen_class(c))
    text.append("#define DSSCF_FOR_EACH_GRAD_CLASS(X) \\")
    text.append(" \\\n".join("    X(%d, %d, %d, %d, %s)" % (*c, "".join("spdf"[l] for l in c)) for c in GRAD_CLASSES))
    # AO transform offsets (as the ERI classes' xform_off) for the wave classes' density transform
    xs = [gen_eri.xform_offset(l) for l in range(gen_eri.MAXL + 1)]
    xc = [gen_eri.xform_offset_cart(l) if l >= 2 else gen_eri.xform_offset(l) for l in range(gen_eri.MAXL + 1)]
    text.append("__device__ __forceinline__ constexpr int grad_xform_off(int l, int C) {")
    text.append("  constexpr int xs[%d] = {%s}, xc[%d] = {%s};" % (len(xs), ",".join(map(str, xs)), len(xc),
                                                                 ",".join(map(str, xc))))
    text.append("  return C && l >= 2 ? xc[l] : xs[l];")
    text.append("}")
    text.append("template <int la, int lb, int lc, int ld> struct GradWaveTab;")
    wc = grad_wave_classes()
    for c in wc:
        text.append(grad_wave_class(c))
    text.append("#define DSSCF_FOR_EACH_GRAD_WAVE_CLASS(X) \\")
    text.append(" \\\n".join("    X(%d, %d, %d, %d, %s)" % (*c, "".join("spdf"[l] for l in c)) for c in wc))
    with open(out, "w") as f:
        f.write("\n".join(text) + "\n")


if __name__ == "__main__":
    main()
