"""Generates md_gen.h: straight-line McMurchie-Davidson code per (la, lb).

For every shell-pair class la >= lb (l <= 3) it emits

  md_prim_<la><lb>_oc / _tc(dx, dy, dz, R0, E, V)  (_p0 / _p1: the two parts
      of the A components -- without z, with z -- of a split class, SPLIT)
      one primitive pair at one grid point: the pruned Hermite Coulomb
      recursion R^0_tuv (t+u+v <= L, L = la+lb) from R0[n] = c_n F_n(T),
      then V[a][b] += sum_tuv Ex[ax][bx][t] Ey[ay][by][u] Ez[az][bz][v] R_tuv
      factorized as z -> y -> x contractions.  E is the primitive pair's
      coefficient record (uniform across the wave, read through the scalar
      cache); V (of type vacc, which may be wider than the arithmetic) is the
      Cartesian block accumulated over the contraction.  _oc is the
      one-centre variant (A = B = P), _tc the two-centre one; they differ in
      which E products are compensated in FP32 and in the one-centre zeros.
  md_fcart_<l>(Fs, Fc)      Fc[c] = sum_m T[m][c] Fs[m]   (sph -> cart)
  md_gsph_<l>(Gc, Gs)       Gs[m] = sum_c T[m][c] Gc[c]   (cart -> sph)
      (in FP32 every non-trivial constant product is compensated with the
      constant's rounding remainder, wb_cm)

with every index a compile-time constant, so all arrays live in registers.
E layout per axis: ((i * (lb+1) + j) * (L+1) + t); axes x, y, z follow.
The math follows the reference's MD recurrences (gpu/shellpairs.py:76-113,
447-485) and its Cartesian order (lx, then ly descending).
"""
import os
import sys

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), '..'))
from pyscf_wb97mv_fast_amd.cart2sph import cart2sph, cart_components  # noqa: E402

LMAX = 3


def _lit(v):
    return repr(float(v))


def gen_r(L):
    """Lines computing R_t_u_v (n = 0) for t+u+v <= L from R0[n]."""
    lines = []
    done = {}

    def name(n, t, u, v):
        return 'r%d_%d_%d_%d' % (n, t, u, v)

    def need(n, t, u, v):
        key = (n, t, u, v)
        if key in done:
            return done[key]
        if t == u == v == 0:
            done[key] = 'R0[%d]' % n
            return done[key]
        if t > 0:
            a = need(n + 1, t - 1, u, v)
            expr = 'dx * %s' % a
            if t > 1:
                b = need(n + 1, t - 2, u, v)
                expr = 'fma(dx, %s, (real)%d * %s)' % (a, t - 1, b)
        elif u > 0:
            a = need(n + 1, t, u - 1, v)
            expr = 'dy * %s' % a
            if u > 1:
                b = need(n + 1, t, u - 2, v)
                expr = 'fma(dy, %s, (real)%d * %s)' % (a, u - 1, b)
        else:
            a = need(n + 1, t, u, v - 1)
            expr = 'dz * %s' % a
            if v > 1:
                b = need(n + 1, t, u, v - 2)
                expr = 'fma(dz, %s, (real)%d * %s)' % (a, v - 1, b)
        nm = name(n, t, u, v)
        lines.append('    const real %s = %s;' % (nm, expr))
        done[key] = nm
        return nm

    out = {}
    for t in range(L + 1):
        for u in range(L + 1 - t):
            for v in range(L + 1 - t - u):
                out[(t, u, v)] = need(0, t, u, v)
    return lines, out


# classes whose Cartesian block is formed in two parts of the A components,
# one pass of the primitive loop each (fewer block registers): the
# components without z (az = 0) and those with z, so that the two parts
# share none of the z-contracted intermediates.  Measured on MI355X: (d, d)
# 249 -> 165 VGPRs and (f, d) 308 -> 222 registers run faster split; (f, f)
# stays at one wave per SIMD either way and is slower split (17.5 -> 20.3 ms
# on a 1642-AO def2-TZVP molecule), (f, p) keeps its occupancy.
SPLIT = {(2, 2), (3, 2)}


def a_part(la, part):
    """indices of the A Cartesian components of a part"""
    ca = cart_components(la)
    return [i for i, a in enumerate(ca) if (a[2] > 0) == (part == 1)]


def gen_prim(la, lb, oc, part=None):
    """oc: one-centre variant (P = A = B): E^{ij}_t vanishes unless i+j-t is
    even, and every remaining E is a function of p alone -- the same for
    every copy of an element -- so each FP32 product carries the E rounding
    remainder.  Two-centre variant: only E^{ij}_{i+j} = (2p)^-(i+j) is
    geometry independent; those products are compensated, the others (which
    depend on the pair's orientation) are plain."""
    L = la + lb
    ca, cb = cart_components(la), cart_components(lb)
    if part is not None:
        ca = [ca[i] for i in a_part(la, part)]
    nb = len(cb)
    naxis = (la + 1) * (lb + 1) * (L + 1)
    erl = 3 * naxis

    def eidx(axis, i, j, t):
        return axis * naxis + (i * (lb + 1) + j) * (L + 1) + t

    lines = ['template <typename real, typename vacc, typename EP>',
             '__device__ __forceinline__ void md_prim_%d%d_%s%s(real dx, real dy, real dz,'
             % (la, lb, 'oc' if oc else 'tc', '' if part is None else '_p%d' % part),
             '    const real* __restrict__ R0, EP E, vacc* __restrict__ V)',
             '{']
    rl, R = gen_r(L)
    if part is not None:
        # only the R values this half's contractions read
        used = set()
        for a in ca:
            for b in cb:
                for t in range(a[0] + b[0] + 1):
                    for u in range(a[1] + b[1] + 1):
                        for v in range(a[2] + b[2] + 1):
                            used.add((t, u, v))
        rl = _prune_r(rl, [R[k] for k in used])
    lines += rl

    def zero(i, j, t):
        return oc and (i + j - t) % 2 == 1

    def comp(i, j, t):
        return oc or t == i + j

    def emul(axis, i, j, t, x):
        if i == 0 and j == 0 and t == 0:
            return x                      # E^{00}_0 = 1
        fn = 'wb_emulc' if comp(i, j, t) else 'wb_emul'
        return '%s<real, %d>(E, %d, %s)' % (fn, erl, eidx(axis, i, j, t), x)

    def efma(axis, i, j, t, x, acc):
        if i == 0 and j == 0 and t == 0:
            return '%s + %s' % (acc, x)
        fn = 'wb_efmac' if comp(i, j, t) else 'wb_efma'
        return '%s<real, %d>(E, %d, %s, %s)' % (fn, erl, eidx(axis, i, j, t), x, acc)

    def chain(axis, i, j, terms):
        """sum over (index, value) of E^{ij}_index * value, zeros skipped"""
        terms = [(t, x) for t, x in terms if not zero(i, j, t)]
        expr = emul(axis, i, j, terms[0][0], terms[0][1])
        for t, x in terms[1:]:
            expr = efma(axis, i, j, t, x, expr)
        return expr

    zpairs = sorted({(a[2], b[2]) for a in ca for b in cb})
    sz = {}
    for (az, bz) in zpairs:
        rem = L - az - bz
        for t in range(rem + 1):
            for u in range(rem + 1 - t):
                nm = 'sz_%d%d_%d_%d' % (az, bz, t, u)
                expr = chain(2, az, bz, [(v, R[(t, u, v)]) for v in range(az + bz + 1)])
                lines.append('    const real %s = %s;' % (nm, expr))
                sz[(az, bz, t, u)] = nm
    yz = {}
    yzpairs = sorted({(a[1], b[1], a[2], b[2]) for a in ca for b in cb})
    for (ay, by, az, bz) in yzpairs:
        tmax = L - ay - by - az - bz
        for t in range(tmax + 1):
            nm = 'syz_%d%d%d%d_%d' % (ay, by, az, bz, t)
            expr = chain(1, ay, by, [(u, sz[(az, bz, t, u)]) for u in range(ay + by + 1)])
            lines.append('    const real %s = %s;' % (nm, expr))
            yz[(ay, by, az, bz, t)] = nm
    # this primitive pair's block is formed on its own and then added to the
    # running sum V (type vacc): started from V, the FMA chain would round
    # every small primitive contribution against the large running sum
    for i, a in enumerate(ca):
        for j, b in enumerate(cb):
            ax, bx = a[0], b[0]
            expr = chain(0, ax, bx, [(t, yz[(a[1], b[1], a[2], b[2], t)])
                                     for t in range(ax + bx + 1)])
            lines.append('    V[%d] += (vacc)(%s);' % (i * nb + j, expr))   # i: within this half
    lines.append('}')
    return lines


def _prune_r(lines, roots):
    """The R recursion lines that the given R values depend on."""
    defs = {}
    for ln in lines:
        nm = ln.split('const real ')[1].split(' =')[0]
        defs[nm] = ln
    need, stack = set(), [r for r in roots if r in defs]
    while stack:
        nm = stack.pop()
        if nm in need:
            continue
        need.add(nm)
        rhs = defs[nm].split('=', 1)[1]
        for tok in rhs.replace('(', ' ').replace(')', ' ').replace(',', ' ').replace(';', ' ').split():
            if tok in defs and tok not in need:
                stack.append(tok)
    return [ln for ln in lines if ln.split('const real ')[1].split(' =')[0] in need]


def _cm(v, x):
    """Constant product v * x: in FP32 as fma(v_hi, x, v_lo * x) with the
    FP32 rounding remainder v_lo, so the constant's rounding (the same at
    every grid point) does not bias the result; plain in FP64."""
    if float(v) in (1.0, -1.0, 2.0, -2.0, 3.0, -3.0, 4.0, -4.0):
        return '(real)%s * %s' % (_lit(v), x)
    import numpy
    hi = float(numpy.float32(v))
    lo = float(numpy.float32(v - hi))
    return 'wb_cm<real>(%sf, %sf, %s, %s)' % (repr(hi), repr(lo), _lit(v), x)


def gen_transforms(l):
    T = cart2sph(l)
    nsph, nc = T.shape
    lines = ['template <typename real>',
             '__device__ __forceinline__ void md_fcart_%d(const real* __restrict__ Fs,'
             ' real* __restrict__ Fc)' % l, '{']
    for c in range(nc):
        terms = [(m, T[m, c]) for m in range(nsph) if T[m, c] != 0.0]
        expr = ' + '.join(_cm(v, 'Fs[%d]' % m) for m, v in terms) or '(real)0'
        lines.append('    Fc[%d] = %s;' % (c, expr))
    lines.append('}')
    lines += ['template <typename real>',
              '__device__ __forceinline__ void md_gsph_%d(const real* __restrict__ Gc,'
              ' real* __restrict__ Gs)' % l, '{']
    for m in range(nsph):
        terms = [(c, T[m, c]) for c in range(nc) if T[m, c] != 0.0]
        expr = ' + '.join(_cm(v, 'Gc[%d]' % c) for c, v in terms)
        lines.append('    Gs[%d] = %s;' % (m, expr))
    lines.append('}')
    return lines


def main(out_path):
    lines = ['// generated by gen_md.py -- do not edit',
             '#pragma once',
             '// c * x for a constant c = hi + lo (FP32) or d (FP64)',
             'template <typename real>',
             '__device__ __forceinline__ real wb_cm(float hi, float lo, double d, real x);',
             'template <> __device__ __forceinline__ float wb_cm<float>(float hi, float lo,'
             ' double, float x) { return fmaf(hi, x, lo * x); }',
             'template <> __device__ __forceinline__ double wb_cm<double>(float, float,'
             ' double d, double x) { return d * x; }',
             '// E-coefficient products.  The FP32 record holds the ERL leading values',
             '// and then their rounding remainders: E values of one-centre pairs are',
             '// the same for every copy of an element, so their rounding would bias',
             '// K coherently; each product is formed with one rounding.  EP: pointer',
             '// to the record\'s E values, in whatever address space the caller reads',
             '// them from.',
             'template <typename real, int ERL, typename EP>',
             '__device__ __forceinline__ real wb_efmac(EP E, int k, real x,'
             ' real acc) {',
             '    if constexpr (sizeof(real) == 4) return fma(E[k], x, fma(E[k + ERL], x, acc));',
             '    else return fma(E[k], x, acc);',
             '}',
             'template <typename real, int ERL, typename EP>',
             '__device__ __forceinline__ real wb_emulc(EP E, int k, real x) {',
             '    if constexpr (sizeof(real) == 4) return fma(E[k], x, E[k + ERL] * x);',
             '    else return E[k] * x;',
             '}',
             'template <typename real, int ERL, typename EP>',
             '__device__ __forceinline__ real wb_efma(EP E, int k, real x,'
             ' real acc) { return fma(E[k], x, acc); }',
             'template <typename real, int ERL, typename EP>',
             '__device__ __forceinline__ real wb_emul(EP E, int k, real x)'
             ' { return E[k] * x; }',
             '// the A Cartesian components of part P of a split class (see gen_md.py)',
             'template <int LA, int LB, int P> struct MdPart;',
             '// number of parts of a class (2: split, 1: whole block)',
             '__host__ __device__ constexpr int md_nparts(int la, int lb) {',
             '    return %s ? 2 : 1;' % ' || '.join('(la == %d && lb == %d)' % c for c in sorted(SPLIT)),
             '}']
    for l in range(LMAX + 1):
        lines += gen_transforms(l)
    for la in range(LMAX + 1):
        for lb in range(la + 1):
            if (la, lb) not in SPLIT:
                lines += gen_prim(la, lb, True)
                lines += gen_prim(la, lb, False)
            else:
                for part in range(2):
                    idx = a_part(la, part)
                    lines.append('template <> struct MdPart<%d, %d, %d> {' % (la, lb, part))
                    lines.append('    static constexpr __host__ __device__ int n() { return %d; }'
                                 % len(idx))
                    lines.append('    static constexpr __host__ __device__ int at(int i) {'
                                 ' constexpr int t[%d] = {%s}; return t[i]; }'
                                 % (len(idx), ', '.join(map(str, idx))))
                    lines.append('};')
                    lines += gen_prim(la, lb, True, part)
                    lines += gen_prim(la, lb, False, part)
    text = '\n'.join(lines) + '\n'
    os.makedirs(os.path.dirname(out_path), exist_ok=True)
    old = open(out_path).read() if os.path.exists(out_path) else None
    if old != text:
        with open(out_path, 'w') as f:
            f.write(text)


if __name__ == '__main__':
    main(sys.argv[1])
