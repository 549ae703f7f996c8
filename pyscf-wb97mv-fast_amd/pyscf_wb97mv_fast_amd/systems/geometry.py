"""Seeded synthetic geometries (Angstrom).

* ``water_cluster(n, seed)``: rigid waters (O-H 0.9572 A, H-O-H 104.52 deg)
  with uniform random orientations (normalized Gaussian quaternions); the O
  centres are placed by random sequential addition inside a sphere of the
  volume n waters occupy at 1 g/cm^3 (0.03343 molecules / A^3), with a
  minimum O-O distance of 2.6 A.
* ``drug_like(n_heavy, seed)``: a random-walk sp3 carbon backbone (C-C 1.54
  A, tetrahedral bond angles, dihedrals from {60, 180, 300} deg); every 5th
  backbone carbon becomes N or O, three substituents are drawn from {F, Cl,
  S-CH3}, and hydrogens fill the remaining valences at 1.09 A.
* ``WATER_DIMER``: the water dimer used by the reference's unit tests
  (``core/testsystems.py:5-12``), as data.
"""
import numpy

WATER_DIMER = [
    ('O', (-1.551007, -0.114520, 0.000000)),
    ('H', (-1.934259, 0.762503, 0.000000)),
    ('H', (-0.599677, 0.040712, 0.000000)),
    ('O', (1.350625, 0.111469, 0.000000)),
    ('H', (1.680398, -0.373741, -0.758561)),
    ('H', (1.680398, -0.373741, 0.758561)),
]

_OH = 0.9572
_HOH = numpy.deg2rad(104.52)
_WATER_DENSITY = 0.03343            # molecules per A^3 at 1 g/cm^3


def _quat_rot(q):
    a, b, c, d = q / numpy.linalg.norm(q)
    return numpy.array([
        [a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
        [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
        [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])


def water_cluster(n, seed=0, min_oo=2.6, max_tries=200000):
    rng = numpy.random.default_rng(seed)
    radius = (3.0 * n / (4.0 * numpy.pi * _WATER_DENSITY)) ** (1.0 / 3.0)
    # local frame: O at origin, H atoms in the xy plane
    local = numpy.array([[0.0, 0.0, 0.0],
                         [_OH, 0.0, 0.0],
                         [_OH * numpy.cos(_HOH), _OH * numpy.sin(_HOH), 0.0]])
    centres = []
    tries = 0
    r_eff = radius
    while len(centres) < n:
        tries += 1
        if tries > max_tries:
            raise RuntimeError('random sequential addition did not converge')
        if tries % 5000 == 0:           # grow the sphere if it is jammed
            r_eff *= 1.02
        p = rng.uniform(-r_eff, r_eff, size=3)
        if numpy.dot(p, p) > r_eff * r_eff:
            continue
        if centres and numpy.min(numpy.linalg.norm(
                numpy.asarray(centres) - p, axis=1)) < min_oo:
            continue
        centres.append(p)
    atoms = []
    for c in centres:
        rot = _quat_rot(rng.normal(size=4))
        xyz = local @ rot.T + c
        atoms += [('O', tuple(xyz[0])), ('H', tuple(xyz[1])), ('H', tuple(xyz[2]))]
    return atoms


def _place(a, b, c, bond, angle, dihedral):
    """Position of d with |cd| = bond, angle(b,c,d), dihedral(a,b,c,d)."""
    bc = c - b
    bc /= numpy.linalg.norm(bc)
    n = numpy.cross(b - a, bc)
    nn = numpy.linalg.norm(n)
    if nn < 1e-8:
        n = numpy.cross(bc, [1.0, 0.0, 0.0])
        if numpy.linalg.norm(n) < 1e-8:
            n = numpy.cross(bc, [0.0, 1.0, 0.0])
        nn = numpy.linalg.norm(n)
    n /= nn
    m = numpy.cross(n, bc)
    d2 = numpy.array([-bond * numpy.cos(angle),
                      bond * numpy.sin(angle) * numpy.cos(dihedral),
                      bond * numpy.sin(angle) * numpy.sin(dihedral)])
    return c + d2[0] * bc + d2[1] * m + d2[2] * n


_TET = numpy.arccos(-1.0 / 3.0)
_BOND = {('C', 'C'): 1.54, ('C', 'N'): 1.47, ('C', 'O'): 1.43, ('C', 'F'): 1.35,
         ('C', 'Cl'): 1.77, ('C', 'S'): 1.82, ('S', 'C'): 1.82}
_VALENCE = {'C': 4, 'N': 3, 'O': 2, 'S': 2, 'F': 1, 'Cl': 1}


def drug_like(n_heavy=40, seed=0, max_tries=500):
    """Returns [(symbol, xyz)].  Backbone atoms that would clash (< 2.4 A
    to a non-neighbour) trigger a re-draw of the dihedral."""
    rng = numpy.random.default_rng(seed)
    n_back = n_heavy - 3 - 1          # three substituents; S-CH3 adds one C
    pos = [numpy.zeros(3), numpy.array([1.54, 0.0, 0.0])]
    pos.append(_place(numpy.array([0.0, 1.0, 0.0]), pos[0], pos[1], 1.54,
                      _TET, numpy.pi))
    while len(pos) < n_back:
        for _ in range(max_tries):
            dih = numpy.deg2rad(rng.choice([60.0, 180.0, 300.0]))
            p = _place(pos[-3], pos[-2], pos[-1], 1.54, _TET, dih)
            if numpy.min(numpy.linalg.norm(numpy.asarray(pos[:-2]) - p,
                                           axis=1)) > 2.4:
                break
        else:
            raise RuntimeError('backbone random walk jammed')
        pos.append(p)
    syms = ['C'] * n_back
    for i in range(4, n_back, 5):
        syms[i] = 'N' if rng.uniform() < 0.5 else 'O'
    atoms = [[s, p] for s, p in zip(syms, pos)]
    nbrs = {i: [] for i in range(n_back)}
    for i in range(n_back - 1):
        nbrs[i].append(i + 1)
        nbrs[i + 1].append(i)
    heavy = list(range(n_back))
    # substituents on backbone carbons with a free valence
    cand = [i for i in range(1, n_back - 1) if syms[i] == 'C']
    subs = rng.choice(['F', 'Cl', 'SCH3'], size=3)
    sites = rng.choice(cand, size=3, replace=False)

    def free_dirs(i):
        """Tetrahedral directions completing atom i's bonds."""
        c = atoms[i][1]
        bonded = [atoms[j][1] for j in nbrs[i]]
        if len(bonded) == 2:
            a, b = bonded
            u = (c - a) / numpy.linalg.norm(c - a) + (c - b) / numpy.linalg.norm(c - b)
            u /= numpy.linalg.norm(u)
            w = numpy.cross(a - c, b - c)
            w /= numpy.linalg.norm(w)
            ang = _TET / 2.0
            return [numpy.cos(ang) * u + numpy.sin(ang) * w,
                    numpy.cos(ang) * u - numpy.sin(ang) * w]
        a = bonded[0]
        ref = atoms[nbrs[nbrs[i][0]][0]][1] if nbrs[nbrs[i][0]] else a + [0, 1, 0]
        return [(_place(ref, a, c, 1.0, _TET, d) - c)
                for d in (numpy.pi / 3, numpy.pi, 5 * numpy.pi / 3)]

    for site, sub in zip(sites, subs):
        d = free_dirs(site)[0]
        c = atoms[site][1]
        if sub == 'SCH3':
            ps = c + _BOND[('C', 'S')] * d
            atoms.append(['S', ps])
            iS = len(atoms) - 1
            nbrs[site].append(iS)
            nbrs[iS] = [site]
            pc = _place(atoms[nbrs[site][0]][1], c, ps, 1.82, numpy.deg2rad(100.0),
                        numpy.pi)
            atoms.append(['C', pc])
            iC = len(atoms) - 1
            nbrs[iS].append(iC)
            nbrs[iC] = [iS]
            heavy += [iS, iC]
        else:
            atoms.append([sub, c + _BOND[('C', sub)] * d])
            ix = len(atoms) - 1
            nbrs[site].append(ix)
            nbrs[ix] = [site]
            heavy.append(ix)
    # hydrogens
    out_h = []
    for i in heavy:
        sym = atoms[i][0]
        need = _VALENCE[sym] - len(nbrs[i])
        if need <= 0:
            continue
        c = atoms[i][1]
        bonded = [atoms[j][1] for j in nbrs[i]]
        if len(bonded) == 1:
            a = bonded[0]
            far = [atoms[j][1] for j in nbrs[nbrs[i][0]] if j != i]
            ref = far[0] if far else a + numpy.array([0.3, 1.0, 0.2])
            dirs = [(_place(ref, a, c, 1.0, _TET, d) - c)
                    for d in (numpy.pi / 3, numpy.pi, 5 * numpy.pi / 3)]
        elif len(bonded) == 2:
            dirs = free_dirs(i)
        else:
            u = sum((c - b) / numpy.linalg.norm(c - b) for b in bonded)
            dirs = [u / numpy.linalg.norm(u)]
        for d in dirs[:need]:
            out_h.append(['H', c + 1.09 * d / numpy.linalg.norm(d)])
    return [(s, tuple(float(x) for x in p)) for s, p in atoms + out_h]
