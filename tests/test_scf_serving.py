"""Staged SCF jobs through the serving loop, on the CPU: the loop's device
cycle is replaced by the FP64 oracle (K for both operators and the VV10
term on host grids), so the stage machine itself is checked here -- coarse
grid (mixed), fine grid (mixed), fine grid FP64 tail, switches on the energy
change and gradient or the stage's cycle limit, the step's history reset at
each switch, convergence in the last stage -- and the staged job reaches the
fixed point of an SCF run on the fine grid alone (same final energy to the
convergence tolerance).  ModelSCF (systems/scf.py) is the SCF step."""
import numpy

import oracle
from pyscf_wb97mv_fast_amd import serving
from pyscf_wb97mv_fast_amd.systems import geometry, grids, guess, mole, scf

OM = (0.0, 0.3)


class OracleServer(serving.CycleServer):
    """The serving loop with the device cycle done by the oracle."""

    def __init__(self, mol):
        super().__init__(slots=1, contexts=[object()], runner=lambda c, r: None)
        self._runner = None
        self.basis = oracle.OracleBasis.from_mol(mol)
        self.calls = []

    def _prepare(self, ctx, req):
        return None, [(st.kgrid, st.nlcgrid, st.precision) for st in self._stages(req)]

    def _need(self, ctx, req, state):
        return 0, 0

    def _cycle(self, ctx, req, basis, kg, vg, precision, dm):
        self.calls.append((len(kg[1]), precision))
        k = numpy.array([oracle.k_build(self.basis, kg[0], kg[1], dm, om) for om in OM])
        n, e, v = oracle.nr_nlc_vxc(self.basis, vg[0], vg[1], dm)
        return k, numpy.array([n, e]), v


def _grid(mol, npts, seed):
    g = grids.Grids(mol, level=0).build()
    sel = numpy.sort(numpy.random.default_rng(seed).choice(g.size, npts, replace=False))
    return g.coords[sel], g.weights[sel] * g.size / npts


def test_staged_scf_job_converges_to_the_fine_grid_fixed_point():
    mol = mole.Mole(geometry.water_cluster(2, seed=1), 'def2-svp')
    coarse, fine = _grid(mol, 700, 1), _grid(mol, 1500, 2)
    srv = OracleServer(mol)
    req = serving.Request(pack=None, dm=guess.huckel_density(mol), scf=scf.ModelSCF(mol),
                          stages=serving.staged((coarse, coarse), (fine, fine)),
                          conv_tol=1e-9, conv_tol_grad=1e-6, max_cycles=60)
    out = srv.submit(req).result(timeout=300)
    srv.close()
    hist = out['history']
    stages = [h['stage'] for h in hist]
    print('cycles %d, stages %s, E %.10f' % (out['cycles'], stages, out['e_tot']))
    assert out['converged'] and stages[0] == 0 and stages[-1] == 2
    assert stages == sorted(stages)                            # coarse -> fine -> FP64
    assert [c[1] for c in srv.calls].count('fp64') == stages.count(2)
    assert {c[0] for c in srv.calls if c[1] == 'fp64'} == {1500}
    assert hist[-1]['de'] < 1e-9 and hist[-1]['gnorm'] < 1e-6
    # the same fixed point as an SCF on the fine grid only, FP64 throughout
    srv2 = OracleServer(mol)
    ref = srv2.submit(serving.Request(
        pack=None, dm=guess.huckel_density(mol), scf=scf.ModelSCF(mol),
        stages=[serving.Stage(fine, fine, 'fp64', 0.0, 0.0, 10 ** 6)], conv_tol=1e-9,
        conv_tol_grad=1e-6, max_cycles=80)).result(timeout=300)
    srv2.close()
    print('fine-only: %d cycles, E %.10f' % (ref['cycles'], ref['e_tot']))
    assert ref['converged'] and abs(ref['e_tot'] - out['e_tot']) < 1e-8


def test_stage_switch_on_cycle_limit_resets_the_step():
    mol = mole.Mole(geometry.water_cluster(2, seed=1), 'def2-svp')
    g = _grid(mol, 600, 3)
    resets = []

    class Step(scf.ModelSCF):
        def reset(self):
            resets.append(True)
            super().reset()
    srv = OracleServer(mol)
    st = [serving.Stage(g, g, 'mixed', 0.0, 0.0, 2), serving.Stage(g, g, 'mixed', 0.0, 0.0, 3),
          serving.Stage(g, g, 'fp64', 0.0, 0.0, 10 ** 6)]
    out = srv.submit(serving.Request(pack=None, dm=guess.huckel_density(mol),
                                     scf=Step(mol, x_scale=-1.0), stages=st,
                                     max_cycles=8)).result(timeout=300)
    srv.close()
    assert [h['stage'] for h in out['history']] == [0, 0, 1, 1, 1, 2, 2, 2]
    assert len(resets) == 2 and not out['converged'] and out['cycles'] == 8
