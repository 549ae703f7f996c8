"""Generally contracted shells (nctr > 1) through the device path: AO values,
the K build (FP64 and mixed, both operators, screening on) and the VV10
flow of a basis with general contraction against the closed relations to
its uncontracted twin on the same device (tests/gc_basis.py):

    X_gc = X_prim C,   K_gc[D] = C^T K_prim[C D C^T] C,   vmat likewise,

and K against the FP64 oracle.  The reference handles such shells in its
generic tile kernel (gpu/sgx_k.py:73-150, gpu/int3c1e.py:66-226); this
build evaluates one segmented shell per contraction column."""
import copy

import numpy
import pytest

import oracle
from gc_basis import molecules
from pyscf_wb97mv_fast_amd import lib
from pyscf_wb97mv_fast_amd.basis import pack_mol
from pyscf_wb97mv_fast_amd.gpu import sgx_k, vv10
from pyscf_wb97mv_fast_amd.systems import geometry, grids, guess, sgx

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def ctx():
    c = lib.Context(0)
    yield c
    c.close()


@pytest.fixture(scope='module')
def system():
    mg, mp, C = molecules(geometry.WATER_DIMER)
    g = grids.Grids(mg, level=0).build()
    return mg, mp, C, g, guess.huckel_density(mg)


def test_ao_values_are_the_contracted_primitives(ctx, system):
    mg, mp, C, g, _ = system
    x = lib.f64(g.coords[:3000])
    out = {}
    for name, m in (('g', mg), ('p', mp)):
        b = lib.DeviceBasis(ctx, pack_mol(m))
        X = numpy.empty((4, len(x), b.nao))
        lib.check(lib.load().wb_ao_eval(ctx.h, b.h, len(x), lib.ptr(x), 1, lib.ptr(X)))
        out[name] = X
    ref = out['p'] @ C
    assert numpy.abs(out['g'] - ref).max() < 1e-12 * numpy.abs(ref).max()


@pytest.mark.parametrize('precision', ['fp64', 'mixed'])
def test_k_is_the_contracted_primitive_k(ctx, system, precision):
    mg, mp, C, g, dm = system
    kg = sgx_k.GpuKBuilder(mg, ctx, tile_tol=1e-11, precision=precision)
    kp = sgx_k.GpuKBuilder(mp, ctx, tile_tol=1e-11, precision=precision)
    Kg = kg.k_raw(g, dm, [0.0, 0.3])
    Kp = kp.k_raw(g, C @ dm @ C.T, [0.0, 0.3])
    for op in range(2):
        ref = C.T @ Kp[op] @ C
        err = numpy.abs(Kg[op] - ref).max() / numpy.abs(ref).max()
        # both sides screen at 1e-11 on their own tasks; FP32 integrals are
        # rounded per primitive pair on each side
        assert err < (1e-9 if precision == 'fp64' else 2e-6), (precision, op, err)


def test_k_matches_the_oracle(ctx, system):
    mg, _, _, g, dm = system
    rng = numpy.random.default_rng(3)
    sel = numpy.sort(rng.choice(g.size, 1500, replace=False))
    gs = copy.copy(g)
    gs.coords, gs.weights = g.coords[sel], g.weights[sel]
    ob = oracle.OracleBasis.from_mol(mg)
    for precision, tol in (('fp64', 1e-10), ('mixed', 1e-6)):
        K = sgx_k.GpuKBuilder(mg, ctx, tile_tol=0.0, precision=precision).k_raw(gs, dm, [0.0, 0.3])
        for op, omega in enumerate((0.0, 0.3)):
            ref = oracle.k_build(ob, gs.coords, gs.weights, dm, omega)
            err = numpy.abs(K[op] - ref).max() / numpy.abs(ref).max()
            assert err < tol, (precision, omega, err)


def test_vv10_vmat_is_the_contracted_primitive_vmat(ctx, system):
    mg, mp, C, g, dm = system
    ni = sgx.SyntheticNumInt()
    ng, eg, vg = vv10.gpu_nr_nlc_vxc(vv10.NlcState(ctx, mg), ni, mg, g, 'wb97m-v', dm,
                                     precision='fp64')
    npr, ep, vp = vv10.gpu_nr_nlc_vxc(vv10.NlcState(ctx, mp), ni, mp, g, 'wb97m-v',
                                      C @ dm @ C.T, precision='fp64')
    assert abs(ng - npr) < 1e-10 * abs(npr) and abs(eg - ep) < 1e-10
    ref = C.T @ vp @ C
    assert numpy.abs(vg - ref).max() < 1e-10 * numpy.abs(ref).max()
