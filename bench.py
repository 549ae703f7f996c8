"""Benchmark of the omega-B97M-V / COSX SCF-cycle hot path on one MI355X.

One step = one SCF cycle of the hot path for the synthetic (H2O)_27 cluster
(def2-SVP, 648 AOs): COSX K(1/r) + K(erf, omega = 0.3) on the SGX grid and
the VV10 term of nr_nlc_vxc on the VV10 grid (both the project's own
level-1 Becke grid, 218 566 points), in the mixed-precision form (FP32
integrals and VV10 pair sums; FP64 AO values, GEMMs, weights, K and vmat),
with the density matrix already resident in HBM (wb_k_build_dev /
wb_nlc_vxc_dev).  Also reported: the all-FP64 form of the same kernels, the
per-cycle error of the mixed form against it, p50/p99 per-cycle latency, the
exchange kernel's FP32 roofline fraction (HIP events on the library stream)
and the FP64 CPU oracle timed on a bounded sample on the host cores.

    python bench.py [--gpus 1] [--steps 10] [--warmup 2]
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, 'pyscf-wb97mv-fast_amd'))

OMEGA = 0.3
NLC = (6.0, 0.01)
K_TOL = 1e-11
# FP32 flops per primitive-pair evaluation (one grid point) by class, FMA = 2:
# Boys, Hermite recursion and E contraction (pruned), and per task and point
# the F / G contraction 4 * nsph_a * nsph_b
FLOPS_PRIM = {(0, 0): 43, (1, 0): 89, (1, 1): 223, (2, 0): 182, (2, 1): 516,
              (2, 2): 1256, (3, 0): 358, (3, 1): 1081, (3, 2): 2674, (3, 3): 5688}
PEAK_FP32_PACKED = 157.3e12        # MI355X published FP32 vector peak
PEAK_FP32_VALU = 78.6e12           # without packed FP32 (not used by this build)
# HBM bytes per operator of the exchange kernels (all classes), measured with
# rocprofv3 FETCH_SIZE / WRITE_SIZE on this workload and scaled by the same
# counters' reading for a plain copy of known size; the summary file holds
# the commands and the calibration
TRAFFIC_FILE = os.path.join(ROOT, 'profiles', 'md_k_kernel_water27_svp_traffic.txt')
PHASES = ['ao', 'fgemm', 'screen', 'exchange', 'kgemm', 'nlc_ao', 'nlc_rho', 'vv10',
          'nlc_fock']


def log(*a):
    print(*a, file=sys.stderr, flush=True)


WORKLOADS = {
    'water27': '(H2O)27 def2-SVP',
    'water8': '(H2O)8 def2-SVP',
    'drug': 'drug-like chain, 40 heavy atoms (seed 0), def2-TZVP',
}


def build_workload(kind='water27', seed=0):
    from pyscf_wb97mv_fast_amd import lib
    from pyscf_wb97mv_fast_amd.systems import geometry, grids, guess, mole
    t = time.perf_counter()
    if kind == 'drug':
        mol = mole.Mole(geometry.drug_like(40, seed=seed), 'def2-tzvp')
    else:
        mol = mole.Mole(geometry.water_cluster(27 if kind == 'water27' else 8, seed=seed),
                        'def2-svp')
    # the grid's cell partition runs on the GPU (wb_grid_partition)
    g = grids.Grids(mol, level=1, device=lib.Context(0)).build()
    dm = guess.huckel_density(mol)
    log('workload: nao %d, %d grid points, setup %.1f s' % (mol.nao, g.size,
                                                            time.perf_counter() - t))
    return mol, g, dm


def measured_traffic():
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if line.startswith('traffic_bytes_per_operator:'):
                    return int(line.split(':')[1])
    except OSError:
        pass
    return None


def _clusters(count, nmax, seed):
    """count water clusters, n ~ U{4..nmax} (seeded), def2-SVP, level-1 grids
    (cell partition on the GPU) and extended-Hueckel densities."""
    from pyscf_wb97mv_fast_amd import lib
    from pyscf_wb97mv_fast_amd.basis import pack_mol
    from pyscf_wb97mv_fast_amd.systems import geometry, grids, guess, mole
    rng = numpy.random.default_rng(seed)
    pctx = lib.Context(0)
    out = []
    for i, n in enumerate(rng.integers(4, nmax + 1, count)):
        mol = mole.Mole(geometry.water_cluster(int(n), seed=seed + i), 'def2-svp')
        g = grids.Grids(mol, level=1, device=pctx).build()
        out.append((int(n), pack_mol(mol), g.coords, g.weights, guess.huckel_density(mol)))
        if (i + 1) % 8 == 0:
            log('  %d/%d molecules set up' % (i + 1, count))
    return out


def bench_setup(args):
    """Per-molecule set-up on the device path: wb_basis_create (shell pairs on
    the host threads), wb_grid_partition (cell weights on the GPU) and
    wb_grid_build (the whole grid on the GPU) for water clusters of 8, 27 and
    64 molecules, level-1 grids; the host numpy partition timed beside it on
    (H2O)8."""
    from pyscf_wb97mv_fast_amd import lib
    from pyscf_wb97mv_fast_amd.basis import pack_mol
    from pyscf_wb97mv_fast_amd.gpu import grids as dgrids
    from pyscf_wb97mv_fast_amd.systems import geometry, grids, mole
    ctx = lib.Context(0)
    rows = []
    for n in (8, 27, 64):
        mol = mole.Mole(geometry.water_cluster(n, seed=0), 'def2-svp')
        pk = pack_mol(mol)
        t = time.perf_counter()
        b = lib.DeviceBasis(ctx, pk)
        t_basis = time.perf_counter() - t
        coords, _, owner = grids.Grids(mol, level=1).raw_points()
        atoms = mol.atom_coords()
        dgrids.ssf_partition(ctx, coords, atoms, owner)          # first call: set-up
        t = time.perf_counter()
        part = dgrids.ssf_partition(ctx, coords, atoms, owner)
        t_part = time.perf_counter() - t
        t = time.perf_counter()
        g = grids.Grids(mol, level=1, device=ctx).build()
        t_build = time.perf_counter() - t
        row = {'waters': n, 'nao': b.nao, 'shell_pairs': b.npairs, 'primitive_pairs': b.nprims,
               'basis_create_s': round(t_basis, 4), 'raw_points': len(coords),
               'partition_s': round(t_part, 4), 'device_grid_build_s': round(t_build, 4),
               'grid_points': g.size}
        if n == 8:
            t = time.perf_counter()
            ref = grids._ssf_partition(coords, atoms, owner)
            row['host_partition_s'] = round(time.perf_counter() - t, 3)
            row['max_abs_diff_vs_host'] = float(abs(ref - part).max())
        log('(H2O)%d: %s' % (n, row))
        rows.append(row)
    print(json.dumps({'metric': 'set-up seconds per molecule (shell pairs, grid-cell partition)',
                      'unit': 's', 'n_gpus': 1, 'higher_is_better': False, 'data': 'synthetic',
                      'host_threads': min(16, os.cpu_count() or 1), 'molecules': rows}), flush=True)


def bench_batch(args):
    """One K(1/r) + K(erf) + VV10 cycle for each of N molecules of uneven
    size: separate single-molecule calls vs one wb_cycle_batch call (1 and
    --lanes lanes) vs one wb_cycle_queue call (the exchange kernels of all
    molecules in one launch per shell-pair class), mixed and FP64."""
    from pyscf_wb97mv_fast_amd import lib
    from pyscf_wb97mv_fast_amd.gpu import batch
    t = time.perf_counter()
    mols = _clusters(args.batch, args.batch_max, 11)
    ctx = lib.Context(0)
    bases = [lib.DeviceBasis(ctx, m[1]) for m in mols]
    grs = [lib.DeviceGrid(ctx, m[2], m[3]) for m in mols]
    log('batch workload: %d clusters, nao %d..%d, %d grid points in all, setup %.1f s'
        % (len(mols), min(b.nao for b in bases), max(b.nao for b in bases),
           sum(len(m[3]) for m in mols), time.perf_counter() - t))
    om = numpy.array([0.0, OMEGA])
    L = lib.load()

    def separate(prec):
        res = []
        for b, g, m in zip(bases, grs, mols):
            k = numpy.empty((2, b.nao, b.nao))
            lib.check(L.wb_k_build(ctx.h, b.h, g.h, lib.ptr(m[4]), 2, lib.ptr(om), K_TOL, prec,
                                   lib.ptr(k)))
            o2, v = numpy.empty(2), numpy.empty((b.nao, b.nao))
            lib.check(L.wb_nlc_vxc(ctx.h, b.h, g.h, lib.ptr(m[4]), NLC[0], NLC[1], 0.0, prec,
                                   lib.ptr(o2), lib.ptr(v)))
            res.append((k, o2))
        return res

    def batched(precision, lanes):
        jobs = [batch.CycleJob(b, g, m[4], g, m[4]) for b, g, m in zip(bases, grs, mols)]
        return batch.run_batch(ctx, jobs, om, K_TOL, NLC, 0.0, precision, lanes)

    def queued(precision):
        jobs = [batch.CycleJob(b, g, m[4], g, m[4]) for b, g, m in zip(bases, grs, mols)]
        return batch.run_queue(ctx, jobs, om, K_TOL, NLC, 0.0, precision)

    def timed(fn, *a):
        fn(*a)                                      # sizes every workspace
        t0 = time.perf_counter()
        r = fn(*a)
        return time.perf_counter() - t0, r

    t_sep, sep = timed(separate, 0)
    log('separate calls (mixed): %.3f s' % t_sep)
    ctx.trim()                                      # the lanes get the device memory
    t_b1, jobs1 = timed(batched, 'mixed', 1)
    log('batch, 1 lane (mixed): %.3f s' % t_b1)
    t_bn, jobsn = timed(batched, 'mixed', args.lanes)
    log('batch, %d lanes (mixed): %.3f s' % (args.lanes, t_bn))
    ctx.trim()
    t_q, jobsq = timed(queued, 'mixed')
    log('queue (mixed): %.3f s' % t_q)
    # the reported configuration: the fastest of the lane counts and the queue
    best = min([(t_b1, 'lanes', 1, jobs1), (t_bn, 'lanes', args.lanes, jobsn),
                (t_q, 'queue', 0, jobsq)], key=lambda r: r[0])
    t_bl, mode, lanes, jobs = best
    ctx.trim()
    if mode == 'queue':
        t_64, jobs64 = timed(queued, 'fp64')
    else:
        t_64, jobs64 = timed(batched, 'fp64', lanes)
    log('%s (fp64): %.3f s' % (mode if mode == 'queue' else '%d lane(s)' % lanes, t_64))
    same = all(numpy.array_equal(j.k, s_[0]) and numpy.array_equal(j.nlc, s_[1])
               for jl in (jobs1, jobsn, jobsq) for j, s_ in zip(jl, sep))
    dex = [abs(-0.25 * float(numpy.einsum('ij,ji->', m[4], 0.15 * (j.k[0] - j6.k[0])
                                           + 0.85 * (j.k[1] - j6.k[1]))))
           for j, j6, m in zip(jobs, jobs64, mols)]
    dnl = [abs(j.nlc[1] - j6.nlc[1]) for j, j6 in zip(jobs, jobs64)]
    ctx.trim()
    print(json.dumps({
        'metric': 'batched COSX K(1/r)+K(erf)+VV10 molecule-cycles/s', 'value': len(mols) / t_bl,
        'unit': 'molecule-cycles/s', 'n_gpus': 1, 'higher_is_better': True,
        'dtype': 'f32 integrals + f64 GEMMs', 'data': 'synthetic',
        'config': {'workload': '%d water clusters, n ~ U{4..%d} (seed 11), def2-SVP, level-1 '
                              'grids' % (len(mols), args.batch_max), 'mode': mode,
                   'lanes': lanes},
        'seconds': {'separate_calls': t_sep, 'batch_1_lane': t_b1,
                    'batch_%d_lanes' % args.lanes: t_bn, 'queue': t_q, 'fp64': t_64},
        'batch_equals_separate': bool(same),
        'max_abs_dE_x_vs_fp64': max(dex), 'max_abs_dE_nl_vs_fp64': max(dnl),
        'mixed_vs_fp64': t_64 / t_bl}), flush=True)


def _serve_jobs(count, max_waters, max_heavy, seed):
    """count SCF jobs: water clusters (n ~ U{8..max_waters}, def2-SVP) and
    drug-like chains (U{10..max_heavy} heavy atoms, def2-SVP or def2-TZVP),
    seeded; level-0 (coarse) and level-1 (fine) grids with the cell
    partition on the GPU; extended-Hueckel densities."""
    from pyscf_wb97mv_fast_amd import lib
    from pyscf_wb97mv_fast_amd.basis import pack_mol
    from pyscf_wb97mv_fast_amd.systems import geometry, grids, guess, mole
    rng = numpy.random.default_rng(seed)
    pctx = lib.Context(0)
    out = []
    while len(out) < count:
        i = len(out)
        try:
            if rng.random() < 0.5:
                n = int(rng.integers(8, max_waters + 1))
                mol = mole.Mole(geometry.water_cluster(n, seed=seed + i), 'def2-svp')
                name = '(H2O)%d def2-SVP' % n
            else:
                n = int(rng.integers(10, max_heavy + 1))
                basis = 'def2-tzvp' if rng.random() < 0.5 else 'def2-svp'
                mol = mole.Mole(geometry.drug_like(n, seed=seed + i), basis)
                name = 'drug-like %d heavy atoms %s' % (n, basis)
        except RuntimeError:                        # a jammed random walk: draw again
            continue
        g0 = grids.Grids(mol, level=0, device=pctx).build()
        g1 = grids.Grids(mol, level=1, device=pctx).build()
        out.append(dict(name=name, mol=mol, pack=pack_mol(mol), coarse=(g0.coords, g0.weights),
                        fine=(g1.coords, g1.weights), dm=guess.huckel_density(mol)))
    return out


def bench_serve(args):
    """Seeded Poisson arrivals of SCF jobs run to convergence through the
    continuous-batching loop: each job level-0 grids (mixed), level-1 grids
    (mixed), level-1 grids in FP64 until |dE| < 1e-9 Ha and |g| < 1e-6, with
    the model SCF step; the same schedule in FP64 throughout beside it."""
    from pyscf_wb97mv_fast_amd.serving import CycleServer, Request, Stage, staged
    from pyscf_wb97mv_fast_amd.systems import scf
    t = time.perf_counter()
    jobs = _serve_jobs(args.serve, args.serve_waters, args.serve_heavy, 23)
    gaps = numpy.random.default_rng(5).exponential(1.0 / args.rate, len(jobs))
    log('serve workload: %d jobs (%s), setup %.1f s' % (
        len(jobs), ', '.join(j['name'] for j in jobs), time.perf_counter() - t))
    out = {}
    for form in ('mixed', 'fp64'):
        srv = CycleServer(slots=args.lanes, hbm_budget=200e9)
        t0 = time.perf_counter()
        futs = []
        for gap, j in zip(gaps, jobs):
            time.sleep(float(gap))
            st = staged((j['coarse'], j['coarse']), (j['fine'], j['fine']))
            if form == 'fp64':
                st = [Stage(x.kgrid, x.nlcgrid, 'fp64', x.tau_e, x.tau_g, x.max_cycles) for x in st]
            futs.append(srv.submit(Request(pack=j['pack'], dm=j['dm'], scf=scf.ModelSCF(j['mol']),
                                           stages=st, conv_tol=1e-9, conv_tol_grad=1e-6,
                                           max_cycles=100)))
        res = [f.result() for f in futs]
        srv.close()
        stt = srv.stats()
        stt['wall_s'] = time.perf_counter() - t0
        stt['converged'] = sum(bool(r['converged']) for r in res)
        stt['cycles'] = [r['cycles'] for r in res]
        stt['energies'] = [r['e_tot'] for r in res]
        log('%s: %d jobs (%d converged), %.2f jobs/s, time to converged energy p50 %.2f s, '
            'p99 %.2f s, at most %d at once' % (form, stt['jobs'], stt['converged'],
                                                stt['jobs_per_s'], stt['p50_s'], stt['p99_s'],
                                                stt['max_active']))
        out[form] = stt
    de = max(abs(a - b) for a, b in zip(out['mixed'].pop('energies'), out['fp64'].pop('energies')))
    print(json.dumps({
        'metric': 'served SCF jobs/s and p50/p99 time to converged energy', 'value':
        out['mixed']['jobs_per_s'], 'unit': 'jobs/s', 'n_gpus': 1, 'higher_is_better': True,
        'data': 'synthetic', 'config': {
            'workload': '%d jobs, Poisson arrivals %.1f/s (seed 5): water clusters n ~ U{8..%d} '
                        'def2-SVP, drug-like chains U{10..%d} heavy atoms def2-SVP/TZVP; '
                        'level-0 -> level-1 grids (mixed) -> level-1 FP64 tail to |dE| < 1e-9 '
                        'Ha; model SCF step (extended-Hueckel h + device K and VV10)'
                        % (len(jobs), args.rate, args.serve_waters, args.serve_heavy),
            'slots': args.lanes, 'hbm_budget_bytes': 200e9,
            'jobs': [j['name'] for j in jobs]},
        'mixed': out['mixed'], 'fp64': out['fp64'],
        'max_abs_final_dE_mixed_vs_fp64': de}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=10)
    ap.add_argument('--warmup', type=int, default=2)
    ap.add_argument('--fp64-steps', type=int, default=2)
    ap.add_argument('--vv10-tol', type=float, default=0.0)
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--mixed-only', action='store_true',
                    help='time the mixed form only, every cycle with the K build and VV10 '
                         'flows in sequence and phase timing (for profiler runs)')
    ap.add_argument('--copy-probe', type=int, default=0, metavar='MIB',
                    help='only run a device copy of MIB MiB (5 times): the known-size '
                         'kernel that calibrates the FETCH_SIZE / WRITE_SIZE counters')
    ap.add_argument('--workload', choices=sorted(WORKLOADS), default='water27',
                    help='the molecule of the cycle measurement (default: the headline '
                         '(H2O)27 def2-SVP)')
    ap.add_argument('--batch', type=int, default=0, metavar='N',
                    help='batched mode: N water clusters (n ~ U{4..--batch-max}), one '
                         'K(1/r) + K(erf) + VV10 cycle each, separate calls vs one '
                         'wb_cycle_batch; prints its own JSON line')
    ap.add_argument('--batch-max', type=int, default=64)
    ap.add_argument('--lanes', type=int, default=4)
    ap.add_argument('--setup', action='store_true',
                    help='set-up mode: shell-pair construction and the grid-cell partition '
                         'for (H2O)8/27/64; prints its own JSON line')
    ap.add_argument('--serve', type=int, default=0, metavar='N',
                    help='serving mode: N SCF jobs with seeded Poisson arrivals (--rate per '
                         's) run to convergence through the continuous-batching loop; prints '
                         'its own JSON line')
    ap.add_argument('--rate', type=float, default=2.0)
    ap.add_argument('--serve-waters', type=int, default=32,
                    help='largest water cluster of the serving mix (8..N)')
    ap.add_argument('--serve-heavy', type=int, default=20,
                    help='largest drug-like chain of the serving mix (10..N heavy atoms)')
    args = ap.parse_args()
    if args.gpus != 1:
        print('bench.py runs on one GPU only: this build has no multi-GPU path '
              '(the grid-sharded K with one nao^2 all-reduce is a design note)',
              file=sys.stderr)
        sys.exit(2)

    import torch
    if args.copy_probe:
        n = args.copy_probe << 18                  # float32 elements
        a = torch.ones(n, dtype=torch.float32, device='cuda:0')
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({'copy_probe_bytes_read': 4 * n, 'copy_probe_bytes_written': 4 * n,
                          'launches': 5}), flush=True)
        return
    if args.setup:
        return bench_setup(args)
    if args.batch:
        return bench_batch(args)
    if args.serve:
        return bench_serve(args)
    from pyscf_wb97mv_fast_amd import lib
    from pyscf_wb97mv_fast_amd.basis import pack_mol

    mol, g, dm = build_workload(args.workload)
    nao = mol.nao
    L = lib.load()
    ctx = lib.Context(0)
    basis = lib.DeviceBasis(ctx, pack_mol(mol))
    grid = lib.DeviceGrid(ctx, g.coords, g.weights)
    dev = torch.device('cuda:0')
    d_dm = torch.tensor(dm, dtype=torch.float64, device=dev)
    d_K = torch.zeros((2, nao, nao), dtype=torch.float64, device=dev)
    d_V = torch.zeros((nao, nao), dtype=torch.float64, device=dev)
    d_e = torch.zeros(2, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    om = numpy.array([0.0, OMEGA])

    def cycle(prec, vv10_tol):
        # K build and VV10 term of one cycle; the VV10 flow runs concurrently
        # on a second stream unless phase timing is on (wb_cycle_dev)
        lib.check(L.wb_cycle_dev(ctx.h, basis.h, grid.h, d_dm.data_ptr(), 2, lib.ptr(om), K_TOL,
                                 grid.h, d_dm.data_ptr(), NLC[0], NLC[1], vv10_tol, prec,
                                 d_K.data_ptr(), d_e.data_ptr(), d_V.data_ptr()))
        ctx.sync()

    def run(prec, steps, warmup, profile=False, vv10_tol=args.vv10_tol, profile_steps=2,
            profile_all=False):
        """Time `steps` cycles (concurrent flows, no phase timing), then, with
        profile, `profile_steps` more with the flows in sequence and HIP
        events between the phases (per-kernel-phase times).  profile_all:
        every cycle in sequence with phase timing (profiler runs)."""
        if profile_all:                 # warm-up cycles in sequence too (profiler runs)
            lib.check(L.wb_ctx_set_profile(ctx.h, 1))
        for _ in range(warmup):
            cycle(prec, vv10_tol)
        torch.cuda.synchronize()
        ctx.sync()
        if profile_all:
            lib.check(L.wb_profile_read(ctx.h, lib.ptr(numpy.zeros(len(PHASES))), len(PHASES), 1))
        times = []
        t0 = time.perf_counter()
        for _ in range(steps):
            t = time.perf_counter()
            cycle(prec, vv10_tol)
            times.append(time.perf_counter() - t)
        ctx.sync()
        torch.cuda.synchronize()
        total = time.perf_counter() - t0
        out = (d_K.cpu().numpy().copy(), d_e.cpu().numpy().copy(), d_V.cpu().numpy().copy())
        ph = numpy.zeros(len(PHASES))
        if profile_all:
            lib.check(L.wb_profile_read(ctx.h, lib.ptr(ph), len(PHASES), 1))
            lib.check(L.wb_ctx_set_profile(ctx.h, 0))
            ph /= max(steps, 1)
        elif profile:
            lib.check(L.wb_ctx_set_profile(ctx.h, 1))
            for _ in range(profile_steps):
                cycle(prec, vv10_tol)
            lib.check(L.wb_profile_read(ctx.h, lib.ptr(ph), len(PHASES), 1))
            lib.check(L.wb_ctx_set_profile(ctx.h, 0))
            ph /= profile_steps
        return total, numpy.array(times), ph, out

    # mixed precision: the measured configuration
    total, times, ph, res32 = run(lib.PREC['mixed'], args.steps, args.warmup, profile=True,
                                  profile_all=args.mixed_only)
    value = args.steps / total
    log('mixed: %.2f ms/cycle (K build and VV10 on two streams); phases with the flows in '
        'sequence (ms) %s, sum %.2f' % (1e3 * total / args.steps,
                                        dict(zip(PHASES, numpy.round(ph, 3))), ph.sum()))
    work = numpy.zeros(20, dtype=numpy.int64)
    lib.check(L.wb_k_work(ctx.h, lib.ptr(work)))
    kept = numpy.zeros(2, dtype=numpy.int64)
    lib.check(L.wb_k_stats(ctx.h, grid.h, lib.ptr(kept)))
    flops = 0.0
    for la in range(4):
        for lb in range(la + 1):
            c = la * (la + 1) // 2 + lb
            ntask, nprim = int(work[2 * c]), int(work[2 * c + 1])
            flops += 64.0 * (nprim * FLOPS_PRIM[(la, lb)] + ntask * 4 * (2 * la + 1) * (2 * lb + 1))
    # the exchange phase covers both operators (the same task lists; the
    # light classes evaluate both in one launch): per operator = half
    md_ms = ph[3] / 2.0
    achieved = flops / (md_ms * 1e-3) if md_ms > 0 else 0.0

    if args.mixed_only:
        print(json.dumps({'mixed_ms_per_step': round(1e3 * total / args.steps, 3),
                          'phases_ms': dict(zip(PHASES, [round(float(v), 4) for v in ph])),
        'phases_note': 'per-phase times from a profiled pass with the K build and VV10 flows '
                       'in sequence (HIP events); the timed cycles run them concurrently',
                          'md_algorithmic_flops_per_operator': flops}), flush=True)
        return
    # the same mixed cycle replayed as one HIP graph (wb_cycle_graph_*)
    graph = lib.CycleGraph(ctx, basis, grid, d_dm.data_ptr(), om, K_TOL, grid, d_dm.data_ptr(),
                           NLC, args.vv10_tol, 'mixed', d_K.data_ptr(), d_e.data_ptr(),
                           d_V.data_ptr())
    for _ in range(args.warmup):
        graph.launch()
    ctx.sync()
    gtimes = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        t = time.perf_counter()
        graph.launch()
        ctx.sync()
        gtimes.append(time.perf_counter() - t)
    gtotal = time.perf_counter() - t0
    gsame = (numpy.array_equal(d_K.cpu().numpy(), res32[0]) and
             numpy.array_equal(d_e.cpu().numpy(), res32[1]) and
             numpy.array_equal(d_V.cpu().numpy(), res32[2]))
    graph.close()
    log('mixed, one HIP graph per cycle: %.2f ms/cycle (outputs identical to the eager '
        'cycle: %s)' % (1e3 * gtotal / args.steps, gsame))

    # all-FP64 form of the same kernels
    # (exact VV10: the reference for the mixed form's error, pruning included)
    total64, times64, ph64, res64 = run(lib.PREC['fp64'], args.fp64_steps, 1, profile=True,
                                        vv10_tol=0.0, profile_steps=1)
    value64 = args.fp64_steps / total64
    log('fp64: %.2f ms/cycle, phases (ms) %s' % (1e3 * total64 / args.fp64_steps,
                                                dict(zip(PHASES, numpy.round(ph64, 3)))))
    dK = res32[0] - res64[0]
    dEk = [0.25 * float(numpy.einsum('ij,ji->', dm, dK[i])) for i in range(2)]
    dEx = -0.25 * float(numpy.einsum('ij,ji->', dm, 0.15 * dK[0] + 0.85 * dK[1]))
    dEnl = float(res32[1][1] - res64[1][1])

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(mol, g, dm)

    line = {
        'metric': 'COSX+VV10 SCF cycles/s (mixed FP32/FP64, |dE| <= 1e-6 Ha vs FP64)',
        'value': round(value, 4), 'unit': 'cycles/s', 'n_gpus': 1,
        'steps': args.steps, 'warmup': args.warmup,
        'ms_per_step': round(1e3 * total / args.steps, 3),
        'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': 'f32+f64', 'data': 'synthetic',
        'config': {'workload': WORKLOADS[args.workload] + ': COSX K(1/r) + K(erf, 0.3) + '
                               'VV10 nr_nlc_vxc per cycle, dm resident in HBM',
                   'nao': nao, 'grid_points': int(g.size), 'grid': 'own Becke level 1',
                   'k_tile_tol': K_TOL, 'vv10_tile_tol': args.vv10_tol,
                   'guess': 'extended Hueckel'},
        'latency_ms': {'p50': round(1e3 * float(numpy.percentile(times, 50)), 3),
                       'p99': round(1e3 * float(numpy.percentile(times, 99)), 3)},
        'graph': {'ms_per_step': round(1e3 * gtotal / args.steps, 3),
                  'p50': round(1e3 * float(numpy.percentile(gtimes, 50)), 3),
                  'p99': round(1e3 * float(numpy.percentile(gtimes, 99)), 3),
                  'identical_to_eager': bool(gsame)},
        'fp64': {'value': round(value64, 4), 'ms_per_step': round(1e3 * total64 / args.fp64_steps, 3),
                 'speedup_mixed_vs_fp64': round(value / value64, 3)},
        'accuracy_vs_fp64': {'dE_K_1/r': dEk[0], 'dE_K_erf': dEk[1], 'dE_x': dEx,
                             'dE_nl': dEnl, 'max_abs_dK': float(numpy.abs(dK).max()),
                             'max_abs_dVnl': float(numpy.abs(res32[2] - res64[2]).max())},
        'phases_ms': dict(zip(PHASES, [round(float(v), 4) for v in ph])),
        'phases_note': 'per-phase times from a profiled pass with the K build and VV10 flows '
                       'in sequence (HIP events); the timed cycles run them concurrently',
        'roofline': {'bound': 'valu',
                     'kernel': 'md_k_kernel (all classes; per operator = both operators / 2)',
                     'achieved': round(achieved / 1e12, 3), 'peak': PEAK_FP32_PACKED / 1e12,
                     'unit': 'TFLOP/s', 'frac': round(achieved / PEAK_FP32_PACKED, 4),
                     'frac_vs_nonpacked_78.6': round(achieved / PEAK_FP32_VALU, 4),
                     'traffic': measured_traffic() if args.workload == 'water27' else None,
                     'traffic_unit': 'bytes per operator (rocprofv3 FETCH_SIZE + WRITE_SIZE, '
                                     'copy-calibrated; profiles/md_k_kernel_water27_svp_traffic.txt)',
                     'tasks_kept': int(kept[1]), 'tasks_total': int(kept[0]),
                     'algorithmic_flops_per_launch': flops},
        'cpu_baseline': cpu,
    }
    print(json.dumps(line), flush=True)


def cpu_baseline(mol, g, dm, n_k=192, n_v=1024, seed=1):
    """FP64 oracle on the host cores, bounded sample: K(1/r)+K(erf) on n_k
    grid points with per-point pair screening at the device build's
    tolerance (1e-11), and the VV10 pair sum for n_v outer points, scaled to
    the full grid (both are linear in the sampled outer points)."""
    sys.path.insert(0, os.path.join(ROOT, 'oracle'))
    import oracle
    ob = oracle.OracleBasis.from_mol(mol)
    rng = numpy.random.default_rng(seed)
    n = g.size
    sel = rng.choice(n, n_k, replace=False)
    t = time.perf_counter()
    kept = 0
    for omega in (0.0, OMEGA):
        kept += oracle.k_build_screened(ob, g.coords[sel], g.weights[sel], dm, omega,
                                        pair_cut=1e-14, tol=K_TOL)[1]
    tk = (time.perf_counter() - t) * n / n_k
    # VV10 inputs from the oracle's own density on the grid (FP64)
    rho = numpy.empty((4, n))
    for i0 in range(0, n, 16384):
        ao = oracle.ao_eval(ob, g.coords[i0:i0 + 16384], deriv=1)
        tmp = ao[0] @ dm
        rho[0, i0:i0 + 16384] = numpy.einsum('pm,pm->p', tmp, ao[0])
        for d in range(3):
            rho[d + 1, i0:i0 + 16384] = 2 * numpy.einsum('pm,pm->p', tmp, ao[d + 1])
    ind = rho[0] >= oracle.THRESH
    q, W0, kappa = oracle.vv10_pointwise(rho[0][ind], rho[1:4][:, ind], g.weights[ind], NLC)[:3]
    m = int(ind.sum())
    outer = rng.choice(m, min(n_v, m), replace=False)
    t = time.perf_counter()
    oracle.vv10_sum_outer(g.coords[ind], q, W0, kappa, outer)
    tv = (time.perf_counter() - t) * m / len(outer)
    cores = len(os.sched_getaffinity(0))
    log('cpu oracle: K %.1f s + VV10 %.1f s per cycle (%d threads)' % (tk, tv, cores))
    return {'value': round(1.0 / (tk + tv), 6), 'unit': 'cycles/s', 'cores': cores,
            'kind': 'port',
            'sample': 'FP64 C oracle (OpenMP), K(1/r)+K(erf) on %d of %d grid points with '
                      'per-point pair screening at tol %g (%d point-pair evaluations kept) and '
                      'the VV10 sum for %d of %d outer points, scaled linearly; stock PySCF '
                      'is not installed' % (n_k, n, K_TOL, kept, len(outer), m),
            'seconds_per_cycle': {'k_both': round(tk, 2), 'vv10': round(tv, 2)}}


if __name__ == '__main__':
    main()
