"""Builds libwb97mv_amd.so in-tree (gfx950, hipcc, ahead of time).

    python pyscf-wb97mv-fast_amd/build_lib.py

Generates md_gen.h (csrc/gen_md.py), compiles the three HIP translation
units in parallel and links pyscf_wb97mv_fast_amd/libwb97mv_amd.so.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, 'csrc')
BUILD = os.path.join(HERE, 'build')
OUT = os.path.join(HERE, 'pyscf_wb97mv_fast_amd', 'libwb97mv_amd.so')
UNITS = ['wb_api', 'wb_kmd', 'wb_kmisc']
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
FLAGS = ['--offload-arch=gfx950', '-O3', '-std=c++17', '-fPIC', '-fvisibility=hidden',
         # no packed FP32 instructions anywhere (correctness policy)
         '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
         '-I', CSRC, '-I', os.path.join(BUILD, 'gen'), '-I', os.path.join(ROOT, 'include')]


def build(verbose=False, out=OUT, defines=()):
    """defines: extra -D macros for variant builds written to another
    ``out`` (objects go to their own build folder)."""
    tag = '_'.join(d.replace('=', '') for d in defines)
    obj_dir = os.path.join(BUILD, 'obj' + ('_' + tag if tag else ''))
    os.makedirs(os.path.join(BUILD, 'gen'), exist_ok=True)
    os.makedirs(obj_dir, exist_ok=True)
    subprocess.check_call([sys.executable, os.path.join(CSRC, 'gen_md.py'),
                           os.path.join(BUILD, 'gen', 'md_gen.h')])

    def compile_unit(u):
        obj = os.path.join(obj_dir, u + '.o')
        cmd = [HIPCC] + FLAGS + ['-D' + d for d in defines] + \
            ['-c', os.path.join(CSRC, u + '.hip'), '-o', obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError('hipcc failed for %s:\n%s' % (u, r.stderr[-4000:]))
        if verbose and r.stderr:
            print(r.stderr)
        return obj

    with concurrent.futures.ThreadPoolExecutor(len(UNITS)) as ex:
        objs = list(ex.map(compile_unit, UNITS))
    tmp = out + '.tmp'
    subprocess.check_call([HIPCC, '--offload-arch=gfx950', '-shared', '-fPIC', '-o', tmp] + objs +
                          ['-L/opt/rocm/lib', '-lrocsolver', '-lrocblas',
                           '-Wl,-rpath,/opt/rocm/lib'])
    os.replace(tmp, out)
    return out


if __name__ == '__main__':
    print(build(verbose='-v' in sys.argv))
