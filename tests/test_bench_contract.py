"""bench.py's command-line contract on the CPU: one GPU only (a request for
more exits with a message and a non-zero status before touching a device),
and the modes it documents parse."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, 'bench.py'), *args],
                          capture_output=True, text=True, timeout=120)


def test_more_than_one_gpu_is_refused():
    r = _run('--gpus', '2')
    assert r.returncode != 0
    assert 'one GPU' in r.stderr
    assert r.stdout.strip() == ''


def test_help_lists_the_modes():
    r = _run('--help')
    assert r.returncode == 0
    for flag in ('--steps', '--warmup', '--gpus', '--workload', '--batch', '--serve', '--setup',
                 '--mixed-only', '--copy-probe'):
        assert flag in r.stdout, flag
