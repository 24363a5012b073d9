"""bench.py command line: one GPU only, and argument errors before any GPU work (CPU test)."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, 'bench.py'), *args], capture_output=True,
                          text=True, timeout=120, cwd=ROOT)


def test_more_than_one_gpu_is_refused_with_a_message():
    r = _run('--gpus', '2')
    assert r.returncode == 2
    assert 'one GPU only' in r.stderr and r.stdout == ''


def test_unknown_extra_is_an_argument_error():
    r = _run('--extras', 'lgs,nonsense')
    assert r.returncode == 2
    assert 'unknown extras' in r.stderr and 'nonsense' in r.stderr
