"""The C-ABI library loads and exports every function include/gpaw_fd.h
declares (no GPU calls)."""
import ctypes
import os
import re

from gpaw_amd import _lib

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), '..'))


def declared_functions():
    text = open(os.path.join(ROOT, 'include', 'gpaw_fd.h')).read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    return sorted(set(re.findall(r'\b(gfd_[a-z0-9_]+)\s*\(', text)))


def test_header_lists_functions():
    names = declared_functions()
    assert 'gfd_operator_apply' in names and len(names) >= 18


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_table_matches_header():
    assert sorted(_lib.SIGNATURES) == declared_functions()


def test_load_and_version_without_gpu():
    lib = _lib.load()
    assert lib.gfd_abi_version() == 1
    assert lib.gfd_set_tuning(b'no_such_key', 0) == _lib.GFD_EINVAL
    assert b'unknown key' in lib.gfd_last_error()


def test_documented_tuning_keys_are_known():
    """Every key the header documents for gfd_set_tuning is recognised (a
    range error at most, never 'unknown key'), and the fused
    preconditioner's tile key rejects out-of-range fields.  Run in a child
    process: the keys are process-wide settings."""
    import subprocess
    import sys
    text = open(os.path.join(ROOT, 'include', 'gpaw_fd.h')).read()
    doc = text[:text.index('int gfd_set_tuning(')]
    doc = doc[doc.rindex('/*'):]
    keys = sorted(set(re.findall(r'"([a-z_0-9]+)"', doc)))
    assert 'precond_fine' in keys and 'batch_tile' in keys
    code = f'''
import ctypes, sys
sys.path.insert(0, {ROOT!r})
from gpaw_amd import _lib
lib = _lib.load()
for k in {keys!r}:
    lib.gfd_set_tuning(k.encode(), -(1 << 40))
    assert b'unknown key' not in lib.gfd_last_error(), k
assert lib.gfd_set_tuning(b'precond_fine_tile', 3) == _lib.GFD_EINVAL
assert lib.gfd_set_tuning(b'precond_fine_tile', 2 | 8 << 8) == 0
assert lib.gfd_set_tuning(b'precond_fine', 0) == 0
'''
    r = subprocess.run([sys.executable, '-c', code], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
