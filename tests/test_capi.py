"""libhipsph.so loads and exports every entry point include/hipsph.h declares.

No compute call is made here (no GPU in this container): this is the
"does the C ABI exist" check. The GPU tests call through the same binding.
"""
import ctypes
import os
import re

from conftest import ROOT

HEADER = os.path.join(ROOT, 'include', 'hipsph.h')


def declared():
    src = open(HEADER).read()
    src = re.sub(r'/\*.*?\*/', '', src, flags=re.S)
    return sorted(set(re.findall(r'\b(sph_[a-z0-9_]+)\s*\(', src)))


def test_header_declares_entry_points():
    names = declared()
    assert 'sph_ctx_create' in names and 'sph_step' in names and 'sph_nn_dump' in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    from hipsph import _lib
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, f'not exported: {missing}'


def test_binding_covers_header():
    from hipsph import _lib
    assert sorted(_lib.EXPORTED) == declared()
    lib = _lib.load()
    assert b'gfx950' in lib.sph_version()


def test_struct_layout_matches_header(tmp_path):
    """Every field offset of sph_params_t / sph_state_t / sph_adjoint_t as gcc lays out the header."""
    import shutil
    import subprocess
    import pytest
    from hipsph import _lib
    if not shutil.which('gcc'):
        pytest.skip('gcc not available')
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "hipsph.h"', 'int main(void) {']
    for st in (_lib.sph_params_t, _lib.sph_state_t, _lib.sph_adjoint_t):
        cname = st.__name__
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f, _ in st._fields_:
            lines.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    lines.append('return 0; }')
    src = tmp_path / 'probe.c'
    src.write_text('\n'.join(lines))
    exe = tmp_path / 'probe'
    subprocess.run(['gcc', '-I', os.path.join(ROOT, 'include'), str(src), '-o', str(exe)], check=True)
    got = dict(line.rsplit(' ', 1) for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                              check=True).stdout.split('\n') if line)
    for st in (_lib.sph_params_t, _lib.sph_state_t, _lib.sph_adjoint_t):
        cname = st.__name__
        assert int(got[cname]) == ctypes.sizeof(st)
        for f, _ in st._fields_:
            assert int(got[f'{cname}.{f}']) == getattr(st, f).offset, f


def test_invalid_params_fail_loudly_without_gpu():
    """Argument validation happens before any device call."""
    import pytest
    from hipsph import _lib
    p = _lib.sph_params_t()
    p.n = 0
    out = ctypes.c_void_p()
    rc = _lib.load().sph_ctx_create(ctypes.byref(p), ctypes.byref(out))
    assert rc == -1
    assert b'n=0' in _lib.load().sph_last_error()
    with pytest.raises(_lib.HipSphError):
        _lib.call('sph_kernel_evaluate', 7, 1, None, None, None, None)
