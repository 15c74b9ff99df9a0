"""Build libhipsph.so (HIP, gfx950) in-tree.

    python offshore-sph_amd/build.py [--jobs N]

Compiles each translation unit of offshore-sph_amd/csrc with hipcc for
gfx950 and links them into offshore-sph_amd/hipsph/libhipsph.so:
  sph_engine.hip    -ffp-contract=off  (neighbour search, integrators, dt, C ABI)
  sph_pair_fp64.hip -ffp-contract=off  (double pair pass, reference operation order)
  sph_pair_fp32.hip -ffp-contract=fast-honor-pragmas (float pair pass; the exact
                    neighbour re-test and the TwoSum steps switch contraction off)
  sph_adjoint.hip   -ffp-contract=fast (tape and reverse step; not bit-matched)
  sph_batch.hip     -ffp-contract=off  (batched ensemble: the engine's arithmetic per instance)
Packed FP32 instructions are disabled for every device compile
(-Xclang -target-feature -Xclang -packed-fp32-ops).
"""
import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
OUT_DIR = os.path.join(HERE, 'hipsph')
BUILD_DIR = os.path.join(HERE, '_build')
LIB = os.path.join(OUT_DIR, 'libhipsph.so')

ARCH = os.environ.get('HIPSPH_ARCH', 'gfx950')
COMMON = ['-O3', '-fPIC', '-std=c++17', f'--offload-arch={ARCH}', '-fno-fast-math',
          '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
          '-Wall', '-Wno-unused-function', '-Wno-unused-command-line-argument', '-Wno-unused-value',
          '-I', os.path.join(HERE, '..', 'include')]
UNITS = [
    ('sph_engine.hip', ['-ffp-contract=off']),
    ('sph_pair_fp64.hip', ['-ffp-contract=off']),
    ('sph_pair_fp32.hip', ['-ffp-contract=fast-honor-pragmas']),
    ('sph_adjoint.hip', ['-ffp-contract=fast-honor-pragmas']),
    ('sph_batch.hip', ['-ffp-contract=off']),
]


def hipcc():
    for cand in (os.environ.get('HIPCC'), shutil.which('hipcc'), '/opt/rocm/bin/hipcc'):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError('hipcc not found')


def source_id():
    """A hash of every source file the library is built from (sph_version()
    reports it, so measurements can be tied to the build they came from)."""
    import hashlib
    h = hashlib.sha256()
    inc = os.path.join(HERE, '..', 'include')
    for d in (CSRC, inc):
        for f in sorted(os.listdir(d)):
            if f.endswith(('.hip', '.h')):
                with open(os.path.join(d, f), 'rb') as fh:
                    h.update(f.encode() + b'\0' + fh.read())
    return h.hexdigest()[:12]


def _compile(src, flags):
    obj = os.path.join(BUILD_DIR, os.path.splitext(src)[0] + '.o')
    cmd = [hipcc(), *COMMON, *flags, f'-DHIPSPH_BUILD_ID="{source_id()}"', '-c', os.path.join(CSRC, src),
           '-o', obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f'compile failed: {" ".join(cmd)}\n{r.stdout}\n{r.stderr}')
    return obj, r.stderr


def build(jobs=4, verbose=False):
    os.makedirs(BUILD_DIR, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(lambda u: _compile(*u), UNITS))
    if verbose:
        for _, err in results:
            if err.strip():
                print(err, file=sys.stderr)
    objs = [o for o, _ in results]
    tmp = LIB + '.tmp'
    cmd = [hipcc(), '-shared', '-fPIC', f'--offload-arch={ARCH}', *objs, '-o', tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f'link failed: {" ".join(cmd)}\n{r.stdout}\n{r.stderr}')
    os.replace(tmp, LIB)
    return LIB


if __name__ == '__main__':
    ap = argparse.ArgumentParser()
    ap.add_argument('--jobs', type=int, default=3)
    ap.add_argument('-v', '--verbose', action='store_true')
    a = ap.parse_args()
    print(build(a.jobs, a.verbose))
