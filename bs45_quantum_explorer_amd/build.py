"""Builds libwzfh.so (HIP kernels + host runtime + C ABI) and the wzfh_match CLI for gfx950.

Run as `python -m bs45_quantum_explorer_amd.build`; __graft_entry__.build() calls build().
Every HIP translation unit is compiled with packed-FP32 instructions disabled and with
FP contraction off (the FP64 hall_ok path must round exactly like the reference).
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / 'csrc'
OBJ = PKG / '_obj'
LIB = PKG / 'libwzfh.so'
CLI = PKG / 'wzfh_match'
ARCH = os.environ.get('WZFH_ARCH', 'gfx950')

HIP_FLAGS = ['-O3', '-std=c++17', f'--offload-arch={ARCH}', '-fPIC', '-ffp-contract=off',
             '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops', '-Wno-unused-result']
SOURCES = ['wzfh_stream.hip', 'wzfh_complete.hip', 'wzfh_api.hip', 'wzfh_join.hip', 'wzfh_setup.cpp', 'wzfh_serve.cpp']


def hipcc():
    for cand in (os.environ.get('HIPCC'), '/opt/rocm/bin/hipcc', shutil.which('hipcc')):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError('hipcc not found')


def _compile(src):
    out = OBJ / (Path(src).stem + '.o')
    cmd = [hipcc(), *HIP_FLAGS, '-c', str(CSRC / src), '-o', str(out)]
    if src.endswith('.cpp'):
        cmd = [hipcc(), '-O3', '-std=c++17', '-fPIC', '-pthread', '-ffp-contract=off', '-c', str(CSRC / src), '-o', str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f'compile failed: {" ".join(cmd)}\n{r.stderr}')
    return out


def build(verbose=True):
    OBJ.mkdir(exist_ok=True)
    with ThreadPoolExecutor(4) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB.with_suffix('.so.tmp')
    cmd = [hipcc(), f'--offload-arch={ARCH}', '-shared', '-fPIC', '-pthread', '-o', str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f'link failed: {" ".join(cmd)}\n{r.stderr}')
    os.replace(tmp, LIB)
    cli_src = CSRC / 'wzfh_match.cpp'
    if cli_src.exists():
        cmd = [hipcc(), '-O2', '-std=c++17', str(cli_src), '-o', str(CLI), f'-L{PKG}', '-lwzfh',
               f'-Wl,-rpath,$ORIGIN']
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f'CLI build failed: {" ".join(cmd)}\n{r.stderr}')
    if verbose:
        print(f'built {LIB}' + (f' and {CLI}' if cli_src.exists() else ''))
    return LIB


if __name__ == '__main__':
    build()
    sys.exit(0)
