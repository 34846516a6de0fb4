"""Builds the CPU emulation of libwzfh for tests: the library's own kernel and host sources,
compiled by the host clang++ against tests/emu/hip/hip_runtime.h instead of HIP, into
tests/emu/_build/libwzfh_emu.so. Test infrastructure only; never used by the product path."""
import shutil
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE.parent.parent / 'bs45_quantum_explorer_amd' / 'csrc'
OUT = HERE / '_build'
LIB = OUT / 'libwzfh_emu.so'
KERNELS = ['wzfh_stream.hip', 'wzfh_complete.hip', 'wzfh_api.hip', 'wzfh_join.hip']
HOST = ['wzfh_setup.cpp', 'wzfh_serve.cpp']


def compiler():
    for c in ('/opt/rocm/llvm/bin/clang++', shutil.which('clang++')):
        if c and Path(c).exists():
            return c
    return None


def build():
    cxx = compiler()
    if cxx is None:
        raise RuntimeError('no clang++ for the emulation build')
    OUT.mkdir(exist_ok=True)
    srcs = []
    for k in KERNELS:   # the kernel sources as host C++ (one attribute the host compiler rejects)
        dst = OUT / (Path(k).stem + '.cpp')
        dst.write_text((CSRC / k).read_text().replace('__attribute__((amdgpu_waves_per_eu(4)))', ''))
        srcs.append(str(dst))
    srcs += [str(CSRC / h) for h in HOST]
    cmd = [cxx, '-O1', '-std=c++20', '-pthread', '-fPIC', '-shared', f'-I{HERE}', f'-I{CSRC}',
           '-Wno-unknown-attributes', '-Wno-ignored-attributes', '-o', str(LIB), *srcs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f'emulation build failed:\n{r.stderr[-4000:]}')
    return LIB


if __name__ == '__main__':
    print(build())
