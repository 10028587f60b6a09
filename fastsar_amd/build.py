"""Build of the HIP library libfsh.so for gfx950, in-tree (fastsar_amd/_lib/), with hipcc directly.

    python -m fastsar_amd.build          # or __graft_entry__.build()

Flags: -O3 for gfx950 only; packed float32 instructions disabled (-packed-fp32-ops), which hipcc would otherwise
form from ordinary float code; hipFFT for the exact former's range compression.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, 'csrc')
OUT = os.path.join(HERE, '_lib')
SOURCES = ['fsh_api.hip', 'host_stage.hip', 'ffbp_kernels.hip', 'range_mfma.hip', 'level_mfma.hip', 'fused_kernels.hip', 'fused_kernels_k2p2.hip', 'fused_kernels_k2p3.hip', 'fused_kernels_k3p2.hip', 'fused_kernels_k3p3.hip', 'exact_kernels.hip', 'gate_kernels.hip', 'batch.hip', 'learned_kernels.hip', 'sim_kernels.hip']
ARCH = os.environ.get('FSH_ARCH', 'gfx950')


def hipcc():
    for c in (os.environ.get('HIPCC'), '/opt/rocm/bin/hipcc', shutil.which('hipcc')):
        if c and os.path.exists(c):
            return c
    raise RuntimeError('hipcc not found (ROCm is needed to build the HIP library)')


def flags():
    return ['-O3', '-std=c++17', f'--offload-arch={ARCH}', '-fPIC', '-Xclang', '-target-feature', '-Xclang',
            '-packed-fp32-ops', '-I', SRC, '-Wno-unused-result']


def build(verbose=False):
    """Compile every source and link libfsh.so; returns its path."""
    os.makedirs(OUT, exist_ok=True)
    objs, procs = [], []
    cc = hipcc()
    for s in SOURCES:                      # the sources compile in parallel
        src = os.path.join(SRC, s)
        if not os.path.exists(src):
            continue
        obj = os.path.join(OUT, s.replace('.hip', '.o'))
        cmd = [cc, *flags(), '-c', src, '-o', obj]
        if verbose:
            print(' '.join(cmd), flush=True)
        procs.append((subprocess.Popen(cmd), cmd))
        objs.append(obj)
    failed = [cmd for pr, cmd in procs if pr.wait() != 0]
    if failed:
        raise subprocess.CalledProcessError(1, failed[0])
    lib = os.path.join(OUT, 'libfsh.so')
    tmp = lib + '.tmp'
    cmd = [cc, f'--offload-arch={ARCH}', '-shared', '-fPIC', '-o', tmp, *objs, '-lhipfft']
    if verbose:
        print(' '.join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, lib)
    for o in objs:
        os.remove(o)
    return lib


if __name__ == '__main__':
    print(build(verbose='-v' in sys.argv))
