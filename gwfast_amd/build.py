"""Build libgwfast_hip.so in-tree for gfx950 (``python -m gwfast_amd.build``)."""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRCS = [os.path.join(HERE, 'csrc', 'gwfast_hip.hip'), os.path.join(HERE, 'csrc', 'gwf_mlp.hip')]
OUT_DIR = os.path.join(HERE, 'lib')
OUT = os.path.join(OUT_DIR, 'libgwfast_hip.so')
ARCH = os.environ.get('GWF_OFFLOAD_ARCH', 'gfx950')


def hipcc():
    exe = shutil.which('hipcc') or '/opt/rocm/bin/hipcc'
    if not os.path.exists(exe):
        raise RuntimeError('hipcc not found')
    return exe


def command(out=OUT):
    return [hipcc(), '-O3', '--offload-arch=%s' % ARCH, '-std=c++17', '-fPIC', '-shared',
            '-I', os.path.join(ROOT, 'include'), '-I', os.path.join(HERE, 'csrc'),
            # no packed-FP32 instructions anywhere in the device code
            '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
            # no machine-level hoisting of loop invariants: in the Fisher kernels it
            # hoisted 64-bit constant materialisations out of the frequency loop and
            # spilled them to scratch (46 VGPRs, 107 MB of spill writes per launch)
            '-mllvm', '-disable-machine-licm',
            *SRCS, '-o', out]


def build(verbose=True):
    os.makedirs(OUT_DIR, exist_ok=True)
    tmp = OUT + '.tmp%d' % os.getpid()
    cmd = command(tmp)
    if verbose:
        print(' '.join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError('hipcc failed (%d)' % r.returncode)
    os.replace(tmp, OUT)
    return OUT


if __name__ == '__main__':
    print(build())
