"""Build the gfx950 library (libhsl2.so) and the droid_backends torch extension in place.

    python hi-slam2_amd/build.py

Both artefacts land next to this file: ``libhsl2.so`` (the C ABI of include/hsl2.h,
HIP kernels compiled by hipcc for gfx950 only) and
``droid_backends<EXT_SUFFIX>`` (a pybind11 torch extension compiled with g++ that
links libhsl2.so with an $ORIGIN rpath).  No JIT cache is involved.
"""
import concurrent.futures as cf
import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, 'csrc')
BUILD = os.path.join(HERE, '_build')
INCLUDE = os.path.join(ROOT, 'include')
ROCM = os.environ.get('ROCM_PATH', '/opt/rocm')
ARCH = 'gfx950'

LIB_SOURCES = ['status.cpp', 'ba_plan.cpp', 'ba.hip', 'gn_step.hip', 'projective.hip', 'chol.hip', 'corr_lookup.hip', 'altcorr_pyramid.hip',
               'corr_build.hip', 'altcorr.hip', 'geometry.hip', 'pgba.hip']

HIP_FLAGS = [
    '-O3', '-std=c++17', '-fPIC', f'--offload-arch={ARCH}', '-mcode-object-version=5',
    # packed FP32 instructions stay off: plain float code only
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-Wall', '-Wno-unused-function', '-Wno-unused-variable',
]


def torch_lib_dir():
    import torch
    return os.path.join(os.path.dirname(torch.__file__), 'lib')


def run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(' '.join(cmd) + '\n' + r.stdout)
        raise RuntimeError(f'build step failed: {cmd[0]} ... {cmd[-1]}')
    return r.stdout


def compile_lib(jobs=8):
    os.makedirs(BUILD, exist_ok=True)
    hipcc = os.path.join(ROCM, 'bin', 'hipcc')
    objs, cmds = [], []
    for src in LIB_SOURCES:
        obj = os.path.join(BUILD, src + '.o')
        objs.append(obj)
        lang = ['-x', 'hip'] if src.endswith('.hip') else []
        cmds.append([hipcc] + HIP_FLAGS + [f'-I{INCLUDE}', f'-I{CSRC}', '-c'] + lang +
                    [os.path.join(CSRC, src), '-o', obj])
    with cf.ThreadPoolExecutor(jobs) as ex:
        list(ex.map(run, cmds))
    tl = torch_lib_dir()
    out = os.path.join(HERE, 'libhsl2.so')
    # Resolve the HIP runtime from torch's bundled ROCm first so that a process that also
    # imports torch loads one runtime only.
    run([hipcc, '-shared', f'--offload-arch={ARCH}', '-o', out + '.tmp'] + objs +
        [f'-L{tl}', f'-Wl,-rpath,{tl}', f'-L{ROCM}/lib', f'-Wl,-rpath,{ROCM}/lib', '-lamdhip64'])
    os.replace(out + '.tmp', out)
    return out


def compile_ext():
    import torch.utils.cpp_extension as ce
    suffix = sysconfig.get_config_var('EXT_SUFFIX')
    out = os.path.join(HERE, 'droid_backends' + suffix)
    inc = ce.include_paths() + [sysconfig.get_paths()['include'], INCLUDE, f'{ROCM}/include']
    tl = torch_lib_dir()
    cmd = ['g++', '-O2', '-std=c++17', '-fPIC', '-shared',
           '-D__HIP_PLATFORM_AMD__=1', '-DUSE_ROCM=1', '-DTORCH_EXTENSION_NAME=droid_backends',
           '-DTORCH_API_INCLUDE_EXTENSION_H', '-D_GLIBCXX_USE_CXX11_ABI=1', '-w']
    cmd += [f'-I{p}' for p in inc]
    cmd += [os.path.join(CSRC, 'ext', 'droid_backends.cpp'), '-o', out + '.tmp']
    cmd += [f'-L{HERE}', '-lhsl2', "-Wl,-rpath,$ORIGIN", f'-L{tl}', f'-Wl,-rpath,{tl}',
            '-lc10', '-ltorch', '-ltorch_cpu', '-ltorch_python', '-lc10_hip', '-ltorch_hip',
            f'-L{ROCM}/lib', '-lamdhip64']
    run(cmd)
    os.replace(out + '.tmp', out)
    return out


def build():
    lib = compile_lib()
    ext = compile_ext()
    return lib, ext


if __name__ == '__main__':
    for p in build():
        print(p)
