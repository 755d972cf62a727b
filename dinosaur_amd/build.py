"""Builds the gfx950 HIP library libdinosaur_amd.so in-tree.

Each csrc/*.hip file is compiled with hipcc for gfx950 (in parallel), then
linked into dinosaur_amd/_lib/libdinosaur_amd.so.  Packed-FP32 instructions
are disabled (-packed-fp32-ops).  Usage: python -m dinosaur_amd.build
"""
from __future__ import annotations

import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
OUT_DIR = os.path.join(HERE, '_lib')
BUILD_DIR = os.path.join(HERE, '_build')
LIB = os.path.join(OUT_DIR, 'libdinosaur_amd.so')
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
ARCH = 'gfx950'

CFLAGS = [
    f'--offload-arch={ARCH}', '-O3', '-fPIC', '-std=c++17',
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-Wno-unused-command-line-argument',
]


def _compile(src: str) -> str:
  obj = os.path.join(BUILD_DIR, os.path.basename(src) + '.o')
  cmd = [HIPCC, *CFLAGS, '-c', src, '-o', obj]
  res = subprocess.run(cmd, capture_output=True, text=True)
  if res.returncode != 0:
    raise RuntimeError(f'hipcc failed for {src}:\n{res.stdout}\n{res.stderr}')
  return obj


def build(verbose: bool = False) -> str:
  os.makedirs(OUT_DIR, exist_ok=True)
  os.makedirs(BUILD_DIR, exist_ok=True)
  srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith('.hip'))
  workers = min(len(srcs), max(1, min(8, os.cpu_count() or 1)))
  with concurrent.futures.ThreadPoolExecutor(workers) as ex:
    objs = list(ex.map(_compile, srcs))
  tmp = LIB + '.tmp'
  cmd = [HIPCC, f'--offload-arch={ARCH}', '-shared', '-fPIC', *objs, '-o', tmp]
  res = subprocess.run(cmd, capture_output=True, text=True)
  if res.returncode != 0:
    raise RuntimeError(f'link failed:\n{res.stdout}\n{res.stderr}')
  os.replace(tmp, LIB)
  if verbose:
    print('built', LIB)
  return LIB


if __name__ == '__main__':
  build(verbose=True)
  sys.exit(0)
