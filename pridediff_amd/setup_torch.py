"""Build entry mirroring deeprecon/setup_torch.py.

The reference builds its ``radontorch`` extension with
``torch.utils.cpp_extension.CUDAExtension`` over ``cuda/**/*.cu``,
``cpu/**/*.cpp`` and ``interface/RadonTorch.cpp`` (deeprecon/setup_torch.py:
6-25).  Here ``python setup_torch.py build_ext`` builds the same module name
in the checkout (this directory plays the reference's deeprecon/ project
root, with the ``deeprecon`` package beside it): ``libdd.so`` (hand-written HIP for gfx950, hipcc) and
``radontorch.so`` (the torch-op boundary, g++ against torch's headers, no
.cu sources, no hipify).  Nothing is installed into site-packages; callers
use ``pridediff_amd.install_dropin()`` to make ``radontorch`` and
``deeprecon`` importable.
"""
import os
import sys

from setuptools import setup
from setuptools.command.build_ext import build_ext

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


class BuildRadontorch(build_ext):
    def run(self):
        sys.path.insert(0, ROOT)
        from pridediff_amd.build import build_all
        build_all()


if __name__ == "__main__":
    setup(name="deeprecon", version="2.0.0", packages=[],
          cmdclass={"build_ext": BuildRadontorch})
