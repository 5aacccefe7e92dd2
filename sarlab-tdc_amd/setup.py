"""Installable package of the MI355X TDC image former: ``pip install
--no-build-isolation ./sarlab-tdc_amd`` (from the repository root; no network
is needed: the build runs this directory's Makefile with hipcc for gfx950 and
packages ``sarlabtdc`` with the three shared objects it builds). The
reference builds the same package name with CMake (py-sarlabtdc/setup.py,
setup.cfg)."""
import os
import subprocess

from setuptools import setup
from setuptools.command.build_py import build_py
from setuptools.dist import Distribution

HERE = os.path.dirname(os.path.abspath(__file__))
INCLUDE = os.path.join(os.path.dirname(HERE), "include")


class BuildWithMake(build_py):
    """Runs the hipcc Makefile (library, Python module, scene simulator) before
    the package files are collected."""

    def run(self):
        if not os.path.isdir(INCLUDE):
            raise RuntimeError(f"the C ABI headers are expected in {INCLUDE}")
        jobs = str(min(8, os.cpu_count() or 1))
        subprocess.run(["make", "-j", jobs, "-C", HERE, f"INC=-I{INCLUDE} -Icsrc"], check=True)
        super().run()


class BinaryDistribution(Distribution):
    """The package carries compiled code: platform wheels, not pure ones."""

    def has_ext_modules(self):
        return True


setup(
    name="sarlabtdc",
    version="2.0.0",
    description="Time-domain correlation SAR image formation on AMD Instinct MI355X (gfx950)",
    packages=["sarlabtdc"],
    package_data={"sarlabtdc": ["*.so"]},
    python_requires=">=3.8",
    install_requires=["numpy"],
    cmdclass={"build_py": BuildWithMake},
    distclass=BinaryDistribution,
    zip_safe=False,
)
