"""The installable package: ``pip wheel --no-build-isolation`` of
sarlab-tdc_amd builds offline with the hipcc Makefile and packages
``sarlabtdc`` with its three shared objects; installed into a fresh directory
it imports without PYTHONPATH pointing into the repository (the reference
ships the same package name, py-sarlabtdc/setup.py, setup.cfg)."""
import os
import subprocess
import sys
import zipfile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_wheel_builds_installs_and_imports(tmp_path):
    whl_dir = tmp_path / "wheels"
    r = subprocess.run([sys.executable, "-m", "pip", "wheel", "--no-build-isolation", "--no-deps",
                        "-q", "-w", str(whl_dir), os.path.join(ROOT, "sarlab-tdc_amd")],
                       capture_output=True, text=True, timeout=600, cwd=str(tmp_path))
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    wheels = list(whl_dir.glob("sarlabtdc-*.whl"))
    assert len(wheels) == 1, wheels
    assert "linux_x86_64" in wheels[0].name  # a platform wheel: it carries compiled code
    names = zipfile.ZipFile(wheels[0]).namelist()
    for want in ("sarlabtdc/__init__.py", "sarlabtdc/libsarlabtdc_amd.so", "sarlabtdc/_sarlabtdc",
                 "sarlabtdc/libtdcscene.so", "sarlabtdc/scene.py"):
        assert any(want in n for n in names), (want, names)
    target = tmp_path / "site"
    r = subprocess.run([sys.executable, "-m", "pip", "install", "--no-deps", "--no-index", "-q",
                        "--target", str(target), str(wheels[0])],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    env = {k: v for k, v in os.environ.items() if k != "PYTHONPATH"}
    env["PYTHONPATH"] = str(target)
    code = ("import sarlabtdc, os; "
            "assert os.path.dirname(sarlabtdc.__file__) == os.path.join(%r, 'sarlabtdc'); "
            "print(sarlabtdc.device_count() >= 0, sarlabtdc.TdcProcessor.start.__doc__ is not None)"
            % str(target))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120,
                       env=env, cwd=str(tmp_path))
    assert r.returncode == 0, r.stderr[-2000:]
    assert r.stdout.strip() == "True True"
