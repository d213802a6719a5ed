"""Soundness of the group bounds that let the GPU clustering kernel skip
members (csrc/ppf_cluster.h): every member the bound rejects fails the exact
pose-pair test of rot_clustering_kernel (reference kernel.cu:735-756) for all
poses of the group, including members built just inside and outside both
thresholds, non-unit, negated, zero and NaN quaternions, with the group's
centre at its first pose or at its mean (as the kernel builds it); and that
the 4-D Hilbert order of the chunks is a curve through neighbouring cells."""
import os
import subprocess

from conftest import PKG, ROOT


def test_cluster_skip_bound_is_sound(tmp_path):
    exe = str(tmp_path / "cluster_bound_test")
    src = os.path.join(ROOT, "tests", "cpp", "cluster_bound_test.cpp")
    r = subprocess.run(["g++", "-std=c++17", "-O2", "-w", "-ffp-contract=off",
                        "-I" + os.path.join(PKG, "csrc"), src, "-o", exe],
                       capture_output=True, text=True, timeout=180)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([exe], capture_output=True, text=True, timeout=180)
    assert r.returncode == 0, r.stdout[-2000:]
    assert "violations 0" in r.stdout and "hilbert ok" in r.stdout
