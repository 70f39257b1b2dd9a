"""Cluster culling and the root interval never change a closest-hit result.

The traversal skips a cluster of a leaf's triangles when cluster_may_hit shows
that no triangle in it can be accepted by the reference's test
(xyz/geom/triangle.cpp:127-165) with t in the current interval, and
root_clip narrows a query's root interval to where the scene's root record
allows an accepted t.  This test
compiles tests/cpp/cull_fuzz.cpp on the host with the production tri_test,
cluster_may_hit and cluster_bounds, and checks millions of adversarial
ray/triangle pairs: grazing rays with |det| just above the 1e-6 threshold,
large and sliver triangles, intervals shrunk to the computed t, zero
direction components, origins inside the box; every accepted t must also lie
in root_clip's interval for a root record holding the cluster's.  A copy
built with the margin removed must fail, which shows that the cases reach the
margin."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "cull_fuzz.cpp")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


def build(out, defines=()):
    subprocess.check_call([HIPCC, "-O2", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
                           "-x", "hip", "--offload-arch=gfx950", *defines, "-o", out, SRC],
                          stderr=subprocess.DEVNULL)


@pytest.fixture(scope="module")
def fuzz_bins(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    d = tmp_path_factory.mktemp("cull")
    good, bad = str(d / "cull_fuzz"), str(d / "cull_fuzz_nomargin")
    build(good)
    build(bad, ["-DXYZ_CULL_MARGIN=0.0f"])
    return good, bad


def run(binary, cases, seed):
    r = subprocess.run([binary, str(cases), str(seed)], capture_output=True, text=True, timeout=600)
    m = re.search(r"checks (\d+) accepted (\d+) root checks (\d+) violations (\d+) culled ([\d.]+) "
                  r"worst distance/margin ([\d.e+-]+)", r.stdout)
    assert m, r.stdout + r.stderr
    return r.returncode, [float(x) for x in m.groups()]


@pytest.mark.parametrize("seed", [11, 12])
def test_culling_keeps_every_accepted_triangle(fuzz_bins, seed):
    rc, (checks, accepted, roots, violations, culled, worst) = run(fuzz_bins[0], 300000, seed)
    print(f"checks {checks:.0f}, accepted {accepted:.0f}, root intervals {roots:.0f}, "
          f"culled clusters {culled:.3f}, largest hit distance from the box / margin {worst:.3e}")
    assert rc == 0 and violations == 0
    assert accepted > 100000      # the cases do hit triangles
    assert roots > 100000         # and check the root interval
    assert 0.2 < culled < 0.9      # and the test does cull
    assert worst < 0.25            # the margin is several times the largest error seen


def test_fuzz_detects_a_missing_margin(fuzz_bins):
    rc, (_, _, _, violations, _, _) = run(fuzz_bins[1], 300000, 11)
    assert rc == 1 and violations > 0
