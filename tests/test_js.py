"""The JavaScript package (three-mesh-bvh_amd/src) and its Node-API addon, host side only."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ADDON = os.path.join(ROOT, "three-mesh-bvh_amd", "lib", "mbvh.node")


@pytest.mark.skipif(shutil.which("node") is None, reason="node is not installed")
def test_js_host_checks():
    assert os.path.exists(ADDON), "build the addon first (__graft_entry__.build())"
    r = subprocess.run(["node", os.path.join(ROOT, "tests", "js", "cpu_checks.js")], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr


@pytest.mark.skipif(shutil.which("node") is None, reason="node is not installed")
def test_addon_exports():
    r = subprocess.run(["node", "-e", f"const m=require({ADDON!r}); console.log(Object.keys(m).sort().join(','))"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    names = set(r.stdout.strip().split(","))
    assert {"build", "uploadMesh", "freeMesh", "raycastFirst", "raycast", "closestPoint", "setSideMask",
            "deviceCount", "raycastFirstBackward", "closestPointBackward", "batchCreate", "freeBatch",
            "batchRaycastFirst"} <= names
