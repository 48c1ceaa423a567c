import os
import sys

# torch ships its own HIP runtime under the same soname as /opt/rocm's. A process that uses both
# torch and libmbvh.so must load torch's first (libmbvh.so then binds to it); if libmbvh.so loads
# first, torch later finds no HIP device. Import torch before any test loads the library.
try:
    import torch  # noqa: F401
except ImportError:
    pass

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "three-mesh-bvh_amd", "python"), os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); run with -m gpu")
