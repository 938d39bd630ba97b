import os
import sys

# The factorisation keeps up to twelve HIP streams per GPU busy, as the
# benchmark does; each needs a hardware queue of its own.  HIP reads this
# once, at its first call, so it is set before any test module loads.
os.environ["GPU_MAX_HW_QUEUES"] = str(min(32, max(16, int(os.environ.get("GPU_MAX_HW_QUEUES") or 0))))

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X / gfx950)")
