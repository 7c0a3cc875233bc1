import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

# One hardware queue for the whole test process.  The quadrature-point kernels keep their per-lane
# matrices in private memory (8.6 KB per lane for 12 slip systems, 25 KB for 24, 5-6 KB for 48
# beside a workspace in global memory; the one-lane pseudo-inverse for n = 48, 37.7 KB), and the
# runtime keeps a queue's scratch sized for
# the largest private segment that queue has run.  With the default four queues, the suite's many
# contexts (one stream each) leave several queues holding such reservations, and a later kernel
# whose queue must grow its scratch fails with HSA_STATUS_ERROR_OUT_OF_RESOURCES.  Set before the
# HIP runtime starts (this file is imported before any test module).
os.environ["GPU_MAX_HW_QUEUES"] = "1"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
