import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "green-mbpt_amd"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); runs the native library")


_torch_gpu_ready = False


def pytest_runtest_setup(item):
    """Bring torch's HIP runtime up before the first GPU test touches the library.

    torch ships its own HIP/HSA runtime next to the one libgwhip links; both serve the same
    process.  Initialising torch's first (as bench.py and the autograd harness do) keeps the
    order the same in every run, whichever test file comes first.
    """
    global _torch_gpu_ready
    if _torch_gpu_ready or item.get_closest_marker("gpu") is None:
        return
    _torch_gpu_ready = True
    import torch
    if torch.cuda.is_available():
        torch.zeros(1, device="cuda:0")
        torch.cuda.synchronize()
