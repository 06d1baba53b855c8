import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")
    config.addinivalue_line("markers", "slow: long CPU-side oracle runs")


def pytest_sessionstart(session):
    # torch ships its own HIP runtime.  In one process, torch must open the GPU
    # before libdragon_gpu.so does: in the other order torch reports "No HIP
    # GPUs are available".  Tests that build large states with torch
    # (test_gpu_full_configs) therefore need torch started first.
    try:
        import torch
    except ImportError:
        return
    if torch.cuda.device_count() > 0:
        torch.cuda.init()
