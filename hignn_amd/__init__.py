"""MI355X implementation of the H-HIGNN mobility matvec and its neighbours.

libhh.so (HIP kernels and the C ABI) and PyTorch each bring a HIP runtime
with the soname libamdhip64.so.7, and a process keeps whichever copy it
loads first. PyTorch cannot initialise its devices on another copy, while
libhh.so runs on PyTorch's. So when PyTorch is installed it is imported
here, before any submodule loads libhh.so, and both work in one process in
any import order.
"""
try:
    import torch  # noqa: F401
except ImportError:  # libhh.so alone: its own runtime
    pass
