"""MI355X implementation of the H-HIGNN mobility matvec and its neighbours.

The package loads libhh.so (HIP kernels and the C ABI) and needs nothing
else on the GPU side: device buffers for the device-pointer entry points
come from the library itself (include/hh_device.h, abi.DeviceArray).

libhh.so and PyTorch each bring a HIP runtime with the soname
libamdhip64.so.7, and a process keeps whichever copy it loads first;
PyTorch only initialises its devices on its own copy, while libhh.so runs
on either.  A process that also uses PyTorch on the GPU imports torch before
this package loads libhh.so.
"""
