"""pisco4cmb_amd: MI355X-native pixel-space beam convolution (PISCO's
GPUConvolver path) -- hand-written gfx950 HIP kernels behind a C-ABI
(include/pisco.h), with pypisco as the Python binding.

    from pisco4cmb_amd import pypisco          # Sky, Scan, PolBeam, GPUConvolver, ...
    from pisco4cmb_amd import torch_ops        # autograd Function over device tensors
"""
__version__ = "0.1.0"
