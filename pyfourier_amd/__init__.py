"""pyfourier_amd: an MI355X-native NUFFT engine with pyfourier's NUFFT API.

Drop-in entry points (same names, signatures and shape rules as
INFN-MRI/pyfourier, src/pyfourier/__init__.py:9-16):

    plan_nufft, nufft, nufft_adj      forward / adjoint NUFFT (+ subspace, B0)
    plan_spgram, sparse_gram          Toeplitz-embedded Gram operator
    pipemenon.pipe_menon_dcf          iterative density compensation (also top level)
    voronoi.voronoi                   Voronoi density compensation (2-D)
    plan_fft, fftn, ifftn             masked Cartesian FFT (+ subspace, B0)
    plan_gram, gram                   dense Cartesian Gram operator
    fft, ifft                         centred FFT utilities
    plan_spfft, sparse_fftn, sparse_ifftn  sparse (Cartesian-sampled) FFT
    grog_interp                       GRAPPA operator gridding

Beyond the reference: plan_nufft_batch / NUFFTBatch run an uneven list of
problems (own matrix sizes and trajectories) as one HIP-graph replay;
autotune_plan picks a plan's sort/spread tile by timing the candidates.

The compute path is libpfhip.so: hand-written HIP kernels for gfx950 behind a
C ABI (include/pfhip.h), called through ctypes on PyTorch-ROCm streams.
"""

from ._lib import PFError
from ._nufft import nufft, nufft_adj
from ._plan import NUFFTPlan, plan_nufft
from ._traj import normalize_coordinates
from ._b0 import mri_exp_approx
from ._sparsegram import GramMatrix, plan_spgram, sparse_gram
from ._batch import NUFFTBatch, plan_nufft_batch
from . import pipemenon
from .pipemenon import pipe_menon_dcf
from . import voronoi
from ._tune import autotune_plan, spread_tiles
from ._fft import FFTPlan, fft, fftn, ifft, ifftn, plan_fft
from ._gram import gram, plan_gram
from ._sparsefft import plan_spfft, sparse_fftn, sparse_ifftn
from ._grog import grog_interp

__all__ = [
    "PFError",
    "NUFFTPlan",
    "plan_nufft",
    "nufft",
    "nufft_adj",
    "normalize_coordinates",
    "mri_exp_approx",
    "GramMatrix",
    "plan_spgram",
    "sparse_gram",
    "NUFFTBatch",
    "plan_nufft_batch",
    "pipe_menon_dcf",
    "pipemenon",
    "voronoi",
    "autotune_plan",
    "spread_tiles",
    "FFTPlan",
    "plan_fft",
    "fftn",
    "ifftn",
    "fft",
    "ifft",
    "plan_gram",
    "gram",
    "plan_spfft",
    "sparse_fftn",
    "sparse_ifftn",
    "grog_interp",
]

__version__ = "0.1.0"
