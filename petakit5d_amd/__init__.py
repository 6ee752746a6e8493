"""petakit5d_amd -- MI355X (gfx950) engine for PetaKit5D's RL / OMW deconvolution hot path.

The compute lives in libpkdecon.so (hand-written HIP kernels behind the C-ABI
declared in include/pkdecon.h); this package is its Python front end with the
reference's MATLAB call signatures.
"""
from ._lib import Context, PkDeconError, default_context, load  # noqa: F401
from .backprojector import BackProjector, omw_backprojector_generation  # noqa: F401
from .batch import decon_batch  # noqa: F401
from .chunks import XR_zarrChunkCoordinatesExtraction, decon_large_in_memory, edge_erosion_mask  # noqa: F401
from .psfgen import psf_gen_new  # noqa: F401
from .rldecon import RLdecon_volume  # noqa: F401
from .tiff_io import readtiff, writetiff  # noqa: F401
from .zarr_io import decon_zarr, read_zarr, write_zarr  # noqa: F401
from .serve import Server  # noqa: F401
from .deskew import deskewFrame3D  # noqa: F401
from .rotate import rotateFrame3D  # noqa: F401
from .deskew_rotate import deskewRotateFrame3D  # noqa: F401
from .decon import (  # noqa: F401
    crop_psf_to_data,
    decon_lucy_function,
    decon_lucy_omw_function,
    decon_psf2otf,
    fft_c2r,
    fft_r2c,
    findGoodFactorNumber,
)

__all__ = [
    "Context",
    "PkDeconError",
    "default_context",
    "load",
    "decon_lucy_function",
    "decon_lucy_omw_function",
    "decon_psf2otf",
    "findGoodFactorNumber",
    "omw_backprojector_generation",
    "BackProjector",
    "decon_batch",
    "XR_zarrChunkCoordinatesExtraction",
    "decon_large_in_memory",
    "edge_erosion_mask",
    "RLdecon_volume",
    "psf_gen_new",
    "read_zarr",
    "write_zarr",
    "decon_zarr",
    "readtiff",
    "deskewFrame3D",
    "rotateFrame3D",
    "deskewRotateFrame3D",
    "writetiff",
    "Server",
    "crop_psf_to_data",
    "fft_r2c",
    "fft_c2r",
]
