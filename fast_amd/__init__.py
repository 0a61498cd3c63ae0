"""fast_amd: MI355X-native tortuosity (`fast tau`) solver and reactive area density (`fast alpha`).

The compute path is libfasttau.so (hand-written HIP kernels for gfx950 behind a C ABI,
include/fast_tau.h); this package is its host-side mirror of the reference interface.
"""
from .tau import (Context, DiffusionSolverType, Dir, Precision, Server, TortuosityParams, Volume, binarize,
                  filterSegmentInfo, generate_ellipsoids, getCCLSegmentRatio, generate_spheres, getDirs, getPorosity, getReactiveAreaDensity,
                  getReactiveAreaDensityTensor, getSubvolume, getTortuosity, load_bin, loadTiffFolder, make_params,
                  rasterizeSpheres, save_bin, solve_batch)

__all__ = [
    "Context", "DiffusionSolverType", "Dir", "Precision", "Server", "TortuosityParams", "Volume", "binarize",
    "filterSegmentInfo", "generate_ellipsoids", "getCCLSegmentRatio", "generate_spheres", "getDirs", "getPorosity", "getReactiveAreaDensity",
    "getReactiveAreaDensityTensor", "getSubvolume", "getTortuosity", "load_bin",
    "loadTiffFolder", "make_params", "rasterizeSpheres", "save_bin", "solve_batch",
]
