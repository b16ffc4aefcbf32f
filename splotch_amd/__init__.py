"""splotch_amd — MI355X (gfx950) particle-splat renderer with Splotch's render interface.

The hot path (camera transform + perspective projection + palette colouring,
depth sort, tile binning and Gaussian splat; splotch/splotch_host.cc:58-697)
runs as hand-written HIP kernels in libsplotch_hip.so behind the C-ABI of
include/splotch_hip.h.  This package holds the sources (csrc/), the in-tree
build (build.py), the ctypes binding (capi.py) and the seeded synthetic
workloads (synth.py).
"""
from . import synth  # noqa: F401
from .capi import (Camera, Norm, Params, Post, Renderer, SplhipError,  # noqa: F401
                   bin_block_count, camera_setup, exptable, norm_from_par, params,
                   params_from_par, post, post_from_par, scene_file_parse, scene_views,
                   tga_header, to_device, to_particle_sim)

__all__ = ["Camera", "Norm", "Params", "Post", "Renderer", "SplhipError", "bin_block_count",
           "camera_setup", "exptable", "norm_from_par", "params", "params_from_par", "post",
           "post_from_par", "scene_file_parse", "scene_views", "synth", "tga_header",
           "to_device", "to_particle_sim"]
