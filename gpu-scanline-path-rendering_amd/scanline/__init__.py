"""MI355X comb-scanline path renderer: Python side of the C ABI (include/scanline.h)."""
from ._abi import (CURVE_CUBIC, CURVE_LINEAR, CURVE_QUADRATIC, CURVE_RATIONAL, FILL_EVEN_ODD,
                   FILL_NONZERO, ScanlineError, lib)
from .batch import BatchRasterizer, batch_workload
from .rasterizer import VGRasterizer, createRasterizer
from .scene import (FT_COLOR, FT_GRADIENT, FT_NONE, ST_COLOR, ST_GRADIENT, ST_NONE, VGContainer, SceneArrays,
                    config_scene, generate, rgba)
from . import animation, rvg, svg, vgfile
from .animation import ChordAnimation, ChordKeyframes, chord_keyframes, refugee_diagrams

__all__ = [
    "CURVE_CUBIC", "CURVE_LINEAR", "CURVE_QUADRATIC", "CURVE_RATIONAL", "FILL_EVEN_ODD",
    "FILL_NONZERO", "ScanlineError", "lib", "VGRasterizer", "createRasterizer", "VGContainer",
    "SceneArrays", "config_scene", "generate", "rgba", "BatchRasterizer", "batch_workload", "rvg", "svg", "vgfile",
    "FT_NONE", "FT_COLOR", "FT_GRADIENT", "ST_NONE", "ST_COLOR", "ST_GRADIENT",
    "animation", "ChordAnimation", "ChordKeyframes", "chord_keyframes", "refugee_diagrams",
]
