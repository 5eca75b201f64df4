"""crender_amd -- MI355X-native path tracer for CRender's render hot path.

The compute path is the HIP library ``libcrt.so`` (C ABI in include/crt.h); this package
holds its Python mirror of the reference's scene / renderer interface and the seeded
scene generators.  Importing the package does not load the library; the first use of
:class:`Context` does, and fails loudly when the library is missing.
"""
from ._abi import CrtError, header_functions, lib  # noqa: F401
from .render import (Context, Nrc, ObjModel, Renderer, Scene, autotune, denoise, post_process,  # noqa: F401
                     render_batch, scene_from_spec, write_image)
from . import scenes  # noqa: F401

__all__ = ["Context", "Scene", "Renderer", "Nrc", "ObjModel", "render_batch", "autotune", "scene_from_spec", "write_image", "post_process", "denoise", "scenes", "CrtError", "lib", "header_functions"]
