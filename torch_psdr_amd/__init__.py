"""torch_psdr_amd: MI355X (gfx950) drop-in for torch_psdr's differentiable path tracer.

    import torch_psdr_amd as dr
    scene = dr.Scene([camera], [mesh], [light])
    img = dr.PathIntegrator(spp_interior=16, max_bounce=3).renderD(scene)[0]
"""
from .api import (AreaLight, BaseCamera, DiffuseBsdfMaterial, DiffuseMaterial, EnvLight, Light, Material, Mesh,
                  OrthographicCamera, PathIntegrator, PerspectiveCamera, PrincipledMaterial, Scene, SH9Light,
                  Texture, VarContex)
from .io import load_obj, read_mtl, read_obj
from .tune import autotune, get_tuning, set_tuning

__all__ = ["AreaLight", "BaseCamera", "DiffuseBsdfMaterial", "DiffuseMaterial", "EnvLight", "Light", "Material",
           "Mesh", "OrthographicCamera", "PathIntegrator", "PerspectiveCamera", "PrincipledMaterial", "Scene",
           "SH9Light", "Texture", "VarContex", "autotune", "get_tuning", "load_obj", "read_mtl", "read_obj", "set_tuning"]
