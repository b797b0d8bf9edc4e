"""`import torch_psdr as dr` (the reference's import name,
example/optimize_texture.py:1) gives this package's API unchanged (no GPU)."""


def test_reference_import_name_reexports_the_api():
    import torch_psdr as dr
    import torch_psdr_amd as amd
    # the classes of the reference module (src/python/torch_psdr.py:84-556)
    ref_classes = ["VarContex", "BaseCamera", "PerspectiveCamera", "OrthographicCamera", "Texture", "Material",
                   "DiffuseMaterial", "DiffuseBsdfMaterial", "PrincipledMaterial", "Mesh", "Light", "AreaLight",
                   "EnvLight", "SH9Light", "Scene", "PathIntegrator"]
    for name in ref_classes:
        assert getattr(dr, name) is getattr(amd, name), name
    for name in amd.__all__:
        assert getattr(dr, name) is getattr(amd, name), name
    assert hasattr(dr.PathIntegrator, "renderC") and hasattr(dr.PathIntegrator, "renderD")


def test_example_style_calls_resolve():
    """The names the reference's example scripts call (example/optimize_texture.py:1-80) exist and
    PathIntegrator takes their keyword arguments."""
    import torch_psdr as dr
    used = {"PerspectiveCamera", "Texture", "DiffuseBsdfMaterial", "DiffuseMaterial", "Mesh", "AreaLight", "Scene",
            "PathIntegrator"}
    for name in used:
        assert hasattr(dr, name), name
    integ = dr.PathIntegrator(spp_interior=4, max_bounce=2, mis_light_samples=3, mis_bsdf_samples=3)
    assert integ.cfg[1] == 4 and integ.cfg[5] == 2
