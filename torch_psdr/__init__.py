"""The reference's import name for this package.

Callers of torch_psdr write `import torch_psdr as dr`
(example/optimize_texture.py:1, example/optimize_points.py); this module
re-exports torch_psdr_amd's public API under that name, so such a script runs
unchanged.  Every name is the torch_psdr_amd object itself.
"""
from torch_psdr_amd import *  # noqa: F401,F403
from torch_psdr_amd import __all__  # noqa: F401
