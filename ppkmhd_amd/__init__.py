"""MI355X (gfx950) MUSCL-Hancock + HLLD + constrained-transport MHD step behind the
reference's solver interface (MHD_Muscl_2D / MHD_Muscl_3D)."""
from . import _abi  # noqa: F401
from .params import make_params, default_problem, f32  # noqa: F401
