"""Model-file generator standing in for the reference's MATLAB preprocessor."""
from .export import Model, write_model, write_static, write_dynamic  # noqa: F401
from .configs import WORKLOADS, tiny, c1, c2  # noqa: F401
