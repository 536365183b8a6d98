"""FA_CP Python interface on the MI355X build (QP/FA_CP/python/fa_cp/__init__.py:17-29).

    from pdal_amd.fa_cp import Model, FA_CP

`Model` and `FA_CP` behave as the reference package's; the module
`pdal_amd.fa_cp._fa_cp_core` provides `solve_once` and `get_default_params`
as the reference's compiled core does, over libfa_cp.so.
"""
from . import FA_CP
from .model import Model

__all__ = ["Model"]
__version__ = "0.1.0"
