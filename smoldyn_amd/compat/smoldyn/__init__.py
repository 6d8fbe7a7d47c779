"""The reference's ``smoldyn`` Python package (source/python/smoldyn/__init__.py)
served by the MI355X engine: ``import smoldyn`` gives the user-level API,
``smoldyn._smoldyn`` the low-level simulation class it derives from.  Put
the directory that holds this package (smoldyn_amd/compat) first on
sys.path to run a reference model script unchanged."""
from .smoldyn import *  # noqa: F401,F403
from .smoldyn import __version__  # noqa: F401
from . import _smoldyn  # noqa: F401
