"""moldy_amd: MI355X-native force evaluation behind Moldy's force_calc()/ewald().

Importing the package gives the system description and builders; the HIP
engine lives in ``moldy_amd.engine`` and raises if its library is not built.
"""
from . import units, system  # noqa: F401

__all__ = ["units", "system"]
