"""assist-amd: the ASSIST force model and IAS15 for batches of test particles on MI355X.

Drop-in names of the reference Python package (reference assist/__init__.py):
``Ephem``, ``Extras``, ``assist_error_messages`` and the version strings, plus
``Simulation``/``Particle`` (REBOUND is not part of this build) and the numpy
batch API in :mod:`assist_amd.batch`.
"""
from ._libassist import (__build__, __githash__, __libpath__, __version__, assist_error_messages,
                         clibassist)
from .ephem import ASSIST_BODY_IDS, Ephem
from .simulation import Particle, Simulation

moduleversion = "1.2.0"
libassistversion = __version__


def __getattr__(name: str):
    if name == "Extras":
        from .extras import Extras
        return Extras
    raise AttributeError(f"module {__name__!r} has no attribute {name!r}")


__all__ = ["__libpath__", "__version__", "__build__", "__githash__", "assist_error_messages", "clibassist",
           "Extras", "Ephem", "Simulation", "Particle", "ASSIST_BODY_IDS"]
