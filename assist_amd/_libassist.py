"""Loads libassist.so (reference assist/_libassist.py:34-67).

``ASSIST_LIBASSIST_PATH`` overrides the library location; a relative path is
taken relative to the repository root.  There is no fallback: a missing or
unloadable library raises immediately.
"""
from __future__ import annotations

import os
from ctypes import c_char_p, c_int, cdll

pymodulespath = os.path.dirname(os.path.abspath(__file__))
_override = os.environ.get("ASSIST_LIBASSIST_PATH")
if _override:
    if not os.path.isabs(_override):
        _override = os.path.join(os.path.dirname(pymodulespath), _override)
    _path = _override
else:
    _path = os.path.join(pymodulespath, "libassist.so")
if not os.path.exists(_path):
    raise ImportError("libassist.so not found at %s; build it with `python assist_amd/build.py`" % _path)
clibassist = cdll.LoadLibrary(_path)
__libpath__ = _path


def _libassist_str(name: str) -> str:
    val = c_char_p.in_dll(clibassist, name).value
    if val is None:
        raise RuntimeError(f"unable to find symbol {name} in libassist")
    return val.decode("ascii")


__version__ = _libassist_str("assist_version_str")
__build__ = _libassist_str("assist_build_str")
__githash__ = _libassist_str("assist_githash_str")


def assist_error_messages(e: int) -> str:
    e_N = c_int.in_dll(clibassist, "assist_error_messages_N").value
    if e >= e_N:
        raise RuntimeError("An error occured while trying to process an ASSIST error message.")
    ccpp = c_char_p * e_N
    message = ccpp.in_dll(clibassist, "assist_error_messages")[e]
    if message is None:
        raise RuntimeError("assist_error_messages is missing from libassist.")
    return message.decode("ascii")
