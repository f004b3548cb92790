"""Engine registry (Engines/__init__.py:6-23 of the reference): every module in
this package is imported and its classes are re-exported, so
``import pyphasefield_amd.Engines as engines; engines.Warren1995(...)`` works
exactly like the reference's plugin discovery.  Unlike the reference, import
errors are not swallowed: a broken engine module fails loudly."""

from importlib import import_module
from inspect import isclass
from pathlib import Path
from pkgutil import iter_modules

from .Diffusion import Diffusion  # noqa: F401

_pkg_dir = Path(__file__).resolve().parent
for (_, _module_name, _) in iter_modules([str(_pkg_dir)]):
    _module = import_module(f"{__name__}.{_module_name}")
    for _attr in dir(_module):
        _obj = getattr(_module, _attr)
        if isclass(_obj) and _obj.__module__ == _module.__name__:
            globals()[_attr] = _obj
