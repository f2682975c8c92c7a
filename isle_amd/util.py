"""Run parameters and their YAML form (src/isle/util.py:100-181, src/isle/yamlio.py:77-120).

``parameters(**kwargs)`` returns a read-only object holding the keyword arguments as
attributes, with ``asdict()``, ``replace(**kwargs)`` and ``tilde(value, nt, beta=None)``
(= value * beta / nt, ``value`` an attribute name or a number, ``nt`` a number or a
Lattice, ``beta`` the object's own ``beta`` unless given).  ``dumpParameters`` /
``loadParameters`` write and read it as a ``!parameters`` YAML mapping; the action enums
HFAHopping, HFABasis and HFAAlgorithm are tagged scalars (``!HFAHopping EXP``), as in the
reference's files.  Loading uses a safe loader: nothing in the text is executed.
"""
import dataclasses

import yaml

from .action import HFAAlgorithm, HFABasis, HFAHopping


def _asdict(self):
    return {f.name: getattr(self, f.name) for f in dataclasses.fields(self)}


def _replace(self, **kwargs):
    return dataclasses.replace(self, **kwargs)


def _tilde(self, value, nt, beta=None):
    if isinstance(value, str):
        value = getattr(self, value)
    if not isinstance(nt, (int, float)):
        nt = nt.nt()  # a Lattice
    if beta is None:
        beta = self.beta
    return value * beta / nt


def parameters(**kwargs):
    """A new read-only parameters object with the given attributes (util.py:143-181)."""
    cls = dataclasses.make_dataclass(
        "Parameters", [(name, object) for name in kwargs], frozen=True,
        namespace={"asdict": _asdict, "replace": _replace, "tilde": _tilde})
    return cls(**kwargs)


class _Dumper(yaml.SafeDumper):
    pass


class _Loader(yaml.SafeLoader):
    pass


def _represent_any(dumper, obj):
    if dataclasses.is_dataclass(obj) and hasattr(obj, "tilde"):
        return dumper.represent_mapping("!parameters", obj.asdict(), flow_style=False)
    return dumper.represent_undefined(obj)


_Dumper.add_multi_representer(object, _represent_any)
for _enum in (HFAHopping, HFABasis, HFAAlgorithm):
    _Dumper.add_representer(_enum, lambda dumper, value, _tag="!" + _enum.__name__:
                            dumper.represent_scalar(_tag, value.name))
    _Loader.add_constructor("!" + _enum.__name__, lambda loader, node, _cls=_enum:
                            _cls[loader.construct_scalar(node)])
_Loader.add_constructor("!parameters", lambda loader, node:
                        parameters(**loader.construct_mapping(node, deep=True)))


def dumpParameters(params):
    """YAML text of a parameters object (or of a plain dict)."""
    return yaml.dump(params, Dumper=_Dumper)


def loadParameters(text):
    """Inverse of dumpParameters."""
    return yaml.load(text, Loader=_Loader)
