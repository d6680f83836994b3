"""INI and dict (de)serialisation mixins for nested configuration dataclasses.

Semantics of the reference's tools/ini_parsable.py:27-90 and
tools/dict_parsable.py:24-63: one INI section per nested dataclass field,
one key per leaf field; only keys the dataclass defines are read; values are
parsed by the type of the current (default) value; unknown keys are ignored.
"""

import configparser
import dataclasses


def _sections(obj):
    return [f.name for f in dataclasses.fields(obj)]


def _keys(section_obj):
    return [f.name for f in dataclasses.fields(section_obj)]


class INIParsable:
    @classmethod
    def from_file(cls, filename):
        obj = cls()
        obj.read(filename)
        return obj

    def read(self, filename):
        cp = configparser.ConfigParser()
        cp.read(filename)
        self._from_parser(cp)

    def write(self, filename):
        cp = configparser.ConfigParser()
        for sec in _sections(self):
            part = getattr(self, sec)
            cp.add_section(sec)
            for key in _keys(part):
                cp[sec][key] = str(getattr(part, key))
        with open(filename, "w") as fh:
            cp.write(fh)

    def _from_parser(self, cp):
        for sec in _sections(self):
            if sec not in cp:
                continue
            part = getattr(self, sec)
            for key in _keys(part):
                if key not in cp[sec]:
                    continue
                kind = type(getattr(part, key))
                if kind is bool:
                    val = cp.getboolean(sec, key)
                elif kind is int:
                    val = cp.getint(sec, key)
                elif kind is float:
                    val = cp.getfloat(sec, key)
                else:
                    val = cp[sec][key]
                setattr(part, key, val)


class DictParsable:
    @classmethod
    def from_dict(cls, data: dict):
        obj = cls()
        for sec in _sections(obj):
            if sec not in data:
                continue
            part = getattr(obj, sec)
            for key in _keys(part):
                if key in data[sec]:
                    setattr(part, key, data[sec][key])
        return obj

    def to_dict(self) -> dict:
        return {sec: {key: getattr(getattr(self, sec), key) for key in _keys(getattr(self, sec))}
                for sec in _sections(self)}
