"""Configuration helpers with the reference's semantics
(src/python/lennardjonesium/tools/{ini_parsable,dict_parsable,linspace}.py)."""

from lennardjonesium.tools.parsable import INIParsable, DictParsable  # noqa: F401
from lennardjonesium.tools.linspace import linspace  # noqa: F401
