"""Name/description mixin shared by the API objects.

Same behaviour as the reference's ``oqupy/base_api.py:21-80``: a name and a
description that default to placeholder strings and must be text.
"""

from typing import Optional, Text

from oqupy_amd.config import SEPERATOR


class BaseAPIClass:
    """Base class of API objects carrying a ``name`` and a ``description``."""

    def __init__(self,
                 name: Optional[Text] = None,
                 description: Optional[Text] = None) -> None:
        self.name = name
        self.description = description

    def __str__(self) -> Text:
        return (SEPERATOR + f"{type(self).__name__} object: {self.name}\n"
                + f" {self.description}\n")

    @property
    def name(self) -> Text:
        """Name of the object."""
        return self._name

    @name.setter
    def name(self, value: Optional[Text] = None) -> None:
        if value is None:
            value = "__unnamed__"
        assert isinstance(value, str), "Name must be text."
        self._name = value

    @name.deleter
    def name(self) -> None:
        self.name = None

    @property
    def description(self) -> Text:
        """Detailed description of the object."""
        return self._description

    @description.setter
    def description(self, value: Optional[Text] = None) -> None:
        if value is None:
            value = "__no_description__"
        assert isinstance(value, str), "Description must be text."
        self._description = value

    @description.deleter
    def description(self) -> None:
        self.description = None
