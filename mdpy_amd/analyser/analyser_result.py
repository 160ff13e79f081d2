"""AnalyserResult (mdpy/analyser/analyser_result.py:14-67): a titled dict of arrays with a
description per key, saved to / loaded from .npz."""
import numpy as np

from ..error import AnalyserPoorDefinedError, FileFormatError


class AnalyserResult:
    def __init__(self, title: str, description: dict, data: dict) -> None:
        if description.keys() != data.keys():
            raise AnalyserPoorDefinedError("description and data must have the same keys")
        self._title = title
        self._description = description
        self._data = data

    def __repr__(self):
        out = "------------\nDescription of AnalyserResult object at %x\nTitle: \n- %s\nKeys: \n" % (id(self),
                                                                                                  self._title)
        out += "".join("- %s: %s\n" % (k, v) for k, v in self._description.items())
        return out + "------------"

    __str__ = __repr__

    def save(self, file_path: str):
        if not file_path.endswith("npz"):
            raise FileFormatError("an AnalyserResult is saved as a .npz file")
        d = {"title": self._title}
        for k, v in self._description.items():
            d["description-%s" % k] = v
        d.update(self._data)
        np.savez(file_path, **d)

    title = property(lambda self: self._title)
    description = property(lambda self: self._description)
    data = property(lambda self: self._data)


def load_analyser_result(file_path: str):
    if not file_path.endswith("npz"):
        raise FileFormatError("an AnalyserResult is loaded from a .npz file")
    raw = np.load(file_path)  # arrays and strings only (no pickles)
    description, data = {}, {}
    for key in raw.keys():
        if key.startswith("description-"):
            k = key[len("description-"):]
            description[k] = raw[key].item()
            data[k] = raw[k]
    return AnalyserResult(title=raw["title"].item(), description=description, data=data)
