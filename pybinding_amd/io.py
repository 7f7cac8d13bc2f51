"""Compressed result files: `save` / `load` (the interface of pybinding/support/pickle.py:40-75).

A file holds one pickled object (protocol 4) inside gzip compression, the reference's `.pbz`
format, so result objects (`Series`, `SpatialLDOS`, `StructureMap`, arrays, tuples of them) can be
written by one session and read by another.  `.pbz` is appended to names without an extension.
Unpickling can run code stored in the file: load only files you trust.
"""
import gzip
import os
import pathlib
import pickle

__all__ = ['save', 'load']


def _as_name(file):
    """str for path-like objects; open file objects pass through"""
    return os.fspath(file) if isinstance(file, (str, pathlib.PurePath)) else file


def _with_extension(file):
    if isinstance(file, str) and not pathlib.Path(file).suffix:
        return file + ".pbz"
    return file


def save(obj, file):
    """Write `obj` to `file` (a name, a path or an open binary file), gzip-compressed pickle."""
    with gzip.open(_with_extension(_as_name(file)), 'wb') as f:
        pickle.dump(obj, f, protocol=4)


def load(file):
    """Read back an object written by :func:`save`; a name without extension also finds the
    `.pbz` file `save` made of it."""
    file = _as_name(file)
    if isinstance(file, str) and not os.path.exists(file) and os.path.exists(_with_extension(file)):
        file = _with_extension(file)
    with gzip.open(file, 'rb') as f:
        return pickle.load(f)
