"""Particle arrays of a drop-in state, resident in device memory.

The reference's state classes hold the census as plain NumPy arrays
(IMC/IMC2D.py:65-84, IMC/IMC1D.py:836-848, MG_IMC/MG_IMC2D.py:116-136,
MG_IMC/MG_IMC1D.py:116-138), and its drivers replace a census one field at a
time (MG_IMC/problems/test_infinite_medium_multigroup_expband.py:187-195,
IMC/CrookedPipe2D.py:532).  Here the census lives on the GPU: reading a field
downloads the census (cached until the engine changes it), assigning a field
uploads it.

Assigned fields are held until every field has the same length: a census
that grows or shrinks is replaced field by field, and the intermediate
states, with fields of two lengths, never reach the device.  The pending
fields are uploaded as soon as the lengths agree; a step (or anything else
that needs the device census) with fields of unequal length pending raises
ValueError instead of cutting the census to the shortest field.
"""
from __future__ import annotations

import numpy as np


class DeviceCensus:
    """Base of the drop-in state classes; subclasses set ``_PARTICLE_FIELDS``
    and may override ``_download`` / ``_upload`` / ``_device_count``."""

    _PARTICLE_FIELDS: tuple = ()

    def __init__(self, engine):
        object.__setattr__(self, "_engine", engine)
        object.__setattr__(self, "_cache", None)
        object.__setattr__(self, "_pending", {})

    # engine access -----------------------------------------------------------
    def _download(self):
        return tuple(self._engine.get_census())

    def _upload(self, arrs):
        self._engine.put_census(*arrs)

    def _device_count(self):
        return self._engine.n

    # census ------------------------------------------------------------------
    def _census(self):
        if self._cache is None:
            object.__setattr__(self, "_cache", self._download())
        return self._cache

    def _overlay(self):
        cen = self._census()
        return [self._pending.get(f, cen[i]) for i, f in enumerate(self._PARTICLE_FIELDS)]

    def _flush(self):
        """Upload pending field assignments; ValueError if their lengths differ."""
        if not self._pending:
            return
        arrs = self._overlay()
        lengths = {f: len(a) for f, a in zip(self._PARTICLE_FIELDS, arrs)}
        if len(set(lengths.values())) != 1:
            raise ValueError("census fields have different lengths after assignment: "
                             + ", ".join(f"{f}={n}" for f, n in lengths.items())
                             + " (assign every particle field before the next step)")
        self._upload(arrs)
        self._pending.clear()
        object.__setattr__(self, "_cache", None)

    def _invalidate(self):
        self._flush()
        object.__setattr__(self, "_cache", None)

    def __getattr__(self, name):
        if name in type(self)._PARTICLE_FIELDS:
            pend = self.__dict__.get("_pending", {})
            if name in pend:
                return pend[name]
            return self._census()[type(self)._PARTICLE_FIELDS.index(name)]
        raise AttributeError(name)

    def __setattr__(self, name, value):
        if name in type(self)._PARTICLE_FIELDS:
            self._pending[name] = np.asarray(value)
            arrs = self._overlay()
            if len({len(a) for a in arrs}) == 1:
                self._flush()
            return
        object.__setattr__(self, name, value)

    @property
    def n_particles(self):
        self._flush()
        return self._device_count()
