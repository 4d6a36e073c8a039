"""Uniform material fields of a particle set.

A field that holds one value, bit for bit, on every particle goes to the
step as that value (``pmpm_particles_in.uniform_mask``) instead of an array,
which saves its per-particle reads in the constitutive kernel.  The
reference's add_object fills every field from one scalar per object
(anisompm.py:366-424), so most scenes have several such fields: config1
keeps vol0, mass, sigma_c, eta, zeta, l0 and residual uniform; its moduli
and fibres vary per particle.
"""
from __future__ import annotations

import torch

from . import _lib


class UniformFields:
    """Which of ``_lib.UNIFORM_FIELDS`` of an owner (an object with those
    attributes as tensors of one row per particle) are uniform, and their
    values.  A field is re-checked -- one device reduction, read together
    with the others that changed -- only when its tensor was replaced or
    edited in place (its version moved)."""

    def __init__(self):
        self._keys: dict = {}
        self._vals: dict = {}
        self._all = None
        self._state = (0, {})

    def state(self, owner):
        """(mask, {field: values}) with bit k of mask for field k; the same
        object while no field changed."""
        ts = [getattr(owner, name) for name in _lib.UNIFORM_FIELDS]
        allkeys = tuple((id(t), t.data_ptr(), t._version, t.shape[0]) for t in ts)
        if allkeys == self._all:
            return self._state
        changed = [(name, t, key) for name, t, key in zip(_lib.UNIFORM_FIELDS, ts, allkeys)
                   if self._keys.get(name) != key]
        flags, firsts = [], []
        for name, t, _ in changed:
            flat = t.reshape(t.shape[0], 1 if t.dim() == 1 else 9)
            if flat.shape[0] == 0:
                flags.append(torch.zeros((), dtype=torch.bool, device=t.device))
                firsts.append(torch.zeros(flat.shape[1], dtype=torch.float64, device=t.device))
                continue
            bits = flat.view(torch.int32 if t.element_size() == 4 else torch.int64)
            flags.append((bits == bits[:1]).all())
            firsts.append(flat[0].double())
        ok = torch.stack(flags).tolist()
        vals = torch.cat(firsts).tolist()
        o = 0
        for (name, t, key), u in zip(changed, ok):
            w = 1 if t.dim() == 1 else 9
            self._vals[name] = tuple(vals[o:o + w]) if u else None
            self._keys[name] = key
            o += w
        mask, out = 0, {}
        for k, name in enumerate(_lib.UNIFORM_FIELDS):
            v = self._vals.get(name)
            if v is not None:
                mask |= 1 << k
                out[name] = v
        self._all, self._state = allkeys, (mask, out)
        return self._state


def key(state):
    """A hashable form of a (mask, values) state, for graph capture keys (the
    values are kernel arguments of a captured step)."""
    mask, vals = state
    return mask, tuple(sorted(vals.items()))


def fill(pin, state):
    """Write a (mask, values) state into a ParticlesIn (skipped when this
    pin already holds this very state object)."""
    if getattr(pin, "_uniform_state", None) is state:
        return
    pin._uniform_state = state
    mask, vals = state
    pin.uniform_mask = mask
    for k, name in enumerate(_lib.UNIFORM_FIELDS):
        if name in vals:
            if name == "A0":
                pin.uniform_A0[:] = vals[name]
            else:
                pin.uniform_value[k] = vals[name][0]
