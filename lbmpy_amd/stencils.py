"""Lattice velocity sets.

``LBStencil`` mirrors lbmpy's ``src/lbmpy/stencils.py:8-128``.  The direction
tables are the 'walberla' orderings of ``_predefined_stencils``
(``stencils.py:133-135`` D2Q9, ``:194-198`` D3Q19, ``:221-227`` D3Q27); those
orderings are the PDF index layout in memory, so they must be identical to
lbmpy's for PDF arrays to be exchangeable.  The HIP kernels are compiled for
exactly these orderings (the same tables live in ``csrc/stencils.h``).
"""
from .enums import Stencil

_WALBERLA = {
    "D2Q9": ((0, 0),
             (0, 1), (0, -1), (-1, 0), (1, 0),
             (-1, 1), (1, 1), (-1, -1), (1, -1)),
    "D3Q7": ((0, 0, 0),
             (0, 1, 0), (0, -1, 0),
             (-1, 0, 0), (1, 0, 0),
             (0, 0, 1), (0, 0, -1)),
    "D3Q15": ((0, 0, 0),
              (0, 1, 0), (0, -1, 0), (-1, 0, 0), (1, 0, 0), (0, 0, 1), (0, 0, -1),
              (1, 1, 1), (-1, 1, 1), (1, -1, 1), (-1, -1, 1), (1, 1, -1), (-1, 1, -1), (1, -1, -1),
              (-1, -1, -1)),
    "D3Q19": ((0, 0, 0),
              (0, 1, 0), (0, -1, 0), (-1, 0, 0), (1, 0, 0), (0, 0, 1), (0, 0, -1),
              (-1, 1, 0), (1, 1, 0), (-1, -1, 0), (1, -1, 0),
              (0, 1, 1), (0, -1, 1), (-1, 0, 1), (1, 0, 1),
              (0, 1, -1), (0, -1, -1), (-1, 0, -1), (1, 0, -1)),
    "D3Q27": ((0, 0, 0),
              (0, 1, 0), (0, -1, 0), (-1, 0, 0), (1, 0, 0), (0, 0, 1), (0, 0, -1),
              (-1, 1, 0), (1, 1, 0), (-1, -1, 0), (1, -1, 0),
              (0, 1, 1), (0, -1, 1), (-1, 0, 1), (1, 0, 1),
              (0, 1, -1), (0, -1, -1), (-1, 0, -1), (1, 0, -1),
              (1, 1, 1), (-1, 1, 1), (1, -1, 1), (-1, -1, 1), (1, 1, -1), (-1, 1, -1), (1, -1, -1),
              (-1, -1, -1)),
}

#: stencils the HIP sweep kernels are instantiated for
GPU_STENCILS = ("D2Q9", "D3Q19", "D3Q27")


class LBStencil:
    """A discrete velocity set in DxQy notation ('walberla' ordering by default)."""

    def __init__(self, stencil, ordering="walberla"):
        if isinstance(stencil, LBStencil):
            self._entries = stencil.stencil_entries
            self._ordering = stencil.ordering
        elif isinstance(stencil, tuple):
            self._entries = tuple(tuple(int(c) for c in d) for d in stencil)
            self._ordering = None
        else:
            name = stencil.name if isinstance(stencil, Stencil) else str(stencil)
            if ordering != "walberla" or name not in _WALBERLA:
                raise ValueError(f"No such stencil available: {name} ({ordering}); "
                                 f"available: {', '.join(_WALBERLA)} (walberla ordering)")
            self._entries = _WALBERLA[name]
            self._ordering = ordering
        if len({len(d) for d in self._entries}) != 1:
            raise ValueError("The stencil you have created is not valid. "
                             "It probably contains elements with different lengths")
        if len(set(self._entries)) < len(self._entries):
            raise ValueError("The stencil you have created is not valid. It contains duplicated elements")

    @property
    def D(self):
        return len(self._entries[0])

    @property
    def Q(self):
        return len(self._entries)

    @property
    def name(self):
        return f"D{self.D}Q{self.Q}"

    @property
    def ordering(self):
        return self._ordering

    @property
    def stencil_entries(self):
        return self._entries

    @property
    def inverse_stencil_entries(self):
        return tuple(tuple(-c for c in d) for d in self._entries)

    def index(self, direction):
        return self._entries.index(tuple(direction))

    def inverse_index(self, direction):
        return self._entries.index(tuple(-c for c in direction))

    def inverse_indices(self):
        return [self.inverse_index(d) for d in self._entries]

    def __getitem__(self, i):
        return self._entries[i]

    def __iter__(self):
        return iter(self._entries)

    def __len__(self):
        return len(self._entries)

    def __eq__(self, other):
        return isinstance(other, LBStencil) and set(self._entries) == set(other._entries) \
            and self._ordering == other._ordering

    def __hash__(self):
        return hash(self._entries)

    def __str__(self):
        return str(self._entries)

    __repr__ = __str__

    def is_gpu_layout(self):
        """True when the direction order equals the 'walberla' table the HIP kernels use."""
        return self.name in GPU_STENCILS and self._entries == _WALBERLA[self.name]
