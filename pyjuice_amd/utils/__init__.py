"""Small host-side helpers shared by the node DAG, the compiler and the runtime."""
from __future__ import annotations

from typing import Iterable

__all__ = ["BitSet", "max_cdf_power_of_2", "next_power_of_2_int", "DecoratorContextManager"]


class BitSet:
    """Variable scope: a set of non-negative variable ids.

    Semantics follow ``pyjuice/utils/bitset.py`` (ascending iteration, ``&``, ``|``,
    equality ignoring trailing zero bytes, hashable).  Stored as one Python integer
    bit mask (bit ``v`` set <=> variable ``v`` in scope), which makes union,
    intersection and equality single integer operations.
    """

    __slots__ = ("_mask", "_len")

    def __init__(self, mask: int = 0):
        self._mask = int(mask)
        self._len = bin(self._mask).count("1")

    @classmethod
    def from_array(cls, arr: Iterable[int]) -> "BitSet":
        m = 0
        for v in arr:
            v = int(v)
            assert v >= 0, "variable ids must be non-negative"
            m |= 1 << v
        return cls(m)

    def to_list(self):
        out = []
        m = self._mask
        v = 0
        while m:
            if m & 1:
                out.append(v)
            m >>= 1
            v += 1
        return out

    def add(self, value: int):
        if not self.hasitem(value):
            self._mask |= 1 << int(value)
            self._len += 1

    def remove(self, value: int):
        if self.hasitem(value):
            self._mask &= ~(1 << int(value))
            self._len -= 1

    def hasitem(self, value: int) -> bool:
        return (self._mask >> int(value)) & 1 == 1

    def contains_all(self, other: "BitSet") -> bool:
        return (self._mask & other._mask) == other._mask

    def contains_any(self, other: "BitSet") -> bool:
        return (self._mask & other._mask) != 0

    def __and__(self, other: "BitSet") -> "BitSet":
        return BitSet(self._mask & other._mask)

    def __or__(self, other: "BitSet") -> "BitSet":
        return BitSet(self._mask | other._mask)

    def __ior__(self, other: "BitSet") -> "BitSet":
        self._mask |= other._mask
        self._len = bin(self._mask).count("1")
        return self

    def __eq__(self, other) -> bool:
        return isinstance(other, BitSet) and self._mask == other._mask

    def __ne__(self, other) -> bool:
        return not self.__eq__(other)

    def __hash__(self):
        return hash(("BitSet", self._mask))

    def __iter__(self):
        return iter(self.to_list())

    def __len__(self):
        return self._len

    def __deepcopy__(self, memo):
        return BitSet(self._mask)

    def __repr__(self):
        if self._len <= 16:
            return "BitSet([" + ",".join(str(v) for v in self) + "])"
        return f"BitSet(num_elements={self._len})"


def max_cdf_power_of_2(val: int) -> int:
    """Largest power of two dividing ``val`` (``pyjuice/utils/util.py:6``)."""
    val = int(val)
    if val == 0:
        return 1
    return val & (-val)


def next_power_of_2_int(x: int) -> int:
    """Smallest power of two >= x for x >= 1 (1 maps to 1)."""
    x = int(x)
    if x <= 1:
        return 1
    return 1 << (x - 1).bit_length()


class DecoratorContextManager:
    """Context manager that can also be used as a function decorator."""

    def __call__(self, func):
        import functools

        @functools.wraps(func)
        def wrapper(*args, **kwargs):
            with self.__class__(*self._ctor_args, **self._ctor_kwargs):
                return func(*args, **kwargs)

        return wrapper

    def __new__(cls, *args, **kwargs):
        obj = super().__new__(cls)
        obj._ctor_args = args
        obj._ctor_kwargs = kwargs
        return obj
