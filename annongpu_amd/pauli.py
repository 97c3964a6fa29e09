"""Sums of Pauli strings in the PauliString encoding of the reference.

The reference builds its Operator from QuantumExpression's PauliExpression
(source/operator/Operator.cpp:18-39), reading for every term the bit masks
``a`` and ``b`` and a complex coefficient.  The encoding per site
(include/basis/PauliString.hpp:37-56,133-143) is

    identity (0,0),  sigma_x (1,0),  sigma_y (0,1),  sigma_z (1,1).

QuantumExpression is not available here, so this module provides the part of
its interface the VMC path needs: ``sigma_x/y/z(i)``, sums, products (with the
Pauli algebra phases), scalar factors, iteration over ``(a, b, coefficient)``.
"""
from numbers import Number


def _type_masks(a, b):
    x = a & ~b
    y = ~a & b
    z = a & b
    return x, y, z


class PauliExpression:
    """Map {(a, b): coefficient} with a, b arbitrary-width integer bit masks."""

    __slots__ = ("terms",)

    def __init__(self, terms=None):
        self.terms = {}
        if terms:
            for k, v in terms.items():
                if v != 0:
                    self.terms[k] = complex(v)

    @staticmethod
    def single(site, kind):
        """kind in 'xyz' (or 1, 2, 3 as in PauliString::set_at)."""
        if isinstance(kind, str):
            kind = {"x": 1, "y": 2, "z": 3}[kind.lower()]
        a = (kind & 1) << site
        b = ((kind >> 1) & 1) << site
        return PauliExpression({(a, b): 1.0})

    @staticmethod
    def identity(value=1.0):
        return PauliExpression({(0, 0): value})

    def __iter__(self):
        return iter(self.terms.items())

    def __len__(self):
        return len(self.terms)

    def size(self):
        return len(self.terms)

    def copy(self):
        return PauliExpression(dict(self.terms))

    def __add__(self, other):
        if isinstance(other, Number):
            other = PauliExpression.identity(other)
        r = dict(self.terms)
        for k, v in other.terms.items():
            r[k] = r.get(k, 0) + v
        return PauliExpression({k: v for k, v in r.items() if v != 0})

    __radd__ = __add__

    def __neg__(self):
        return PauliExpression({k: -v for k, v in self.terms.items()})

    def __sub__(self, other):
        return self + (-other)

    def __rsub__(self, other):
        return (-self) + other

    def __mul__(self, other):
        if isinstance(other, Number):
            return PauliExpression({k: v * other for k, v in self.terms.items()})
        r = {}
        for (a1, b1), c1 in self.terms.items():
            for (a2, b2), c2 in other.terms.items():
                phase = _product_phase(a1, b1, a2, b2)
                key = (a1 ^ a2, b1 ^ b2)
                r[key] = r.get(key, 0) + c1 * c2 * phase
        return PauliExpression({k: v for k, v in r.items() if v != 0})

    def __rmul__(self, other):
        if isinstance(other, Number):
            return self * other
        return NotImplemented

    def __truediv__(self, other):
        return self * (1.0 / other)

    def __eq__(self, other):
        return isinstance(other, PauliExpression) and self.terms == other.terms

    def __repr__(self):
        names = "IXYZ"
        parts = []
        for (a, b), c in self.terms.items():
            n = max(a.bit_length(), b.bit_length())
            ops = [f"{names[((a >> i) & 1) | (((b >> i) & 1) << 1)]}{i}" for i in range(n)
                   if ((a | b) >> i) & 1]
            parts.append(f"{c:.6g}*" + ("".join(ops) or "I"))
        return " + ".join(parts) or "0"

    def max_site(self):
        m = -1
        for (a, b) in self.terms:
            m = max(m, (a | b).bit_length() - 1)
        return m

    def to_json(self):
        """{"type": "PauliExpression", "terms": [[a, b, re, im], ...]} with the masks a, b as
        decimal strings (any width)"""
        return {"type": "PauliExpression",
                "terms": [[str(a), str(b), c.real, c.imag] for (a, b), c in self.terms.items()]}

    @staticmethod
    def from_json(obj):
        if obj.get("type") != "PauliExpression":
            raise ValueError("not a PauliExpression JSON object")
        return PauliExpression({(int(a), int(b)): complex(re, im) for a, b, re, im in obj["terms"]})

    def term_list(self):
        """one single-string PauliExpression per term (iteration order)"""
        return [PauliExpression({k: v}) for k, v in self.terms.items()]

    def roll(self, shift, n):
        """Cyclic shift of every string by `shift` sites on a ring of n sites."""
        mask = (1 << n) - 1

        def rot(x):
            return ((x << shift) | (x >> (n - shift))) & mask

        return PauliExpression({(rot(a), rot(b)): c for (a, b), c in self.terms.items()})


def _product_phase(a1, b1, a2, b2):
    """Phase of (string1)(string2) per site: s_p s_q = i eps_pqr s_r for p != q."""
    x1, y1, z1 = _type_masks(a1, b1)
    x2, y2, z2 = _type_masks(a2, b2)
    cyc = bin((x1 & y2) | (y1 & z2) | (z1 & x2)).count("1")
    anti = bin((y1 & x2) | (z1 & y2) | (x1 & z2)).count("1")
    return 1j ** ((cyc - anti) % 4)


def sigma_x(i):
    return PauliExpression.single(i, "x")


def sigma_y(i):
    return PauliExpression.single(i, "y")


def sigma_z(i):
    return PauliExpression.single(i, "z")


def identity(value=1.0):
    return PauliExpression.identity(value)
