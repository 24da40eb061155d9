"""Result records of the drop-in API: the same attribute names and to_dict() as the
reference's representation classes (python/coffe/representation.py:1-198)."""


class Representation:
    _fields = ()

    def __init__(self, **kwargs):
        missing = [f for f in self._fields if f not in kwargs]
        extra = [k for k in kwargs if k not in self._fields]
        if missing or extra:
            raise TypeError(f"{type(self).__name__}: missing {missing}, unexpected {extra}")
        for f in self._fields:
            object.__setattr__(self, "_" + f, kwargs[f])

    def to_dict(self):
        return {f: getattr(self, "_" + f) for f in self._fields}

    def __repr__(self):
        return f"<{type(self).__name__}({self.to_dict()})>"

    def __str__(self):
        return f"{type(self).__name__}({self.to_dict()})"

    def __eq__(self, other):
        return type(self) is type(other) and self.to_dict() == other.to_dict()


def _ro(name):
    return property(lambda self: getattr(self, "_" + name))


class Multipoles(Representation):
    """One multipole point: mean redshift z, separation r (Mpc), multipole l, value."""
    _fields = ("z", "r", "l", "value")
    z, r, l, value = _ro("z"), _ro("r"), _ro("l"), _ro("value")


class AverageMultipoles(Representation):
    """One redshift-averaged multipole point: bin [z_min, z_max], separation r (Mpc), l."""
    _fields = ("z_min", "z_max", "r", "l", "value")
    z_min, z_max, r, l, value = _ro("z_min"), _ro("z_max"), _ro("r"), _ro("l"), _ro("value")


class Covariance(Representation):
    """One covariance entry between (l1, r1) and (l2, r2) at mean redshift z."""
    _fields = ("r1", "r2", "l1", "l2", "z", "value")
    r1, r2, l1, l2, z, value = _ro("r1"), _ro("r2"), _ro("l1"), _ro("l2"), _ro("z"), _ro("value")


class AverageCovariance(Representation):
    """One covariance entry of the redshift-averaged multipoles of bin [z_min, z_max]."""
    _fields = ("r1", "r2", "l1", "l2", "z_min", "z_max", "value")
    r1, r2, l1, l2 = _ro("r1"), _ro("r2"), _ro("l1"), _ro("l2")
    z_min, z_max, value = _ro("z_min"), _ro("z_max"), _ro("value")


class Corrfunc(Representation):
    """One point of the correlation function: mean redshift z, separation r (Mpc), mu."""
    _fields = ("r", "mu", "z", "value")
    r, mu, z, value = _ro("r"), _ro("mu"), _ro("z"), _ro("value")
