"""Configuration objects with the reference's names, defaults and validation.

``SolverConfig``  solveras.py:27-169  (epsilon, incr_f, decr_f, safety_f, adapt_cutoff, minh)
``ETDConfig``     etd.py:46-131       (modecutoff, contour_points, contour_radius)
``SolverType``    util/solver_type.py:21-84
"""
from enum import Enum, auto


class SolverConfig:
    """Adaptive step-size configuration (solveras.py:71-169)."""

    def __init__(self, epsilon=1e-4, incr_f=1.25, decr_f=0.85, safety_f=0.8, adapt_cutoff=0.01,
                 minh=1e-16):
        self._epsilon = self._incr_f = self._decr_f = None
        self._safety_f = self._adapt_cutoff = self._minh = None
        self.epsilon = epsilon
        self.incr_f = incr_f
        self.decr_f = decr_f
        self.safety_f = safety_f
        self.adapt_cutoff = adapt_cutoff
        self.minh = minh

    @property
    def epsilon(self):
        return self._epsilon

    @epsilon.setter
    def epsilon(self, value):
        if value <= 0:
            raise ValueError(f"epsilon must be positive but is {value}")
        self._epsilon = float(value)

    @property
    def incr_f(self):
        return self._incr_f

    @incr_f.setter
    def incr_f(self, value):
        if value <= 1.0:
            raise ValueError(f"incr_f must be > 1.0 but is {value}")
        self._incr_f = float(value)

    @property
    def decr_f(self):
        return self._decr_f

    @decr_f.setter
    def decr_f(self, value):
        if value >= 1.0:
            raise ValueError(f"decr_f must be < 1.0 but is {value}")
        self._decr_f = float(value)

    @property
    def safety_f(self):
        return self._safety_f

    @safety_f.setter
    def safety_f(self, value):
        if value > 1.0:
            raise ValueError(f"safety_f must be <= 1.0 but is {value}")
        self._safety_f = float(value)

    @property
    def adapt_cutoff(self):
        return self._adapt_cutoff

    @adapt_cutoff.setter
    def adapt_cutoff(self, value):
        if value >= 1.0:
            raise ValueError(f"adapt_cutoff must be < 1.0 but is {value}")
        self._adapt_cutoff = float(value)

    @property
    def minh(self):
        return self._minh

    @minh.setter
    def minh(self, value):
        if value <= 0:
            raise ValueError(f"minh must be positive but is {value}")
        self._minh = float(value)


class ETDConfig:
    """psi-function evaluation settings (etd.py:81-131)."""

    def __init__(self, modecutoff=0.01, contour_points=32, contour_radius=1.0):
        self._modecutoff = 0.01
        self._contour_points = 32
        self._contour_radius = 1.0
        self.modecutoff = modecutoff
        self.contour_points = contour_points
        self.contour_radius = contour_radius

    @property
    def modecutoff(self):
        return self._modecutoff

    @modecutoff.setter
    def modecutoff(self, value):
        if (value > 1.0) or (value <= 0):
            raise ValueError(f"modecutoff must be between 0.0 and 1.0 but is {value}")
        self._modecutoff = value

    @property
    def contour_points(self):
        return self._contour_points

    @contour_points.setter
    def contour_points(self, value):
        if not isinstance(value, int):
            raise TypeError(f"contour_points must be an integer but is {value}")
        if value <= 1:
            raise ValueError(f"contour_points must be an integer greater than 1 but is {value}")
        self._contour_points = value

    @property
    def contour_radius(self):
        return self._contour_radius

    @contour_radius.setter
    def contour_radius(self, value):
        if value <= 0:
            raise ValueError(f"contour_radius must greater than 0 but is {value}")
        self._contour_radius = value


class SolverType(Enum):
    """Constant-step or adaptive-step solver (util/solver_type.py:21-46)."""

    CONSTANT_STEP = auto()
    ADAPTIVE_STEP = auto()
    CS = CONSTANT_STEP
    AS = ADAPTIVE_STEP

    def __str__(self):
        return self.name.replace("_", " ").title()

    @classmethod
    def from_solver(cls, solver):
        from .solvers import BaseSolverAS, BaseSolverCS
        if isinstance(solver, BaseSolverCS):
            return cls.CONSTANT_STEP
        if isinstance(solver, BaseSolverAS):
            return cls.ADAPTIVE_STEP
        raise TypeError(f"Cannot determine solver type for {type(solver).__name__}")
