"""FP64 NumPy restatement of rkstiff's diagonal stepping path (oracle).

Every expression keeps the reference's operation order, so that on the same
inputs this module reproduces the reference bitwise (NumPy is deterministic):
that is what the golden fixtures under ``tests/golden`` check.

Reference map (paths relative to the reference tree):

* ``SolverConfig``                 solveras.py:27-169
* ``ETDConfig``                    etd.py:46-131
* ``psi1/psi2/psi3``               etd.py:134-182
* controller (``AdaptiveDriver``)  solveras.py:255-257 (constants), 336-410
  (step), 412-455 (``_compute_s``), 457-554 (reject/accept), 556-650 (evolve)
* ``ConstantDriver``               solvercs.py:163-279
* ETD35 coefficients/stages        etd35.py:169-288 / 290-345, wrapper 842-889
* ETD34                            etd34.py:96-191, wrapper 560-621
* ETD4                             etd4.py:87-175, wrapper 380-422
* ETD5                             etd5.py:115-261, wrapper 600-655
* IF34                             if34.py:81-131, wrapper 308-338
* IF45DP                           if45dp.py:112-249
* IF4                              if4.py:73-122, wrapper 255-275

Pinned by: ``tests/golden/*.npz`` (bitwise trajectories, attempt logs,
coefficients and single attempts generated from the reference in the CPU
container), the KdV soliton closed form, the nl=0 exact-exponential identity,
psi closed forms vs mpmath, and order-of-convergence slopes
(``tests/test_oracle_*.py``).
"""
import numpy as np


# ----------------------------------------------------------------------------
# configuration objects
# ----------------------------------------------------------------------------
class SolverConfig:
    """Adaptive-step tolerances; validation rules of solveras.py:99-169."""

    def __init__(self, epsilon=1e-4, incr_f=1.25, decr_f=0.85, safety_f=0.8,
                 adapt_cutoff=0.01, minh=1e-16):
        if epsilon <= 0:
            raise ValueError(f"epsilon must be positive but is {epsilon}")
        if incr_f <= 1.0:
            raise ValueError(f"incr_f must be > 1.0 but is {incr_f}")
        if decr_f >= 1.0:
            raise ValueError(f"decr_f must be < 1.0 but is {decr_f}")
        if safety_f > 1.0:
            raise ValueError(f"safety_f must be <= 1.0 but is {safety_f}")
        if adapt_cutoff >= 1.0:
            raise ValueError(f"adapt_cutoff must be < 1.0 but is {adapt_cutoff}")
        if minh <= 0:
            raise ValueError(f"minh must be positive but is {minh}")
        self.epsilon = float(epsilon)
        self.incr_f = float(incr_f)
        self.decr_f = float(decr_f)
        self.safety_f = float(safety_f)
        self.adapt_cutoff = float(adapt_cutoff)
        self.minh = float(minh)


class ETDConfig:
    """psi-function evaluation settings; validation rules of etd.py:99-131."""

    def __init__(self, modecutoff=0.01, contour_points=32, contour_radius=1.0):
        if modecutoff > 1.0 or modecutoff <= 0:
            raise ValueError(f"modecutoff must be between 0.0 and 1.0 but is {modecutoff}")
        if not isinstance(contour_points, int):
            raise TypeError(f"contour_points must be an integer but is {contour_points}")
        if contour_points <= 1:
            raise ValueError(f"contour_points must be an integer greater than 1 but is {contour_points}")
        if contour_radius <= 0:
            raise ValueError(f"contour_radius must greater than 0 but is {contour_radius}")
        self.modecutoff = modecutoff
        self.contour_points = contour_points
        self.contour_radius = contour_radius


# ----------------------------------------------------------------------------
# psi functions (psi_r = r! phi_r), etd.py:134-182
# ----------------------------------------------------------------------------
def psi1(z):
    return (np.exp(z) - 1) / z


def psi2(z):
    return 2 * (np.exp(z) - 1 - z) / z**2


def psi3(z):
    return 6 * (np.exp(z) - 1 - z - z**2 / 2) / z**3


def contour_nodes(etd):
    """Quadrature nodes R*exp(2*pi*i*(j+1/2)/N), j=0..N-1 (etd35.py:259)."""
    npts = etd.contour_points
    return etd.contour_radius * np.exp(2j * np.pi * np.arange(0.5, npts) / npts)


def _split(z, etd):
    small = np.abs(z) < etd.modecutoff
    return small, ~small


class SolverError(RuntimeError):
    pass


class MaxLoopsExceeded(SolverError):
    pass


class MinimumStepReached(SolverError):
    pass


# ----------------------------------------------------------------------------
# coefficient builders: (L, h[, etd]) -> dict of arrays / scalars
# ----------------------------------------------------------------------------
def _etd_fifth_order_rows(p1_14, p2_14, p1_12, p2_12, p1_34, p2_34, p1, p2, p3):
    """Linear combinations shared by ETD35 (a7x) and ETD5 (b_x); etd35.py:213-237."""
    return {
        "a21": p1_14 / 4.0,
        "a31": (p1_14 - p2_14 / 2.0) / 4.0,
        "a32": p2_14 / 8.0,
        "a41": (p1_12 - p2_12) / 2.0,
        "a43": p2_12 / 2.0,
        "a51": 3.0 * (p1_34 - 3.0 * p2_34 / 4.0) / 4.0,
        "a52": -3 * p1_34 / 8.0,
        "a54": 9 * p2_34 / 16.0,
        "a61": (-77 * p1 + 59 * p2) / 42.0,
        "a62": 8 * p1 / 7.0,
        "a63": (111 * p1 - 87 * p2) / 28.0,
        "a65": (-47 * p1 + 143 * p2) / 84.0,
        "a71": 7 * (257 * p1 - 497 * p2 + 270 * p3) / 2700,
        "a73": (1097 * p1 - 467 * p2 - 150 * p3) / 1350,
        "a74": 2 * (-49 * p1 + 199 * p2 - 135 * p3) / 225,
        "a75": (-313 * p1 + 883 * p2 - 90 * p3) / 1350,
        "a76": (509 * p1 - 2129 * p2 + 1830 * p3) / 2700,
    }


FIFTH_NAMES = ["a21", "a31", "a32", "a41", "a43", "a51", "a52", "a54",
               "a61", "a62", "a63", "a65", "a71", "a73", "a74", "a75", "a76"]


def etd35_coeffs(L, h, etd, cast=True):
    """ETD(3,5) coefficients, etd35.py:169-288 (contour uses factor*Z)."""
    if cast:  # the diagonal strategies cast lin_op (etd35.py:124); the diagonalized ones
        L = L.astype(np.complex128, copy=False)  # keep the eigenvalues' dtype (etd35.py:432)
    z = h * L
    out = {
        "EL14": np.exp(z / 4.0),
        "EL12": np.exp(z / 2.0),
        "EL34": np.exp(3.0 * z / 4.0),
        "EL": np.exp(z),
    }
    n = L.shape[0]
    for name in FIFTH_NAMES:
        out[name] = np.zeros(n, dtype=np.complex128)
    small, big = _split(z, etd)
    if np.any(big):
        zb = z[big]
        rows = _etd_fifth_order_rows(
            h * psi1(zb / 4), h * psi2(zb / 4),
            h * psi1(zb / 2), h * psi2(zb / 2),
            h * psi1(3 * zb / 4), h * psi2(3 * zb / 4),
            h * psi1(zb), h * psi2(zb), h * psi3(zb))
        for name in FIFTH_NAMES:
            out[name][big] = rows[name]
    if np.any(small):
        npts = etd.contour_points
        rr, zz = np.meshgrid(contour_nodes(etd), z[small])
        Z = zz + rr

        def avg(func, factor=1):
            return h * np.sum(func(factor * Z), axis=1) / npts

        rows = _etd_fifth_order_rows(
            avg(psi1, 1 / 4), avg(psi2, 1 / 4),
            avg(psi1, 1 / 2), avg(psi2, 1 / 2),
            avg(psi1, 3 / 4), avg(psi2, 3 / 4),
            avg(psi1), avg(psi2), avg(psi3))
        for name in FIFTH_NAMES:
            out[name][small] = rows[name]
    return out


def etd5_coeffs(L, h, etd):
    """ETD5 coefficients, etd5.py:115-203 (contour uses Z/4, Z/2, 3*Z/4)."""
    L = L.astype(np.complex128, copy=False)
    z = h * L
    out = {
        "EL14": np.exp(z / 4.0),
        "EL12": np.exp(z / 2.0),
        "EL34": np.exp(3.0 * z / 4.0),
        "EL": np.exp(z),
    }
    n = L.shape[0]
    for name in FIFTH_NAMES:
        out[name] = np.zeros(n, dtype=np.complex128)
    small, big = _split(z, etd)
    zb = z[big]
    rows = _etd_fifth_order_rows(
        h * psi1(zb / 4), h * psi2(zb / 4),
        h * psi1(zb / 2), h * psi2(zb / 2),
        h * psi1(3 * zb / 4), h * psi2(3 * zb / 4),
        h * psi1(zb), h * psi2(zb), h * psi3(zb))
    for name in FIFTH_NAMES:
        out[name][big] = rows[name]
    npts = etd.contour_points
    rr, zz = np.meshgrid(contour_nodes(etd), z[small])
    Z = zz + rr

    def avg(v):
        return h * np.sum(v, axis=1) / npts

    rows = _etd_fifth_order_rows(
        avg(psi1(Z / 4)), avg(psi2(Z / 4)),
        avg(psi1(Z / 2)), avg(psi2(Z / 2)),
        avg(psi1(3 * Z / 4)), avg(psi2(3 * Z / 4)),
        avg(psi1(Z)), avg(psi2(Z)), avg(psi3(Z)))
    for name in FIFTH_NAMES:
        out[name][small] = rows[name]
    # ETD5 names its final-row weights b1, b3..b6 (etd5.py:166-171)
    for a, b in (("a71", "b1"), ("a73", "b3"), ("a74", "b4"), ("a75", "b5"), ("a76", "b6")):
        out[b] = out.pop(a)
    return out


def _fourth_order_rows(p1_12, p2_12, p1, p2, p3):
    """Krogstad rows shared by ETD34 and ETD4, etd34.py:117-124."""
    return {
        "a21": 0.5 * p1_12,
        "a31": 0.5 * (p1_12 - p2_12),
        "a32": 0.5 * p2_12,
        "a41": p1 - p2,
        "a43": p2,
        "a51": p1 - (3.0 / 2) * p2 + (2.0 / 3) * p3,
        "a52": p2 - (2.0 / 3) * p3,
        "a54": -(1.0 / 2) * p2 + (2.0 / 3) * p3,
    }


FOURTH_NAMES = ["a21", "a31", "a32", "a41", "a43", "a51", "a52", "a54"]


def etd34_coeffs(L, h, etd, cast=True):
    """ETD(3,4) coefficients, etd34.py:96-149; ETD4 uses the same (etd4.py:87-139)."""
    if cast:  # the diagonal strategies cast lin_op (etd35.py:124); the diagonalized ones
        L = L.astype(np.complex128, copy=False)  # keep the eigenvalues' dtype (etd35.py:432)
    z = h * L
    out = {"EL": np.exp(z), "EL2": np.exp(z / 2)}
    n = L.shape[0]
    for name in FOURTH_NAMES:
        out[name] = np.zeros(n, dtype=np.complex128)
    small, big = _split(z, etd)
    zb = z[big]
    rows = _fourth_order_rows(h * psi1(zb / 2), h * psi2(zb / 2),
                              h * psi1(zb), h * psi2(zb), h * psi3(zb))
    for name in FOURTH_NAMES:
        out[name][big] = rows[name]
    npts = etd.contour_points
    rr, zz = np.meshgrid(contour_nodes(etd), z[small])
    Z = zz + rr
    rows = _fourth_order_rows(h * np.sum(psi1(Z / 2), axis=1) / npts,
                              h * np.sum(psi2(Z / 2), axis=1) / npts,
                              h * np.sum(psi1(Z), axis=1) / npts,
                              h * np.sum(psi2(Z), axis=1) / npts,
                              h * np.sum(psi3(Z), axis=1) / npts)
    for name in FOURTH_NAMES:
        out[name][small] = rows[name]
    return out


def etd4_coeffs(L, h, etd):
    """ETD4 (Krogstad): ETD34's rows with a51/a52/a54 named b1/b2/b4 (etd4.py:113-120)."""
    out = etd34_coeffs(L, h, etd)
    for a, b in (("a51", "b1"), ("a52", "b2"), ("a54", "b4")):
        out[b] = out.pop(a)
    return out


def if_exp_coeffs(L, h):
    """IF34/IF4 integrating factors, if34.py:81-93 / if4.py:73-83 (L keeps its dtype)."""
    z = h * L
    return {"EL": np.exp(z), "EL2": np.exp(z / 2)}


def if45dp_coeffs(L, h):
    """IF Dormand-Prince coefficients, if45dp.py:183-238 (L keeps its dtype)."""
    z = h * L
    c = {}
    c["EL15"] = np.exp(z / 5)
    c["EL310"] = np.exp(3 * z / 10)
    c["EL45"] = np.exp(4 * z / 5)
    c["EL89"] = np.exp(8 * z / 9)
    c["EL"] = np.exp(z)
    EL710 = np.exp(7 * z / 10)
    EL19 = np.exp(z / 9)
    c["a21"] = h * c["EL15"] / 5.0
    c["a31"] = 3 * h * c["EL310"] / 40.0
    c["a32"] = 9 * h * np.exp(z / 10) / 40.0
    c["a41"] = 44 * h * c["EL45"] / 45.0
    c["a42"] = -56 * h * np.exp(3 * z / 5) / 15.0
    c["a43"] = 32 * h * np.exp(z / 2) / 9.0
    c["a51"] = 19372 * h * c["EL89"] / 6561.0
    c["a52"] = -25360 * h * np.exp(31 * z / 45) / 2187.0
    c["a53"] = 64448.0 * h * np.exp(53 * z / 90) / 6561.0
    c["a54"] = -212 * h * np.exp(4 * z / 45) / 729.0
    c["a61"] = 9017 * h * c["EL"] / 3168.0
    c["a62"] = -355 * h * c["EL45"] / 33.0
    c["a63"] = 46732 * h * EL710 / 5247.0
    c["a64"] = 49 * h * c["EL15"] / 176.0
    c["a65"] = -5103 * h * EL19 / 18656.0
    c["a71"] = 35 * h * c["EL"] / 384.0
    c["a73"] = 500 * h * EL710 / 1113.0
    c["a75"] = -2187 * h * EL19 / 6784.0
    c["a74"] = 125 * h * c["EL15"] / 192.0
    c["a76"] = 11 * h / 84.0
    c["r1"] = h * 71 * c["EL"] / 57600.0
    c["r3"] = -71 * h * EL710 / 16695.0
    c["r4"] = 17 * h * c["EL15"] / 1920.0
    c["r5"] = -17253 * h * EL19 / 339200.0
    c["r6"] = 22 * h / 525.0
    c["r7"] = -h / 40.0
    return c


# ----------------------------------------------------------------------------
# methods: coefficient cache + stage pass (the reference's strategy objects)
# ----------------------------------------------------------------------------
class _Method:
    adaptive = True
    q = 4

    def __init__(self, L, nl, etd=None):
        self.L = L
        self.nl = nl
        self.etd = etd if etd is not None else ETDConfig()
        self.reset()
        self.nl_calls = 0

    def reset(self):
        self.h_coeff = None
        self.c = None
        self.NL1 = None
        self.started = False

    def _nl(self, v):
        self.nl_calls += 1
        return self.nl(v)

    def coeffs(self, h):
        if h != self.h_coeff:
            self.h_coeff = h
            self.c = self._build(h)
        return self.c


class ETD35(_Method):
    """etd35.py:301-345 stages; NL1 recomputed after an accepted attempt."""
    name = "ETD35"

    def _build(self, h):
        return etd35_coeffs(self.L, h, self.etd, cast=not getattr(self, "keep_dtype", False))

    def attempt(self, u, h, accept):
        c = self.coeffs(h)
        if not self.started:
            self.NL1 = self._nl(u)
            self.started = True
        if accept:
            self.NL1 = self._nl(u)
        NL1 = self.NL1
        k = c["EL14"] * u + c["a21"] * NL1
        NL2 = self._nl(k)
        k = c["EL14"] * u + c["a31"] * NL1 + c["a32"] * NL2
        NL3 = self._nl(k)
        k = c["EL12"] * u + c["a41"] * NL1 + c["a43"] * NL3
        NL4 = self._nl(k)
        k = c["EL34"] * u + c["a51"] * NL1 + c["a52"] * (NL2 - NL3) + c["a54"] * NL4
        NL5 = self._nl(k)
        k = (c["EL"] * u + c["a61"] * NL1 + c["a62"] * (NL2 - 3 * NL4 / 2.0)
             + c["a63"] * NL3 + c["a65"] * NL5)
        NL6 = self._nl(k)
        k = (c["EL"] * u + c["a71"] * NL1 + c["a73"] * NL3 + c["a74"] * NL4
             + c["a75"] * NL5 + c["a76"] * NL6)
        err = c["a75"] * (-NL1 + 4 * NL3 - 6 * NL4 + 4 * NL5 - NL6)
        return k, err


class ETD34(_Method):
    """etd34.py:162-191 stages; FSAL NL1 <- NL5 after an accepted attempt."""
    name = "ETD34"

    def _build(self, h):
        return etd34_coeffs(self.L, h, self.etd, cast=not getattr(self, "keep_dtype", False))

    def reset(self):
        super().reset()
        self.NL5 = None

    def attempt(self, u, h, accept):
        c = self.coeffs(h)
        if not self.started:
            self.NL1 = self._nl(u)
            self.started = True
        if accept:
            self.NL1 = self.NL5.copy()
        NL1 = self.NL1
        k = c["EL2"] * u + c["a21"] * NL1
        NL2 = self._nl(k)
        k = c["EL2"] * u + c["a31"] * NL1 + c["a32"] * NL2
        NL3 = self._nl(k)
        k = c["EL"] * u + c["a41"] * NL1 + c["a43"] * NL3
        NL4 = self._nl(k)
        k = c["EL"] * u + c["a51"] * NL1 + c["a52"] * (NL2 + NL3) + c["a54"] * NL4
        NL5 = self._nl(k)
        self.NL5 = NL5
        err = c["a54"] * (NL4 - NL5)
        return k, err


class IF34(_Method):
    """if34.py:99-131 stages; FSAL NL1 <- NL5 after an accepted attempt."""
    name = "IF34"

    def _build(self, h):
        return if_exp_coeffs(self.L, h)

    def reset(self):
        super().reset()
        self.NL5 = None

    def attempt(self, u, h, accept):
        c = self.coeffs(h)
        if not self.started:
            self.NL1 = self._nl(u)
            self.started = True
        if accept:
            self.NL1 = self.NL5.copy()
        NL1 = self.NL1
        EL, EL2 = c["EL"], c["EL2"]
        k = EL2 * u + h * EL2 * NL1 / 2.0
        NL2 = self._nl(k)
        k = EL2 * u + h * NL2 / 2.0
        NL3 = self._nl(k)
        k = EL * u + h * EL2 * NL3
        NL4 = self._nl(k)
        k = EL * u + h * (EL * NL1 / 6.0 + EL2 * NL2 / 3.0 + EL2 * NL3 / 3.0 + NL4 / 6.0)
        NL5 = self._nl(k)
        self.NL5 = NL5
        err = h * (NL4 - NL5) / 6.0
        return k, err


class IF45DP(_Method):
    """if45dp.py:112-181 stages; FSAL NL1 <- NL7 (``elif`` after first init)."""
    name = "IF45DP"
    q = 5

    def __init__(self, L, nl, etd=None):
        if len(L.shape) > 1:
            raise ValueError("IF45DP only handles 1D linear operators (diagonal systems): "
                             "try IF34,ETD34, or ETD35")
        super().__init__(L, nl, etd)

    def _build(self, h):
        return if45dp_coeffs(self.L, h)

    def reset(self):
        super().reset()
        self.NL7 = None

    def attempt(self, u, h, accept):
        c = self.coeffs(h)
        if not self.started:
            self.NL1 = self._nl(u)
            self.started = True
        elif accept:
            self.NL1 = self.NL7.copy()
        NL1 = self.NL1
        k = c["EL15"] * u + c["a21"] * NL1
        NL2 = self._nl(k)
        k = c["EL310"] * u + c["a31"] * NL1 + c["a32"] * NL2
        NL3 = self._nl(k)
        k = c["EL45"] * u + c["a41"] * NL1 + c["a42"] * NL2 + c["a43"] * NL3
        NL4 = self._nl(k)
        k = (c["EL89"] * u + c["a51"] * NL1 + c["a52"] * NL2 + c["a53"] * NL3
             + c["a54"] * NL4)
        NL5 = self._nl(k)
        k = (c["EL"] * u + c["a61"] * NL1 + c["a62"] * NL2 + c["a63"] * NL3
             + c["a64"] * NL4 + c["a65"] * NL5)
        NL6 = self._nl(k)
        k = (c["EL"] * u + c["a71"] * NL1 + c["a73"] * NL3 + c["a74"] * NL4
             + c["a75"] * NL5 + c["a76"] * NL6)
        NL7 = self._nl(k)
        self.NL7 = NL7
        err = (c["r1"] * NL1 + c["r3"] * NL3 + c["r4"] * NL4 + c["r5"] * NL5
               + c["r6"] * NL6 + c["r7"] * NL7)
        return k, err


class ETD4(_Method):
    """etd4.py:152-175 stages; constant step, FSAL NL1 = nl(k) every step."""
    name = "ETD4"
    adaptive = False

    def _build(self, h):
        return etd4_coeffs(self.L, h, self.etd)

    def advance(self, u, h):
        c = self.coeffs(h)
        if not self.started:
            self.NL1 = self._nl(u)
            self.started = True
        NL1 = self.NL1
        k = c["EL2"] * u + c["a21"] * NL1
        NL2 = self._nl(k)
        k = c["EL2"] * u + c["a31"] * NL1 + c["a32"] * NL2
        NL3 = self._nl(k)
        k = c["EL"] * u + c["a41"] * NL1 + c["a43"] * NL3
        NL4 = self._nl(k)
        k = c["EL"] * u + c["b1"] * NL1 + c["b2"] * (NL2 + NL3) + c["b4"] * NL4
        self.NL1 = self._nl(k)
        return k


class ETD5(_Method):
    """etd5.py:219-261 stages; constant step, FSAL NL1 = nl(k) every step."""
    name = "ETD5"
    adaptive = False

    def _build(self, h):
        return etd5_coeffs(self.L, h, self.etd)

    def advance(self, u, h):
        c = self.coeffs(h)
        if not self.started:
            self.NL1 = self._nl(u)
            self.started = True
        NL1 = self.NL1
        k = c["EL14"] * u + c["a21"] * NL1
        NL2 = self._nl(k)
        k = c["EL14"] * u + c["a31"] * NL1 + c["a32"] * NL2
        NL3 = self._nl(k)
        k = c["EL12"] * u + c["a41"] * NL1 + c["a43"] * NL3
        NL4 = self._nl(k)
        k = c["EL34"] * u + c["a51"] * NL1 + c["a52"] * (NL2 - NL3) + c["a54"] * NL4
        NL5 = self._nl(k)
        k = (c["EL"] * u + c["a61"] * NL1 + c["a62"] * (NL2 - 3 * NL4 / 2.0)
             + c["a63"] * NL3 + c["a65"] * NL5)
        NL6 = self._nl(k)
        k = (c["EL"] * u + c["b1"] * NL1 + c["b3"] * NL3 + c["b4"] * NL4
             + c["b5"] * NL5 + c["b6"] * NL6)
        self.NL1 = self._nl(k)
        return k


class IF4(_Method):
    """if4.py:96-122 stages; constant step, FSAL NL1 = nl(k) every step."""
    name = "IF4"
    adaptive = False

    def _build(self, h):
        return if_exp_coeffs(self.L, h)

    def advance(self, u, h):
        c = self.coeffs(h)
        if not self.started:
            self.NL1 = self._nl(u)
            self.started = True
        NL1 = self.NL1
        EL, EL2 = c["EL"], c["EL2"]
        k = EL2 * u + h * EL2 * NL1 / 2.0
        NL2 = self._nl(k)
        k = EL2 * u + h * NL2 / 2.0
        NL3 = self._nl(k)
        k = EL * u + h * EL2 * NL3
        NL4 = self._nl(k)
        k = EL * u + h * (EL * NL1 / 6.0 + EL2 * NL2 / 3.0 + EL2 * NL3 / 3.0 + NL4 / 6.0)
        self.NL1 = self._nl(k)
        return k


METHODS = {m.name: m for m in (ETD35, ETD34, ETD4, ETD5, IF34, IF45DP, IF4)}


class Diagonalized:
    """Diagonalized strategies (etd35.py:348-495, etd34.py:194-304, if34.py:137-196).

    lin_op = S diag(lam) S^-1 (np.linalg.eig); the diagonal scheme runs on
    v = S^-1 u with the eigenvalues, every stage's NL_k = S^-1 nl(S k); v and
    NL1 = S^-1 nl(u) are set from u at the first attempt and after an accept
    (NL1 <- NL5 instead for the FSAL methods).  Returns (S k, err) with err
    left in the eigenbasis, as the reference does.
    """

    adaptive = True

    def __init__(self, base, L, nl, etd=None):
        L = np.asarray(L)
        if L.ndim != 2:
            raise ValueError("Cannot diagonalize a 1D system")
        if np.linalg.cond(L) > 1e16:
            raise ValueError("Linear operator is non-invertible")
        self.eig, self.S = np.linalg.eig(L)
        self.Sinv = np.linalg.inv(self.S)
        self.user_nl = nl
        self.m = base(self.eig, self._nl_v, etd)
        self.m.keep_dtype = True
        self.q = self.m.q
        self.name = base.name
        self.recompute = base is ETD35
        self.v = None

    def _nl_v(self, k):
        return self.Sinv.dot(self.user_nl(self.S.dot(k)))

    def reset(self):
        self.m.reset()
        self.v = None

    def attempt(self, u, h, accept):
        m = self.m
        inner_accept = accept
        if not m.started:
            m.NL1 = self.Sinv.dot(self.user_nl(u))
            self.v = self.Sinv.dot(u)
            m.started = True
            inner_accept = False
        elif accept:
            self.v = self.Sinv.dot(u)
            if self.recompute:
                m.NL1 = self.Sinv.dot(self.user_nl(u))
                inner_accept = False
        k, err = m.attempt(self.v, h, inner_accept)
        return self.S.dot(k), err


DIAGONALIZABLE = ("ETD35", "ETD34", "IF34")


# ----------------------------------------------------------------------------
# drivers
# ----------------------------------------------------------------------------
class AdaptiveDriver:
    """Adaptive accept/reject controller and evolve loop (solveras.py:255-650).

    ``log`` records every attempt as ``(h, s, accepted)``.
    """

    MAX_LOOPS = 50
    MAX_S = 4.0
    MIN_S = 0.25

    def __init__(self, method, config=None):
        self.m = method
        self.config = config if config is not None else SolverConfig()
        self.accept = False
        self.t, self.u = [], []
        self.log = []

    def reset(self):
        self.t, self.u = [], []
        self.accept = False
        self.m.reset()

    def compute_s(self, u, err):
        """Embedded error norm, solveras.py:449-455."""
        cfg = self.config
        with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
            magu = np.abs(u)
            idx = magu / magu.max() > cfg.adapt_cutoff
            tol = cfg.epsilon * np.linalg.norm(u[idx])
            return cfg.safety_f * np.power(tol / np.linalg.norm(err[idx]), 1.0 / self.m.q)

    def reject_h(self, s, h):
        """solveras.py:490-506."""
        self.accept = False
        if np.isinf(s) or np.isnan(s):
            return self.MIN_S * h
        s = np.max([s, self.MIN_S])
        s = np.min([s, self.config.decr_f])
        return s * h

    def accept_h(self, s, h):
        """solveras.py:543-554."""
        self.accept = True
        s = np.min([s, self.MAX_S])
        if s > self.config.incr_f:
            return s * h
        return h

    def step(self, u, h_suggest):
        """One accepted step; solveras.py:378-410."""
        h = h_suggest
        assert h >= 0.0
        numloops = 0
        while True:
            with np.errstate(over="ignore", invalid="ignore", divide="ignore"):
                unew, err = self.m.attempt(u, h, self.accept)
            s = self.compute_s(unew, err)
            if np.isinf(s) or np.isnan(s) or s < 1.0:
                self.log.append((h, float(s), False))
                h = self.reject_h(s, h)
            else:
                self.log.append((h, float(s), True))
                h_suggest = self.accept_h(s, h)
                return unew, h, h_suggest
            numloops += 1
            if numloops > self.MAX_LOOPS:
                raise MaxLoopsExceeded("adaptive step made too many attempts")
            if h < self.config.minh:
                raise MinimumStepReached("adaptive step reached minimum step size")

    def evolve(self, u, t0, tf, h_init=None, store_data=True, store_freq=1):
        """solveras.py:556-650."""
        self.reset()
        tc = t0
        if store_data:
            self.t.append(t0)
            self.u.append(u)
        if h_init is None:
            h_init = (tf - t0) / 100.0
        h = h_init
        if tc + h > tf:
            h = tf - tc
        step_count = 0
        while tc < tf:
            u, h, h_suggest = self.step(u, h)
            tc += h
            step_count += 1
            if tc + h_suggest > tf:
                h = tf - tc
            else:
                h = h_suggest
            if store_data and (step_count % store_freq == 0):
                self.t.append(tc)
                self.u.append(u)
        self.tc = tc
        self.steps = step_count
        return u


class ConstantDriver:
    """Constant-step driver, solvercs.py:163-279."""

    def __init__(self, method):
        self.m = method
        self.t, self.u = [], []

    def reset(self):
        self.t, self.u = [], []
        self.m.reset()

    def step(self, u, h):
        assert h >= 0.0
        return self.m.advance(u, h)

    def evolve(self, u, t0, tf, h, store_data=True, store_freq=1):
        self.reset()
        tc = t0
        if store_data:
            self.t.append(t0)
            self.u.append(u)
        if tc + h > tf:
            raise ValueError("Step size h must be <= (tf - t0); reduce h or extend tf.")
        step_count = 0
        while tc < tf:
            u = self.step(u, h)
            tc += h
            step_count += 1
            if store_data and (step_count % store_freq == 0):
                self.t.append(tc)
                self.u.append(u)
        self.tc = tc
        self.steps = step_count
        return u


def make_solver(name, L, nl, config=None, etd=None, diagonalize=False):
    """Method + driver for ``name`` in METHODS (``diagonalize``: a 2-D lin_op in its eigenbasis)."""
    if diagonalize:
        if name not in DIAGONALIZABLE:
            raise ValueError(f"{name} has no diagonalized strategy")
        m = Diagonalized(METHODS[name], L, nl, etd)
    else:
        m = METHODS[name](L, nl, etd)
    if m.adaptive:
        return AdaptiveDriver(m, config)
    return ConstantDriver(m)


def rel_l2(a, b):
    """||a-b||_2 / ||b||_2 (the parity metric of every test and bench line)."""
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def ulp_distance(a, b):
    """Max distance in units of the last place between two float64 arrays."""
    a = np.ascontiguousarray(a, dtype=np.float64).view(np.int64)
    b = np.ascontiguousarray(b, dtype=np.float64).view(np.int64)
    a = np.where(a < 0, np.int64(-(2**63)) - a, a)
    b = np.where(b < 0, np.int64(-(2**63)) - b, b)
    return int(np.max(np.abs(a - b))) if a.size else 0

