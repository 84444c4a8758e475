"""Dense (non-diagonal) strategies of the reference, restated in NumPy FP64.

Test infrastructure, like the rest of ``oracle``: the product never imports
it.  A square 2-D lin_op with ``diagonalize=False`` selects these strategies
in the reference (etd35.py:823-830 and the other wrappers).  Coefficients are
matrices: exponentials by ``scipy.linalg.expm`` and psi functions by the
Cauchy integral over the contour r_m = R exp(2 pi i (m + 1/2) / N),
psi_k(fZ) ~ (1/N) sum_m r_m psi_k(r_m) (r_m I - fZ)^-1; stages use
matrix-vector products.

  ETD35  etd35.py:498-720     ETD34  etd34.py:307-436     IF34  if34.py:202-254
  ETD5   etd5.py:264-482      ETD4   etd4.py:178-308      IF4   if4.py:125-199
"""
import numpy as np
from scipy.linalg import expm

from .solvers import ETDConfig, psi1, psi2, psi3


def _points(etd):
    return etd.contour_radius * np.exp(2j * np.pi * np.arange(0.5, etd.contour_points) / etd.contour_points)


def fifth_order(lin_op, h, etd):
    """ETD35 / ETD5 matrix coefficients (etd35.py:582-646, etd5.py:345-421)."""
    z = h * lin_op
    c = {"EL14": expm(z / 4.0), "EL12": expm(z / 2.0), "EL34": expm(3.0 * z / 4.0), "EL": expm(z)}
    shape = lin_op.shape
    p1_14, p2_14, p1_12, p2_12, p1_34, p2_34 = [np.zeros(shape, dtype=np.complex128) for _ in range(6)]
    p1_1, p2_1, p3_1 = [np.zeros(shape, dtype=np.complex128) for _ in range(3)]
    npts = etd.contour_points
    for point in _points(etd):
        Q14 = np.linalg.inv(point * np.eye(*shape) - z / 4.0)
        Q12 = np.linalg.inv(point * np.eye(*shape) - z / 2.0)
        Q34 = np.linalg.inv(point * np.eye(*shape) - 3 * z / 4.0)
        Q = np.linalg.inv(point * np.eye(*shape) - z)
        p1_14 += point * psi1(point) * Q14 / npts
        p2_14 += point * psi2(point) * Q14 / npts
        p1_12 += point * psi1(point) * Q12 / npts
        p2_12 += point * psi2(point) * Q12 / npts
        p1_34 += point * psi1(point) * Q34 / npts
        p2_34 += point * psi2(point) * Q34 / npts
        p1_1 += point * psi1(point) * Q / npts
        p2_1 += point * psi2(point) * Q / npts
        p3_1 += point * psi3(point) * Q / npts
    c["a21"] = h * p1_14 / 4.0
    c["a31"] = h * (p1_14 - p2_14 / 2.0) / 4.0
    c["a32"] = h * p2_14 / 8.0
    c["a41"] = h * (p1_12 - p2_12) / 2.0
    c["a43"] = h * p2_12 / 2.0
    c["a51"] = h * 3.0 * (p1_34 - 3.0 * p2_34 / 4.0) / 4.0
    c["a52"] = -3 * h * p1_34 / 8.0
    c["a54"] = h * 9 * p2_34 / 16.0
    c["a61"] = h * (-77 * p1_1 + 59 * p2_1) / 42.0
    c["a62"] = h * 8 * p1_1 / 7.0
    c["a63"] = h * (111 * p1_1 - 87 * p2_1) / 28.0
    c["a65"] = h * (-47 * p1_1 + 143 * p2_1) / 84.0
    c["a71"] = h * 7 * (257 * p1_1 - 497 * p2_1 + 270 * p3_1) / 2700
    c["a73"] = h * (1097 * p1_1 - 467 * p2_1 - 150 * p3_1) / 1350
    c["a74"] = h * 2 * (-49 * p1_1 + 199 * p2_1 - 135 * p3_1) / 225
    c["a75"] = h * (-313 * p1_1 + 883 * p2_1 - 90 * p3_1) / 1350
    c["a76"] = h * (509 * p1_1 - 2129 * p2_1 + 1830 * p3_1) / 2700
    return c


def fourth_order(lin_op, h, etd):
    """ETD34 / ETD4 matrix coefficients (etd34.py:355-400, etd4.py:222-268)."""
    z = h * lin_op
    c = {"EL": expm(z), "EL2": expm(z / 2)}
    shape = lin_op.shape
    p1_12, p2_12, p1_1, p2_1, p3_1 = [np.zeros(shape, dtype=np.complex128) for _ in range(5)]
    npts = etd.contour_points
    for point in _points(etd):
        Q = np.linalg.inv(point * np.eye(*shape) - z)
        Q2 = np.linalg.inv(point * np.eye(*shape) - z / 2)
        p1_12 += point * psi1(point) * Q2 / npts
        p2_12 += point * psi2(point) * Q2 / npts
        p1_1 += point * psi1(point) * Q / npts
        p2_1 += point * psi2(point) * Q / npts
        p3_1 += point * psi3(point) * Q / npts
    c["a21"] = 0.5 * h * p1_12
    c["a31"] = 0.5 * h * (p1_12 - p2_12)
    c["a32"] = 0.5 * h * p2_12
    c["a41"] = h * (p1_1 - p2_1)
    c["a43"] = h * p2_1
    c["a51"] = h * (p1_1 - (3.0 / 2) * p2_1 + (2.0 / 3) * p3_1)
    c["a52"] = h * (p2_1 - (2.0 / 3) * p3_1)
    c["a54"] = h * (-(1.0 / 2) * p2_1 + (2.0 / 3) * p3_1)
    return c


def if_matrices(lin_op, h):
    """IF34 / IF4 (if34.py:227-231, if4.py:160-170): lin_op keeps its dtype."""
    z = h * lin_op
    return {"EL": expm(z), "EL2": expm(z / 2)}


class _Dense:
    adaptive = True
    q = 4

    def __init__(self, L, nl, etd=None):
        L = np.asarray(L)
        if L.ndim != 2 or L.shape[0] != L.shape[1]:
            raise ValueError("dense strategies need a square 2-D lin_op")
        self.L = L.astype(np.complex128, copy=False) if self.cast else L
        self.nl = nl
        self.etd = etd if etd is not None else ETDConfig()
        self.reset()

    def reset(self):
        self.h_coeff = None
        self.c = None
        self.NL1 = None
        self.started = False

    def coeffs(self, h):
        if h != self.h_coeff:
            self.h_coeff = h
            self.c = self._build(h)
        return self.c


class ETD35(_Dense):
    """etd35.py:648-720; NL1 = nl(u) after an accepted attempt."""
    name, cast = "ETD35", True

    def _build(self, h):
        return fifth_order(self.L, h, self.etd)

    def attempt(self, u, h, accept):
        c = self.coeffs(h)
        if not self.started:
            self.NL1 = self.nl(u)
            self.started = True
        if accept:
            self.NL1 = self.nl(u)
        NL1 = self.NL1
        k = c["EL14"].dot(u) + c["a21"].dot(NL1)
        NL2 = self.nl(k)
        k = c["EL14"].dot(u) + c["a31"].dot(NL1) + c["a32"].dot(NL2)
        NL3 = self.nl(k)
        k = c["EL12"].dot(u) + c["a41"].dot(NL1) + c["a43"].dot(NL3)
        NL4 = self.nl(k)
        k = c["EL34"].dot(u) + c["a51"].dot(NL1) + c["a52"].dot(NL2 - NL3) + c["a54"].dot(NL4)
        NL5 = self.nl(k)
        k = (c["EL"].dot(u) + c["a61"].dot(NL1) + c["a62"].dot(NL2 - 3 * NL4 / 2.0) + c["a63"].dot(NL3)
             + c["a65"].dot(NL5))
        NL6 = self.nl(k)
        k = (c["EL"].dot(u) + c["a71"].dot(NL1) + c["a73"].dot(NL3) + c["a74"].dot(NL4) + c["a75"].dot(NL5)
             + c["a76"].dot(NL6))
        err = c["a75"].dot(-NL1 + 4 * NL3 - 6 * NL4 + 4 * NL5 - NL6)
        return k, err


class ETD34(_Dense):
    """etd34.py:402-436; FSAL NL1 <- NL5 after an accepted attempt."""
    name, cast = "ETD34", True

    def _build(self, h):
        return fourth_order(self.L, h, self.etd)

    def attempt(self, u, h, accept):
        c = self.coeffs(h)
        if not self.started:
            self.NL1 = self.nl(u)
            self.started = True
        if accept:
            self.NL1 = self.NL5.copy()
        NL1 = self.NL1
        k = c["EL2"].dot(u) + c["a21"].dot(NL1)
        NL2 = self.nl(k)
        k = c["EL2"].dot(u) + c["a31"].dot(NL1) + c["a32"].dot(NL2)
        NL3 = self.nl(k)
        k = c["EL"].dot(u) + c["a41"].dot(NL1) + c["a43"].dot(NL3)
        NL4 = self.nl(k)
        k = c["EL"].dot(u) + c["a51"].dot(NL1) + c["a52"].dot(NL2 + NL3) + c["a54"].dot(NL4)
        NL5 = self.nl(k)
        self.NL5 = NL5
        return k, c["a54"].dot(NL4 - NL5)


class IF34(_Dense):
    """if34.py:233-254; FSAL NL1 <- NL5 after an accepted attempt."""
    name, cast = "IF34", False

    def _build(self, h):
        return if_matrices(self.L, h)

    def attempt(self, u, h, accept):
        c = self.coeffs(h)
        if not self.started:
            self.NL1 = self.nl(u)
            self.started = True
        if accept:
            self.NL1 = self.NL5.copy()
        EL, EL2, NL1 = c["EL"], c["EL2"], self.NL1
        k = EL2.dot(u) + h * EL2.dot(NL1 / 2.0)
        NL2 = self.nl(k)
        k = EL2.dot(u) + h * NL2 / 2.0
        NL3 = self.nl(k)
        k = EL.dot(u) + h * EL2.dot(NL3)
        NL4 = self.nl(k)
        k = EL.dot(u) + h * (EL.dot(NL1 / 6.0) + EL2.dot(NL2 / 3.0) + EL2.dot(NL3 / 3.0) + NL4 / 6.0)
        NL5 = self.nl(k)
        self.NL5 = NL5
        return k, h * (NL4 - NL5) / 6.0


class _DenseConstant(_Dense):
    adaptive = False

    def advance(self, u, h):
        if not self.started:
            self.NL1 = self.nl(u)
            self.started = True
        return self._stages(self.coeffs(h), u, h)


class ETD4(_DenseConstant):
    """etd4.py:276-308; NL1 = nl(k) at the end of every step."""
    name, cast = "ETD4", True

    def _build(self, h):
        return fourth_order(self.L, h, self.etd)

    def _stages(self, c, u, h):
        NL1 = self.NL1
        k = c["EL2"].dot(u) + c["a21"].dot(NL1)
        NL2 = self.nl(k)
        k = c["EL2"].dot(u) + c["a31"].dot(NL1) + c["a32"].dot(NL2)
        NL3 = self.nl(k)
        k = c["EL"].dot(u) + c["a41"].dot(NL1) + c["a43"].dot(NL3)
        NL4 = self.nl(k)
        k = c["EL"].dot(u) + c["a51"].dot(NL1) + c["a52"].dot(NL2 + NL3) + c["a54"].dot(NL4)
        self.NL1 = self.nl(k)
        return k


class ETD5(_DenseConstant):
    """etd5.py:423-482; NL1 = nl(k) at the end of every step."""
    name, cast = "ETD5", True

    def _build(self, h):
        return fifth_order(self.L, h, self.etd)

    def _stages(self, c, u, h):
        NL1 = self.NL1
        k = c["EL14"].dot(u) + c["a21"].dot(NL1)
        NL2 = self.nl(k)
        k = c["EL14"].dot(u) + c["a31"].dot(NL1) + c["a32"].dot(NL2)
        NL3 = self.nl(k)
        k = c["EL12"].dot(u) + c["a41"].dot(NL1) + c["a43"].dot(NL3)
        NL4 = self.nl(k)
        k = c["EL34"].dot(u) + c["a51"].dot(NL1) + c["a52"].dot(NL2 - NL3) + c["a54"].dot(NL4)
        NL5 = self.nl(k)
        k = (c["EL"].dot(u) + c["a61"].dot(NL1) + c["a62"].dot(NL2 - 3 * NL4 / 2.0) + c["a63"].dot(NL3)
             + c["a65"].dot(NL5))
        NL6 = self.nl(k)
        k = (c["EL"].dot(u) + c["a71"].dot(NL1) + c["a73"].dot(NL3) + c["a74"].dot(NL4) + c["a75"].dot(NL5)
             + c["a76"].dot(NL6))
        self.NL1 = self.nl(k)
        return k


class IF4(_DenseConstant):
    """if4.py:172-199; NL1 = nl(k) at the end of every step."""
    name, cast = "IF4", False

    def _build(self, h):
        return if_matrices(self.L, h)

    def _stages(self, c, u, h):
        EL, EL2, NL1 = c["EL"], c["EL2"], self.NL1
        k = EL2.dot(u) + h * EL2.dot(NL1 / 2.0)
        NL2 = self.nl(k)
        k = EL2.dot(u) + h * NL2 / 2.0
        NL3 = self.nl(k)
        k = EL.dot(u) + h * EL2.dot(NL3)
        NL4 = self.nl(k)
        k = EL.dot(u) + h * (EL.dot(NL1 / 6.0) + EL2.dot(NL2 / 3.0) + EL2.dot(NL3 / 3.0) + NL4 / 6.0)
        self.NL1 = self.nl(k)
        return k


METHODS = {m.name: m for m in (ETD35, ETD34, IF34, ETD4, ETD5, IF4)}


def make_solver(name, L, nl, config=None, etd=None):
    """Dense method + the oracle's drivers."""
    from .solvers import AdaptiveDriver, ConstantDriver
    m = METHODS[name](L, nl, etd)
    return AdaptiveDriver(m, config) if m.adaptive else ConstantDriver(m)
