"""``PairwiseInteractionABFS``: the QC/MM and QC/QC electrostatic interactions of the ABFS model.

Options, validation and defaults as pMolecule-1.9.0/extensions/pyrex/pMolecule.PairwiseInteraction.pyx:107-240;
the electrostatic spline ordinates as its ``MakeElectrostaticSplines`` (pyx:24-105,182-194): the same
function, switch and grid, with the switched-region integrals done by a panelled Gauss-Legendre rule
(the reference uses an adaptive Simpson rule; both are converged to the FP64 rounding of the ordinates).
Like the reference, the ordinates are made here (host code) and the cubic spline itself is built by the
native library (``nbabfs_set_qcmm_spline``, CubicSpline_MakeFromReal1DArrays with zero end slopes).

Only the Gaussian models need a spline on this path: NBModelABFS's QC/MM interaction is
"Delta/Gaussian" and its QC/QC interaction "Gaussian/Gaussian" (pMolecule.NBModelABFS.pyx:72-97).
The MM/MM interaction uses the analytic form.
"""

import math

import numpy as np

from .LogFile import LogFileActive, logFile

UNITS_LENGTH_ANGSTROMS_TO_BOHRS = 1.0e-10 / 5.291772083e-11       # pCore-1.9.0/pCore/Units.py:43
UNITS_LENGTH_BOHRS_TO_ANGSTROMS = 5.291772083e-11 / 1.0e-10       # pCore-1.9.0/pCore/Units.py:44
_MODELS = ("Delta/Delta", "Delta/Gaussian", "Gaussian/Gaussian")


def _gaussian_potential(p):
    """erf(p r) / r with its small-argument series (pyx:32-39)."""
    rpi = math.sqrt(math.pi)

    def value(r):
        u = p * r
        u2 = u * u
        if u > 1.0e-2:
            v = math.erf(u) / u
        else:
            v = (2.0 - u2 * (2.0 / 3.0 - u2 * (0.2 - u2 * (1.0 / 21.0 - u2 * (1.0 / 108.0 - u2 / 660.0))))) / rpi
        return v * p
    return value


def _gaussian_derivative(p):
    """d/dr of erf(p r) / r with its small-argument series (pyx:41-48)."""
    rpi = math.sqrt(math.pi)

    def value(r):
        u = p * r
        u2 = u * u
        if u > 1.0e-2:
            v = 2.0 * math.exp(-u2) / (rpi * u) - math.erf(u) / u2
        else:
            v = -u * (4.0 / 3.0 - u2 * (0.8 - u2 * (2.0 / 7.0 - u2 * (2.0 / 27.0 - u2 / 66.0)))) / rpi
        return v * p * p
    return value


def _switch(r, r_on, r_off):
    return (r_off ** 2 - r ** 2) ** 2 * (r_off ** 2 + 2.0 * r ** 2 - 3.0 * r_on ** 2) / (r_off ** 2 - r_on ** 2) ** 3


# Gauss-Legendre rule of 10 points on [-1, 1] (Golub-Welsch nodes and weights, exact for polynomials of
# degree 19).  The integrand of the switched region is analytic, and each spline interval (1/50 A by
# default) is split into two panels, so the rule's error is far below the FP64 rounding of the sums.
_GL_NODES, _GL_WEIGHTS = np.polynomial.legendre.leggauss(10)
_GL_PANELS = 2


def _gauss_legendre(f, a, b):
    """Integral of f over [a, b] (a <= b) by the panelled Gauss-Legendre rule."""
    total = 0.0
    width = (b - a) / _GL_PANELS
    for k in range(_GL_PANELS):
        lo = a + k * width
        half = 0.5 * width
        mid = lo + half
        total += half * math.fsum(w * f(mid + half * t) for t, w in zip(_GL_NODES, _GL_WEIGHTS))
    return total


def spline_abscissae(rOff, pointDensity):
    """The spline's radii: ceil(density * rOff + 1) points, evenly spaced from 0, the last exactly rOff (the
    layout of the reference's spline arrays, pMolecule.PairwiseInteraction.pyx:57-69)."""
    n = int(math.ceil(float(pointDensity) * rOff + 1.0))
    step = rOff / float(n - 1)
    r = step * np.arange(n, dtype=np.float64)
    r[-1] = rOff
    return r


def Integrate(F, G, rDamp, rOn, rOff, pointDensity):
    """Spline ordinates (x = r^2, y) of a potential F with derivative G, force-switched between rOn and rOff
    (the construction of pMolecule.PairwiseInteraction.pyx:57-105):

    * rOn <= r <= rOff: y(r) = -int_r^rOff G(t) S(t) dt, S the force switch, so that y' = G S and y(rOff) = 0;
    * rDamp <= r < rOn: y(r) = F(r) + c, with c making y continuous at rOn;
    * r < rDamp: the parabola a r^2 + b whose value and slope match F + c at rDamp.

    The switched-region integrals are summed per interval from rOff inward (each interval by the panelled
    Gauss-Legendre rule, the running total in FP64)."""
    r = spline_abscissae(rOff, pointDensity)
    n = r.size
    y = np.zeros(n)

    def GS(t):
        return G(t) * _switch(t, rOn, rOff)

    # Interval k = [r_k, r_k+1] contributes its part above rOn.
    above = [k for k in range(n - 1) if r[k + 1] > rOn]
    tail = 0.0
    for k in reversed(above):
        tail -= _gauss_legendre(GS, max(r[k], rOn), r[k + 1])
        if r[k] >= rOn:
            y[k] = tail
    shift = tail - F(rOn)          # tail is now -int_rOn^rOff G S
    if rDamp > 0.0:
        fd, gd = F(rDamp) + shift, G(rDamp)
        a, b = 0.5 * gd / rDamp, fd - 0.5 * gd * rDamp
    else:
        a = b = 0.0
    for i in range(n):
        if r[i] >= rOn:
            break
        y[i] = a * r[i] * r[i] + b if r[i] < rDamp else F(r[i]) + shift
    return r * r, y


class PairwiseInteractionABFS:
    """An ABFS pairwise interaction (options and electrostatic spline ordinates)."""

    def __init__(self, **options):
        self.dampingCutoff = 0.5
        self.electrostaticModel = "Delta/Delta"
        self.innerCutoff = 8.0
        self.outerCutoff = 12.0
        self.splinePointDensity = 50
        self.useAnalyticForm = True
        self.width1 = 0.0
        self.width2 = 0.0
        self.splineX = None
        self.splineY = None
        self.SetOptions(**options)

    @classmethod
    def FromOptions(cls, **options):
        return cls(**options)

    def __getstate__(self):
        return {"dampingCutoff": self.dampingCutoff, "electrostaticModel": self.electrostaticModel,
                "innerCutoff": self.innerCutoff, "outerCutoff": self.outerCutoff,
                "splinePointDensity": self.splinePointDensity, "useAnalyticForm": self.useAnalyticForm,
                "width1": self.width1, "width2": self.width2}

    def __setstate__(self, state):
        self.__init__(**state)

    def __copy__(self):
        return self.__class__(**self.__getstate__())

    def __deepcopy__(self, memo):
        return self.__copy__()

    def CheckOptions(self):
        if self.electrostaticModel is None:
            self.electrostaticModel = "Delta/Delta"
        elif self.electrostaticModel not in _MODELS:
            raise ValueError("Invalid pairwise interaction electrostatic model: " + str(self.electrostaticModel) + ".")

    def SetOptions(self, **keywordArguments):
        for k in ("dampingCutoff", "innerCutoff", "outerCutoff", "width1", "width2"):
            if k in keywordArguments:
                setattr(self, k, float(keywordArguments.pop(k)))
        if "electrostaticModel" in keywordArguments:
            self.electrostaticModel = keywordArguments.pop("electrostaticModel")
        if "splinePointDensity" in keywordArguments:
            self.splinePointDensity = int(keywordArguments.pop("splinePointDensity"))
        if "useAnalyticForm" in keywordArguments:
            self.useAnalyticForm = bool(keywordArguments.pop("useAnalyticForm"))
        if len(keywordArguments) > 0:
            raise ValueError("Invalid options: " + ", ".join(sorted(keywordArguments.keys())) + ".")
        self.CheckOptions()
        self.splineX = self.splineY = None

    def MakeElectrostaticSplines(self, useAtomicUnits=False):
        """Ordinates of the Gaussian electrostatic spline (pyx:182-194): kept in splineX (= r^2) and
        splineY, in atomic units (1/bohr) or kJ/mol per e^2."""
        if self.electrostaticModel == "Delta/Gaussian":
            p = 2.0 / math.sqrt(math.pi) / self.width2
        elif self.electrostaticModel == "Gaussian/Gaussian":
            p = 2.0 / math.sqrt(math.pi) / math.sqrt(self.width1 ** 2 + self.width2 ** 2)
        else:
            raise NotImplementedError("the Delta/Delta spline is not used on this path (the analytic form is)")
        x, y = Integrate(_gaussian_potential(p), _gaussian_derivative(p), self.dampingCutoff, self.innerCutoff,
                         self.outerCutoff, self.splinePointDensity)
        if useAtomicUnits:
            y = y * (1.0 / UNITS_LENGTH_ANGSTROMS_TO_BOHRS)
        else:
            from .Parameters import UNITS_ENERGY_E2ANGSTROMS_TO_KILOJOULES_PER_MOLE
            y = y * UNITS_ENERGY_E2ANGSTROMS_TO_KILOJOULES_PER_MOLE
        self.splineX, self.splineY = x, y

    def MakeSplines(self, electrostatic=True, lennardJones=True, useAtomicUnits=False):
        """The splines this path needs: the Gaussian electrostatic ones (the LJ and Delta/Delta
        interactions use the analytic form)."""
        if electrostatic and self.electrostaticModel != "Delta/Delta":
            self.MakeElectrostaticSplines(useAtomicUnits=useAtomicUnits)

    def Summary(self, log=logFile):
        if LogFileActive(log):
            s = log.GetSummary()
            s.Start("Pairwise Interaction Summary")
            s.Entry("Damping Cutoff", "{:.3f}".format(self.dampingCutoff))
            s.Entry("Electrostatic Model", self.electrostaticModel)
            s.Entry("Inner Cutoff", "{:.3f}".format(self.innerCutoff))
            s.Entry("Outer Cutoff", "{:.3f}".format(self.outerCutoff))
            s.Entry("Spline Point Density", "{:d}".format(self.splinePointDensity))
            s.Entry("Use Analytic Form", "{!r}".format(self.useAnalyticForm))
            s.Entry("Width 1", "{:.3f}".format(self.width1))
            s.Entry("Width 2", "{:.3f}".format(self.width2))
            s.Stop()
