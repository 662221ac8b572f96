"""``PairwiseInteractionABFS``: the QC/MM and QC/QC electrostatic interactions of the ABFS model.

Options, validation and defaults as pMolecule-1.9.0/extensions/pyrex/pMolecule.PairwiseInteraction.pyx:107-240;
the electrostatic spline ordinates as its ``MakeElectrostaticSplines`` / ``Integrate`` (pyx:24-105,182-194)
with the adaptive Simpson rule of pCore-1.9.0/pCore/NumericalIntegration.py:21-70.  Like the
reference, the ordinates are made here (host code) and the cubic spline itself is built by the
native library (``nbabfs_set_qcmm_spline``, CubicSpline_MakeFromReal1DArrays with zero end slopes).

Only the Gaussian models need a spline on this path: NBModelABFS's QC/MM interaction is
"Delta/Gaussian" and its QC/QC interaction "Gaussian/Gaussian" (pMolecule.NBModelABFS.pyx:72-97).
The MM/MM interaction uses the analytic form.
"""

import heapq
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


def AdaptiveSimpsonsRule(F, a, b, tolerance, maximumIterations):
    """Integral of F over [a, b]: intervals refined largest-error first from a heap until each
    meets its share of the tolerance or the iterations run out.  Returns (integral, state)."""
    width = math.fabs(a - b)
    per_length = tolerance / width
    total = 0.0
    heap = []
    count = 0
    mid = 0.5 * (a + b)
    fa, fm, fb = F(a), F(mid), F(b)
    heapq.heappush(heap, (0.0, width, a, mid, b, fa, fm, fb, (fa + 4.0 * fm + fb) * width / 6.0))
    while count < maximumIterations and heap:
        _, w, x0, x2, x4, f0, f2, f4, whole = heapq.heappop(heap)
        x1 = 0.5 * (x0 + x2)
        x3 = 0.5 * (x2 + x4)
        f1, f3 = F(x1), F(x3)
        left = (f0 + 4.0 * f1 + f2) * w / 12.0
        right = (f2 + 4.0 * f3 + f4) * w / 12.0
        excess = left + right - whole
        err = math.fabs(excess)
        if err < 15.0 * w * per_length:
            total += left + right + excess / 15.0
        else:
            w *= 0.5
            heapq.heappush(heap, (-err, w, x0, x1, x2, f0, f1, f2, left))
            heapq.heappush(heap, (-err, w, x2, x3, x4, f2, f3, f4, right))
        count += 1
    state = {"Iterations": count, "Is Converged": not heap}
    if heap:
        worst = 0.0
        for (e, w, _, _, _, _, _, _, whole) in heap:
            total += whole
            worst = max(worst, math.fabs(e) * width / (30.0 * w))
        state["Error"] = worst
    if a > b:
        total *= -1.0
    return total, state


def Integrate(F, G, rDamp, rOn, rOff, epsilon, maximumIterations, pointDensity):
    """Spline ordinates of F on [0, rOff] (pyx:57-105): in the switching region minus the integral of
    G S from r to rOff, below rOn F plus the offset that keeps it continuous, below rDamp a parabola.
    Returns (x = r^2, y)."""
    n = int(math.ceil(float(pointDensity) * rOff + 1.0))
    r = np.zeros(n)
    y = np.zeros(n)
    step = rOff / float(n - 1)
    for i in range(1, n - 1):
        r[i] = step * float(i)
    r[-1] = rOff

    def GS(t):
        return G(t) * _switch(t, rOn, rOff)

    acc = 0.0
    top = n - 2
    for i in range(n - 2, -1, -1):
        lo, hi = r[i], r[i + 1]
        if lo < rOn:
            top = i
            if hi > rOn:
                acc -= AdaptiveSimpsonsRule(GS, rOn, hi, epsilon, maximumIterations)[0]
            break
        acc -= AdaptiveSimpsonsRule(GS, lo, hi, epsilon, maximumIterations)[0]
        y[i] = acc
    shift = acc - F(rOn)
    if rDamp > 0:
        f_d, g_d = F(rDamp) + shift, G(rDamp)
        alpha, beta = 0.5 * g_d / rDamp, -0.5 * g_d * rDamp + f_d
    else:
        alpha = beta = 0.0
    for i in range(top + 1):
        t = r[i]
        y[i] = alpha * t ** 2 + beta if t < rDamp else F(t) + shift
    return r ** 2, y


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
                         self.outerCutoff, 1.0e-15, 1000, self.splinePointDensity)
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
