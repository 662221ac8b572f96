"""Crystal symmetry for the NBModelABFS image path.

Mirrors the pieces of pDynamo that a periodic NBModelABFS calculation
consumes:

* ``Transformation3``, ``Transformation3Container`` (pCore-1.9.0/extensions/
  pyrex/pCore.Transformation3.pyx, pCore.Transformation3Container.pyx):
  fractional rotations and translations, the identity and the inverse pairs
  (Transformation3Container.c:52-141);
* ``Transformation3_FromSymmetryOperationString`` (pCore.Transformation3.pyx:
  127-163), with a small parser in place of ``eval``;
* ``SymmetryParameters`` (pMolecule-1.9.0/extensions/csource/
  SymmetryParameters.c): a, b, c, alpha, beta, gamma and the matrix M whose
  columns are the lattice vectors;
* the crystal classes of pMolecule-1.9.0/pMolecule/CrystalClass.py;
* ``Symmetry`` (pMolecule-1.9.0/pMolecule/Symmetry.py) and
  ``SymmetryParameterGradients`` (SymmetryParameterGradients.c:50-82).

M and M^-1 are formed with the reference's operation order so that the device
library, which forms them again in C++, sees the same matrices.
"""

import math
import re

import numpy as np

_DEGREES_TO_RADIANS = math.pi / 180.0e+00


# ------------------------------------------------------------------------------------------------
# Transformations.
# ------------------------------------------------------------------------------------------------
class Transformation3:
    """x -> R x + t in fractional coordinates."""

    def __init__(self, rotation=None, translation=None):
        self.rotation = np.eye(3) if rotation is None else np.array(rotation, dtype=np.float64).reshape(3, 3)
        self.translation = np.zeros(3) if translation is None else np.array(translation, dtype=np.float64).reshape(3)

    @classmethod
    def Identity(cls):
        return cls()

    @classmethod
    def Null(cls):
        return cls(np.zeros((3, 3)), np.zeros(3))

    def IsIdentity(self):
        """Transformation3_IsIdentity (pCore-1.9.0/extensions/csource/Transformation3.c:79-86), tolerance 1e-6."""
        return bool(np.all(np.abs(self.rotation - np.eye(3)) <= 1.0e-6) and np.all(np.abs(self.translation) <= 1.0e-6))


_TOKEN = re.compile(r"\s*([+-]?)\s*(\d+(?:\.\d*)?(?:/\d+(?:\.\d*)?)?)?\s*\*?\s*([xyz])?")


def _parse_component(text):
    """Linear form c_x x + c_y y + c_z z + t of one component such as "-x+y+1/3"."""
    coef = {"x": 0.0, "y": 0.0, "z": 0.0}
    const = 0.0
    s = text.replace(" ", "").lower()
    pos = 0
    while pos < len(s):
        m = _TOKEN.match(s, pos)
        if m is None or m.end() == pos:
            raise ValueError("Invalid symmetry operation component: " + text)
        sign, num, var = m.groups()
        if num is None and var is None:
            raise ValueError("Invalid symmetry operation component: " + text)
        value = 1.0
        if num is not None:
            if "/" in num:
                p, q = num.split("/")
                value = float(p) / float(q)
            else:
                value = float(num)
        if sign == "-":
            value = -value
        if var is None:
            const += value
        else:
            coef[var] += value
        pos = m.end()
    return coef, const


def Transformation3_FromSymmetryOperationString(ostring):
    """Transformation from a crystallographic operation such as "(-x+1/2,y+1/2,-z)"."""
    if not isinstance(ostring, str):
        return None
    t = ostring.replace(" ", "")
    if t.startswith("("):
        t = t[1:]
    if t.endswith(")"):
        t = t[:-1]
    parts = t.split(",")
    if len(parts) != 3:
        raise ValueError("Invalid symmetry operation: " + ostring)
    self = Transformation3.Null()
    for i, p in enumerate(parts):
        coef, const = _parse_component(p)
        self.rotation[i] = [coef["x"], coef["y"], coef["z"]]
        self.translation[i] = const
    return self


class Transformation3Container:
    """Transformations plus their identity and rotational inverses
    (Transformation3Container_FindIdentity / _FindInverses)."""

    def __init__(self, transformations):
        self.items = list(transformations)
        self.identity = -1
        for i, t in enumerate(self.items):
            if t.IsIdentity():
                self.identity = i
                break
        n = len(self.items)
        self.inverses = [-1] * n
        for i in range(n):
            if self.inverses[i] >= 0:
                continue
            m = _invert(self.items[i].rotation)
            for j in range(i + 1):
                if self.inverses[j] < 0 and np.all(np.abs(m - self.items[j].rotation) <= 1.0e-6):
                    self.inverses[i] = j
                    self.inverses[j] = i
                    break

    @classmethod
    def WithTransformations(cls, transformations):
        return cls(transformations)

    def __len__(self):
        return len(self.items)

    def __getitem__(self, i):
        return self.items[i]

    def IsIdentity(self):
        return len(self.items) == 1 and self.items[0].IsIdentity()

    def Arrays(self):
        """(rotations (k, 9), translations (k, 3)) as contiguous float64."""
        rot = np.ascontiguousarray(np.array([t.rotation.reshape(9) for t in self.items], np.float64))
        tra = np.ascontiguousarray(np.array([t.translation for t in self.items], np.float64))
        return rot, tra


def Transformation3Container_Identity():
    return Transformation3Container([Transformation3.Identity()])


def _det(a):
    """Matrix33_Determinant (pCore-1.9.0/extensions/csource/Matrix33.c:96-111)."""
    return (a[0, 0] * (a[1, 1] * a[2, 2] - a[2, 1] * a[1, 2]) - a[0, 1] * (a[1, 0] * a[2, 2] - a[2, 0] * a[1, 2]) +
            a[0, 2] * (a[1, 0] * a[2, 1] - a[2, 0] * a[1, 1]))


def _invert(a):
    """Matrix33_Invert (Matrix33.c:165-195): cofactors scaled by 1/det, in the reference's order."""
    a = [[float(v) for v in row] for row in a]
    r = [[a[1][1] * a[2][2] - a[1][2] * a[2][1], a[0][2] * a[2][1] - a[2][2] * a[0][1],
          a[0][1] * a[1][2] - a[1][1] * a[0][2]],
         [a[1][2] * a[2][0] - a[1][0] * a[2][2], a[0][0] * a[2][2] - a[0][2] * a[2][0],
          a[0][2] * a[1][0] - a[0][0] * a[1][2]],
         [a[1][0] * a[2][1] - a[1][1] * a[2][0], a[0][1] * a[2][0] - a[0][0] * a[2][1],
          a[0][0] * a[1][1] - a[1][0] * a[0][1]]]
    det = (a[0][0] * (a[1][1] * a[2][2] - a[2][1] * a[1][2]) - a[0][1] * (a[1][0] * a[2][2] - a[2][0] * a[1][2]) +
           a[0][2] * (a[1][0] * a[2][1] - a[2][0] * a[1][1]))
    s = 1.0 / det
    return np.array([[v * s for v in row] for row in r])


# ------------------------------------------------------------------------------------------------
# Symmetry parameters.
# ------------------------------------------------------------------------------------------------
class SymmetryParameters:
    """Unit cell a, b, c (A), alpha, beta, gamma (degrees), M and M^-1."""

    def __init__(self, a=0.0, b=0.0, c=0.0, alpha=0.0, beta=0.0, gamma=0.0):
        self.SetCrystalParameters(a, b, c, alpha, beta, gamma)

    def SetCrystalParameters(self, a, b, c, alpha, beta, gamma):
        """SymmetryParameters_SetCrystalParameters / _MakeM (SymmetryParameters.c:392-425,452-470)."""
        self.a, self.b, self.c = float(a), float(b), float(c)
        self.alpha, self.beta, self.gamma = float(alpha), float(beta), float(gamma)
        al, be, ga = self.alpha * _DEGREES_TO_RADIANS, self.beta * _DEGREES_TO_RADIANS, self.gamma * _DEGREES_TO_RADIANS
        ca, cb, cg, sg = math.cos(al), math.cos(be), math.cos(ga), math.sin(ga)
        M = np.zeros((3, 3))
        if sg != 0.0:
            M[0, 0] = self.a
            M[0, 1] = self.b * cg
            M[1, 1] = self.b * sg
            M[0, 2] = self.c * cb
            M[1, 2] = self.c * (ca - cb * cg) / sg
            M[2, 2] = self.c * math.sqrt(1.0e+00 - ca * ca - cb * cb - cg * cg + 2.0e+00 * ca * cb * cg) / sg
        self.M = M
        self.inverseM = _invert(M) if _det(M) != 0.0 else np.zeros((3, 3))

    def AsArray(self):
        return np.array([self.a, self.b, self.c, self.alpha, self.beta, self.gamma], np.float64)

    def Displacement(self, a, b, c):
        """SymmetryParameters_Displacement (SymmetryParameters.c:166-178)."""
        return np.array([a * self.M[i, 0] + b * self.M[i, 1] + c * self.M[i, 2] for i in range(3)])

    def Volume(self):
        return self.M[0, 0] * self.M[1, 1] * self.M[2, 2]

    def IsMinimumImageConventionSatisfied(self, length):
        d = 2.0 * length
        return (d < self.a) and (d < self.b) and (d < self.c)

    def __getstate__(self):
        return {"a": self.a, "b": self.b, "c": self.c, "alpha": self.alpha, "beta": self.beta, "gamma": self.gamma}

    def __setstate__(self, state):
        self.SetCrystalParameters(state["a"], state["b"], state["c"], state["alpha"], state["beta"], state["gamma"])


class SymmetryParameterGradients:
    """dE/dM (row-major 3x3, incremented by the energy models) and its crystal form
    (SymmetryParameterGradients.c:50-82)."""

    def __init__(self):
        self.dEdM = np.zeros((3, 3))
        self.dEdA = self.dEdB = self.dEdC = 0.0
        self.dEdAlpha = self.dEdBeta = self.dEdGamma = 0.0

    def CrystalDerivatives(self, sp):
        d = self.dEdM
        al, be, ga = sp.alpha * _DEGREES_TO_RADIANS, sp.beta * _DEGREES_TO_RADIANS, sp.gamma * _DEGREES_TO_RADIANS
        ca, cb, cg, sg = math.cos(al), math.cos(be), math.cos(ga), math.sin(ga)
        f12 = ca - cb * cg
        f22 = math.sqrt(1.0e+00 - ca * ca - cb * cb - cg * cg + 2.0e+00 * ca * cb * cg)
        self.dEdA = d[0, 0]
        self.dEdB = cg * d[0, 1] + sg * d[1, 1]
        self.dEdC = cb * d[0, 2] + f12 * d[1, 2] / sg + f22 * d[2, 2] / sg
        self.dEdAlpha = -(d[1, 2] - f12 * d[2, 2] / f22) * (sp.c * math.sin(al) * _DEGREES_TO_RADIANS) / sg
        self.dEdBeta = -(sg * d[0, 2] - cg * d[1, 2] + (ca * cg - cb) * d[2, 2] / f22) * \
            (sp.c * math.sin(be) * _DEGREES_TO_RADIANS) / sg
        self.dEdGamma = (sp.b * (-sg * d[0, 1] + cg * d[1, 1]) +
                         sp.c * ((cb - f12 * cg / (sg * sg)) * d[1, 2] +
                                 ((cg - ca * cb) / f22 - f22 * cg / (sg * sg)) * d[2, 2])) * _DEGREES_TO_RADIANS
        return (self.dEdA, self.dEdB, self.dEdC, self.dEdAlpha, self.dEdBeta, self.dEdGamma)


# ------------------------------------------------------------------------------------------------
# Crystal classes (pMolecule-1.9.0/pMolecule/CrystalClass.py): parameter completion and checks.
# ------------------------------------------------------------------------------------------------
class CrystalClass:
    _DefaultAlpha = _DefaultBeta = _DefaultGamma = 0.0

    def AssignCrystalParameters(self, a, b, c, alpha, beta, gamma):
        return (a, b, c, alpha, beta, gamma)

    def CreateSymmetryParameters(self, a=None, b=None, c=None, alpha=None, beta=None, gamma=None):
        if alpha is None:
            alpha = self._DefaultAlpha
        if beta is None:
            beta = self._DefaultBeta
        if gamma is None:
            gamma = self._DefaultGamma
        a, b, c, alpha, beta, gamma = self.AssignCrystalParameters(a, b, c, alpha, beta, gamma)
        if (a <= 0.0) or (b <= 0.0) or (c <= 0.0):
            raise ValueError("Unit cell lengths zero or negative.")
        for v in (alpha, beta, gamma):
            if v <= 0.0 or v >= 180.0:
                raise ValueError("Unit cell angles not in range [0,180].")
        return SymmetryParameters(a, b, c, alpha, beta, gamma)

    def Label(self):
        return None


class CrystalClassCubic(CrystalClass):
    _DefaultAlpha = _DefaultBeta = _DefaultGamma = 90.0

    def AssignCrystalParameters(self, a, b, c, alpha, beta, gamma):
        b = a if b is None else b
        c = a if c is None else c
        if not ((a is not None) and a == b and a == c and alpha == 90.0 and beta == 90.0 and gamma == 90.0):
            raise ValueError("Invalid crystal parameters: a = b = c ; alpha = beta = gamma = 90.")
        return (a, b, c, alpha, beta, gamma)

    def Label(self):
        return "Cubic"


class CrystalClassTetragonal(CrystalClass):
    _DefaultAlpha = _DefaultBeta = _DefaultGamma = 90.0

    def AssignCrystalParameters(self, a, b, c, alpha, beta, gamma):
        b = a if b is None else b
        if not ((a is not None) and a == b and (c is not None) and alpha == 90.0 and beta == 90.0 and gamma == 90.0):
            raise ValueError("Invalid crystal parameters: a = b != c ; alpha = beta = gamma = 90.")
        return (a, b, c, alpha, beta, gamma)

    def Label(self):
        return "Tetragonal"


class CrystalClassOrthorhombic(CrystalClass):
    _DefaultAlpha = _DefaultBeta = _DefaultGamma = 90.0

    def AssignCrystalParameters(self, a, b, c, alpha, beta, gamma):
        if not ((a is not None) and (b is not None) and (c is not None) and alpha == 90.0 and beta == 90.0 and
                gamma == 90.0):
            raise ValueError("Invalid crystal parameters: a != b != c ; alpha = beta = gamma = 90.")
        return (a, b, c, alpha, beta, gamma)

    def Label(self):
        return "Orthorhombic"


class CrystalClassHexagonal(CrystalClass):
    _DefaultAlpha = _DefaultBeta = 90.0
    _DefaultGamma = 120.0

    def AssignCrystalParameters(self, a, b, c, alpha, beta, gamma):
        b = a if b is None else b
        if not ((a is not None) and a == b and (c is not None) and alpha == 90.0 and beta == 90.0 and gamma == 120.0):
            raise ValueError("Invalid crystal parameters: a = b != c ; alpha = beta = 90 ; gamma = 120.")
        return (a, b, c, alpha, beta, gamma)

    def Label(self):
        return "Hexagonal"


class CrystalClassMonoclinic(CrystalClass):
    _DefaultAlpha = _DefaultGamma = 90.0
    _DefaultBeta = None

    def AssignCrystalParameters(self, a, b, c, alpha, beta, gamma):
        if not ((a is not None) and (b is not None) and (c is not None) and (beta is not None) and alpha == 90.0 and
                gamma == 90.0):
            raise ValueError("Invalid crystal parameters: a != b != c ; alpha = gamma = 90 != beta.")
        return (a, b, c, alpha, beta, gamma)

    def Label(self):
        return "Monoclinic"


class CrystalClassRhombohedral(CrystalClass):
    _DefaultAlpha = _DefaultBeta = _DefaultGamma = None

    def AssignCrystalParameters(self, a, b, c, alpha, beta, gamma):
        b = a if b is None else b
        c = a if c is None else c
        beta = alpha if beta is None else beta
        gamma = alpha if gamma is None else gamma
        if not ((a is not None) and a == b and a == c and (alpha is not None) and alpha == beta and alpha == gamma):
            raise ValueError("Invalid crystal parameters: a = b = c ; alpha = beta = gamma != 90.")
        return (a, b, c, alpha, beta, gamma)

    def Label(self):
        return "Rhombohedral"


class CrystalClassTriclinic(CrystalClass):
    _DefaultAlpha = _DefaultBeta = _DefaultGamma = None

    def AssignCrystalParameters(self, a, b, c, alpha, beta, gamma):
        if None in (a, b, c, alpha, beta, gamma):
            raise ValueError("Invalid crystal parameters: a != b != c ; alpha != beta != gamma.")
        return (a, b, c, alpha, beta, gamma)

    def Label(self):
        return "Triclinic"


class Symmetry:
    """Crystal class plus transformations (pMolecule-1.9.0/pMolecule/Symmetry.py)."""

    def __init__(self, crystalClass=None, transformations=None):
        self.crystalClass = crystalClass
        self.transformations = transformations
