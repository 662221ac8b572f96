"""pdynamo_amd -- pDynamo's NBModelABFS MM/MM energy + gradient and its cell-grid
pair list, rebuilt for AMD Instinct MI355X (gfx950).

Drop-in names: ``System``, ``NBModelABFS``, ``NBModelABFSState``,
``PairListGenerator``, ``LJParameterContainer`` (+ ``_FromEpsilonSigma`` /
``_FromTableCoefficients``), ``MMAtomContainer``, ``SelfPairList``,
``EnergyTerms``; crystal symmetry for the periodic image lists:
``System.DefineSymmetry`` with the ``CrystalClass*`` classes,
``Transformation3Container`` and ``Transformation3_FromSymmetryOperationString``.
Batches of images / replicas: ``SGOFProcessPoolGPU`` (the pool interface of pDynamo's
chain-of-states optimizer) and ``BatchEnergy``.  QC/MM systems: ``QCAtomContainer``
(``QCAtomContainer_FromAtomContainer``) passed to ``System.FromMMParameters`` or
``System.DefineQCModel``, ``PairwiseInteractionABFS``, ``QCMMInteractionState`` and the
``QCModelPointCharges`` stand-in for a QC program.
"""

from ._native import CLibraryError
from .BatchPool import BatchEnergy, SGOFProcessPoolGPU
from .LogFile import LogFileActive, TextLogFile, logFile
from .NBModelABFS import NBModelABFS, NBModelABFSState, PairListGenerator
from .PairwiseInteraction import PairwiseInteractionABFS
from .QCAtoms import QCAtomContainer, QCAtomContainer_FromAtomContainer, QCMMInteractionState, QCModelPointCharges
from .Parameters import (LJParameterContainer, LJParameterContainer_FromEpsilonSigma,
                         LJParameterContainer_FromTableCoefficients, MMAtomContainer, SelfPairList,
                         SelfPairList_FromIntegerPairs, UNITS_ENERGY_E2ANGSTROMS_TO_KILOJOULES_PER_MOLE)
from .Symmetry import (CrystalClassCubic, CrystalClassHexagonal, CrystalClassMonoclinic, CrystalClassOrthorhombic,
                       CrystalClassRhombohedral, CrystalClassTetragonal, CrystalClassTriclinic, Symmetry,
                       SymmetryParameterGradients, SymmetryParameters, Transformation3, Transformation3Container,
                       Transformation3Container_Identity, Transformation3_FromSymmetryOperationString)
from .System import Configuration, EnergyTerms, System

__all__ = ["CLibraryError", "LogFileActive", "TextLogFile", "logFile", "NBModelABFS", "NBModelABFSState",
           "PairListGenerator", "LJParameterContainer", "LJParameterContainer_FromEpsilonSigma",
           "LJParameterContainer_FromTableCoefficients", "MMAtomContainer", "SelfPairList",
           "SelfPairList_FromIntegerPairs", "UNITS_ENERGY_E2ANGSTROMS_TO_KILOJOULES_PER_MOLE", "Configuration",
           "EnergyTerms", "System", "CrystalClassCubic", "CrystalClassHexagonal", "CrystalClassMonoclinic",
           "CrystalClassOrthorhombic", "CrystalClassRhombohedral", "CrystalClassTetragonal", "CrystalClassTriclinic",
           "Symmetry", "SymmetryParameterGradients", "SymmetryParameters", "Transformation3",
           "Transformation3Container", "Transformation3Container_Identity",
           "Transformation3_FromSymmetryOperationString", "PairwiseInteractionABFS", "QCAtomContainer",
           "QCAtomContainer_FromAtomContainer", "QCMMInteractionState", "QCModelPointCharges", "BatchEnergy",
           "SGOFProcessPoolGPU"]
