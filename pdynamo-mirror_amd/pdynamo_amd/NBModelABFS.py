"""The atom-based force-switching (ABFS) non-bonding model, MI355X edition.

Same classes, option names, validation and labels as pDynamo 1.9.0:

* ``PairListGenerator`` -- pCore-1.9.0/extensions/pyrex/pCore.PairListGenerator.pyx:15-125
  (defaults pCore-1.9.0/extensions/csource/PairListGenerator.c:360-378).
* ``NBModelABFS``       -- pMolecule-1.9.0/extensions/pyrex/pMolecule.NBModelABFS.pyx:15-302.
* ``NBModelABFSState``  -- pMolecule-1.9.0/extensions/pyrex/pMolecule.NBModelABFSState.pyx:13-127.

QC/MM systems (a ``QCAtomContainer`` passed to ``SetUp``): the MM/MM lists leave out the pure QC
atoms, the QC/MM LJ terms join the energy, and ``QCMMPotentials`` / ``QCMMGradients`` give the
QC model the electrostatic potentials at the QC atoms and add the gradients of its charges
(pMolecule.NBModelABFS.pyx:72-97,108-138; NBModelABFS.c:306-498).  QC/MM image lists are not built.

Energies and gradients are computed by libnbabfs.so on the GPU (cluster-pair
list + FP32 pair kernel with fixed-point sums by default, or the plain FP64
kernel with ``precision="fp64"``).  There is no CPU compute path.
"""

import copy

import numpy as np

from . import _native
from ._native import CLibraryError
from .LogFile import LogFileActive, logFile
from .PairwiseInteraction import PairwiseInteractionABFS, UNITS_LENGTH_BOHRS_TO_ANGSTROMS
from .QCAtoms import QCMMInteractionState

_STATENAMES = ("nbState", "qcmmstate")


class PairListGenerator:
    """Options of the cell-grid self pair list generator."""

    _DEFAULTS = {"sortIndices": True, "useGridByCell": False, "minimumCellExtent": 2, "minimumPoints": 500,
                 "cellSize": 0.0, "cutoff": 0.0, "cutoffCellSizeFactor": 0.5, "minimumCellSize": 1.0,
                 "minimumExtentFactor": 1.5}

    def __init__(self, **options):
        for k, v in self._DEFAULTS.items():
            setattr(self, k, v)
        self.SetOptions(**options)

    @classmethod
    def FromOptions(cls, **options):
        return cls(**options)

    def SetOptions(self, **keywordArguments):
        for k in ("sortIndices", "useGridByCell"):
            if k in keywordArguments:
                setattr(self, k, bool(keywordArguments.pop(k)))
        for k in ("minimumCellExtent", "minimumPoints"):
            if k in keywordArguments:
                setattr(self, k, int(keywordArguments.pop(k)))
        for k in ("cellSize", "cutoff", "cutoffCellSizeFactor", "minimumCellSize", "minimumExtentFactor"):
            if k in keywordArguments:
                setattr(self, k, float(keywordArguments.pop(k)))
        if len(keywordArguments) > 0:
            raise ValueError("Invalid options: " + ", ".join(sorted(keywordArguments.keys())) + ".")
        # The cell size always follows the cutoff (pCore.PairListGenerator.pyx:104).
        self.cellSize = self.cutoffCellSizeFactor * self.cutoff

    def __getstate__(self):
        return {k: getattr(self, k) for k in self._DEFAULTS}

    def __setstate__(self, state):
        self.__init__(**state)

    def __copy__(self):
        return self.__class__(**self.__getstate__())

    def __deepcopy__(self, memo):
        return self.__copy__()

    def Summary(self, log=logFile):
        if LogFileActive(log):
            s = log.GetSummary()
            s.Start("Pairlist Generator Summary")
            s.Entry("Cell Size", "{:.3f}".format(self.cellSize))
            s.Entry("Cutoff", "{:.3f}".format(self.cutoff))
            s.Entry("Cutoff Cell Size Factor", "{:.3f}".format(self.cutoffCellSizeFactor))
            s.Entry("Grid Cell/Cell Method", "{!r}".format(self.useGridByCell))
            s.Entry("Minimum Cell Extent", "{:d}".format(self.minimumCellExtent))
            s.Entry("Minimum Cell Size", "{:.3f}".format(self.minimumCellSize))
            s.Entry("Minimum Extent Factor", "{:.3f}".format(self.minimumExtentFactor))
            s.Entry("Minimum Points", "{:d}".format(self.minimumPoints))
            s.Entry("Sort Indices", "{!r}".format(self.sortIndices))
            s.Stop()


class NBModelABFSState:
    """Per-configuration cache: the device lists and the energies of the last call."""

    def __init__(self, native, mmAtoms):
        self.native = native
        self.natoms = mmAtoms.size
        self.emmel = self.emmlj = self.emmel14 = self.emmlj14 = 0.0
        self.eimmmel = self.eimmmlj = 0.0
        self.eqcmmlj = self.eqcmmlj14 = 0.0
        self.present = (False, False)
        self.qcPresent = (False, False)
        self.imagesPresent = False
        self.lastRebuilt = False

    def GetEnergies(self, energies):
        """Append the energy terms; a term is present only for a non-empty list."""
        if self.present[0]:
            energies.append(("MM/MM Elect.", self.emmel))
            energies.append(("MM/MM LJ", self.emmlj))
        if self.present[1]:
            energies.append(("MM/MM 1-4 Elect.", self.emmel14))
            energies.append(("MM/MM 1-4 LJ", self.emmlj14))
        if self.qcPresent[0]:
            energies.append(("QC/MM LJ", self.eqcmmlj))
        if self.qcPresent[1]:
            energies.append(("QC/MM 1-4 LJ", self.eqcmmlj14))
        if self.imagesPresent:
            energies.append(("MM/MM Image Elect.", self.eimmmel))
            energies.append(("MM/MM Image LJ", self.eimmmlj))

    def Images(self):
        """[(a, b, c, transformation, scale, pairs)] of the MM/MM image list."""
        return self.native.images()

    def QCMMPairs(self, which):
        """A QC/MM list ('el', 'lj', 'el14' or 'lj14') as sorted keys row * 2^32 + column."""
        return self.native.export_qcmm_pairs(which)

    def ImagePairs(self, image):
        """Pairs of one image as sorted keys i * n + j (i primary, j in the image)."""
        return self.native.export_image_pairs(image)

    def Statistics(self):
        return self.native.stats()

    def Summary(self, log=logFile):
        if LogFileActive(log):
            st = self.native.stats()
            s = log.GetSummary()
            s.Start("ABFS NB Model State Summary")
            if st["npairs"] > 0:
                s.Entry("MM/MM Pairs", "{:d}".format(st["npairs"]))
            if st["npairs14"] > 0:
                s.Entry("MM/MM 1-4 Pairs", "{:d}".format(st["npairs14"]))
            if st["qcmmElPairs"] > 0:
                s.Entry("QC/MM Elect. Pairs", "{:d}".format(st["qcmmElPairs"]))
            if st["qcmmLJPairs"] > 0:
                s.Entry("QC/MM LJ Pairs", "{:d}".format(st["qcmmLJPairs"]))
            if st["images"] > 0:
                s.Entry("MM/MM Image Images", "{:d}".format(st["images"]))
                s.Entry("MM/MM Image Pairs", "{:d}".format(st["imagePairs"]))
            s.Stop()

    def StatisticsSummary(self, log=logFile):
        if LogFileActive(log):
            st = self.native.stats()
            s = log.GetSummary()
            s.Start("ABFS NB Model State Statistics Summary")
            s.Entry("Number of Calls", "{:d}".format(st["numberOfCalls"]))
            s.Entry("Number of Updates", "{:d}".format(st["numberOfUpdates"]))
            n = st["numberOfUpdates"]
            if n > 0:
                if st["numberOfMMMMPairs"] != 0.0:
                    s.Entry("<MM/MM Pairs>", "{:.1f}".format(st["numberOfMMMMPairs"] / n))
                if st["npairs14"] > 0:
                    s.Entry("MM/MM 1-4 Pairs", "{:d}".format(st["npairs14"]))
                if st["numberOfQCMMElPairs"] != 0.0:
                    s.Entry("<QC/MM Elect. Pairs>", "{:.1f}".format(st["numberOfQCMMElPairs"] / n))
                if st["numberOfQCMMLJPairs"] != 0.0:
                    s.Entry("<QC/MM LJ Pairs>", "{:.1f}".format(st["numberOfQCMMLJPairs"] / n))
                if st["numberOfMMMMImageImages"] != 0.0:
                    s.Entry("<MM/MM Image Images>", "{:.1f}".format(st["numberOfMMMMImageImages"] / n))
                if st["numberOfMMMMImagePairs"] != 0.0:
                    s.Entry("<MM/MM Image Pairs>", "{:.1f}".format(st["numberOfMMMMImagePairs"] / n))
                s.Entry("Calls per Update", "{:.1f}".format(float(st["numberOfCalls"]) / n))
                s.Entry("Grid Updating", "{!r}".format(bool(st["useGridSearch"])))
            s.Stop()

    def Virial(self):
        """3x3 virial of the last call's MM/MM gradient (sum_i x_i (x) g_i; an extension)."""
        return self.native.virial()

    def Pairs(self, which=0):
        """The list as sorted canonical keys max(i,j)*n + min(i,j)."""
        return self.native.export_pairs(which)


class NBModelABFS:
    """Define an ABFS NB model."""

    _PRECISIONS = {"mixed": _native.MIXED, "fp64": _native.PLAIN_FP64, "tiled_fp64": _native.TILED_FP64}

    def __init__(self, **options):
        self.label = "ABFS"
        self.checkForInverses = True
        self.dampingCutoff = 0.5
        self.dielectric = 1.0
        self.electrostaticScale14 = 1.0
        self.imageExpandFactor = 0
        self.innerCutoff = 8.0
        self.listCutoff = 13.5
        self.outerCutoff = 12.0
        self.qcmmCoupling = "RC"
        self.useCentering = False
        self.generator = None
        self.mmmmPairwiseInteraction = None
        self.qcmmPairwiseInteraction = None
        self.qcqcPairwiseInteraction = None
        # Extensions of this implementation (tuning; reference defaults retained above).
        self.precision = "mixed"
        self.entriesPerItem = 16
        self.wavesPerCU = 0
        self.SetOptions(**options)

    @classmethod
    def FromOptions(cls, **options):
        return cls(**options)

    def SetOptions(self, **keywordArguments):
        """Set options for the model (pMolecule.NBModelABFS.pyx:140-179)."""
        for k in ("dampingCutoff", "dielectric", "electrostaticScale14", "innerCutoff", "listCutoff", "outerCutoff"):
            if k in keywordArguments:
                setattr(self, k, float(keywordArguments.pop(k)))
        for k in ("generator", "mmmmPairwiseInteraction", "qcmmPairwiseInteraction", "qcqcPairwiseInteraction"):
            if k in keywordArguments:
                setattr(self, k, keywordArguments.pop(k))
        if "imageExpandFactor" in keywordArguments:
            self.imageExpandFactor = int(keywordArguments.pop("imageExpandFactor"))
        for k in ("checkForInverses", "useCentering"):
            if k in keywordArguments:
                setattr(self, k, bool(keywordArguments.pop(k)))
        if "qcmmCoupling" in keywordArguments:
            coupling = keywordArguments.pop("qcmmCoupling")
            if coupling not in ("MM", "RC", "RD"):
                raise TypeError("Unrecognized QC/MM coupling option: " + str(coupling) + ".")
            self.qcmmCoupling = coupling
        if "precision" in keywordArguments:
            p = keywordArguments.pop("precision")
            if p not in self._PRECISIONS:
                raise ValueError("Unrecognized precision: {!r} (use 'mixed', 'fp64' or 'tiled_fp64').".format(p))
            self.precision = p
        if "entriesPerItem" in keywordArguments:
            v = int(keywordArguments.pop("entriesPerItem"))
            if v <= 0 or v > 128:
                raise ValueError("entriesPerItem must be in [1, 128].")
            self.entriesPerItem = v
        if "wavesPerCU" in keywordArguments:
            v = int(keywordArguments.pop("wavesPerCU"))
            if v < 0:
                raise ValueError("wavesPerCU must be >= 0 (0 = one workgroup per work item).")
            self.wavesPerCU = v
        if len(keywordArguments) > 0:
            raise ValueError("Invalid options: " + ", ".join(sorted(keywordArguments.keys())) + ".")
        if (self.dampingCutoff < 0.0) or (self.innerCutoff < self.dampingCutoff) or \
           (self.outerCutoff < self.innerCutoff) or (self.listCutoff < self.outerCutoff):
            raise TypeError("Invalid cutoff values: damping - {:.3f}; inner - {:.3f}; outer - {:.3f}; list - {:.3f}.".format(
                self.dampingCutoff, self.innerCutoff, self.outerCutoff, self.listCutoff))
        self.CheckGenerator()

    def CheckPairwiseInteractions(self):
        """The QC/MM ("Delta/Gaussian", widths 2 bohr) and QC/QC ("Gaussian/Gaussian") interactions of
        pMolecule.NBModelABFS.pyx:72-97 with their splines in atomic units."""
        width = 2.0 * UNITS_LENGTH_BOHRS_TO_ANGSTROMS
        if self.qcmmPairwiseInteraction is None:
            self.qcmmPairwiseInteraction = PairwiseInteractionABFS.FromOptions(
                dampingCutoff=0.0, electrostaticModel="Delta/Gaussian", innerCutoff=self.innerCutoff,
                outerCutoff=self.outerCutoff, width2=width)
        if self.qcmmPairwiseInteraction.splineX is None:
            self.qcmmPairwiseInteraction.MakeSplines(lennardJones=False, useAtomicUnits=True)
        if self.qcqcPairwiseInteraction is None:
            self.qcqcPairwiseInteraction = PairwiseInteractionABFS.FromOptions(
                dampingCutoff=0.0, electrostaticModel="Gaussian/Gaussian", innerCutoff=self.innerCutoff,
                outerCutoff=self.outerCutoff, width1=width, width2=width)

    def CheckGenerator(self):
        """The ABFS default generator (pMolecule.NBModelABFS.pyx:59-71)."""
        if self.generator is None:
            self.generator = PairListGenerator.FromOptions(cutoff=self.listCutoff, cutoffCellSizeFactor=0.5,
                                                           minimumCellExtent=2, minimumCellSize=3.0,
                                                           minimumExtentFactor=1.5, minimumPoints=500,
                                                           sortIndices=False, useGridByCell=True)
        else:
            self.generator.SetOptions(cutoff=self.listCutoff)

    def __getstate__(self):
        state = {k: getattr(self, k) for k in ("checkForInverses", "imageExpandFactor", "dampingCutoff", "dielectric",
                                                "electrostaticScale14", "innerCutoff", "listCutoff", "outerCutoff",
                                                "qcmmCoupling", "useCentering")}
        if self.generator is not None:
            state["generator"] = self.generator
        if self.precision != "mixed":
            state["precision"] = self.precision
        return state

    def __setstate__(self, state):
        self.__init__(**state)

    def __copy__(self):
        return self.__class__(**copy.deepcopy(self.__getstate__()))

    def __deepcopy__(self, memo):
        return self.__copy__()

    def NativeOptions(self):
        o = _native.default_options()
        o.dampingCutoff = self.dampingCutoff
        o.innerCutoff = self.innerCutoff
        o.outerCutoff = self.outerCutoff
        o.listCutoff = self.listCutoff
        o.dielectric = self.dielectric
        o.electrostaticScale14 = self.electrostaticScale14
        o.cutoffCellSizeFactor = self.generator.cutoffCellSizeFactor
        o.minimumCellExtent = self.generator.minimumCellExtent
        o.minimumPoints = self.generator.minimumPoints
        o.useGridByCell = 1 if self.generator.useGridByCell else 0
        o.precision = self._PRECISIONS[self.precision]
        o.entriesPerItem = self.entriesPerItem
        o.wavesPerCU = self.wavesPerCU
        return o

    def Clear(self, configuration):
        """Clear up temporary data (pMolecule.NBModel.pyx:74-78)."""
        if configuration is not None:
            for name in _STATENAMES:
                if hasattr(configuration, name):
                    state = getattr(configuration, name)
                    if hasattr(state, "native"):
                        state.native.close()
                    delattr(configuration, name)

    def SetUp(self, mmAtoms, qcAtoms, ljParameters, ljParameters14, fixedAtoms, interactions14, exclusions,
              symmetry, isolates, configuration, log=logFile):
        """Create the state on first use, then run the list update (pyx:181-273)."""
        if configuration is None:
            return
        hasQC = qcAtoms is not None and getattr(qcAtoms, "size", 0) > 0
        if hasQC and symmetry is not None:
            raise NotImplementedError("QC/MM image lists are not part of this implementation")
        if self.useCentering and symmetry is not None:
            raise NotImplementedError("useCentering is not part of this implementation")
        transformations = None if symmetry is None else getattr(symmetry, "transformations", None)
        if not hasattr(configuration, "nbState"):
            if ljParameters is None:
                raise CLibraryError("Unable to create NB state: LJ parameters are required.")
            fixed = None
            if fixedAtoms is not None:
                fixed = np.zeros(mmAtoms.size, np.uint8)
                fixed[np.asarray(list(fixedAtoms), dtype=np.int64)] = 1
            lj14 = ljParameters14 if ljParameters14 is not None else ljParameters
            native = _native.NativeState(self.NativeOptions(), mmAtoms.charges, mmAtoms.ljtypes, ljParameters.ntypes,
                                         ljParameters.tableA, ljParameters.tableB, lj14.tableA, lj14.tableB,
                                         None if exclusions is None else exclusions.pairs,
                                         None if interactions14 is None else interactions14.pairs, fixed)
            if transformations is not None:
                rot, tra = transformations.Arrays()
                native.set_transformations(rot, tra, self.checkForInverses, self.imageExpandFactor)
            if hasQC:
                index, boundary, partner, lf, off, mmp = qcAtoms.Arrays()
                coupling = {"MM": _native.QCMM_MM, "RC": _native.QCMM_RC, "RD": _native.QCMM_RD}[self.qcmmCoupling]
                native.set_qc_atoms(index, boundary, partner, lf, qcAtoms.useLinkRatio, off, mmp, coupling)
                self.CheckPairwiseInteractions()
                qi = self.qcmmPairwiseInteraction
                native.set_qcmm_spline(qi.splineX, qi.splineY)
                configuration.qcmmstate = QCMMInteractionState.WithExtent(qcAtoms.size)
            configuration.nbState = NBModelABFSState(native, mmAtoms)
        nbState = configuration.nbState
        xyz = configuration.coordinates3
        grad = getattr(configuration, "gradients3", None)
        if transformations is not None:
            sp = getattr(configuration, "symmetryParameters", None)
            if sp is None:
                raise CLibraryError("Symmetry without symmetry parameters.")
            nbState.native.set_symmetry_parameters(sp.AsArray())
        e, present, rebuilt = nbState.native.energy_host(xyz, grad)
        nbState.emmel, nbState.emmlj, nbState.emmel14, nbState.emmlj14 = (float(x) for x in e)
        nbState.present = (bool(present[0]), bool(present[1]))
        if hasattr(configuration, "qcmmstate"):
            # NBModelABFSState_Initialize zeroes the QC charges and potentials of each call.
            configuration.qcmmstate.Initialize()
            qe, qp = nbState.native.qcmm_terms()
            nbState.eqcmmlj, nbState.eqcmmlj14 = float(qe[0]), float(qe[1])
            nbState.qcPresent = (bool(qp[0]), bool(qp[1]))
        if transformations is not None:
            spg = getattr(configuration, "symmetryParameterGradients", None)
            ie, nbState.imagesPresent = nbState.native.image_terms(None if spg is None else spg.dEdM)
            nbState.eimmmel, nbState.eimmmlj = float(ie[0]), float(ie[1])
        nbState.lastRebuilt = rebuilt
        if rebuilt:
            nbState.Summary(log=log)

    def Energy(self, configuration):
        """Energy terms of the last set-up call (pyx:108-122)."""
        energies = []
        if hasattr(configuration, "nbState"):
            configuration.nbState.GetEnergies(energies)
        return energies

    def Autotune(self, system, log=logFile):
        """Choose entriesPerItem and wavesPerCU by timing the pair kernel on the system's list.

        An extension of this implementation (the reference has no work items): the
        system's state keeps the fastest cut, and the model's options are set to it
        for states created later.  The pair list itself is unchanged."""
        configuration = system.configuration
        if not hasattr(configuration, "nbState"):
            system.Energy(log=None, doGradients=False)
        native = configuration.nbState.native
        x = _native.DeviceArray.from_numpy(np.ascontiguousarray(configuration.coordinates3, dtype=np.float64), native)
        try:
            epi, wpc, ms = native.autotune(x.ptr)
        finally:
            _native.synchronize()
            x.free()
        self.entriesPerItem, self.wavesPerCU = epi, wpc
        if LogFileActive(log):
            s = log.GetSummary()
            s.Start("ABFS Work-Item Autotuner")
            s.Entry("Entries per Item", "{:d}".format(epi))
            s.Entry("Waves per CU", "{:d}".format(wpc))
            s.Entry("Pair Kernel (ms)", "{:.4f}".format(ms))
            s.Stop()
        return {"entriesPerItem": epi, "wavesPerCU": wpc, "kernelMs": ms}

    def QCMMGradients(self, configuration):
        """Gradients of the QC/MM electrostatics for the QC model's charges, added into
        configuration.gradients3 (pyx:124-130)."""
        if hasattr(configuration, "nbState") and hasattr(configuration, "qcmmstate"):
            grad = getattr(configuration, "gradients3", None)
            if grad is not None:
                configuration.nbState.native.qcmm_gradients(configuration.qcmmstate.qcCharges, grad)

    def QCMMPotentials(self, configuration):
        """Electrostatic potentials of the MM charges at the QC atoms (atomic units), added into
        configuration.qcmmstate.qcmmPotentials (pyx:132-138)."""
        if hasattr(configuration, "nbState") and hasattr(configuration, "qcmmstate"):
            configuration.nbState.native.qcmm_potentials(configuration.qcmmstate.qcmmPotentials)

    def Summary(self, log=logFile):
        if LogFileActive(log):
            s = log.GetSummary()
            s.Start("ABFS NB Model Summary")
            s.Entry("Dielectric", "{:.6f}".format(self.dielectric))
            s.Entry("El. 1-4 Scaling", "{:.6f}".format(self.electrostaticScale14))
            s.Entry("Damping Cutoff", "{:.6f}".format(self.dampingCutoff))
            s.Entry("Inner Cutoff", "{:.6f}".format(self.innerCutoff))
            s.Entry("List Cutoff", "{:.6f}".format(self.listCutoff))
            s.Entry("Outer Cutoff", "{:.6f}".format(self.outerCutoff))
            s.Entry("QC/MM Coupling", self.qcmmCoupling)
            s.Entry("Use Centering", "{!r}".format(self.useCentering))
            s.Entry("Precision", self.precision)
            s.Stop()
            if self.generator is not None:
                self.generator.Summary(log=log)
