"""Many configurations of one system in one batched launch.

``SGOFProcessPoolGPU`` has the interface of pDynamo's system-geometry objective function pools
(pMoleculeScripts-1.9.0/pMoleculeScripts/SGOFProcessPool.py:20-52, ``FunctionGradients(x, g)`` over a
list of coordinate sets, as ChainOfStatesOptimizer.FunctionGradients calls it,
ChainOfStatesOptimizer.py:168-183), for the non-bonding energy this package computes: instead of
one process (or one serial call) per image, every image gets its own list state and all images
are evaluated by one persistent launch (``nbabfs_batch_energy``), replayed from a hipGraph while no
image's list needs rebuilding.  Each image keeps its list between calls and rebuilds it by the
half-skin rule, as its own NBModelABFSState would (MD and MC replicas reuse their lists the same way).

``BatchEnergy(system, coordinates)`` is the one-call form.
"""

import numpy as np

from . import _native
from .NBModelABFS import NBModelABFS


class SGOFProcessPoolGPU:
    """A pool of images of one system evaluated together on the GPU."""

    def __init__(self, system, nbModel=None, useGraph=True):
        em = system.energyModel
        if em is None or em.mmAtoms is None:
            raise ValueError("the system has no MM parameters")
        if system.symmetry is not None or (em.qcAtoms is not None and getattr(em.qcAtoms, "size", 0) > 0):
            raise ValueError("batched images take non-periodic MM systems")
        self.system = system
        self.nbModel = nbModel if nbModel is not None else (em.nbModel if em.nbModel is not None else NBModelABFS())
        self.useGraph = bool(useGraph)
        self.natoms = system.natoms
        self.states = []
        self.batch = None
        self.x = self.g = None
        self.rebuilt = None
        self.lastEnergies = np.zeros((0, 4))

    @classmethod
    def FromSystem(cls, system, **options):
        return cls(system, **options)

    def _state(self):
        em = self.system.energyModel
        lj = em.ljParameters
        lj14 = em.ljParameters14 if em.ljParameters14 is not None else lj
        fixed = None
        if self.system.fixedAtoms is not None:
            fixed = np.zeros(self.natoms, np.uint8)
            fixed[np.asarray(self.system.fixedAtoms, np.int64)] = 1
        st = _native.NativeState(self.nbModel.NativeOptions(), em.mmAtoms.charges, em.mmAtoms.ljtypes, lj.ntypes,
                                 lj.tableA, lj.tableB, lj14.tableA, lj14.tableB,
                                 None if em.exclusions is None else em.exclusions.pairs,
                                 None if em.interactions14 is None else em.interactions14.pairs, fixed)
        # The pool's gradient rows get this model's gradient only: written, not added (no zero fill).
        st.set_gradient_mode(_native.GRADIENT_ASSIGN)
        return st

    def _resize(self, m):
        if self.batch is not None and len(self.states) == m:
            return
        self.Finalize()
        self.states = [self._state() for _ in range(m)]
        self.batch = _native.NativeBatch(self.states, use_graph=self.useGraph)
        self.x = _native.DeviceArray((m * self.natoms, 3))
        self.g = _native.DeviceArray((m * self.natoms, 3))

    def FunctionGradients(self, x, g):
        """Energies of the images x[k] (each (natoms, 3)); their gradients are written into g[k] (g may be
        None).  Returns the list of energies (kJ/mol), in image order."""
        m = len(x)
        if m == 0:
            return []
        X = np.concatenate([np.asarray(xi, np.float64).reshape(self.natoms, 3) for xi in x])
        self._resize(m)
        self.x.upload(X)
        stride = 24 * self.natoms
        xp = [self.x.ptr + k * stride for k in range(m)]
        gp = [self.g.ptr + k * stride for k in range(m)] if g is not None else None
        e, self.rebuilt = self.batch.energy(xp, gp)
        self.lastEnergies = e
        if g is not None:
            G = self.g.numpy().reshape(m, self.natoms, 3)
            for k in range(m):
                g[k][...] = G[k]
        return [float(v) for v in e.sum(axis=1)]

    def Autotune(self, x):
        """Choose the work-item cut (entriesPerItem) of the batched pair launch by timing it on the images x[k]
        (the library's batch autotuner; an extension: the reference has no work items).  No gradient is written."""
        m = len(x)
        if m == 0:
            return None
        X = np.concatenate([np.asarray(xi, np.float64).reshape(self.natoms, 3) for xi in x])
        self._resize(m)
        self.x.upload(X)
        stride = 24 * self.natoms
        epi, ms = self.batch.autotune([self.x.ptr + k * stride for k in range(m)])
        return {"entriesPerItem": epi, "kernelMs": ms}

    def Energies(self):
        """Per-image energy terms of the last call as (MM/MM Elect., LJ, 1-4 Elect., 1-4 LJ) rows."""
        return self.lastEnergies

    def Finalize(self):
        if self.batch is not None:
            self.batch.close()
        for st in self.states:
            st.close()
        for d in (self.x, self.g):
            if d is not None:
                d.free()
        self.states, self.batch, self.x, self.g = [], None, None, None

    def __del__(self):
        try:
            self.Finalize()
        except Exception:
            pass


def BatchEnergy(system, coordinates, doGradients=True, pool=None):
    """Energies (and gradients) of several configurations of a system in one launch.  Returns
    (energies, gradients or None, pool); pass the pool back in to keep the images' lists."""
    pool = pool if pool is not None else SGOFProcessPoolGPU(system)
    g = [np.zeros((system.natoms, 3)) for _ in coordinates] if doGradients else None
    e = pool.FunctionGradients(coordinates, g)
    return e, g, pool
