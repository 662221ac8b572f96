"""ctypes binding of libnbabfs.so (the C ABI declared in include/nbabfs.h).

The library is the only compute path: there is no CPU fallback.  Loading it on
a machine without a gfx950 GPU succeeds (so the symbol table can be checked),
but every entry point that needs the device raises ``CLibraryError``.
"""

import ctypes as C
import os

import numpy as np

from .build import LIB

STATUS_CONTINUE = 16
GRADIENT_ADD, GRADIENT_ASSIGN = 0, 1           # nbabfs_set_gradient_mode
MIXED, PLAIN_FP64, TILED_FP64 = 0, 1, 2
QCMM_MM, QCMM_RC, QCMM_RD = 0, 1, 2

# Every symbol include/nbabfs.h declares.
EXPORTED = ["nbabfs_default_options", "nbabfs_init", "nbabfs_last_error", "nbabfs_device_cu_count",
            "nbabfs_state_create", "nbabfs_state_destroy", "nbabfs_set_stream", "nbabfs_energy_host",
            "nbabfs_update", "nbabfs_mmmm_energy_async", "nbabfs_energies", "nbabfs_stats_get",
            "nbabfs_set_timing", "nbabfs_export_pairs", "nbabfs_batch_create", "nbabfs_batch_destroy",
            "nbabfs_batch_energy", "nbabfs_device_alloc", "nbabfs_device_free", "nbabfs_memcpy_htod",
            "nbabfs_memcpy_dtoh", "nbabfs_zero_async", "nbabfs_synchronize", "nbabfs_autotune",
            "nbabfs_virial", "nbabfs_set_transformations", "nbabfs_set_symmetry_parameters", "nbabfs_image_terms",
            "nbabfs_images", "nbabfs_export_image_pairs", "nbabfs_set_qc_atoms", "nbabfs_set_qcmm_spline",
            "nbabfs_qcmm_terms", "nbabfs_qcmm_potentials", "nbabfs_qcmm_gradients_host", "nbabfs_export_qcmm_pairs",
            "nbabfs_qcmm_sites", "nbabfs_set_gradient_mode", "nbabfs_qcmm_potentials_async",
            "nbabfs_qcmm_gradients_async", "nbabfs_batch_info", "nbabfs_batch_set_timing", "nbabfs_batch_autotune"]


class CLibraryError(Exception):
    """Raised when the native library reports a failure (pCore.CLibraryError)."""


class Options(C.Structure):
    _fields_ = [("dampingCutoff", C.c_double), ("innerCutoff", C.c_double), ("outerCutoff", C.c_double),
                ("listCutoff", C.c_double), ("dielectric", C.c_double), ("electrostaticScale14", C.c_double),
                ("cutoffCellSizeFactor", C.c_double), ("minimumCellExtent", C.c_int), ("minimumPoints", C.c_int),
                ("precision", C.c_int), ("entriesPerItem", C.c_int), ("wavesPerCU", C.c_int),
                ("imageEntriesPerItem", C.c_int), ("useGridByCell", C.c_int)]


class Stats(C.Structure):
    _fields_ = [("numberOfCalls", C.c_longlong), ("numberOfUpdates", C.c_longlong),
                ("numberOfMMMMPairs", C.c_double), ("npairs", C.c_longlong), ("npairs14", C.c_longlong),
                ("useGridSearch", C.c_int), ("clusters", C.c_int), ("entries", C.c_longlong),
                ("workItems", C.c_longlong), ("forceKernelMs", C.c_double), ("forceKernelLaunches", C.c_longlong),
                ("images", C.c_int), ("reserved2", C.c_int), ("imagePairs", C.c_longlong),
                ("numberOfMMMMImageImages", C.c_double), ("numberOfMMMMImagePairs", C.c_double),
                ("qcmmElPairs", C.c_longlong), ("qcmmLJPairs", C.c_longlong), ("numberOfQCMMElPairs", C.c_double),
                ("numberOfQCMMLJPairs", C.c_double)]


_lib = None


def library_path():
    return LIB


def load():
    """Load libnbabfs.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        raise CLibraryError("libnbabfs.so is not built (run __graft_entry__.build() or python -m pdynamo_amd.build)")
    L = C.CDLL(LIB)
    vp = C.c_void_p
    L.nbabfs_default_options.argtypes = [C.POINTER(Options)]
    L.nbabfs_default_options.restype = None
    L.nbabfs_init.restype = C.c_int
    L.nbabfs_last_error.restype = C.c_char_p
    L.nbabfs_device_cu_count.restype = C.c_int
    L.nbabfs_state_create.argtypes = [C.POINTER(Options), C.c_int, vp, vp, C.c_int, vp, vp, vp, vp, C.c_int, vp,
                                      C.c_int, vp, vp, C.POINTER(vp)]
    L.nbabfs_state_destroy.argtypes = [vp]
    L.nbabfs_state_destroy.restype = None
    L.nbabfs_set_stream.argtypes = [vp, vp]
    L.nbabfs_energy_host.argtypes = [vp, vp, vp, vp, vp, vp]
    L.nbabfs_update.argtypes = [vp, vp, vp]
    L.nbabfs_mmmm_energy_async.argtypes = [vp, vp, vp]
    L.nbabfs_energies.argtypes = [vp, vp, vp]
    L.nbabfs_stats_get.argtypes = [vp, C.POINTER(Stats)]
    L.nbabfs_set_timing.argtypes = [vp, C.c_int]
    L.nbabfs_export_pairs.argtypes = [vp, C.c_int, vp, C.c_longlong]
    L.nbabfs_export_pairs.restype = C.c_longlong
    L.nbabfs_batch_create.argtypes = [vp, C.c_int, C.c_int, C.POINTER(vp)]
    L.nbabfs_batch_destroy.argtypes = [vp]
    L.nbabfs_batch_destroy.restype = None
    L.nbabfs_batch_energy.argtypes = [vp, vp, vp, vp, vp]
    L.nbabfs_batch_autotune.argtypes = [vp, vp, C.POINTER(C.c_int), C.POINTER(C.c_double)]
    L.nbabfs_device_alloc.argtypes = [C.c_size_t, C.POINTER(vp)]
    L.nbabfs_device_free.argtypes = [vp]
    L.nbabfs_memcpy_htod.argtypes = [vp, vp, vp, C.c_size_t]
    L.nbabfs_memcpy_dtoh.argtypes = [vp, vp, vp, C.c_size_t]
    L.nbabfs_zero_async.argtypes = [vp, vp, C.c_size_t]
    L.nbabfs_synchronize.argtypes = [vp]
    L.nbabfs_autotune.argtypes = [vp, vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_double)]
    L.nbabfs_virial.argtypes = [vp, vp]
    L.nbabfs_set_transformations.argtypes = [vp, C.c_int, vp, vp, C.c_int, C.c_int]
    L.nbabfs_set_symmetry_parameters.argtypes = [vp, vp]
    L.nbabfs_image_terms.argtypes = [vp, vp, vp, vp]
    L.nbabfs_images.argtypes = [vp, C.c_int, vp, vp, vp]
    L.nbabfs_export_image_pairs.argtypes = [vp, C.c_int, vp, C.c_longlong]
    L.nbabfs_export_image_pairs.restype = C.c_longlong
    L.nbabfs_set_qc_atoms.argtypes = [vp, C.c_int, vp, vp, vp, vp, C.c_int, vp, vp, C.c_int]
    L.nbabfs_set_qcmm_spline.argtypes = [vp, C.c_int, vp, vp]
    L.nbabfs_qcmm_terms.argtypes = [vp, vp, vp]
    L.nbabfs_qcmm_potentials.argtypes = [vp, vp]
    L.nbabfs_qcmm_gradients_host.argtypes = [vp, vp, vp]
    L.nbabfs_export_qcmm_pairs.argtypes = [vp, C.c_int, vp, C.c_longlong]
    L.nbabfs_export_qcmm_pairs.restype = C.c_longlong
    L.nbabfs_qcmm_sites.argtypes = [vp, vp, vp, vp, C.c_longlong]
    L.nbabfs_qcmm_sites.restype = C.c_longlong
    L.nbabfs_set_gradient_mode.argtypes = [vp, C.c_int]
    L.nbabfs_qcmm_potentials_async.argtypes = [vp, vp]
    L.nbabfs_qcmm_gradients_async.argtypes = [vp, vp, vp]
    L.nbabfs_batch_info.argtypes = [vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong), C.POINTER(C.c_longlong),
                                    C.POINTER(C.c_double), C.POINTER(C.c_longlong)]
    L.nbabfs_batch_set_timing.argtypes = [vp, C.c_int]
    _lib = L
    return L


def check(status):
    if status != STATUS_CONTINUE:
        msg = load().nbabfs_last_error()
        raise CLibraryError("libnbabfs status %d: %s" % (status, msg.decode() if msg else ""))


def init():
    check(load().nbabfs_init())


def default_options():
    o = Options()
    load().nbabfs_default_options(C.byref(o))
    return o


def _ptr(a):
    return None if a is None else a.ctypes.data


class DeviceArray:
    """A float64 device buffer owned through libnbabfs (one HIP runtime in the process)."""

    def __init__(self, shape):
        self.shape = tuple(int(x) for x in shape)
        self.nbytes = int(np.prod(self.shape)) * 8
        p = C.c_void_p()
        check(load().nbabfs_device_alloc(self.nbytes, C.byref(p)))
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, a, state=None):
        a = np.ascontiguousarray(a, dtype=np.float64)
        d = cls(a.shape)
        d.upload(a, state)
        return d

    def upload(self, a, state=None):
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.nbytes != self.nbytes:
            raise ValueError("size mismatch")
        check(load().nbabfs_memcpy_htod(None if state is None else state.h, self.ptr, a.ctypes.data, self.nbytes))

    def numpy(self, state=None):
        out = np.empty(self.shape)
        check(load().nbabfs_memcpy_dtoh(None if state is None else state.h, out.ctypes.data, self.ptr, self.nbytes))
        return out

    def zero(self, state=None):
        check(load().nbabfs_zero_async(None if state is None else state.h, self.ptr, self.nbytes))

    def free(self):
        if getattr(self, "ptr", None):
            load().nbabfs_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize(state=None):
    check(load().nbabfs_synchronize(None if state is None else state.h))


class NativeState:
    """Owns one nbabfs_state (the NBModelABFSState of one configuration)."""

    def __init__(self, options, charges, ljtypes, ntypes, tableA, tableB, tableA14=None, tableB14=None,
                 exclusions=None, interactions14=None, fixed=None):
        L = load()
        init()
        self.n = int(len(charges))
        q = np.ascontiguousarray(charges, dtype=np.float64)
        t = np.ascontiguousarray(ljtypes, dtype=np.int32)
        tA = np.ascontiguousarray(tableA, dtype=np.float64)
        tB = np.ascontiguousarray(tableB, dtype=np.float64)
        tA14 = None if tableA14 is None else np.ascontiguousarray(tableA14, dtype=np.float64)
        tB14 = None if tableB14 is None else np.ascontiguousarray(tableB14, dtype=np.float64)
        ex = np.zeros((0, 2), np.int32) if exclusions is None else np.ascontiguousarray(exclusions, np.int32).reshape(-1, 2)
        p14 = np.zeros((0, 2), np.int32) if interactions14 is None else np.ascontiguousarray(interactions14, np.int32).reshape(-1, 2)
        fx = None if fixed is None else np.ascontiguousarray(fixed, dtype=np.uint8)
        h = C.c_void_p()
        check(L.nbabfs_state_create(C.byref(options), self.n, _ptr(q), _ptr(t), int(ntypes), _ptr(tA), _ptr(tB),
                                    _ptr(tA14), _ptr(tB14), ex.shape[0], _ptr(ex) if ex.size else None,
                                    p14.shape[0], _ptr(p14) if p14.size else None, _ptr(fx), C.byref(h)))
        self.h = h.value
        self.options = options

    def energy_host(self, xyz, grad=None):
        """One Initialize + Update + MMMMEnergy on host arrays; grad is incremented (overwritten in
        assign mode)."""
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        if xyz.shape != (self.n, 3):
            raise ValueError("coordinates must be (%d, 3)" % self.n)
        if grad is not None and (grad.dtype != np.float64 or not grad.flags.c_contiguous or grad.shape != (self.n, 3)):
            raise ValueError("gradients must be a C-contiguous (%d, 3) float64 array" % self.n)
        e = np.zeros(4)
        present = np.zeros(2, np.int32)
        rebuilt = C.c_int()
        check(load().nbabfs_energy_host(self.h, _ptr(xyz), _ptr(grad), _ptr(e), _ptr(present), C.byref(rebuilt)))
        return e, present, bool(rebuilt.value)

    def update(self, d_xyz):
        rebuilt = C.c_int()
        check(load().nbabfs_update(self.h, C.c_void_p(d_xyz), C.byref(rebuilt)))
        return bool(rebuilt.value)

    def energy_async(self, d_xyz, d_grad=None):
        check(load().nbabfs_mmmm_energy_async(self.h, C.c_void_p(d_xyz), None if d_grad is None else C.c_void_p(d_grad)))

    def energies(self):
        e = np.zeros(4)
        present = np.zeros(2, np.int32)
        check(load().nbabfs_energies(self.h, _ptr(e), _ptr(present)))
        return e, present

    def set_transformations(self, rotations, translations, check_for_inverses=True, image_expand_factor=0):
        """Fractional transformations (k, 3, 3) / (k, 3) of the crystal symmetry."""
        rot = np.ascontiguousarray(rotations, dtype=np.float64).reshape(-1, 9)
        tra = np.ascontiguousarray(translations, dtype=np.float64).reshape(-1, 3)
        if rot.shape[0] != tra.shape[0]:
            raise ValueError("rotations and translations differ in number")
        self._sym = (rot, tra)
        check(load().nbabfs_set_transformations(self.h, rot.shape[0], _ptr(rot), _ptr(tra),
                                                1 if check_for_inverses else 0, int(image_expand_factor)))

    def set_symmetry_parameters(self, crystal):
        """Crystal parameters a, b, c, alpha, beta, gamma of the next calls."""
        c = np.ascontiguousarray(crystal, dtype=np.float64).reshape(6)
        check(load().nbabfs_set_symmetry_parameters(self.h, _ptr(c)))

    def image_terms(self, dEdM=None):
        """(energies (Image Elect., Image LJ), present) of the last energy call; dEdM (3x3) is incremented."""
        e = np.zeros(2)
        present = C.c_int()
        dm = None
        if dEdM is not None:
            dm = np.ascontiguousarray(dEdM, dtype=np.float64).reshape(9).copy()
        check(load().nbabfs_image_terms(self.h, _ptr(e), _ptr(dm), C.byref(present)))
        if dEdM is not None:
            dEdM[...] = dm.reshape(dEdM.shape)
        return e, bool(present.value)

    def images(self):
        """[(a, b, c, t, scale, npairs)] of the image list, in list order."""
        L = load()
        m = L.nbabfs_images(self.h, 0, None, None, None)
        if m < 0:
            check(int(-m))
        abct = np.zeros(4 * max(m, 1), np.int32)
        sc = np.zeros(max(m, 1))
        npr = np.zeros(max(m, 1), np.int64)
        L.nbabfs_images(self.h, m, _ptr(abct), _ptr(sc), _ptr(npr))
        return [(int(abct[4 * k]), int(abct[4 * k + 1]), int(abct[4 * k + 2]), int(abct[4 * k + 3]), float(sc[k]),
                 int(npr[k])) for k in range(m)]

    def export_image_pairs(self, image):
        """Sorted keys i * n + j of image `image` (i primary, j image)."""
        L = load()
        m = L.nbabfs_export_image_pairs(self.h, int(image), None, 0)
        if m < 0:
            check(int(-m))
        keys = np.zeros(max(m, 1), np.int64)
        m2 = L.nbabfs_export_image_pairs(self.h, int(image), _ptr(keys), m)
        if m2 < 0:
            check(int(-m2))
        return keys[:m]

    # QC/MM ------------------------------------------------------------------------------------------
    def set_qc_atoms(self, index, boundary=None, qcpartner=None, linkfactor=None, link_ratio=False,
                     mmpartner_offsets=None, mmpartners=None, coupling=QCMM_RC):
        """QC atoms in ascending atom order; boundary (link) atoms with their QC partner, link factor and MM
        partners (CSR offsets per QC atom); coupling QCMM_MM, QCMM_RC or QCMM_RD."""
        idx = np.ascontiguousarray(index, np.int32)
        nqc = idx.size
        b = np.zeros(nqc, np.int32) if boundary is None else np.ascontiguousarray(boundary, np.int32)
        p = -np.ones(nqc, np.int32) if qcpartner is None else np.ascontiguousarray(qcpartner, np.int32)
        lf = np.zeros(nqc) if linkfactor is None else np.ascontiguousarray(linkfactor, np.float64)
        off = np.zeros(nqc + 1, np.int32) if mmpartner_offsets is None else \
            np.ascontiguousarray(mmpartner_offsets, np.int32)
        mmp = np.zeros(1, np.int32) if mmpartners is None or len(mmpartners) == 0 else \
            np.ascontiguousarray(mmpartners, np.int32)
        self.nqc = nqc
        check(load().nbabfs_set_qc_atoms(self.h, nqc, _ptr(idx), _ptr(b), _ptr(p), _ptr(lf), 1 if link_ratio else 0,
                                         _ptr(off), _ptr(mmp), int(coupling)))

    def set_qcmm_spline(self, x, y):
        """The QC/MM electrostatic spline through (x = r^2, y), clamped ends, atomic units."""
        x = np.ascontiguousarray(x, np.float64)
        y = np.ascontiguousarray(y, np.float64)
        if x.size != y.size:
            raise ValueError("spline x and y differ in length")
        check(load().nbabfs_set_qcmm_spline(self.h, x.size, _ptr(x), _ptr(y)))

    def qcmm_terms(self):
        """((QC/MM LJ, QC/MM 1-4 LJ), (present, present 1-4)) of the last energy call."""
        e = np.zeros(2)
        present = np.zeros(2, np.int32)
        check(load().nbabfs_qcmm_terms(self.h, _ptr(e), _ptr(present)))
        return e, present.astype(bool)

    def qcmm_potentials(self, potentials=None):
        """QC/MM potentials (atomic units) at the QC atoms on the last update's coordinates, added into
        potentials (a new zero array by default)."""
        pot = np.zeros(self.nqc) if potentials is None else potentials
        if pot.dtype != np.float64 or not pot.flags.c_contiguous or pot.size != self.nqc:
            raise ValueError("potentials must be a contiguous float64 array of %d" % self.nqc)
        check(load().nbabfs_qcmm_potentials(self.h, _ptr(pot)))
        return pot

    def qcmm_gradients(self, qc_charges, grad):
        """QC/MM electrostatic gradients for the QC charges, added into grad (n, 3)."""
        q = np.ascontiguousarray(qc_charges, np.float64)
        if q.size != self.nqc:
            raise ValueError("qc_charges must hold %d values" % self.nqc)
        if grad.dtype != np.float64 or not grad.flags.c_contiguous or grad.shape != (self.n, 3):
            raise ValueError("gradients must be a C-contiguous (%d, 3) float64 array" % self.n)
        check(load().nbabfs_qcmm_gradients_host(self.h, _ptr(q), _ptr(grad)))

    def qcmm_potentials_async(self, d_potentials):
        """Add the QC/MM potentials into a device array (nqc FP64), on the state's stream."""
        check(load().nbabfs_qcmm_potentials_async(self.h, C.c_void_p(d_potentials)))

    def qcmm_gradients_async(self, d_qc_charges, d_grad):
        """Add the QC/MM electrostatic gradients of the device QC charges into a device gradient array."""
        check(load().nbabfs_qcmm_gradients_async(self.h, C.c_void_p(d_qc_charges), C.c_void_p(d_grad)))

    def export_qcmm_pairs(self, which):
        """Sorted keys row * 2^32 + column of a QC/MM list: 'el', 'lj', 'el14' or 'lj14'."""
        k = {"el": 0, "lj": 1, "el14": 2, "lj14": 3}[which]
        L = load()
        m = L.nbabfs_export_qcmm_pairs(self.h, k, None, 0)
        if m < 0:
            check(int(-m))
        keys = np.zeros(max(m, 1), np.int64)
        L.nbabfs_export_qcmm_pairs(self.h, k, _ptr(keys), m)
        return keys[:m]

    def qcmm_sites(self):
        """(QC atom coordinates, MM site coordinates, MM site charges) of the last update."""
        L = load()
        m = L.nbabfs_qcmm_sites(self.h, None, None, None, 0)
        if m < 0:
            check(int(-m))
        qx = np.zeros((self.nqc, 3))
        sx = np.zeros((max(m, 1), 3))
        sq = np.zeros(max(m, 1))
        L.nbabfs_qcmm_sites(self.h, _ptr(qx), _ptr(sx), _ptr(sq), m)
        return qx, sx[:m], sq[:m]

    def set_stream(self, stream_handle):
        check(load().nbabfs_set_stream(self.h, None if not stream_handle else C.c_void_p(stream_handle)))

    def set_gradient_mode(self, mode):
        """GRADIENT_ADD (default: energy calls add into the gradient array) or GRADIENT_ASSIGN
        (they overwrite it, for a caller whose first gradient term this is)."""
        check(load().nbabfs_set_gradient_mode(self.h, int(mode)))

    def virial(self):
        """3x3 virial W_ab = sum_i x_i,a g_i,b of the last energy call's gradient."""
        w = np.zeros(9)
        check(load().nbabfs_virial(self.h, _ptr(w)))
        return w.reshape(3, 3)

    def autotune(self, d_xyz):
        """Time the pair kernel per work-item cut on the list for d_xyz; keep the fastest.

        Returns (entriesPerItem, wavesPerCU, kernel ms) of the chosen cut."""
        epi, wpc, ms = C.c_int(), C.c_int(), C.c_double()
        check(load().nbabfs_autotune(self.h, C.c_void_p(d_xyz), C.byref(epi), C.byref(wpc), C.byref(ms)))
        return epi.value, wpc.value, ms.value

    def stats(self):
        s = Stats()
        check(load().nbabfs_stats_get(self.h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in Stats._fields_}

    def set_timing(self, enable):
        check(load().nbabfs_set_timing(self.h, 1 if enable else 0))

    def export_pairs(self, which=0):
        L = load()
        m = L.nbabfs_export_pairs(self.h, which, None, 0)
        if m < 0:
            check(int(-m))
        keys = np.zeros(max(m, 1), np.int64)
        m2 = L.nbabfs_export_pairs(self.h, which, _ptr(keys), m)
        if m2 < 0:
            check(int(-m2))
        return keys[:m]

    def close(self):
        if getattr(self, "h", None):
            load().nbabfs_state_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class NativeBatch:
    """One persistent launch over many states (images / replicas)."""

    def __init__(self, states, use_graph=True):
        L = load()
        self.states = list(states)
        arr = (C.c_void_p * len(self.states))(*[s.h for s in self.states])
        h = C.c_void_p()
        check(L.nbabfs_batch_create(arr, len(self.states), 1 if use_graph else 0, C.byref(h)))
        self.h = h.value

    def _pointer_array(self, ptrs, slot):
        # The ctypes array of the last call is reused while the pointers are the same.
        key = tuple(ptrs)
        cached = self._arrays.get(slot)
        if cached is None or cached[0] != key:
            cached = (key, (C.c_void_p * len(key))(*key))
            self._arrays[slot] = cached
        return cached[1]

    def energy(self, d_xyz_ptrs, d_grad_ptrs=None):
        ns = len(self.states)
        if not hasattr(self, "_arrays"):
            self._arrays = {}
        xa = self._pointer_array(d_xyz_ptrs, 0)
        ga = None if d_grad_ptrs is None else self._pointer_array(d_grad_ptrs, 1)
        e = np.zeros(4 * ns)
        rebuilt = np.zeros(ns, np.int32)
        check(load().nbabfs_batch_energy(self.h, xa, ga, _ptr(e), _ptr(rebuilt)))
        return e.reshape(ns, 4), rebuilt.astype(bool)

    def info(self):
        """{"graph_captures", "launches", "layouts"} of the batch so far."""
        c, l, r = C.c_longlong(), C.c_longlong(), C.c_longlong()
        check(load().nbabfs_batch_info(self.h, C.byref(c), C.byref(l), C.byref(r), None, None))
        return {"graph_captures": c.value, "launches": l.value, "layouts": r.value}

    def set_timing(self, enable):
        """Time the batched pair launch with dispatch events (calls run without the graph meanwhile)."""
        check(load().nbabfs_batch_set_timing(self.h, 1 if enable else 0))

    def autotune(self, d_xyz_ptrs):
        """Time the batched pair launch per entriesPerItem on the systems' lists for d_xyz_ptrs and keep the
        fastest in every state.  Returns (entriesPerItem, pair launch ms)."""
        if not hasattr(self, "_arrays"):
            self._arrays = {}
        xa = self._pointer_array(d_xyz_ptrs, 0)
        epi, ms = C.c_int(), C.c_double()
        check(load().nbabfs_batch_autotune(self.h, xa, C.byref(epi), C.byref(ms)))
        return epi.value, ms.value

    def kernel_timing(self):
        """(mean ms of the batched pair launch, timed launches) since timing was enabled."""
        ms, k = C.c_double(), C.c_longlong()
        check(load().nbabfs_batch_info(self.h, None, None, None, C.byref(ms), C.byref(k)))
        return ms.value, k.value

    def close(self):
        if getattr(self, "h", None):
            load().nbabfs_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
