"""Non-finite inputs and out-of-range sums are reported, not returned as finite wrong values.

The reference propagates NaN through the pair loop (PairwiseInteraction.c:392-425) and returns it with
Status_Continue.  The GPU path instead returns NBABFS_STATUS_VALUE_ERROR (CLibraryError in Python): a NaN or
infinite coordinate is found by the list build (an infinite bounding box) or by the per-call pass that reads
every coordinate (a NaN fails no distance test, so its pairs would otherwise drop out silently); a gradient or
energy contribution beyond the +-2^66 range of the fixed-point sums is flagged where it is added.  Every case
here is an ordinary input: all kernels run to completion, and a later call on good coordinates gives the same
bits as before.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _state(mm, precision, charges=None):
    from pdynamo_amd import _native
    from pdynamo_amd.NBModelABFS import NBModelABFS
    q = mm.charges if charges is None else charges
    return _native.NativeState(NBModelABFS(precision=precision).NativeOptions(), q, mm.ljtypes, mm.ntypes,
                               mm.tableA, mm.tableB, exclusions=mm.exclusions)


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_nonfinite_coordinates(precision):
    from pdynamo_amd import _native, synthetic
    mm = synthetic.random_small_system(1500, 45.0, 8)
    st = _state(mm, precision)
    x = _native.DeviceArray.from_numpy(mm.coordinates3, st)
    g = _native.DeviceArray(mm.coordinates3.shape)
    assert st.update(x.ptr)
    g.zero(st)
    st.energy_async(x.ptr, g.ptr)
    e0, _ = st.energies()
    g0 = g.numpy(st)
    for bad_value in (np.nan, np.inf):
        bad = mm.coordinates3.copy()
        bad[17, 1] = bad_value
        xb = _native.DeviceArray.from_numpy(bad, st)
        if np.isnan(bad_value):
            # With a list in place a NaN fails the half-skin test (no rebuild); the per-call pass reports it.
            assert not st.update(xb.ptr)
            st.energy_async(xb.ptr, g.ptr)
            with pytest.raises(_native.CLibraryError, match="NaN or infinite"):
                st.energies()
        else:
            # An infinite displacement triggers a rebuild, whose bounding box is infinite.
            with pytest.raises(_native.CLibraryError, match="NaN or infinite"):
                st.update(xb.ptr)
        xb.free()
        # The error word was cleared: good coordinates give the earlier bits again.
        st.update(x.ptr)
        g.zero(st)
        st.energy_async(x.ptr, g.ptr)
        e1, _ = st.energies()
        assert np.array_equal(e0, e1)
        assert np.array_equal(g.numpy(st), g0)
        # A new state: the list build finds it.
        st2 = _state(mm, precision)
        with pytest.raises(_native.CLibraryError, match="NaN or infinite"):
            st2.energy_host(bad, np.zeros_like(bad))
        st2.close()
    st.close()


def test_contributions_beyond_fixed_point_range():
    """Charges of 1e30 e: pair energies of ~1e63 kJ/mol cannot be held by the fixed-point sums of the tiled
    path, which reports a range error (the plain FP64 path holds them in FP64 and returns them)."""
    from pdynamo_amd import _native, synthetic
    mm = synthetic.random_small_system(600, 30.0, 3)
    st = _state(mm, "mixed", charges=mm.charges * 1e30)
    with pytest.raises(_native.CLibraryError, match="beyond"):
        st.energy_host(mm.coordinates3, np.zeros_like(mm.coordinates3))
    st.close()
    st = _state(mm, "fp64", charges=mm.charges * 1e30)
    e, _, _ = st.energy_host(mm.coordinates3, np.zeros_like(mm.coordinates3))
    assert np.isfinite(e).all() and abs(e[0]) > 1e50
    st.close()


def test_nonfinite_parameters_rejected():
    from pdynamo_amd import _native, synthetic
    mm = synthetic.random_small_system(100, 20.0, 4)
    q = mm.charges.copy()
    q[3] = np.nan
    with pytest.raises(_native.CLibraryError, match="not finite"):
        _state(mm, "mixed", charges=q)


def test_batch_reports_the_system():
    from pdynamo_amd import _native, synthetic
    systems = [synthetic.random_small_system(n, b, sd) for n, b, sd in [(600, 35.0, 1), (1200, 40.0, 3)]]
    states = [_state(m, "mixed") for m in systems]
    xs = [_native.DeviceArray.from_numpy(m.coordinates3) for m in systems]
    gs = [_native.DeviceArray(m.coordinates3.shape) for m in systems]
    batch = _native.NativeBatch(states, use_graph=True)
    e0, _ = batch.energy([x.ptr for x in xs], [g.ptr for g in gs])
    bad = systems[1].coordinates3.copy()
    bad[7, 2] = np.nan
    xs[1].upload(bad)
    with pytest.raises(_native.CLibraryError, match="batch system 1"):
        batch.energy([x.ptr for x in xs], [g.ptr for g in gs])
    xs[1].upload(systems[1].coordinates3)
    e1, _ = batch.energy([x.ptr for x in xs], [g.ptr for g in gs])
    assert np.array_equal(e0, e1)
    batch.close()


def test_system_energy_raises():
    from pdynamo_amd import CLibraryError, NBModelABFS, System, synthetic
    mm = synthetic.config0()
    s = System.FromSynthetic(mm)
    s.DefineNBModel(NBModelABFS())
    s.Energy(log=None, doGradients=True)
    bad = mm.coordinates3.copy()
    bad[0, 0] = np.nan
    s.coordinates3 = bad
    with pytest.raises(CLibraryError):
        s.Energy(log=None, doGradients=True)
