"""Benchmark of the NBModelABFS energy + gradient on one MI355X.

The default workload is the largest single-GPU configuration: 16 images of the
30,000-atom solute-in-water system (9,667 waters + 999 random solute atoms, 67 A
cube, 8 Lorentz-Berthelot solute LJ types), each with independent N(0, 0.1 A)
perturbations, 480,000 atoms evaluated as one batch.  One step = the half-skin
rebuild checks of all images + one batched energy/gradient evaluation (tiled
mixed-precision kernel), coordinates and gradients resident in HBM.  The other
configurations are named workloads (--workload): per system, one step =
NBModelABFS_Update (rebuilding when the half-skin rule fires) + one E+grad.

Prints ONE JSON line: pair interactions/s (reference list pairs, d <= 13.5 A,
per second), ms per E+grad, the roofline of the pair kernel (algorithmic
flops per launch over the launch time measured with HIP dispatch events on the
launch stream), the plain FP64 comparisons, and the reference C library timed
on this host's cores.
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "pdynamo-mirror_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

PEAK_FP32_TFLOPS = 157.3    # MI355X published FP32 vector peak
PEAK_FP64_TFLOPS = 78.6     # MI355X published FP64 vector peak
# Special-function issue (v_rsq_f32 and the like), taken as a quarter of the FP32 lane rate (the rate of the
# CDNA transcendental unit): 1,024 SIMDs x 4 lanes x 2.4 GHz.
PEAK_SPECIAL_TOPS = 9.83


def in_cutoff_terms(n_plain, n_switch, kms, ms):
    """The in-cutoff (d <= 12 A) pair rate of a step and the reference's two special-function operations per
    in-cutoff pair (1/r^2 and the square root) over the pair launch, against the special-function issue rate."""
    n_in = n_plain + n_switch
    ops = 2.0 * n_in
    achieved = ops / (kms * 1e-3) / 1e12 if kms else None
    return {"in_cutoff_pairs": n_in, "in_cutoff_pairs_per_s": n_in / (ms * 1e-3),
            "special_ops_per_launch": ops, "special_achieved_Tops": achieved, "special_peak_Tops": PEAK_SPECIAL_TOPS,
            "special_frac": (achieved / PEAK_SPECIAL_TOPS) if achieved else None}


DTYPE = {"mixed": "f64 pair core + f32 rsqrt seed/force vectors + int64 fixed-point sums",
         "tiled_fp64": "f64 pair core and force vectors + int64 fixed-point sums", "fp64": "f64"}


def algorithmic_flops(xyz, keys, n):
    """8 N_list + 37 N_(d<=8) + 63 N_(8<d<=12): the reference's per-pair operations
    (displacement and r^2; plain Coulomb+LJ+gradient; switched Coulomb+LJ+gradient)."""
    i = keys // n
    j = keys % n
    d = xyz[i] - xyz[j]
    r2 = np.einsum("ij,ij->i", d, d)
    n_list = keys.size
    n_plain = int(np.count_nonzero(r2 <= 64.0))
    n_switch = int(np.count_nonzero((r2 > 64.0) & (r2 <= 144.0)))
    return 8.0 * n_list + 37.0 * n_plain + 63.0 * n_switch, n_list, n_plain, n_switch


def image_flops(xyz, st, n, transformations, cell):
    """algorithmic_flops over the image pairs (i primary, j image) of the state's image list."""
    from pdynamo_amd.Symmetry import SymmetryParameters
    sp = SymmetryParameters(*cell)
    rot, tra = transformations
    tot = [0.0, 0, 0, 0]
    for q, (a, b, c, t, scale, npr) in enumerate(st.images()):
        keys = st.export_image_pairs(q)
        R = sp.M @ np.asarray(rot[t]).reshape(3, 3) @ sp.inverseM
        T = sp.M @ (np.asarray(tra[t]) + np.array([a, b, c], np.float64))
        d = xyz[keys // n] - (xyz[keys % n] @ R.T + T)
        r2 = np.einsum("ij,ij->i", d, d)
        n_plain = int(np.count_nonzero(r2 <= 64.0))
        n_switch = int(np.count_nonzero((r2 > 64.0) & (r2 <= 144.0)))
        tot[0] += 8.0 * keys.size + 37.0 * n_plain + 63.0 * n_switch
        tot[1] += int(keys.size)
        tot[2] += n_plain
        tot[3] += n_switch
    return tuple(tot)


def copy_calibration(mib):
    """Device-to-device copies of a known size (the runtime's copy kernel), so that a
    profiler's FETCH_SIZE / WRITE_SIZE of the pair kernel can be scaled by what the same
    counters read for a copy of exactly mib MiB read and mib MiB written."""
    import ctypes
    from pdynamo_amd import _native
    hip = ctypes.CDLL("libamdhip64.so")
    n = mib * (1 << 20) // 8
    src = _native.DeviceArray((n,))
    dst = _native.DeviceArray((n,))
    src.zero()
    for _ in range(3):
        rc = hip.hipMemcpy(ctypes.c_void_p(dst.ptr), ctypes.c_void_p(src.ptr), ctypes.c_size_t(n * 8), 3)
        if rc != 0:
            raise RuntimeError("hipMemcpy device-to-device failed: %d" % rc)
    _native.synchronize()
    src.free()
    dst.free()


def traffic_from_summary(path):
    """HBM bytes per launch recorded in a rocprofv3 traffic summary under profiles/."""
    if not path or not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("TRAFFIC_BYTES_PER_LAUNCH:"):
            return float(line.split(":", 1)[1])
    return None


def load_dhfr():
    """The reference's DHFR benchmark system (CHARMM22, 23,558 atoms), non-bonding inputs from the committed
    fixture tests/golden/dhfr_nb.npz, with its 1-4 LJ tables."""
    from pdynamo_amd.synthetic import MMSystem
    with np.load(os.path.join(ROOT, "tests", "golden", "dhfr_nb.npz"), allow_pickle=False) as z:
        mm = MMSystem("dhfr", z["coordinates3"], z["charges"], z["ljtypes"], int(z["ntypes"]), z["tableA"],
                      z["tableB"], z["exclusions"], z["interactions14"])
        return mm, {"tableA14": z["tableA14"], "tableB14": z["tableB14"]}


def make_workload(name):
    """(system, description, extra state arguments)."""
    from pdynamo_amd import synthetic
    if name == "solute30000":
        return synthetic.config1(), "C1 solute-in-water 30,000 atoms (9,667 waters + 999 solute, 67 A cube)", {}
    if name == "water200001":
        return synthetic.config2(), "C2 water box 200,001 atoms (66,667 waters, 126 A cube)", {}
    if name == "water648":
        return synthetic.config0(), "C0 water 648 atoms (216 waters, 18.6 A cube)", {}
    if name == "dhfr":
        mm, extra = load_dhfr()
        return mm, "DHFR CHARMM22 23,558 atoms (reference benchmark system, primary + 1-4 lists, no images)", extra
    if name == "dhfr_periodic":
        mm, extra = load_dhfr()
        with np.load(os.path.join(ROOT, "tests", "golden", "dhfr_nb.npz"), allow_pickle=False) as z:
            cell = tuple(float(v) for v in z["cell"])
        extra["symmetry"] = ((np.eye(3)[None], np.zeros((1, 3))), cell)
        return mm, ("DHFR CHARMM22 23,558 atoms in its cubic cell a = 62.23 A (the reference's benchmark "
                    "configuration: primary, 1-4 and 13 P1 image lists)"), extra
    raise SystemExit("unknown workload " + name)


def run_qcmm(args):
    """QC/MM workload: the branched chain (6 pure QC atoms, 2 link atoms, RC coupling) plus the 40 nearest
    waters as QC atoms, in 7,000 waters.  One step = a System.Energy of a QC/MM system without the QC
    program: the MM/MM and QC/MM LJ energy and gradients (device arrays), the QC/MM potentials at the
    QC atoms, then the QC/MM electrostatic gradients for the QC charges (host arrays, as a QC program
    exchanges them)."""
    from pdynamo_amd import _native, synthetic
    from pdynamo_amd.NBModelABFS import NBModelABFS
    mm, qd = synthetic.qcmm_chain_in_water(7000, 59.4, qc_waters=40)
    lf = np.where(qd["boundary"], 1.09, 0.0)
    model = NBModelABFS(precision=args.precision, qcmmCoupling="RC")
    model.CheckPairwiseInteractions()
    st = _native.NativeState(model.NativeOptions(), mm.charges, mm.ljtypes, mm.ntypes, mm.tableA, mm.tableB,
                             exclusions=mm.exclusions, interactions14=mm.interactions14)
    st.set_qc_atoms(qd["index"], qd["boundary"], qd["partner"], lf, False, qd["mmp_off"], qd["mmp"],
                    _native.QCMM_RC)
    qi = model.qcmmPairwiseInteraction
    st.set_qcmm_spline(qi.splineX, qi.splineY)
    nqc = qd["index"].size
    qcq = np.linspace(-0.8, 0.8, nqc)
    x = _native.DeviceArray.from_numpy(mm.coordinates3, st)
    g = _native.DeviceArray(mm.coordinates3.shape)
    hg = np.zeros_like(mm.coordinates3)
    pot = np.zeros(nqc)

    st.set_gradient_mode(_native.GRADIENT_ASSIGN)
    dpot = _native.DeviceArray((nqc,))
    dq = _native.DeviceArray.from_numpy(qcq, st)

    # qc: "device" -- the QC charges and potentials in GPU memory (the device-array calls, no host
    # round trip), "host" -- host arrays as QCMMInteractionState holds them (copies and a
    # synchronisation per call), None -- the MM/MM part only.
    def step(qc="device"):
        st.update(x.ptr)
        st.energy_async(x.ptr, g.ptr)
        if qc == "device":
            dpot.zero(st)
            st.qcmm_potentials_async(dpot.ptr)
            st.qcmm_gradients_async(dq.ptr, g.ptr)
        elif qc == "host":
            pot[:] = 0.0
            st.qcmm_potentials(pot)
            st.qcmm_gradients(qcq, hg)

    def timed(qc):
        settle(lambda: step(qc), args.settle)
        for _ in range(args.warmup):
            step(qc)
        _native.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step(qc)
        _native.synchronize()
        return (time.perf_counter() - t0) / args.steps * 1e3

    ms = timed("device")
    ms_host = timed("host")
    ms_mm = timed(None)
    pot = dpot.numpy(st)
    s = st.stats()
    e, _ = st.energies()
    qe, _ = st.qcmm_terms()
    n_list = int(s["npairs"]) + int(s["npairs14"]) + int(s["qcmmElPairs"]) + int(s["qcmmLJPairs"])
    result = {
        "metric": "pair interactions/s (reference list pairs d<=13.5 A per E+grad second)",
        "value": n_list / (ms * 1e-3), "unit": "pair interactions/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 (QC/MM) + the MM/MM path's precision", "data": "synthetic (seeded generator, pdynamo_amd.synthetic)",
        "config": {"clock_settle_s": args.settle, "workload": "QC/MM chain + 40 QC waters in 7,000 waters (RC coupling, 2 link atoms)",
                   "atoms": mm.natoms, "qc_atoms": nqc, "precision": args.precision, "list_pairs": n_list,
                   "mmmm_pairs": int(s["npairs"]), "qcmm_el_pairs": int(s["qcmmElPairs"]),
                   "qcmm_lj_pairs": int(s["qcmmLJPairs"]), "ms_mm_part_only": ms_mm,
                   "ms_qcmm_potentials_and_gradients": ms - ms_mm,
                   "qcmm_arrays": "device (potentials and QC charges in GPU memory)",
                   "ms_per_step_host_arrays": ms_host,
                   "energies_kJ_mol": [float(v) for v in list(e) + list(qe)],
                   "potential_sum_au": float(pot.sum())},
    }
    if not args.no_cpu_baseline:
        import oracle as O
        if O.ref_available():
            import qcmm_oracle as Q
            threads = len(os.sched_getaffinity(0))
            if os.environ.get("OMP_NUM_THREADS", "").isdigit():
                threads = min(threads, int(os.environ["OMP_NUM_THREADS"]))
            qc = Q.QCAtoms(mm.natoms, qd["index"], qd["boundary"], qd["partner"], lf, qd["mmp_off"], qd["mmp"])
            ref = Q.ReferenceQCMM(mm, qc, 1, qi.splineX, qi.splineY, nthreads=threads)
            ref.call(mm.coordinates3, qcq)
            calls = max(1, args.cpu_calls // 10)
            t0 = time.perf_counter()
            for _ in range(calls):
                ref.call(mm.coordinates3, qcq)
            dt = (time.perf_counter() - t0) / calls
            ref.close()
            result["cpu_baseline"] = {"value": n_list / dt, "unit": "pair interactions/s", "cores": threads,
                                      "kind": "reference",
                                      "sample": "%d QC/MM calls (Update, MMMMEnergy, QCMMEnergyLJ, QCMMPotentials, "
                                                "QCMMGradients) after the first list build, %.1f ms each" % (calls, dt * 1e3)}
    print(json.dumps(result))


def settle(fn, seconds):
    """Untimed calls for `seconds` of wall time before the warmup steps.  The GPU's clocks ramp up under load:
    back-to-back pair launches get about a fifth faster over their first ~30 ms, so a handful of warmup steps
    alone would time the ramp rather than the kernel."""
    from pdynamo_amd import _native
    n = 0
    t_end = time.perf_counter() + seconds
    while time.perf_counter() < t_end:
        fn()
        n += 1
        if n % 16 == 0:
            _native.synchronize()
    _native.synchronize()
    return n


def make_batch(name):
    from pdynamo_amd import synthetic
    if name == "droplets64":
        return synthetic.config3(), "C3 64 water droplets, log-uniform 500-20,000 atoms, one batched launch"
    if name == "images16":
        return synthetic.config4(), "C4 16 images of the 30,000-atom solute system (0.1 A perturbations), one batched launch"
    return None, None


def run_batch(args, systems, wdesc, workload):
    """One step = the rebuild checks of every system + one batched E+grad of all of them (hipGraph replay of
    prepare, pair, 1-4 and finalize passes over the union of the systems' work items)."""
    from pdynamo_amd import _native
    from pdynamo_amd.NBModelABFS import NBModelABFS

    def states_for(precision):
        model = NBModelABFS(precision=precision, **({"entriesPerItem": args.entries_per_item}
                                                    if args.entries_per_item else {}))
        sts = [_native.NativeState(model.NativeOptions(), m.charges, m.ljtypes, m.ntypes, m.tableA, m.tableB,
                                   exclusions=m.exclusions) for m in systems]
        for st in sts:
            st.set_gradient_mode(_native.GRADIENT_ASSIGN)   # the batch writes every gradient row
        return sts

    sizes = [m.natoms for m in systems]
    offs = np.concatenate([[0], np.cumsum(sizes)])
    allx = np.concatenate([m.coordinates3 for m in systems])
    X = _native.DeviceArray.from_numpy(allx)
    G = _native.DeviceArray(allx.shape)
    xp = [X.ptr + 24 * int(o) for o in offs[:-1]]
    gp = [G.ptr + 24 * int(o) for o in offs[:-1]]

    def timed_batch(precision, steps, warmup, kernel_pass=True):
        states = states_for(precision)
        batch = _native.NativeBatch(states, use_graph=True)
        settle(lambda: batch.energy(xp, gp), args.settle)
        tuned = None
        if not args.no_autotune and not args.entries_per_item and precision != "fp64":
            # After the clock settling: the tuner's timings compare cuts, not clock states.
            epi, akms = batch.autotune(xp)
            tuned = {"entriesPerItem": epi, "kernel_ms": akms}
            # The tuner's host-side re-cuts leave the GPU idle now and then: settle the clocks again.
            settle(lambda: batch.energy(xp, gp), args.settle)
        for _ in range(warmup):
            batch.energy(xp, gp)
        _native.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            e, rebuilt = batch.energy(xp, gp)
        _native.synchronize()
        ms = (time.perf_counter() - t0) / steps * 1e3
        kms = None
        if kernel_pass:
            # The pair launch's own duration: dispatch events in a pass of its own (no graph meanwhile), so
            # that the events do not add to the step time above.
            batch.set_timing(True)
            for _ in range(min(steps, 20)):
                batch.energy(xp, gp)
            kms, nk = batch.kernel_timing()
            batch.set_timing(False)
        return ms, kms, e, states, batch, tuned

    ms, kms, e, states, batch, tuned = timed_batch(args.precision, args.steps, args.warmup)
    stats = [st.stats() for st in states]
    npairs = int(sum(s_["npairs"] for s_ in stats))

    def sequential(precision):
        # The same systems as separate single-state calls, one after another, on one stream.
        seq = states_for(precision)
        for st, x, g in zip(seq, xp, gp):
            st.update(x)
            st.energy_async(x, g)
        _native.synchronize()
        reps = max(1, min(args.steps, 20) // 4)
        t0 = time.perf_counter()
        for _ in range(reps):
            for st, x, g in zip(seq, xp, gp):
                st.update(x)
                st.energy_async(x, g)
            _native.synchronize()
        out = (time.perf_counter() - t0) / reps * 1e3
        for st in seq:
            st.close()
        return out

    cmp = {}
    if not args.no_compare:
        # The same batch in the other arithmetic paths, then the systems as sequential single-state calls.
        for other in ("tiled_fp64", "fp64", "mixed"):
            if other == args.precision:
                continue
            ms_o, kms_o, _, st_o, b_o, _ = timed_batch(other, max(5, args.steps // 5), 2)
            b_o.close()
            for st in st_o:
                st.close()
            cmp.update({"batched_%s_ms" % other: ms_o, "pair_kernel_ms_%s" % other: kms_o,
                        "speedup_vs_batched_%s" % other: ms_o / ms})
        ms_seq_plain = sequential("fp64")
        ms_seq_mixed = sequential("mixed")
        cmp.update({"sequential_plain_fp64_ms": ms_seq_plain, "sequential_mixed_ms": ms_seq_mixed,
                    "speedup_vs_sequential_plain_fp64": ms_seq_plain / ms})
    result = {
        "metric": "pair interactions/s (reference list pairs d<=13.5 A per E+grad second)",
        "value": npairs / (ms * 1e-3), "unit": "pair interactions/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": DTYPE[args.precision], "data": "synthetic (seeded generator, pdynamo_amd.synthetic)",
        "config": {"clock_settle_s": args.settle, "workload": wdesc, "systems": len(systems), "atoms": int(offs[-1]), "list_pairs": npairs,
                   "min_atoms": int(min(sizes)), "max_atoms": int(max(sizes)), "precision": args.precision,
                   "work_items": int(sum(s_["workItems"] for s_ in stats)),
                   "cluster_entries": int(sum(s_["entries"] for s_ in stats)), "autotuned": tuned,
                   "energies_kJ_mol_system0": [float(v) for v in e[0]], **cmp},
    }
    if not args.no_roofline and kms:
        flops = n_plain = n_switch = 0
        for st, m in zip(states, systems):
            f_, _, p_, s2_ = algorithmic_flops(m.coordinates3, st.export_pairs(0), m.natoms)
            flops, n_plain, n_switch = flops + f_, n_plain + p_, n_switch + s2_
        achieved = flops / (kms * 1e-3) / 1e12
        peak = PEAK_FP32_TFLOPS if args.precision == "mixed" else PEAK_FP64_TFLOPS
        tpath = args.traffic_summary
        if tpath is None and args.precision == "mixed":
            tpath = os.path.join(ROOT, "profiles", "traffic_k_force_tiled_%s.md" % workload)
        result["roofline"] = {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                              "frac": achieved / peak, "traffic": traffic_from_summary(tpath),
                              "kernel": "k_force_plain" if args.precision == "fp64" else "k_force_tiled",
                              "kernel_ms": kms, "launches": min(args.steps, 20), "flops_per_launch": flops,
                              "pairs_d_le_8": n_plain, "pairs_8_12": n_switch,
                              **in_cutoff_terms(n_plain, n_switch, kms, ms)}
    batch.close()
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline_batch(systems, npairs, args.cpu_image_calls)
    if args.copy_calibration_mib > 0:
        copy_calibration(args.copy_calibration_mib)
    print(json.dumps(result))


def host_threads():
    threads = len(os.sched_getaffinity(0))
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        threads = min(threads, int(os.environ["OMP_NUM_THREADS"]))
    return threads


def cpu_baseline_batch(systems, npairs, calls):
    """The reference pDynamo C library (oracle/_ref, -O3 -fopenmp) over the same systems, one configuration
    after another as pDynamo's serial image pool evaluates them, in the two buckets of SystemWithTimings
    ("NB Set Up": the list build of the first call; "NB Evaluation": the E+grad of the calls after it), at
    the host's thread count and at one thread.  Bounded sample: per system one build and `calls` E+grad
    calls (one at one thread)."""
    import oracle as O
    kind = "reference" if O.ref_available() else "port"

    def run(nthreads, ncalls):
        t_build = t_eval = 0.0
        for m in systems:
            if kind == "reference":
                eng = O.ReferenceABFS(m, nthreads=nthreads)
            else:
                eng = O.OracleABFS(m, nthreads=nthreads)
            t0 = time.perf_counter()
            eng.energy(m.coordinates3)
            t1 = time.perf_counter()
            for _ in range(ncalls):
                eng.energy(m.coordinates3)
            t2 = time.perf_counter()
            per = (t2 - t1) / ncalls
            t_eval += per
            t_build += max((t1 - t0) - per, 0.0)
            if kind == "reference":
                eng.close()
        return t_build, t_eval

    threads = host_threads()
    b_n, e_n = run(threads, calls)
    b_1, e_1 = run(1, 1)
    return {"value": npairs / e_n, "unit": "pair interactions/s", "cores": threads, "kind": kind,
            "sample": "each of the %d systems: one list build and %d E+grad calls at %d threads, one build and one "
                      "E+grad call at 1 thread" % (len(systems), calls, threads),
            "eval_ms_per_step": e_n * 1e3, "list_build_ms_per_step": b_n * 1e3,
            "one_thread": {"value": npairs / e_1, "unit": "pair interactions/s", "cores": 1,
                           "eval_ms_per_step": e_1 * 1e3, "list_build_ms_per_step": b_1 * 1e3}}


def cpu_baseline(mm, calls, threads, extra=None):
    """The reference pDynamo C (oracle/_ref, -O3 -fopenmp) or, if absent, the
    FP64 C restatement, timed for `calls` E+grad calls after the first build."""
    import oracle as O
    extra = dict(extra or {})
    sym = extra.pop("symmetry", None)
    kind = "reference" if O.ref_available() else "port"
    if sym is not None:
        extra["transformations"] = sym[0]
    if kind == "reference":
        eng = O.ReferenceABFS(mm, nthreads=threads, **extra)
    else:
        eng = O.OracleABFS(mm, nthreads=threads, **extra)

    def call():
        if sym is None:
            eng.energy(mm.coordinates3)
        else:
            eng.energy_sym(mm.coordinates3, sym[1])

    call()
    t0 = time.perf_counter()
    for _ in range(calls):
        call()
    dt = (time.perf_counter() - t0) / calls
    if kind == "reference":
        npairs = eng.canonical(0).size + (sum(x[5] for x in eng.images()) if sym is not None else 0)
        eng.close()
    else:
        npairs = eng.records.npairs + (eng.images.npairs if sym is not None else 0)
    return {"value": npairs / dt, "unit": "pair interactions/s", "cores": threads, "kind": kind,
            "sample": "%d E+grad calls on the same workload after the first list build (%.1f ms each, %d list pairs)"
                      % (calls, dt * 1e3, npairs)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--settle", type=float, default=0.5,
                    help="seconds of untimed calls before the warmup steps (the GPU's clocks ramp up under load)")
    ap.add_argument("--workload", default="images16",
                    help="images16 (default: 16 perturbed images of the 30,000-atom system, one batch), droplets64, "
                         "solute30000, water200001, water648, dhfr, dhfr_periodic, qcmm_solvated")
    ap.add_argument("--precision", default="mixed", choices=["mixed", "tiled_fp64", "fp64"])
    ap.add_argument("--entries-per-item", type=int, default=0)
    ap.add_argument("--waves-per-cu", type=int, default=-1)
    ap.add_argument("--image-entries-per-item", type=int, default=0)
    ap.add_argument("--cpu-calls", type=int, default=100)
    ap.add_argument("--cpu-image-calls", type=int, default=3,
                    help="E+grad calls per system after its list build in the batch workloads' CPU baseline")
    ap.add_argument("--moves", type=float, default=0.0,
                    help="Gaussian displacement (A) applied before every call, cumulative; rebuilds follow the "
                         "half-skin rule (the 200,001-atom protocol uses 0.05)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-compare", action="store_true",
                    help="batch workloads: skip the other precision and the sequential single-state comparisons "
                         "(profiler runs)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="skip the library's work-item autotuner (entriesPerItem / wavesPerCU) before timing")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="after the run, copy this many MiB device to device three times (counter calibration)")
    ap.add_argument("--traffic-summary", default=None,
                    help="rocprofv3 traffic summary giving roofline.traffic (default: the committed one for the "
                         "default workload)")
    ap.add_argument("--no-roofline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (--gpus %d requested)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    from pdynamo_amd import _native
    from pdynamo_amd.NBModelABFS import NBModelABFS

    if args.workload == "qcmm_solvated":
        run_qcmm(args)
        return
    systems, bdesc = make_batch(args.workload)
    if systems is not None:
        run_batch(args, systems, bdesc, args.workload)
        return
    mm, wdesc, extra = make_workload(args.workload)
    opts = {"precision": args.precision}
    if args.entries_per_item:
        opts["entriesPerItem"] = args.entries_per_item
    if args.waves_per_cu >= 0:
        opts["wavesPerCU"] = args.waves_per_cu
    model = NBModelABFS(**opts)
    sym = extra.get("symmetry")
    nopt = model.NativeOptions()
    if args.image_entries_per_item > 0:
        nopt.imageEntriesPerItem = args.image_entries_per_item
    st = _native.NativeState(nopt, mm.charges, mm.ljtypes, mm.ntypes, mm.tableA, mm.tableB,
                             exclusions=mm.exclusions, interactions14=mm.interactions14,
                             **{k: v for k, v in extra.items() if k != "symmetry"})
    if sym is not None:
        st.set_transformations(sym[0][0], sym[0][1])
        st.set_symmetry_parameters(sym[1])
    g = _native.DeviceArray(mm.coordinates3.shape)
    if args.moves > 0.0:
        from pdynamo_amd import synthetic
        seq = synthetic.displacement_sequence(mm, args.warmup + args.steps, args.moves, seed=7)
        xs = [_native.DeviceArray.from_numpy(c, st) for c in seq]
    else:
        xs = [_native.DeviceArray.from_numpy(mm.coordinates3, st)]
    _native.synchronize()
    tuned = None
    if not args.no_autotune and args.entries_per_item == 0 and args.waves_per_cu < 0:
        # After the clock settling: the tuner's timings compare cuts, not clock states.
        def tune_settle():
            st.update(xs[0].ptr)
            st.energy_async(xs[0].ptr, g.ptr)
        settle(tune_settle, args.settle)
        epi, wpc, kms = st.autotune(xs[0].ptr)
        tuned = {"entriesPerItem": epi, "wavesPerCU": wpc, "kernel_ms": kms}
    counter = [0]
    rebuilds = [0]

    # System.Energy zeroes the gradients before the model adds to them; with the model as
    # the first gradient term that zero fill is folded into its finalize pass (assign mode).
    st.set_gradient_mode(_native.GRADIENT_ASSIGN)

    def step():
        x = xs[counter[0] % len(xs)]
        counter[0] += 1
        rebuilds[0] += st.update(x.ptr)
        st.energy_async(x.ptr, g.ptr)

    # Clock settling with the first coordinates (the list stays valid for them), then the warmup steps.
    def settle_step():
        st.update(xs[0].ptr)
        st.energy_async(xs[0].ptr, g.ptr)
    settle(settle_step, args.settle)
    for _ in range(args.warmup):
        step()
    rebuilds[0] = 0
    _native.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    _native.synchronize()
    t1 = time.perf_counter()
    ms = (t1 - t0) / args.steps * 1e3
    timed_rebuilds = rebuilds[0]
    # Pair-kernel duration from HIP events on the state's stream, in a pass of
    # its own so that the events do not add to the step time above.
    st.set_timing(True)
    for _ in range(min(args.steps, 20)):
        step()
    _native.synchronize()
    stats = st.stats()
    st.set_timing(False)
    e, present = st.energies()
    e = list(e)
    if sym is not None:
        ie, _ = st.image_terms()
        e += list(ie)
    keys = st.export_pairs(0)
    x_last = xs[(counter[0] - 1) % len(xs)].numpy(st)
    n_list = int(keys.size) + (int(stats["imagePairs"]) if sym is not None else 0)
    value = n_list / (ms * 1e-3)

    result = {
        "metric": "pair interactions/s (reference list pairs d<=13.5 A per E+grad second)",
        "value": value,
        "unit": "pair interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE[args.precision],
        "data": "synthetic (seeded generator, pdynamo_amd.synthetic)",
        "config": {"clock_settle_s": args.settle, "workload": wdesc, "atoms": mm.natoms, "list_pairs": n_list, "precision": args.precision,
                   "list_cutoff": 13.5, "cutoffs": [0.5, 8.0, 12.0],
                   "energies_kJ_mol": [float(v) for v in e], "work_items": int(stats["workItems"]),
                   "cluster_entries": int(stats["entries"]), "moves_sigma_A": args.moves,
                   "rebuilds_in_timed_steps": int(timed_rebuilds), "autotuned": tuned},
    }
    if sym is not None:
        result["config"]["images"] = int(stats["images"])
        result["config"]["image_pairs"] = int(stats["imagePairs"])
    if not args.no_roofline:
        flops, n_list2, n_plain, n_switch = algorithmic_flops(x_last, keys, mm.natoms)
        if sym is not None:
            f2, _, p2, s2 = image_flops(x_last, st, mm.natoms, sym[0], sym[1])
            flops, n_plain, n_switch = flops + f2, n_plain + p2, n_switch + s2
        kms = stats["forceKernelMs"]
        achieved = flops / (kms * 1e-3) / 1e12 if kms > 0 else None
        peak = PEAK_FP32_TFLOPS if args.precision == "mixed" else PEAK_FP64_TFLOPS
        tpath = args.traffic_summary
        if tpath is None and args.workload == "solute30000" and args.precision == "mixed":
            tpath = os.path.join(ROOT, "profiles", "traffic_k_force_tiled_solute30000.md")
        result["roofline"] = {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                              "frac": (achieved / peak) if achieved else None,
                              "traffic": traffic_from_summary(tpath),
                              "kernel": "k_force_plain" if args.precision == "fp64" else "k_force_tiled",
                              "kernel_ms": kms, "launches": int(stats["forceKernelLaunches"]),
                              "flops_per_launch": flops, "pairs_d_le_8": n_plain, "pairs_8_12": n_switch,
                              **in_cutoff_terms(n_plain, n_switch, kms, ms)}
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(mm, args.cpu_calls, host_threads(), extra)
    if args.copy_calibration_mib > 0:
        copy_calibration(args.copy_calibration_mib)
    print(json.dumps(result))


if __name__ == "__main__":
    main()
