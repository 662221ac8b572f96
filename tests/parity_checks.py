"""Parity checks of the GPU path against the FP64 oracle, shared by the GPU tests.

Energies are checked term by term, each against its own magnitude:
|E_k - E_k,oracle| <= etol * max(|E_k,oracle|, 1 kJ/mol), so that a large term (the
Lennard-Jones overlaps of randomly placed atoms reach 1e11 kJ/mol) cannot hide an error
in a small one (Coulomb, the 1-4 terms).  Where the oracle's sums of pair-term
magnitudes S_k = sum |e_pair| are given, the bound is at least 4 u S_k (u = 2^-53): two
FP64 sums of the same terms in different orders differ by up to about that much, which
exceeds 1e-12 |E_k| for a strongly cancelling total (a neutral water droplet's Coulomb
energy is 1e-5 of S_k).  Gradients are checked as a relative RMS over
all atoms and, separately, over the atoms whose oracle gradient is below 1e4 kJ/mol/A
(atoms in no overlapping pair): the overlap forces (up to 1e13) would otherwise hide the
ordinary atoms from the global norm.  Every call prints its per-term absolute errors.
"""

import numpy as np

ORDINARY_FORCE = 1.0e4      # kJ/mol/A: atoms above this take part in an overlap
FP64_SUM_ROUNDING = 4.0 * 2.0 ** -53


def check_terms(terms, rterms, etol, label="", scales=None):
    rterms = dict(rterms)
    assert set(terms) == set(rterms), (label, sorted(terms), sorted(rterms))
    out = []
    tot_tol = etol * max(abs(sum(rterms.values())), 1.0)
    for k in sorted(rterms):
        err = abs(terms[k] - rterms[k])
        tol = etol * max(abs(rterms[k]), 1.0)
        if scales is not None:
            tol = max(tol, FP64_SUM_ROUNDING * scales[k])
            tot_tol += FP64_SUM_ROUNDING * scales[k]
        out.append("%s=%.10g (abs err %.3g, tol %.3g)" % (k, rterms[k], err, tol))
        assert err <= tol, (label, k, terms[k], rterms[k], err, tol)
    etot = sum(rterms.values())
    err = abs(sum(terms.values()) - etot)
    assert err <= tot_tol, (label, "total", sum(terms.values()), etot, err, tot_tol)
    print("%s terms: %s" % (label, "; ".join(out)), flush=True)


def rel_rms(g, gref, mask=None):
    if mask is not None:
        g, gref = g[mask], gref[mask]
    if gref.size == 0:
        return 0.0
    return float(np.linalg.norm(g - gref) / max(np.linalg.norm(gref), 1e-300))


def check_gradients(grad, rgrad, gtol, label=""):
    rel = rel_rms(grad, rgrad)
    ordinary = np.linalg.norm(rgrad, axis=1) < ORDINARY_FORCE
    rel_o = rel_rms(grad, rgrad, ordinary)
    print("%s gradients: rel RMS %.3g (all %d atoms), %.3g (%d ordinary atoms)"
          % (label, rel, rgrad.shape[0], rel_o, int(ordinary.sum())), flush=True)
    assert rel <= gtol, (label, rel)
    assert rel_o <= gtol, (label, "ordinary atoms", rel_o)


def check(terms, grad, rterms, rgrad, etol, gtol, label="", scales=None):
    check_terms(terms, rterms, etol, label, scales)
    check_gradients(grad, rgrad, gtol, label)
