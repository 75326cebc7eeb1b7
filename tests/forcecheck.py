"""Force error measures used by the GPU parity tests.

The bar is the FP32-mode one: max relative force error <= 1e-5 against the
FP64 oracle, per atom, as |dF_i| / max(|F_i|, F_rms).  The floor F_rms keeps
atoms whose net force nearly cancels (|F_i| -> 0) from dividing by zero;
every other atom is measured against its own force.  The rms error and the
max error in units of F_rms are printed beside it.
"""
import numpy as np

REL_MAX = 1e-5   # max_i |dF_i| / max(|F_i|, F_rms)
RMS_MAX = 5e-6   # rms |dF| / F_rms


def force_errors(f, ref, band=None):
    """(rms |dF| / F_rms, max |dF| / F_rms, max |dF_i| / max(|F_i|, F_rms)),
    over the atoms outside `band` (atoms of a pair whose r^2 lies within
    rounding of a cutoff^2: their inclusion depends on rounding order)."""
    ref = np.asarray(ref, np.float64)
    d = np.linalg.norm(np.asarray(f, np.float64) - ref, axis=1)
    fi = np.linalg.norm(ref, axis=1)
    frms = np.sqrt(np.mean(fi ** 2))
    ok = np.ones(len(d), bool) if band is None else ~band
    rel = d[ok] / np.maximum(fi[ok], frms)
    return (float(np.sqrt(np.mean(d[ok] ** 2)) / frms), float(d[ok].max() / frms),
            float(rel.max()))


def check_forces(f, ref, band=None, tag="", say=print, rel_max=REL_MAX, rms_max=RMS_MAX):
    rms, mx, rel = force_errors(f, ref, band)
    say(f"{tag} force error: rms {rms:.3e}, max {mx:.3e} (of F_rms), "
        f"max per-atom relative {rel:.3e}")
    assert rms <= rms_max, f"{tag}: rms force error {rms:.3e} > {rms_max:g}"
    assert rel <= rel_max, f"{tag}: max relative force error {rel:.3e} > {rel_max:g}"
    return rms, mx, rel
