"""Writes spme_ewald_fixtures.json: SPME reciprocal-energy cases with the
explicit Ewald sum as expected value.

Inputs follow the reference's own fixtures (tests/spme_pipeline.rs:102-145:
1 nm box, alpha = 4e9 1/m, 16^3 grid, order 4, k_max = 8, tolerance 5e-3);
"four_charges_fine" is the same four charges on a 64^3 grid with k_max =
16, where SPME must agree far more tightly.  Run from the repo root:
    python tests/golden/make_spme_fixtures.py
"""
import json
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))
from heddlemd_amd import units as U  # noqa: E402
from tests.test_oracle_spme import explicit_ewald  # noqa: E402

NM = 1e-9


def case(pos_nm, q, grid, kmax, tol):
    L = U.length(1.0 * NM)
    pos = np.asarray(pos_nm, float) * U.length(NM)
    p32 = pos.astype(np.float32).astype(np.float64)
    q32 = np.asarray(q, np.float32).astype(np.float64)
    alpha = 4.0e9 * U.LENGTH_BOHR_TO_M
    e = explicit_ewald(p32, q32, [L, L, L], alpha, kmax)
    return dict(pos_bohr=pos.tolist(), q=list(q), box_bohr=L, alpha=alpha, grid=grid,
                kmax=kmax, e_recip_explicit=e, rel_tol=tol)


two = case([[0.2, 0, 0], [-0.2, 0, 0]], [1.0, -1.0], [16, 16, 16], 8, 5e-3)
four_pos = [[0.1, 0, 0], [-0.1, 0.2, 0], [0, -0.2, 0.15], [-0.15, 0.1, -0.1]]
four = case(four_pos, [1.0, -1.0, 2.0, -2.0], [16, 16, 16], 8, 5e-3)
fine = case(four_pos, [1.0, -1.0, 2.0, -2.0], [64, 64, 64], 16, 3e-6)
out = dict(two_charges=two, four_charges=four, four_charges_fine=fine)
with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "spme_ewald_fixtures.json"), "w") as f:
    json.dump(out, f, indent=1)
print({k: v["e_recip_explicit"] for k, v in out.items()})
