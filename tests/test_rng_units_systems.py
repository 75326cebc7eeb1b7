"""Set-up arithmetic: ChaCha known-answer vectors, Maxwell-Boltzmann
velocities (src/runner.rs:3194-3283), Philox-4x32-10 known-answer vectors
(kernels/philox.cuh, Random123), unit constants and the benchmark systems."""
import math

import numpy as np
import pytest

from heddlemd_amd import rng, systems, units as U


def test_chacha_rfc7539_block_vector():
    # RFC 7539 section 2.3.2 (20 rounds): key 00..1f, counter 1, nonce 0:0:0:9:0:0:0:4a:0:0:0:0
    key = [int.from_bytes(bytes(range(4 * i, 4 * i + 4)), "little") for i in range(8)]
    out = rng.chacha_blocks(key, [0], rounds=20, nonce_words=[1, 0x09000000, 0x4A000000, 0])
    assert out[0].tolist() == [0xE4E7F110, 0x15593BD1, 0x1FDD0F50, 0xC47120A3, 0xC7F4D1C7,
                               0x0368C033, 0x9AAA2204, 0x4E6CD4C3, 0x466482D2, 0x09AA9F07,
                               0x05D7C214, 0xA2028BD9, 0xD19C12B5, 0xB94E16DE, 0xE883D0CB,
                               0x4E3C50A2]


def test_chacha20_and_chacha8_zero_key_keystreams():
    z20 = rng.chacha_blocks([0] * 8, [0], rounds=20)[0].astype("<u4").tobytes().hex()
    assert z20.startswith("76b8e0ada0f13d90405d6ae55386bd28bdd219b8a08ded1aa836efcc8b770dc7")
    z8 = rng.chacha_blocks([0] * 8, [0], rounds=8)[0].astype("<u4").tobytes().hex()
    assert z8.startswith("3e00ef2f895f40d67f5bb8e81f09a5a12c840ec3ce9a7f3b181be188ef711a1e")


def test_chacha_stream_is_counter_ordered():
    s = rng.ChaCha8Stream(42)
    a = s.next_u64_array(20)
    b = rng.ChaCha8Stream(42).next_u64_array(40)
    assert np.array_equal(a, b[:20])
    u = rng.ChaCha8Stream(42).next_f64_array(10000)
    assert u.min() >= 0.0 and u.max() < 1.0 and abs(u.mean() - 0.5) < 0.02


def test_maxwell_boltzmann_invariants():
    s = systems.spc_water(n_mol=64, n_side=4, grid=(16, 16, 16))
    kt = U.temperature(300.0)
    v = rng.maxwell_boltzmann(s.mass, kt, seed=7, n_constraints=s.n_constraints)
    m = s.mass.astype(np.float64)
    p = (m[None, :] * v.astype(np.float64)).sum(1)
    assert np.abs(p).max() < 1e-6 * np.abs(m[None, :] * v).sum()
    ke = 0.5 * float((m * (v.astype(np.float64) ** 2).sum(0)).sum())
    ndof = 3 * s.n - s.n_constraints - 3
    assert ke == pytest.approx(0.5 * ndof * kt, rel=1e-6)
    v2 = rng.maxwell_boltzmann(s.mass, kt, seed=7, n_constraints=s.n_constraints)
    assert np.array_equal(v, v2)
    v3 = rng.maxwell_boltzmann(s.mass, kt, seed=8, n_constraints=s.n_constraints)
    assert not np.array_equal(v, v3)


def philox4x32_10(key, ctr):
    """Random123 Philox-4x32-10 (Salmon et al. SC11)."""
    M0, M1, W0, W1 = 0xD2511F53, 0xCD9E8D57, 0x9E3779B9, 0xBB67AE85
    c = list(ctr)
    k0, k1 = key
    for _ in range(10):
        p0 = c[0] * M0
        p1 = c[2] * M1
        c = [((p1 >> 32) ^ c[1] ^ k0) & 0xFFFFFFFF, p1 & 0xFFFFFFFF,
             ((p0 >> 32) ^ c[3] ^ k1) & 0xFFFFFFFF, p0 & 0xFFFFFFFF]
        k0 = (k0 + W0) & 0xFFFFFFFF
        k1 = (k1 + W1) & 0xFFFFFFFF
    return c


def test_philox_known_answer_vectors():
    # Random123 kat_vectors, philox4x32_10
    assert philox4x32_10((0, 0), (0, 0, 0, 0)) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert philox4x32_10((0xFFFFFFFF, 0xFFFFFFFF), (0xFFFFFFFF,) * 4) == \
        [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert philox4x32_10((0xA4093822, 0x299F31D0), (0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344)) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_units_roundtrip():
    assert U.length(U.LENGTH_BOHR_TO_M) == pytest.approx(1.0)
    assert U.temperature(U.TEMPERATURE_AU_TO_K) == pytest.approx(1.0)
    # E_h / a0 = force unit
    assert U.ENERGY_HARTREE_TO_J / U.LENGTH_BOHR_TO_M == pytest.approx(U.FORCE_AU_TO_N, rel=1e-9)


def test_benchmark_systems_geometry():
    a = systems.config_argon()
    assert a.n == 10000 and a.lattice[0] * U.LENGTH_BOHR_TO_M == pytest.approx(7.48e-9)
    w = systems.spc_water(n_mol=512, n_side=8, grid=(32, 32, 32))
    mass_kg = w.mass.astype(np.float64).sum() * U.MASS_ME_TO_KG
    vol_m3 = (w.lattice[0] * U.LENGTH_BOHR_TO_M) ** 3
    assert mass_kg / vol_m3 == pytest.approx(1000.0, rel=1e-6)  # 1.0 g/cm^3
    d_oh, d_hh = w.water_geom
    for g in w.water_groups[:20]:
        o, h1, h2 = w.posq[g, :3].astype(np.float64)
        assert np.linalg.norm(o - h1) == pytest.approx(d_oh, rel=1e-5)
        assert np.linalg.norm(h1 - h2) == pytest.approx(d_hh, rel=1e-5)
    L = w.lattice[0]
    assert np.all(w.posq[:, :3] >= -L / 2) and np.all(w.posq[:, :3] < L / 2)
    # erfc(alpha r_c) = 1e-5
    assert math.erfc(w.spme.alpha * w.spme.r_cut) == pytest.approx(1e-5, rel=1e-6)
