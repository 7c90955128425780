"""Pins the CPU oracle against the reference's own known-answer tests and
against what the mathematics fixes (closed forms, finite differences,
invariants). Runs on the CPU; the oracle is the checker of the GPU tests.

Reference KATs cited as E/test/unit/<file>:<line> (E/ = unidock/unidock_engine/).
"""
import math

import numpy as np
import pytest

from oracle import oracle as O

TOL = 1e-4  # the reference KATs' tolerance


# ------------------------------------------------------------ Vina terms
@pytest.mark.parametrize("d,e,f", [(-1, 0.018315, 0.146525), (0, 1.0, 0.0), (1, 0.018315, -0.146525),
                                   (1.5, 0.000123409804087, -0.00148091764904)])
def test_gaussian1_kat(d, e, f):  # E/test/unit/score/test_vina.cpp:18-39
    ee, ff = O.vina_terms(d)
    assert abs(ee[0] - e) < TOL and abs(ff[0] - f) < TOL


@pytest.mark.parametrize("d,e,f", [(0, 0.105399224562, 0.158098836843), (3, 1, 0), (6, 0.105399224562,
                                                                                        -0.158098836843)])
def test_gaussian2_kat(d, e, f):  # test_vina.cpp:41-57
    ee, ff = O.vina_terms(d)
    assert abs(ee[1] - e) < TOL and abs(ff[1] - f) < TOL


@pytest.mark.parametrize("d,e,f", [(-1, 1, -2), (0, 0, 0), (1, 0, 0)])
def test_repulsion_kat(d, e, f):  # test_vina.cpp:59-77
    ee, ff = O.vina_terms(d)
    assert abs(ee[2] - e) < TOL and abs(ff[2] - f) < TOL


@pytest.mark.parametrize("d,e,f", [(0.2, 1, 0), (1, 0.5, -1), (2, 0, 0)])
def test_hydrophobic_kat(d, e, f):  # test_vina.cpp:80-97
    ee, ff = O.vina_terms(d)
    assert abs(ee[3] - e) < TOL and abs(ff[3] - f) < TOL


@pytest.mark.parametrize("d,e,f", [(-1, 1, 0), (-0.5, 0.714285714286, -1.42857142857), (0.1, 0, 0)])
def test_hbond_kat(d, e, f):  # test_vina.cpp:99-116
    ee, ff = O.vina_terms(d)
    assert abs(ee[4] - e) < TOL and abs(ff[4] - f) < TOL


def test_eval_ef_kat():  # test_vina.cpp:119-135: C_P / C_P at d = 0
    e, f = O.eval_ef(0.0, 3, 3)
    assert abs(e - (-0.036122)) < TOL and abs(f - (-0.000815)) < TOL


def test_eval_ef_structure():
    # hydrogens score nothing; hydrophobic only for both-hydrophobic pairs; H-bond for donor/acceptor
    assert O.eval_ef(-0.5, 0, 3) == (0.0, 0.0)
    w = [-0.035579, -0.005156, 0.840245, -0.035069, -0.587439]
    for d in (-1.2, -0.3, 0.0, 0.7, 2.5):
        t, _ = O.vina_terms(d)
        base = w[0] * t[0] + w[1] * t[1] + w[2] * t[2]
        assert abs(O.eval_ef(d, 3, 3)[0] - base) < 1e-6                      # C_P/C_P
        assert abs(O.eval_ef(d, 2, 2)[0] - (base + w[3] * t[3])) < 1e-6      # C_H/C_H
        assert abs(O.eval_ef(d, 5, 10)[0] - (base + w[4] * t[4])) < 1e-6     # N_D donor / O_A acceptor
        assert abs(O.eval_ef(d, 10, 10)[0] - base) < 1e-6                    # two acceptors
        assert abs(O.eval_ef(d, 11, 11)[0] - (base + w[4] * t[4])) < 1e-6    # O_DA is both
    # derivative = finite difference of the energy (away from the kinks)
    for d in (-1.1, -0.4, 0.9, 1.2, 3.3):
        for t1, t2 in ((2, 2), (5, 10), (3, 15)):
            h = 1e-6
            fd = (O.eval_ef(d + h, t1, t2)[0] - O.eval_ef(d - h, t1, t2)[0]) / (2 * h)
            assert abs(fd - O.eval_ef(d, t1, t2)[1]) < 1e-6


# ------------------------------------------------------------------ math
def test_math_kats():  # E/test/unit/myutils/test_math.cpp
    pi = float(np.float32(3.1415927))
    assert abs(O.normalize_angle(0.0)) < TOL
    assert abs(O.normalize_angle(pi) - pi) < TOL
    assert abs(O.normalize_angle(-pi) + pi) < TOL
    assert abs(O.normalize_angle(2 * pi)) < TOL
    assert abs(O.normalize_angle(3 * pi) - pi) < TOL
    assert abs(O.normalize_angle(-3 * pi) + pi) < TOL
    assert abs(O.normalize_angle_2pi(0.3 + 13 * pi) - 3.441592) < TOL
    assert abs(O.normalize_angle_2pi(0.3 - 13 * pi) - 3.441592) < TOL
    e, f = O.gaussian(0, 1, 4)  # test_math.cpp:93-95
    assert abs(e - 0.939413071) < TOL and abs(f - 0.117426634) < TOL
    rng = [-180, -160, 20, 90, 110, 150]  # test_math.cpp:42-55
    assert abs(O.radian_in_ranges(rng, 0.54, 0.1) - 27) < TOL
    assert abs(O.radian_in_ranges(rng, 0.89, 1e-6) - 110) < TOL
    assert abs(O.radian_in_ranges(rng, 0.21, 1.0) - (-160)) < TOL


def test_clamp_by_ranges():
    r = [-1.0, -0.5, 0.2, 0.9]
    assert O.clamp_by_ranges(-0.7, r) == -0.7          # inside a range: unchanged
    assert abs(O.clamp_by_ranges(0.0, r) - 0.2) < 1e-7  # nearest boundary
    assert abs(O.clamp_by_ranges(-0.4, r) - (-0.5)) < 1e-7
    assert abs(O.clamp_by_ranges(2.0, r) - 0.9) < 1e-7


# ------------------------------------------------------------ rotations
def test_quaternion_kats():  # E/test/unit/geometry/test_quaternion.cpp
    q = O.quat_mul([0.9848078, 0, 0, 0.1736482], [0.9961947, 0, 0, 0.0871557])
    assert np.allclose(q, [0.9659258, 0, 0, 0.258819], atol=TOL)
    assert np.allclose(O.quat_rotate([0.7071068, 0, 0, 0.7071068], [2, 0, 0]), [0, 2, 0], atol=TOL)
    assert np.allclose(O.quat_rotate([0.8660254, 0, 0, 0.5], [2, 0, 0]), [1, 1.7320508, 0], atol=TOL)
    q3 = [0.51421413, 0.21784632, 0.32676949, 0.76246214]
    assert np.allclose(O.quat_rotate(q3, [1, 0, 0]), [-0.37625361, 0.92650867, -0.00385983], atol=TOL)
    assert np.allclose(O.quat_rotate(q3, [1, 2.3, -9.23]), [-8.02033915, -2.19741106, -4.72529836], atol=TOL)
    pi = float(np.float32(3.1415927))
    assert np.allclose(O.quat_to_rotvec([1, 0, 0, 0]), 0, atol=TOL)
    assert np.allclose(O.quat_to_rotvec([0, 0, 0, 1]), [0, 0, pi], atol=TOL)
    assert np.allclose(O.quat_to_rotvec([0.7071068, 0.7071068, 0, 0]), [pi / 2, 0, 0], atol=TOL)
    assert np.allclose(O.quat_to_rotvec([0.9586831, 0.2370836, 0.0848071, -0.1323853]),
                       [0.4808074, 0.1719894, -0.2684785], atol=TOL)
    assert np.allclose(O.quat_to_rotvec([0.9999999, 0.0001, 0, 0]), [0.0002, 0, 0], atol=TOL)


def test_rodrigues_kats():  # E/test/unit/geometry/test_rotation.cpp:9-49
    assert np.allclose(O.rodrigues([1, 0, 0], [0, 0, 1], math.pi / 2), [0, 1, 0], atol=TOL)
    assert np.allclose(O.rodrigues([1, 0, 0], [0, 0, 1], -7 * math.pi / 2), [0, 1, 0], atol=TOL)
    assert np.allclose(O.rodrigues([0, 1, 0], [1, 0, 0], math.pi / 4), [0, 0.7071, 0.7071], atol=TOL)
    assert np.allclose(O.rodrigues([0, 0, 0], [0, 0, 1], math.pi / 2), 0, atol=TOL)


def _rot_matrix(v):
    """R(v) from three rotated basis vectors (quaternion path of the oracle)."""
    q = O.rotvec_to_quat(v)
    return np.stack([O.quat_rotate(q, e) for e in np.eye(3)], axis=1)


@pytest.mark.parametrize("v", [[0.0, 0.0, 0.0], [-1.1, 0.2, 0.3], [0.3, -0.8, 1.9], [1e-3, 2e-3, -1e-3]])
def test_drot_dv_is_derivative_of_rotation(v):
    """dR/dv_k (rotation.h:80-144) against central differences of R(v) (|v| <= pi)."""
    v = np.asarray(v, float)
    h = 1e-6
    for k in range(3):
        dv = np.zeros(3)
        dv[k] = h
        fd = (_rot_matrix(v + dv) - _rot_matrix(v - dv)) / (2 * h)
        assert np.abs(O.drot_dv(v, k) - fd).max() < 1e-6


def test_rotvec_quaternion_round_trip():
    rng = np.random.default_rng(0)
    for _ in range(50):
        v = rng.normal(size=3)
        v *= rng.uniform(0.01, 3.0) / np.linalg.norm(v)
        assert np.allclose(O.quat_to_rotvec(O.rotvec_to_quat(v)), v, atol=1e-9)


def test_super_fibonacci_is_unit_rotation_spread():
    vs = np.asarray([O.super_fibonacci(i, 64) for i in range(64)])
    qs = np.asarray([O.rotvec_to_quat(v) for v in vs])
    assert np.allclose(np.linalg.norm(qs, axis=1), 1, atol=1e-12)
    # low discrepancy: no two samples are the same rotation
    dots = np.abs(qs @ qs.T) - np.eye(64)
    assert dots.max() < 0.995


# --------------------------------------------------------------- Philox
def test_philox_known_answers():
    """Philox4x32-10 known-answer vectors of the Random123 distribution (Salmon et al., SC'11)."""
    assert list(O.philox([0, 0, 0, 0], [0, 0])) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert list(O.philox([0xffffffff] * 4, [0xffffffff] * 2)) == [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert list(O.philox([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0])) == \
        [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def test_chain_stream_layout():
    b = O.chain_blocks(7, 3, 5, 4)
    for d in range(4):
        assert list(b[d]) == list(O.philox([5, d, 0x55443244, 0], [7, 3]))
