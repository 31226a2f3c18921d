"""Pins the oracle's Feautrier solve (orc_raypair_solve, a restatement of
RayPair::solve, src/Simulation/Raypair/raypair_solver.tpp:200-420).

* Known answer of tests/unit/Simulation/Raypair/test_raypair.cpp:26-293: with
  5 + 5 points, eta = d+1, chi = 0.3 d^2 (dZ = 0.1) or 1e9 d^2 (dZ = 1) and
  n_off_diag = 1, L_diag, L_upper and L_lower equal the diagonal and first
  off-diagonals of the inverse of the Feautrier matrix assembled from A, C and
  dtau.  (The reference compares with exact equality; Eigen's LU rounds
  differently, so a relative 1e-12 bar is used.)
* u = T^-1 S' for random inputs (the second-order system of
  docs/background.tex:303-423) for n_off_diag = 0 and > 0, every origin index.
* Constant-source slab: J = S + (B - S)/2 (e^-tau(x) + e^-tau(L-x))
  (tests/integration/test_0_1D_all_constant.py:173-250) up to the scheme's
  second-order discretisation error.
* Two-sided elimination used by the GPU kernel (u and Lambda at the origin
  from two inward eliminations) agrees with the reference elimination.
"""
import numpy as np
import pytest

from oracle.py_oracle import raypair_solve


def feautrier_matrix(A, C, dtau):
    n = len(A)
    M = np.zeros((n, n))
    for d in range(n - 1):
        M[d, d + 1] = -C[d]
    for d in range(1, n):
        M[d, d - 1] = -A[d]
    for d in range(1, n - 1):
        M[d, d] = 1.0 + A[d] + C[d]
    M[0, 0] = 1 + 2.0 / dtau[0] + 2.0 / (dtau[0] * dtau[0])
    M[n - 1, n - 1] = 1 + 2.0 / dtau[n - 2] + 2.0 / (dtau[n - 2] * dtau[n - 2])
    return M


def reference_setup(chi, dZ, n_r=5, n_ar=5):
    ndep = n_r + n_ar + 1
    eta = 1.0
    term1 = np.zeros(ndep)
    dtau = np.zeros(ndep)
    with np.errstate(divide="ignore"):
        for d in range(ndep):
            # eta/chi is infinite at d = 0 (chi = 0), as in the reference test
            term1[d] = np.float64((d + 1) * eta) / np.float64((d * d) * chi)
            dtau[d] = 0.5 * ((d * d) * chi + ((d - 1) * (d - 1)) * chi) * dZ
    return term1, dtau, n_ar


@pytest.mark.parametrize("chi,dZ", [(0.3, 0.1), (1.0e9, 1.0)])
def test_known_answer_test_raypair(chi, dZ):
    term1, dtau, n_ar = reference_setup(chi, dZ)
    with np.errstate(all="ignore"):
        res = raypair_solve(term1, dtau, 0.0, 0.0, n_ar, n_off_diag=1)
    M = feautrier_matrix(res["A"], res["C"], dtau)
    Minv = np.linalg.inv(M)
    ndep = len(term1)
    for d in range(ndep):
        assert res["L_diag"][d] == pytest.approx(Minv[d, d], rel=1e-12)
    for d in range(ndep - 1):
        assert res["L_upper"][0][d] == pytest.approx(Minv[d, d + 1], rel=1e-12)
        assert res["L_lower"][0][d] == pytest.approx(Minv[d + 1, d], rel=1e-12)


def _random_case(rng, ndep):
    dtau = 10.0 ** rng.uniform(-3, 2, size=ndep)
    term1 = rng.uniform(0.1, 2.0, size=ndep)
    return term1, dtau


@pytest.mark.parametrize("n_off_diag", [0, 2])
def test_solution_is_inverse_times_source(n_off_diag):
    rng = np.random.default_rng(1)
    for ndep in (2, 3, 7, 20):
        for n_ar in range(1, ndep):
            if ndep - 1 - n_ar > n_ar:
                continue       # the longer ray is on the ar side
            term1, dtau = _random_case(rng, ndep)
            I0, In = rng.uniform(0, 1, 2)
            res = raypair_solve(term1, dtau, I0, In, n_ar, n_off_diag=n_off_diag)
            M = feautrier_matrix(res["A"], res["C"], dtau)
            S = term1.copy()
            S[0] += 2.0 * I0 / dtau[0]
            S[-1] += 2.0 * In / dtau[ndep - 2]
            u = np.linalg.solve(M, S)
            Minv = np.linalg.inv(M)
            assert res["Su"][n_ar] == pytest.approx(u[n_ar], rel=1e-11)
            assert res["L_diag"][n_ar] == pytest.approx(Minv[n_ar, n_ar], rel=1e-11)
            if n_off_diag > 0:
                assert np.allclose(res["Su"], u, rtol=1e-11)
                assert np.allclose(res["L_diag"], np.diag(Minv), rtol=1e-11)
                for m in range(n_off_diag):
                    for d in range(ndep - m - 1):
                        assert res["L_upper"][m][d] == pytest.approx(Minv[d, d + m + 1], rel=1e-10)
                        assert res["L_lower"][m][d] == pytest.approx(Minv[d + m + 1, d], rel=1e-10)


def test_wrong_matrix_is_detected():
    """A dropped term in the assembled matrix must fail the comparison."""
    rng = np.random.default_rng(2)
    term1, dtau = _random_case(rng, 9)
    res = raypair_solve(term1, dtau, 0.3, 0.7, 4)
    M = feautrier_matrix(res["A"], res["C"], dtau)
    M[4, 4] -= res["A"][4]          # plant a mistake
    Minv = np.linalg.inv(M)
    assert abs(res["L_diag"][4] - Minv[4, 4]) > 1e-6 * abs(Minv[4, 4])


def two_sided(term1, dtau, I0, In, n_ar):
    """The two-sided elimination of the GPU kernel (see
    magritte_amd/csrc/mgr_radiation.hip), in FP64 numpy.

    Each side is eliminated from its far end towards the origin in the
    variables phi = F_n / b_n and zeta = Z_n / b_n, where b_n is the optical
    depth between point n and the next point inwards, so that no 1/b appears:
        h = (a + b)/2,  phi_n = h + phi_{n-1} q_{n-1},  q_n = 1/(1 + b phi_n),
        zeta_n = (S_n h + zeta_{n-1}) q_n,
    from the virtual start phi = q = 1, zeta = I_far, a = 0.  At the origin
        u = (S h_o + zeta_A + zeta_B) / (h_o + phi_A q_A + phi_B q_B),
        Lambda = h_o / (h_o + phi_A q_A + phi_B q_B),  h_o = (b_A + b_B)/2,
    and an empty side enters as the virtual start (phi q = 1, zeta = I, b = 0).
    """
    ndep = len(term1)

    def side(points, I_far):   # points from the far end inwards: (S, b to the next point inwards)
        phi, q, zeta, a = 1.0, 1.0, I_far, 0.0
        for S, b in points:
            h = 0.5 * (a + b)
            phi = h + phi * q
            q = 1.0 / (1.0 + b * phi)
            zeta = (S * h + zeta) * q
            a = b
        return phi * q, zeta, a

    So = term1[n_ar]
    A = side([(term1[n], dtau[n]) for n in range(0, n_ar)], I0)
    if n_ar == ndep - 1:
        B = (1.0, In, 0.0)
    else:
        B = side([(term1[n], dtau[n - 1]) for n in range(ndep - 1, n_ar, -1)], In)
    h = 0.5 * (A[2] + B[2])
    den = h + A[0] + B[0]
    return (So * h + A[1] + B[1]) / den, h / den


def test_two_sided_elimination_matches_reference():
    rng = np.random.default_rng(3)
    for ndep in (2, 3, 5, 12, 40):
        for n_ar in range((ndep - 1 + 1) // 2, ndep):
            term1, dtau = _random_case(rng, ndep)
            I0, In = rng.uniform(0, 1, 2)
            res = raypair_solve(term1, dtau, I0, In, n_ar)
            u, lam = two_sided(term1, dtau, I0, In, n_ar)
            assert u == pytest.approx(res["Su"][n_ar], rel=1e-12)
            assert lam == pytest.approx(res["L_diag"][n_ar], rel=1e-12)


def test_constant_source_slab():
    """J = S + (B-S)/2 (exp(-tau(x)) + exp(-tau(L-x))) for a slab of constant
    opacity and source function, two rays (+-x)."""
    chi, dx, n = 1.0, 0.05, 81       # total optical depth 4
    S, B = 1.0, 0.2
    x = np.arange(n) * dx
    J = np.zeros(n)
    for o in range(n):
        # ray pair through o: ar side = cells o-1..0, r side = o+1..n-1
        left = list(range(o - 1, -1, -1))
        right = list(range(o + 1, n))
        if len(right) > len(left):
            ar_cells, r_cells = right, left
        else:
            ar_cells, r_cells = left, right
        n_ar = len(ar_cells)
        ndep = n_ar + len(r_cells) + 1
        term1 = np.full(ndep, S)
        dtau = np.full(ndep, chi * dx)
        if ndep == 1:
            J[o] = B
            continue
        res = raypair_solve(term1, dtau, B, B, n_ar, first=0, last=ndep - 1)
        J[o] = res["Su"][n_ar]
    tau_l = chi * (x - x[0])
    tau_r = chi * (x[-1] - x)
    J_exact = S + 0.5 * (B - S) * (np.exp(-tau_l) + np.exp(-tau_r))
    assert np.max(np.abs(J - J_exact)) < 2e-3 * abs(B - S)
