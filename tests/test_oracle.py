"""Pins of the oracle (the CPU restatement of the reference kernels).

Each check is chosen so a plausible mistake in the restatement (a dropped
term, a wrong sign or index, a transposed operand) fails: discrete closed
forms, conservation laws, finite-difference gradients and first-order volume
changes.  None of them re-types the kernel's own expression.
"""
import math

import numpy as np
import pytest

F32, F64 = np.float32, np.float64


def _cnt(n):
    return np.zeros(n, np.int32)


@pytest.mark.parametrize("dtype", [F32, F64])
def test_free_fall_closed_form(orc, dtype):
    """gravity + PBDStep: x_n = x0 + n dt v0 + g w dt^2 (1-d) n(n+1)/2 (integrator.py:23-63)."""
    x = np.array([[0.1, 2.0, -0.3]], dtype)
    v = np.array([[0.5, 0.0, 0.25]], dtype)
    w = np.array([1.0], dtype)
    p = np.zeros_like(x)
    g = np.array([0.0, -9.81, 0.0], dtype)
    dt, damp, n = 1.0 / 600.0, 0.1, 50
    for _ in range(n):
        orc.call("oracle_gravity", dtype, 1, x, p, v, w, g, damp, dt)
        orc.call("oracle_pbd_step", dtype, 1, x, p, v, dt)
    expect = (np.array([0.1, 2.0, -0.3]) + n * dt * np.array([0.5, 0.0, 0.25])
              + np.array([0.0, -9.81, 0.0]) * dt * dt * (1 - damp) * n * (n + 1) / 2)
    # fp32: v = (p - x)/dt cancels ~log2(|x|/|v dt|) bits each substep, so the
    # fp32 trajectory drifts by ~1e-5 over 50 substeps; fp64 stays at 1e-12.
    tol = 1e-12 if dtype == F64 else 1e-4
    assert np.allclose(x[0], expect, atol=tol, rtol=0)


@pytest.mark.parametrize("dtype", [F32, F64])
def test_gravity_scales_with_inverse_mass_and_pins(orc, dtype):
    """Q1: the gravity increment is multiplied by the inverse mass; w = 0 pins."""
    x = np.zeros((3, 3), dtype)
    v = np.zeros((3, 3), dtype)
    v[0] = [1.0, 1.0, 1.0]
    w = np.array([0.0, 1.0, 2.0], dtype)
    p = np.zeros_like(x)
    g = np.array([0.0, -10.0, 0.0], dtype)
    orc.call("oracle_gravity", dtype, 3, x, p, v, w, g, 0.0, 0.1)
    assert np.array_equal(p[0], [0, 0, 0])            # pinned: velocity zeroed
    assert np.isclose(p[1, 1], -0.1, atol=1e-6)       # -10 * 1 * 0.1 * 0.1
    assert np.isclose(p[2, 1], -0.2, atol=1e-6)       # twice for w = 2


@pytest.mark.parametrize("dtype", [F32, F64])
def test_distance_pbd_two_particles(orc, dtype):
    """One PBD step scales the length error by (1 - kS) (integrator.py:65-109)."""
    rng = np.random.default_rng(0)
    for kS in (1.0, 0.5, 0.2):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        L, r = 1.7, 1.2
        p = np.stack([np.zeros(3), L * d]).astype(dtype) + np.array([0.3, -0.2, 0.1], dtype)
        dP, cnt = np.zeros_like(p), _cnt(2)
        edge = np.array([0, 1], np.int32)
        orc.call("oracle_distance_pbd", dtype, 1, p, dP, cnt, edge, np.array([r], dtype), np.ones(2, dtype), None, kS)
        assert list(cnt) == [1, 1]
        assert np.allclose(dP.sum(axis=0), 0.0, atol=1e-6)        # momentum (equal masses)
        orc.call("oracle_apply", dtype, 2, p, dP, cnt)
        newL = np.linalg.norm(p[1].astype(np.float64) - p[0])
        assert np.isclose(newL - r, (1 - kS) * (L - r), atol=1e-6)
        assert list(cnt) == [0, 0] and not dP.any()                # apply clears


@pytest.mark.parametrize("dtype", [F32, F64])
def test_distance_pbd_mass_weighting_and_pin(orc, dtype):
    """The correction splits by inverse mass; a pinned end (w=0) does not move."""
    p = np.array([[0, 0, 0], [2.0, 0, 0]], dtype)
    dP, cnt = np.zeros_like(p), _cnt(2)
    orc.call("oracle_distance_pbd", dtype, 1, p, dP, cnt, np.array([0, 1], np.int32), np.array([1.0], dtype),
             np.array([0.0, 1.0], dtype), None, 1.0)
    orc.call("oracle_apply", dtype, 2, p, dP, cnt)
    assert np.allclose(p, [[0, 0, 0], [1.0, 0, 0]], atol=1e-6)


@pytest.mark.parametrize("dtype", [F32, F64])
def test_distance_chain_count_averaging(orc, dtype):
    """Jacobi averaging on a path a-b-c: the middle vertex divides by 2 (Q3)."""
    L1, L2, r = 1.3, 0.8, 1.0
    p = np.array([[0, 0, 0], [L1, 0, 0], [L1 + L2, 0, 0]], dtype)
    dP, cnt = np.zeros_like(p), _cnt(3)
    edge = np.array([0, 1, 1, 2], np.int32)
    orc.call("oracle_distance_pbd", dtype, 2, p, dP, cnt, edge, np.array([r, r], dtype), np.ones(3, dtype), None, 1.0)
    assert list(cnt) == [1, 2, 1]
    orc.call("oracle_apply", dtype, 3, p, dP, cnt)
    c1, c2 = L1 - r, L2 - r
    expect = [c1 / 2, L1 + (-c1 / 2 + c2 / 2) / 2, L1 + L2 - c2 / 2]
    assert np.allclose(p[:, 0], expect, atol=1e-6)


@pytest.mark.parametrize("dtype", [F32, F64])
def test_distance_xpbd_compliance(orc, dtype):
    """XPBD: dlambda = -(C + a~ lam)/(w + a~), lambda accumulates (integrator.py:191-240)."""
    for alpha in (0.0, 0.36, 5.0):
        L, r = 1.5, 1.0
        p = np.array([[0, 0, 0], [0, L, 0]], dtype)
        dP, cnt, lam = np.zeros_like(p), _cnt(2), np.zeros(1, dtype)
        orc.call("oracle_distance_xpbd", dtype, 1, p, dP, lam, cnt, np.array([0, 1], np.int32),
                 np.array([r], dtype), np.ones(2, dtype), None, alpha)
        C = L - r
        assert np.isclose(lam[0], -C / (2 + alpha), atol=1e-6)
        orc.call("oracle_apply", dtype, 2, p, dP, cnt)
        assert np.isclose(p[1, 1] - p[0, 1], L - 2 * C / (2 + alpha), atol=1e-6)
        # second call with the stored lambda: -(C' + a~ lam)/(2 + a~)
        lam0 = float(lam[0])
        Lp = float(p[1, 1] - p[0, 1])
        orc.call("oracle_distance_xpbd", dtype, 1, p, dP, lam, cnt, np.array([0, 1], np.int32),
                 np.array([r], dtype), np.ones(2, dtype), None, alpha)
        assert np.isclose(lam[0] - lam0, -((Lp - r) + alpha * lam0) / (2 + alpha), atol=1e-6)


def _tet(orient=1, scale=1.0):
    P = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float64) * scale
    if orient < 0:
        P = P[[0, 2, 1, 3]]
    return P


def _signed_vol(P):
    return np.dot(np.cross(P[1] - P[0], P[2] - P[0]), P[3] - P[0]) / 6.0


@pytest.mark.parametrize("dtype", [F32, F64])
@pytest.mark.parametrize("orient", [1, -1])
def test_volume_pbd_first_order(orc, dtype, orient):
    """Q4: PBD volume restores |V| at stiffness kS/6 for negative orientation and
    pushes away from V0 for positive orientation (integrator.py:444-504)."""
    P0 = _tet(orient)
    V0 = abs(_signed_vol(P0))
    s = 1.0 - 1e-3
    P = (P0 * s).astype(dtype)                  # slightly compressed
    Vc = abs(_signed_vol(P.astype(np.float64)))
    kS = 0.8
    dP, cnt = np.zeros_like(P), _cnt(4)
    tet = np.arange(4, dtype=np.int32)
    orc.call("oracle_volume_pbd", dtype, 1, P, dP, cnt, tet, np.array([V0], dtype), np.ones(4, dtype), None, kS)
    assert np.allclose(dP.sum(axis=0), 0.0, atol=1e-9)          # translation invariance
    orc.call("oracle_apply", dtype, 4, P, dP, cnt)
    dV = abs(_signed_vol(P.astype(np.float64))) - Vc
    expect = kS * (V0 - Vc) / 6.0 * (1 if orient < 0 else -1)
    assert np.isclose(dV, expect, rtol=2e-2 if dtype == F32 else 1e-3)


@pytest.mark.parametrize("dtype", [F32, F64])
def test_volume_xpbd_first_order(orc, dtype):
    """XPBD volume (alpha 0): signed V moves by -(V - V0)/6 to first order."""
    P0 = _tet(1)
    V0 = _signed_vol(P0)
    P = (P0 * (1 + 2e-3)).astype(dtype)
    Vc = _signed_vol(P.astype(np.float64))
    dP, cnt, lam = np.zeros_like(P), _cnt(4), np.zeros(1, dtype)
    orc.call("oracle_volume_xpbd", dtype, 1, P, dP, lam, cnt, np.arange(4, dtype=np.int32), np.array([V0], dtype),
             np.ones(4, dtype), None, 0.0)
    orc.call("oracle_apply", dtype, 4, P, dP, cnt)
    dV = _signed_vol(P.astype(np.float64)) - Vc
    assert np.isclose(dV, -(Vc - V0) / 6.0, rtol=2e-2 if dtype == F32 else 1e-3)


def _nh_terms(orc, P, M, volC=2e-5, devC=1e-4, dtype=F64):
    out = np.zeros(26, dtype)
    orc.call("oracle_nh_terms", dtype, np.ascontiguousarray(P, dtype), np.ascontiguousarray(M, dtype), volC, devC, out)
    return out[0], out[1:13].reshape(4, 3), out[13], out[14:26].reshape(4, 3)


def _F(P, M):
    Ds = np.stack([P[1] - P[0], P[2] - P[0], P[3] - P[0]], axis=1)
    return Ds @ M


def _fd_grad(f, P, h=1e-6):
    g = np.zeros_like(P)
    for i in range(4):
        for k in range(3):
            Pp, Pm = P.copy(), P.copy()
            Pp[i, k] += h
            Pm[i, k] -= h
            g[i, k] = (f(Pp) - f(Pm)) / (2 * h)
    return g


def _rand_tet(seed):
    rng = np.random.default_rng(seed)
    R = _tet(1) + rng.uniform(-0.1, 0.1, (4, 3))
    Dm = np.stack([R[1] - R[0], R[2] - R[0], R[3] - R[0]], axis=1)
    M = np.linalg.inv(Dm)
    P = R + rng.uniform(-0.15, 0.15, (4, 3))
    return R, M, P


def test_neohookean_values_at_rest(orc):
    """F = I at rest: C_H = -mu/lambda (= -devC^-1 / volC^-1), C_D = sqrt(3)."""
    R, M, _ = _rand_tet(3)
    CH, gH, CD, gD = _nh_terms(orc, R, M, volC=2e-5, devC=1e-4)
    assert np.isclose(CH, -0.2, atol=1e-9)
    assert np.isclose(CD, math.sqrt(3.0), atol=1e-9)
    # gradient rows at rest: rows of M (cof(I) = I), deviatoric / sqrt(3)
    assert np.allclose(gH[1:], M, atol=1e-9)
    assert np.allclose(gD[1:], M / math.sqrt(3.0), atol=1e-9)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_neohookean_deviatoric_gradient_fd(orc, seed):
    """The deviatoric rows are the true gradient of |F|_F (finite differences)."""
    _, M, P = _rand_tet(seed)
    _, _, CD, gD = _nh_terms(orc, P, M)
    assert np.isclose(CD, np.linalg.norm(_F(P, M)), atol=1e-10)
    fd = _fd_grad(lambda Q: np.linalg.norm(_F(Q, M)), P)
    assert np.allclose(gD, fd, atol=1e-6)


def test_neohookean_hydrostatic_gradient_fd_symmetric_F(orc):
    """For a symmetric F (axis stretch of an axis-aligned rest tet) the
    hydrostatic rows equal the finite-difference gradient of det F."""
    R = _tet(1)
    M = np.linalg.inv(np.stack([R[1] - R[0], R[2] - R[0], R[3] - R[0]], axis=1))
    P = R * np.array([1.1, 0.9, 1.05])
    CH, gH, _, _ = _nh_terms(orc, P, M, volC=1.0, devC=1.0)
    assert np.isclose(CH, np.linalg.det(_F(P, M)) - 2.0, atol=1e-10)
    fd = _fd_grad(lambda Q: np.linalg.det(_F(Q, M)), P)
    assert np.allclose(gH, fd, atol=1e-6)


def test_neohookean_hydrostatic_convention_general_F(orc):
    """Q7: for a general F the reference's hydrostatic rows are those of
    M * cof(F) (its extra transpose), i.e. the FD gradient of det F taken with
    the rows of cof(F) in place of its columns.  Documents the convention."""
    _, M, P = _rand_tet(5)
    F = _F(P, M)
    cof = np.linalg.det(F) * np.linalg.inv(F).T
    _, gH, _, _ = _nh_terms(orc, P, M)
    true_grad = _fd_grad(lambda Q: np.linalg.det(_F(Q, M)), P)
    assert np.allclose(true_grad[1:], (cof @ M.T).T, atol=1e-6)     # true: columns of cof(F) M^T
    assert np.allclose(gH[1:], M @ cof, atol=1e-9)                   # reference: rows of M cof(F)
    assert not np.allclose(gH[1:], true_grad[1:], atol=1e-3)


@pytest.mark.parametrize("dtype", [F32, F64])
def test_neohookean_step_at_rest(orc, dtype):
    """Full kernel at rest (F = I): dP_i = m_i g_i (dl_H + dl_D / sqrt 3) with
    dl_H = gamma / (w + aH/V0), dl_D = -sqrt3 / (w/3 + aD/V0), w = sum m|M_i|^2."""
    R, M, _ = _rand_tet(7)
    V0 = abs(_signed_vol(R))
    volC, devC, dT = 2e-5, 1e-4, 1.0 / 600.0
    P = R.astype(dtype)
    dP, cnt = np.zeros_like(P), _cnt(4)
    orc.call("oracle_neohookean", dtype, 1, P, dP, cnt, np.arange(4, dtype=np.int32), np.array([V0], dtype),
             np.ones(4, dtype), np.ascontiguousarray(M.reshape(-1), dtype), volC, devC, dT)
    g = np.vstack([-M.sum(axis=0), M])
    w = (g * g).sum()
    dlH = 0.2 / (w + volC / dT / dT / V0)
    dlD = -math.sqrt(3.0) / (w / 3.0 + devC / dT / dT / V0)
    expect = g * (dlH + dlD / math.sqrt(3.0))
    assert list(cnt) == [1, 1, 1, 1]
    assert np.allclose(dP, expect, rtol=1e-4 if dtype == F32 else 1e-10, atol=1e-12)


@pytest.mark.parametrize("dtype", [F32, F64])
def test_neohookean_rotation_invariance_of_values(orc, dtype):
    """det F and |F|_F do not change when the deformed tet is rotated."""
    _, M, P = _rand_tet(11)
    th = 0.7
    Rz = np.array([[math.cos(th), -math.sin(th), 0], [math.sin(th), math.cos(th), 0], [0, 0, 1]])
    a = _nh_terms(orc, P, M, dtype=dtype)
    b = _nh_terms(orc, P @ Rz.T, M, dtype=dtype)
    tol = 1e-5 if dtype == F32 else 1e-12
    assert np.isclose(a[0], b[0], atol=tol) and np.isclose(a[2], b[2], atol=tol)
    assert np.allclose(b[3], a[3] @ Rz.T, atol=10 * tol)     # deviatoric rows rotate with the tet


@pytest.mark.parametrize("dtype", [F32, F64])
def test_grasp_pull_midpoint(orc, dtype):
    """dragConstraintsDO with ks = 1 moves a grasped vertex halfway to tip[env*4+2] (Q9)."""
    p = np.array([[1.0, 2.0, 3.0], [5.0, 5.0, 5.0]], dtype)
    drag = np.array([1.0, 0.0], dtype)
    v2e = np.zeros(2, np.int32)
    tips = np.zeros((4, 3), dtype)
    tips[2] = [3.0, 0.0, -1.0]
    tips[3] = [100.0, 100.0, 100.0]
    orc.call("oracle_drag", dtype, 2, p, drag, v2e, tips, 1.0)
    assert np.allclose(p[0], [2.0, 1.0, 1.0], atol=1e-6)
    assert np.allclose(p[1], [5.0, 5.0, 5.0])
    drag[0] = 0.5
    orc.call("oracle_drag", dtype, 2, p, drag, v2e, tips, 1.0)  # quarter of the remaining way
    assert np.allclose(p[0], [2.25, 0.75, 0.5], atol=1e-6)


@pytest.mark.parametrize("dtype", [F32, F64])
def test_bounds_and_apply_identity(orc, dtype):
    p = np.array([[0, -0.5, 0], [1, 0.5, 1]], dtype)
    orc.call("oracle_bounds", dtype, 2, p, 0.0)
    assert p[0, 1] == 0.0 and p[1, 1] == dtype(0.5)
    dP = np.array([[3.0, 6.0, 9.0], [1.0, 1.0, 1.0]], dtype)
    cnt = np.array([3, 0], np.int32)
    orc.call("oracle_apply", dtype, 2, p, dP, cnt)
    assert np.allclose(p, [[1, 2, 3], [1, 0.5, 1]])


@pytest.mark.parametrize("dtype", [F32, F64])
def test_grasp_select_radius(orc, dtype):
    """updateClampsDragKernel: radius test around tip[env*4+3] when closed (modelDO.py:247-279)."""
    pos = np.array([[0, 0, 0], [0.1, 0, 0], [1, 0, 0]], dtype)
    v2e = np.zeros(3, np.int32)
    tips = np.zeros((4, 3), dtype)
    drag = np.full(3, 7.0, dtype)
    state = np.array([1, 0, 0], np.int32)
    update = np.array([1, 0, 0], np.int32)
    orc.call("oracle_grasp_select", dtype, 3, pos, v2e, tips, drag, 0.15, state, update)
    assert list(drag) == [1.0, 1.0, 0.0]
    state[0] = 0
    orc.call("oracle_grasp_select", dtype, 3, pos, v2e, tips, drag, 0.15, state, update)
    assert list(drag) == [0.0, 0.0, 0.0]
    update[0] = 0
    drag[:] = 5.0
    orc.call("oracle_grasp_select", dtype, 3, pos, v2e, tips, drag, 0.15, state, update)
    assert list(drag) == [5.0, 5.0, 5.0]


def test_cube_settles_f32_close_to_f64(pkg, orc):
    """Whole step (config 1 scheme, 20 frames): fp32 stays close to fp64 and the
    pinned bottom face does not move."""
    mesh = pkg.scenes.cube_grid(6)
    cfg = orc.StepConfig(substeps=10, iterations=1, fps=60, gravity=(0, -9.81, 0), damping=0.1,
                         dist_mode=2, vol_mode=2, dist_compliance=1e-6, vol_compliance=0.0)
    a = orc.OracleSim([mesh], [0], F32)
    b = orc.OracleSim([mesh], [0], F64)
    a.step(cfg, 20)
    b.step(cfg, 20)
    pinned = mesh.inv_mass == 0
    assert np.array_equal(a.x[pinned], mesh.verts[pinned])
    assert np.abs(a.x - b.x).max() < 1e-4 * mesh.bbox_diagonal()
    assert np.isfinite(b.x).all() and b.x[~pinned, 1].min() >= 0.0
