"""The similarity gauge that the mixed form's CG deflation uses (k_cg.hip,
k_cg_gauge): a rotation Q = exp([th]x), translation tau or scaling s of the
whole scene changes no projection when camera (a, t) becomes R(a') = R(a) Q^T,
t' = s t - R(a') tau, so the directions

    angle-axis rows  -J_r(a)^-1   (rotation columns)
    translation rows -R(a)        (translation columns), t (scale column)
    focal / distortion rows 0

are null vectors of the undamped reduced camera system S (the camera part of
a null vector of J).  Pinned here against the oracle's S (the restated
src/lm.cu:515-594 at lambda = 0) and against finite differences of the
transformed cameras; the kernel evaluates the same closed forms."""
import numpy as np
import pytest


def skew(a):
    return np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])


def rot(a):  # Rodrigues, the Snavely model's angle-axis convention
    th = np.linalg.norm(a)
    K = skew(a)
    if th < 1e-12:
        return np.eye(3) + K
    return np.eye(3) + np.sin(th) / th * K + (1 - np.cos(th)) / th**2 * K @ K


def log_rot(R):
    th = np.arccos(np.clip((np.trace(R) - 1) / 2, -1, 1))
    w = np.array([R[2, 1] - R[1, 2], R[0, 2] - R[2, 0], R[1, 0] - R[0, 1]])
    return w / 2 if th < 1e-12 else th / (2 * np.sin(th)) * w


def jr_inv(a):
    th2 = a @ a
    K = skew(a)
    if th2 < 1e-4:
        c = 1 / 12 + th2 / 720 + th2 * th2 / 30240
    else:
        th = np.sqrt(th2)
        c = 1 / th2 - (1 + np.cos(th)) / (2 * th * np.sin(th))
    return np.eye(3) + 0.5 * K + c * K @ K


def gauge(cams):
    C = cams.shape[0]
    Z = np.zeros((9 * C, 7))
    for c in range(C):
        a, t = cams[c, :3], cams[c, 3:6]
        Z[9 * c:9 * c + 3, 0:3] = -jr_inv(a)
        Z[9 * c + 3:9 * c + 6, 3:6] = -rot(a)
        Z[9 * c + 3:9 * c + 6, 6] = t
    return Z


def test_gauge_matches_transformed_cameras():
    rng = np.random.default_rng(3)
    cams = np.zeros((40, 9))
    cams[:, :3] = rng.normal(0, 0.8, (40, 3))
    cams[:, 3:6] = rng.normal(0, 1, (40, 3))
    cams[0, :3] = 1e-7  # the small-angle series
    Z = gauge(cams)
    eps = 1e-6
    for g in range(7):
        moved = []
        for e in (eps, -eps):
            th, tau, s = np.zeros(3), np.zeros(3), 1.0
            if g < 3:
                th[g] = e
            elif g < 6:
                tau[g - 3] = e
            else:
                s += e
            out = cams.copy()
            for c in range(cams.shape[0]):
                Rn = rot(cams[c, :3]) @ rot(th).T
                out[c, :3] = log_rot(Rn)
                out[c, 3:6] = s * cams[c, 3:6] - Rn @ tau
            moved.append(out)
        fd = ((moved[0] - moved[1]) / (2 * eps)).reshape(-1)
        assert np.abs(fd - Z[:, g]).max() < 1e-8


@pytest.mark.parametrize("cfg", [1, 2])
def test_gauge_is_null_space_of_oracle_S(gba, orc, cfg):
    prob = gba.generate(cfg, 1)
    lin = orc.linearize(prob)
    o = orc.schur_step(prob, lin, 0.0, want_S=True)
    S = np.tril(o["S"]) + np.tril(o["S"], -1).T
    Z = gauge(prob.cameras)
    Z /= np.linalg.norm(Z, axis=0)
    rel = np.linalg.norm(S @ Z, axis=0) / np.linalg.norm(S, 2)
    # the gauge is exactly null; a random unit direction is not (about 1e-3)
    v = np.random.default_rng(0).normal(size=S.shape[0])
    v /= np.linalg.norm(v)
    print(f"cfg {cfg}: |S z| / |S| per gauge direction {np.array2string(rel, precision=1)}; random "
          f"{np.linalg.norm(S @ v) / np.linalg.norm(S, 2):.1e}")
    assert rel.max() < 1e-9
    assert np.linalg.norm(S @ v) / np.linalg.norm(S, 2) > 1e3 * rel.max()
    # planted errors are not null: a sign flipped in one row of the rotation
    # block, J_r^-1 replaced by its transpose (J_l^-1), the scale column taken
    # from the rotation instead of the translation
    def rel_of(z):
        return np.linalg.norm(S @ (z / np.linalg.norm(z))) / np.linalg.norm(S, 2)

    bad = Z[:, 0].copy()
    bad[0::9] *= -1
    assert rel_of(bad) > 1e3 * rel.max()
    left = gauge(prob.cameras)
    for c in range(prob.num_cameras):
        left[9 * c:9 * c + 3, 0:3] = left[9 * c:9 * c + 3, 0:3].T
    assert max(rel_of(left[:, g]) for g in range(3)) > 1e3 * rel.max()
    scale = np.zeros(S.shape[0])
    for c in range(prob.num_cameras):
        scale[9 * c:9 * c + 3] = prob.cameras[c, :3]
    assert rel_of(scale) > 1e3 * rel.max()
