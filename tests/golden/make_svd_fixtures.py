"""Generate LAPACK-derived golden vectors for the 3x3 SVD / polar / FCR stress.

Run from the repository root:  python tests/golden/make_svd_fixtures.py
Writes tests/golden/svd_fixtures.npz.  Matrices are FP32-representable so
the same inputs can be fed to the FP32 device path.

Conventions of the reference's SVD (McAdams et al. 2011, used through
PolarSVD at src/simulator/MPMCUDA.cu:87-134): U, V in SO(3), singular values
sorted by magnitude, the sign of det F carried by the last one.  LAPACK's
(numpy.linalg.svd) output is brought to that convention by flipping the last
column of U and/or V.  The polar rotation is cross-checked against
scipy.linalg.polar for det F > 0.
"""
import os

import numpy as np
import scipy.linalg

HERE = os.path.dirname(os.path.abspath(__file__))
MU, LAM = 1e4 / (2 * 1.3), 1e4 * 0.3 / (1.3 * 0.4)   # E = 1e4, nu = 0.3


def rot(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def cases(rng):
    out, kind = [], []
    def add(F, k):
        out.append(np.asarray(F, np.float32).astype(np.float64)); kind.append(k)
    for _ in range(60): add(rng.normal(size=(3, 3)), "normal")
    for _ in range(60): add(np.eye(3) + 0.05 * rng.normal(size=(3, 3)), "near_identity")
    for _ in range(40): add(rot(rng) @ np.diag(rng.uniform(0.7, 1.3, 3)) @ rot(rng).T, "rotated_stretch")
    for _ in range(30):
        s = np.array([1.0, rng.uniform(0.1, 1.0), 10 ** rng.uniform(-6, -3)])
        add(rot(rng) @ np.diag(s) @ rot(rng).T, "near_singular")
    for _ in range(30):
        a = rng.uniform(0.5, 1.5); s = np.array([a, a, rng.uniform(0.5, 1.5)])
        add(rot(rng) @ np.diag(rng.permutation(s)) @ rot(rng).T, "repeated")
    for _ in range(30):
        F = rot(rng) @ np.diag(rng.uniform(0.6, 1.4, 3)) @ rot(rng).T
        F[:, 0] *= -1.0
        add(F, "inverted")
    for _ in range(20): add(rot(rng), "rotation")
    for _ in range(20): add(np.diag(rng.uniform(-1.5, 1.5, 3)), "diagonal")
    add(np.eye(3), "identity")
    add(np.diag([2.0, 2.0, 2.0]), "isotropic")
    add(np.zeros((3, 3)), "zero")
    return np.array(out), np.array(kind)


def so3_svd(F):
    U, s, Vt = np.linalg.svd(F)
    V = Vt.T.copy()
    s = s.copy()
    if np.linalg.det(U) < 0:
        U[:, 2] *= -1; s[2] *= -1
    if np.linalg.det(V) < 0:
        V[:, 2] *= -1; s[2] *= -1
    return U, s, V


def main():
    rng = np.random.default_rng(20261202)
    F, kind = cases(rng)
    n = len(F)
    U = np.zeros((n, 3, 3)); S = np.zeros((n, 3)); V = np.zeros((n, 3, 3)); R = np.zeros((n, 3, 3))
    tau = np.zeros((n, 3, 3))
    for i in range(n):
        U[i], S[i], V[i] = so3_svd(F[i])
        R[i] = U[i] @ V[i].T
        if np.linalg.det(F[i]) > 1e-9:
            Rp, _ = scipy.linalg.polar(F[i])
            assert np.allclose(Rp, R[i], atol=1e-6 / max(1e-12, abs(S[i][2]) / S[i][0])), i
        J = np.linalg.det(F[i])
        tau[i] = 2 * MU * (F[i] - R[i]) @ F[i].T + LAM * J * (J - 1) * np.eye(3)
    np.savez_compressed(os.path.join(HERE, "svd_fixtures.npz"), F=F, kind=kind, U=U, S=S, V=V, R=R, tau=tau,
                        mu=MU, lam=LAM)
    print(f"wrote {n} cases")


if __name__ == "__main__":
    main()
