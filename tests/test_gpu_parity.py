"""HIP path vs the FP64 oracle on identical FP32 initial states (GPU only).

Bars: grid fields after P2G and after the grid update within 1e-5 normwise
relative error; particle fields after one step within 1e-5 (x, v, C, F,
J_p / J); integer work (base indices, sort) bit-exact.
"""
import numpy as np
import pytest

import oracle as O
from mpm_cuda_amd import capi
from tests.scenes import Pair, rel

pytestmark = pytest.mark.gpu

TOL_GRID = 1e-5
TOL_PART = 1e-5

MATS = [capi.JELLO, capi.SNOW, capi.WATER]
FORMS = [capi.FORM_PLAIN, capi.FORM_FAST, capi.FORM_FUSED]
MAT_ID = {capi.JELLO: "jello", capi.SNOW: "snow", capi.WATER: "water"}
FORM_ID = {capi.FORM_PLAIN: "plain", capi.FORM_FAST: "fast", capi.FORM_FUSED: "fused"}


def _ids(p):
    return f"{MAT_ID[p[0]]}-{FORM_ID[p[1]]}"


CASES = [(m, f) for m in MATS for f in FORMS]


def test_base_index_bitexact(gpu):
    """Device base-node index == the oracle's FP32 formula, bit for bit,
    including values a few ulps around every cell threshold."""
    rng = np.random.default_rng(7)
    n_grid = 192
    inv_h = np.float32(n_grid)
    x = rng.random(200000).astype(np.float32)
    thr = ((np.arange(1, n_grid - 1) + 0.5) / n_grid).astype(np.float32)
    near = [thr]
    for k in range(1, 6):
        near.append(np.nextafter(near[-1], np.float32(2)))
    down = [thr]
    for k in range(1, 6):
        down.append(np.nextafter(down[-1], np.float32(-1)))
    x = np.concatenate([x, *near, *down, np.float32([0.0, 1e-8, 0.5 / n_grid])])
    got = capi.probe_base_index(x, 0.0, inv_h)
    want = np.array([O.base_f32(v, 0.0, inv_h) for v in x], np.int32)
    assert np.array_equal(got, want)
    # disagreement with the reference's FP64 formula is confined to thresholds
    ref = np.array([O.base_ref(float(v), 0.0, 1.0 / n_grid) for v in x], np.int32)
    diff = np.nonzero(ref != got)[0]
    assert np.all(np.abs(ref[diff] - got[diff]) == 1)
    frac_random = np.mean(ref[:200000] != got[:200000])
    assert frac_random < 1e-4, frac_random


def test_radix_sort_stable(gpu):
    rng = np.random.default_rng(3)
    for n, bits in [(1, 8), (1000, 12), (123457, 17), (200003, 18), (300000, 24), (70000, 32)]:
        keys = rng.integers(0, 2 ** min(bits, 31), size=n, dtype=np.uint64).astype(np.uint32)
        if bits == 32:
            keys = rng.integers(0, 2 ** 32, size=n, dtype=np.uint64).astype(np.uint32)
        keys[: n // 3] = keys[0]   # many duplicates: stability matters
        srt, perm = capi.probe_sort(keys, bits)
        ref = np.argsort(keys, kind="stable")
        assert np.array_equal(perm, ref.astype(np.uint32)), (n, bits)
        assert np.array_equal(srt, keys[ref])
    # only the low `bits` bits order the keys (9-bit digits: the last one is narrower)
    keys = rng.integers(0, 2 ** 20, size=50000, dtype=np.uint64).astype(np.uint32)
    srt, perm = capi.probe_sort(keys, 17)
    ref = np.argsort(keys & ((1 << 17) - 1), kind="stable")
    assert np.array_equal(perm, ref.astype(np.uint32))
    # a re-sort's keys: sorted runs of ~300 equal block keys with 3% of the
    # particles moved to a neighbouring block (runs of equal digits fill
    # whole waves)
    for n, bits in [(1_000_003, 17), (262_144, 15), (4097, 9)]:
        keys = np.sort(rng.integers(0, 1 << bits, size=n // 300 + 1, dtype=np.int64))
        keys = np.repeat(keys, 300)[:n]
        moved = rng.random(n) < 0.03
        keys[moved] += rng.choice([-1, 1, -48, 48, -2304, 2304], size=int(moved.sum()))
        keys = np.clip(keys, 0, (1 << bits) - 1).astype(np.uint32)
        srt, perm = capi.probe_sort(keys, bits)
        ref = np.argsort(keys, kind="stable")
        assert np.array_equal(perm, ref.astype(np.uint32)), (n, bits)
        assert np.array_equal(srt, keys[ref])


def test_svd_vs_lapack_fixtures(gpu):
    d = np.load("tests/golden/svd_fixtures.npz")
    F, S, R, kind = d["F"], d["S"], d["R"], d["kind"]
    U, s, V = capi.probe_svd(F.astype(np.float32), sweeps=4)
    U = U.astype(np.float64); V = V.astype(np.float64); s = s.astype(np.float64)
    for i in range(len(F)):
        if kind[i] == "zero":
            continue
        scale = abs(S[i][0])
        assert abs(np.linalg.det(U[i]) - 1) < 1e-5 and abs(np.linalg.det(V[i]) - 1) < 1e-5, (i, kind[i])
        assert np.abs(s[i] - S[i]).max() <= 2e-6 * scale + 1e-7, (i, kind[i], s[i], S[i])
        rec = U[i] @ np.diag(s[i]) @ V[i].T
        assert np.abs(rec - F[i]).max() <= 3e-6 * scale, (i, kind[i])
        cond = abs(S[i][2]) / scale
        if kind[i] not in ("repeated", "isotropic", "identity", "diagonal", "rotation") and cond > 1e-3:
            Rg = U[i] @ V[i].T
            assert np.abs(Rg - R[i]).max() < 1e-4 / cond, (i, kind[i])


def _p2g_compare(p: Pair):
    p.gpu.phase(capi.PHASE_P2G, p.dt)
    p.orc.clear_grid()
    flags = p.orc.p2g(p.dt)
    assert flags == 0
    m_g = p.gpu.grid(capi.G_MASS)
    mom_g = p.gpu.grid(capi.G_MOMENTUM)
    m_o = p.orc.grid_mass()
    mom_o = p.orc.grid_vec()
    return rel(m_g, m_o), rel(mom_g, mom_o)


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_p2g_grid_parity(gpu, case):
    mat, form = case
    p = Pair(type_=mat, form=form, seed=11)
    try:
        em, ep = _p2g_compare(p)
        print(MAT_ID[mat], FORM_ID[form], f"grid mass {em:.2e} momentum {ep:.2e}")
        assert em < TOL_GRID and ep < TOL_GRID, (em, ep)
        # mass conservation
        assert abs(p.gpu.grid(capi.G_MASS).astype(np.float64).sum() - p.orc.groups[0].mat.mass * p.n) \
            < 1e-5 * p.orc.groups[0].mat.mass * p.n
        p.gpu.phase(capi.PHASE_GRID, p.dt)
        p.orc.grid_update(p.dt)
        vg = p.gpu.grid(capi.G_VELOCITY).astype(np.float64)
        vo = p.orc.grid_vec()
        mo = p.orc.grid_mass()[..., None]
        # mass-weighted (tiny-mass nodes amplify FP32 error in v = p/m)
        ev = rel(vg * mo, vo * mo)
        print(MAT_ID[mat], FORM_ID[form], f"grid velocity (mass-weighted) {ev:.2e}")
        assert ev < TOL_GRID, ev
    finally:
        p.close()


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_one_step_particle_parity(gpu, case):
    mat, form = case
    p = Pair(type_=mat, form=form, seed=21)
    try:
        p.gpu.step(p.dt)
        assert p.orc.step(p.dt) == 0
        g, o = p.gpu_particles(), p.orc_particles()
        fields = {capi.JELLO: ("x", "v", "C", "F"), capi.SNOW: ("x", "v", "C", "F", "S"),
                  capi.WATER: ("x", "v", "C", "S")}[mat]
        errs = {}
        for k in fields:
            if k == "F":
                errs[k] = rel(g[k] - np.eye(3), o[k] - np.eye(3))
            elif k == "S":
                errs[k] = rel(g[k] - 1.0, o[k] - 1.0)
            else:
                errs[k] = rel(g[k], o[k])
        print(MAT_ID[mat], FORM_ID[form], {k: f"{v:.2e}" for k, v in errs.items()})
        for k, e in errs.items():
            assert e < (TOL_PART if k != "S" else 1e-4), (k, e)
        assert p.gpu.health() & ~capi.H_ARENA_SPILL == 0
    finally:
        p.close()


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_p2g_force_term_parity(gpu, case):
    """P2G with v = C = 0: the node momenta are the stress / pressure force
    alone (snow with J_p in [0.75, 1.05], so the hardening factor
    exp(10 (1 - J_p)) spans 0.6-12; water with J in [0.7, 1.3], pressure of
    both signs), device against the oracle.  Bar 1e-5 normwise on the
    momenta, whose FP32 error is that of the stress."""
    mat, form = case
    p = Pair(type_=mat, form=form, seed=31)
    try:
        n = p.n
        rng = np.random.default_rng(32)
        zero3 = np.zeros((n, 3), np.float32)
        S = rng.uniform(0.75, 1.05, n) if mat == capi.SNOW else rng.uniform(0.7, 1.3, n)
        S = S.astype(np.float32)
        p.gpu.write(0, capi.P_VEL, zero3)
        p.gpu.write(0, capi.P_C, np.zeros((n, 9), np.float32))
        p.gpu.write(0, capi.P_JP, S.reshape(n, 1))
        o = p.orc.groups[0]
        o.v[:] = 0.0
        o.C[:] = 0.0
        if mat == capi.SNOW:
            o.Fp[:] = np.cbrt(S.astype(np.float64))[:, None, None] * np.eye(3)
        elif mat == capi.WATER:
            o.J[:] = S.astype(np.float64)
        em, ep = _p2g_compare(p)
        print(MAT_ID[mat], FORM_ID[form], f"force-only grid mass {em:.2e} momentum {ep:.2e}")
        assert em < TOL_GRID and ep < TOL_GRID, (em, ep)
        assert np.linalg.norm(p.orc.grid_vec()) > 0
    finally:
        p.close()


ROLLOUT = [(m, f, s) for m in MATS for f in FORMS for s in ("rest", "perturbed")]


@pytest.mark.parametrize("case", ROLLOUT, ids=lambda c: f"{MAT_ID[c[0]]}-{FORM_ID[c[1]]}-{c[2]}")
def test_rollout_parity(gpu, case):
    """50 steps, FP32 device vs FP64 oracle.  'rest': uniform initial
    velocity, F = I (gravity only); 'perturbed': random v, C, F, J_p / J so
    elastic / plastic / pressure dynamics run.  Bars: x within 5e-6 (FP32
    rounding of x += dt v accumulates ~ulp/2 per step), v within 1e-4, F
    within 1e-5 of ||F||, J_p / J within 1e-5 absolute."""
    mat, form, state = case
    p = Pair(type_=mat, form=form, seed=5, randomize=(state == "perturbed"))
    try:
        for _ in range(50):
            p.gpu.step(p.dt)
            assert p.orc.step(p.dt) == 0
        g, o = p.gpu_particles(), p.orc_particles()
        ex, ev = rel(g["x"], o["x"]), rel(g["v"], o["v"])
        eF = rel(g["F"], o["F"]) if mat != capi.WATER else 0.0
        eS = float(np.abs(g["S"] - o["S"]).max()) if mat != capi.JELLO else 0.0
        print(MAT_ID[mat], FORM_ID[form], state, f"x {ex:.2e} v {ev:.2e} F {eF:.2e} S {eS:.2e}")
        assert ex < 5e-6 and ev < 1e-4 and eF < 1e-5 and eS < 1e-5
        assert p.gpu.health() & ~capi.H_ARENA_SPILL == 0
    finally:
        p.close()


@pytest.mark.parametrize("form", [capi.FORM_FAST, capi.FORM_FUSED], ids=["fast", "fused"])
def test_form_matches_plain(gpu, form):
    """The sorted forms agree with the plain form after 20 snow steps with
    re-sorting every 3 steps (particles drift between sorts)."""
    a = Pair(type_=capi.SNOW, form=capi.FORM_PLAIN, seed=9)
    b = Pair(type_=capi.SNOW, form=form, seed=9, sort_every=3)
    try:
        for _ in range(20):
            a.gpu.step(a.dt)
            b.gpu.step(b.dt)
        ga, gb = a.gpu_particles(), b.gpu_particles()
        errs = {k: rel(gb[k], ga[k]) for k in ("x", "v", "C", "F")}
        print(FORM_ID[form], "vs plain", {k: f"{v:.2e}" for k, v in errs.items()})
        for k, e in errs.items():
            assert e < 1e-5, (k, e)
        st = b.gpu.stats()
        assert st["sorts"] >= 6
    finally:
        a.close()
        b.close()


@pytest.mark.parametrize("form", [capi.FORM_FAST, capi.FORM_FUSED], ids=["fast", "fused"])
def test_two_groups_match_plain(gpu, form):
    """A jello box and a snow box moving toward each other (two particle
    groups, so the sorted forms scatter spilled particles after each group's
    P2G) agree with the plain form after 30 steps."""
    from tests.scenes import MAT_PARAMS, jittered_box
    n_grid = 64
    h = 1.0 / n_grid
    sims = []
    try:
        for f in (capi.FORM_PLAIN, form):
            cfg = capi.default_config(h=h, nx=n_grid, ny=n_grid, nz=n_grid, dt=1e-4, form=f, sort_every=24)
            s = capi.Sim(cfg)
            sims.append(s)
            for k, (type_, lo, hi, vx) in enumerate(((capi.JELLO, 0.25, 0.45, 3.0), (capi.SNOW, 0.5, 0.7, -3.0))):
                E, nu, rho = MAT_PARAMS[type_]
                x = jittered_box((lo, 0.3, 0.3), (hi, 0.5, 0.5), h, 4, 3 + k)
                v = np.tile(np.array([[vx, 0.0, 0.0]], np.float32), (x.shape[0], 1))
                s.add_particles(capi.material(E, nu, rho, type_), x, vel=v, volume=h ** 3)
        a, b = sims
        for _ in range(30):
            a.step(1e-4)
            b.step(1e-4)
        for g in (0, 1):
            for field in (capi.P_POS, capi.P_VEL, capi.P_F):
                e = rel(b.read(g, field), a.read(g, field))
                print(FORM_ID[form], "group", g, "field", field, f"{e:.2e}")
                assert e < 1e-5, (g, field, e)
    finally:
        for s in sims:
            s.close()


def test_resorts_applied_by_g2p_three_groups_warm(gpu):
    """Jello, snow and water groups on one grid with the learned warm start,
    re-sorted every second step: the FAST form's re-sorts are taken in the
    G2P phase and applied by the G2P itself (every group written in its new
    order into the other buffer, with ids, J_p / J and quaternions).  After
    16 steps every field of every group, read back in the original particle
    order, agrees with the plain form; the J_p plane of elastic snow
    particles (not rewritten by a plain G2P) moved with them."""
    from mpm_cuda_amd import warm_start
    from tests.scenes import MAT_PARAMS, jittered_box
    n_grid = 64
    h = 1.0 / n_grid
    sims = []
    try:
        for f in (capi.FORM_PLAIN, capi.FORM_FAST):
            cfg = capi.default_config(h=h, nx=n_grid, ny=n_grid, nz=n_grid, dt=1e-4, form=f, sort_every=2,
                                      warm_start=capi.WARM_MLP, warm_sweeps=2)
            s = capi.Sim(cfg)
            sims.append(s)
            s.set_warm_start_weights(warm_start.load_weights())
            boxes = ((capi.JELLO, 0.22, 0.38, 3.0), (capi.SNOW, 0.42, 0.58, -2.0), (capi.WATER, 0.62, 0.78, -3.0))
            for k, (type_, lo, hi, vx) in enumerate(boxes):
                E, nu, rho = MAT_PARAMS[type_]
                x = jittered_box((lo, 0.3, 0.3), (hi, 0.5, 0.5), h, 4, 7 + k)
                v = np.tile(np.array([[vx, 0.5, 0.0]], np.float32), (x.shape[0], 1))
                s.add_particles(capi.material(E, nu, rho, type_), x, vel=v, volume=h ** 3)
            # snow: a third of the particles compressed past the clamp band, so
            # both the SVD (warm) path and the elastic path run; J_p away from 1
            n = s.group_size(1)
            rng = np.random.default_rng(5)
            F = np.tile(np.eye(3, dtype=np.float32), (n, 1, 1))
            pick = rng.random(n) < 1 / 3
            F[pick] *= np.float32(0.96)
            s.write(1, capi.P_F, F.reshape(n, 9))
            s.write(1, capi.P_JP, (0.9 + 0.2 * rng.random((n, 1))).astype(np.float32))
        a, b = sims
        for _ in range(16):
            a.step(1e-4)
            b.step(1e-4)
        assert b.stats()["sorts"] >= 8
        for g in (0, 1, 2):
            fields = {0: (capi.P_POS, capi.P_VEL, capi.P_C, capi.P_F), 1: (capi.P_POS, capi.P_VEL, capi.P_C, capi.P_F,
                      capi.P_JP), 2: (capi.P_POS, capi.P_VEL, capi.P_C, capi.P_JP)}[g]
            for field in fields:
                ga, gb = a.read(g, field), b.read(g, field)
                if field == capi.P_C:
                    e = np.abs(gb - ga).max() * h / max(np.abs(a.read(g, capi.P_VEL)).max(), 1e-6)
                else:
                    e = rel(gb, ga)
                print("group", g, "field", field, f"{e:.2e}")
                assert e < 1e-5, (g, field, e)
        sv = b.svd_stats()
        print("fast", sv, "plain", a.svd_stats())
        assert sv["svd"] > 0 and sv["warm"] > 0
    finally:
        for s in sims:
            s.close()


def test_fused_dt_change_and_rewrite(gpu):
    """FUSED form re-primes its scatter when dt changes or the state is
    rewritten; results equal the FAST form doing the same."""
    a = Pair(type_=capi.WATER, form=capi.FORM_FAST, seed=4)
    b = Pair(type_=capi.WATER, form=capi.FORM_FUSED, seed=4)
    try:
        for k in range(6):
            dt = 1e-4 if k < 3 else 5e-5
            a.gpu.step(dt)
            b.gpu.step(dt)
            if k == 4:
                v = a.gpu.read(0, capi.P_VEL) * 0.5
                a.gpu.write(0, capi.P_VEL, v)
                b.gpu.write(0, capi.P_VEL, v)
        ga, gb = a.gpu_particles(), b.gpu_particles()
        for k in ("x", "v", "C", "S"):
            assert rel(gb[k] - (1.0 if k == "S" else 0.0), ga[k] - (1.0 if k == "S" else 0.0)) < 1e-5, k
    finally:
        a.close()
        b.close()


def test_positions_roundtrip_order(gpu):
    """Reads return particles in their original order after re-sorting."""
    p = Pair(type_=capi.JELLO, form=capi.FORM_FAST, seed=2, sort_every=1, randomize=False)
    try:
        x0 = p.gpu.read(0, capi.P_POS).copy()
        p.gpu.step(1e-9)
        x1 = p.gpu.read(0, capi.P_POS)
        assert np.abs(x1 - x0).max() < 1e-6
    finally:
        p.close()


@pytest.mark.parametrize("form", [capi.FORM_FAST, capi.FORM_FUSED], ids=["fast", "fused"])
def test_dense_blocks_several_chunks(gpu, form):
    """8 particles per cell put ~512 particles in a node block, more than one
    LDS chunk of the block kernels: grid and one-step parity still hold."""
    p = Pair(type_=capi.SNOW, form=form, seed=13, ppc=8)
    try:
        em, ep = _p2g_compare(p)
        assert em < TOL_GRID and ep < TOL_GRID, (em, ep)
        p.gpu.phase(capi.PHASE_GRID, p.dt)
        p.orc.grid_update(p.dt)
        p.gpu.phase(capi.PHASE_G2P, p.dt)
        p.orc.g2p(p.dt)
        # two full steps: the FUSED form runs its G2P2G kernel over the chunks
        for _ in range(2):
            p.gpu.step(p.dt)
            assert p.orc.step(p.dt) == 0
        g, o = p.gpu_particles(), p.orc_particles()
        for k in ("x", "v", "C"):
            assert rel(g[k], o[k]) < TOL_PART, k
        assert rel(g["F"] - np.eye(3), o["F"] - np.eye(3)) < TOL_PART
    finally:
        p.close()


@pytest.mark.parametrize("tile", [256, 384, 512])
def test_fast_tile_shapes(gpu, tile):
    """Every FAST tile shape (particles per pass over a tile; 8 particles per
    cell put ~512 in a block, so 256 and 384 take several passes): grid parity
    after P2G and particle parity after three steps against the oracle."""
    p = Pair(type_=capi.SNOW, form=capi.FORM_FAST, seed=19, ppc=8, tile=tile)
    try:
        em, ep = _p2g_compare(p)
        assert em < TOL_GRID and ep < TOL_GRID, (em, ep)
        p.gpu.phase(capi.PHASE_GRID, p.dt)
        p.orc.grid_update(p.dt)
        p.gpu.phase(capi.PHASE_G2P, p.dt)
        p.orc.g2p(p.dt)
        for _ in range(2):
            p.gpu.step(p.dt)
            assert p.orc.step(p.dt) == 0
        g, o = p.gpu_particles(), p.orc_particles()
        for k in ("x", "v", "C"):
            assert rel(g[k], o[k]) < TOL_PART, k
        assert rel(g["F"] - np.eye(3), o["F"] - np.eye(3)) < TOL_PART
    finally:
        p.close()


def test_resort_forced_by_spills(gpu):
    """With the interval cap out of reach, particles moving 0.13 cells per
    step leave their block arenas within ~10 steps; the spill count forces
    re-sorts, and the result equals the plain form's."""
    kw = dict(type_=capi.JELLO, seed=17, randomize=False, dt=1e-4, gravity=(0.0, 0.0, 0.0))
    a = Pair(form=capi.FORM_PLAIN, **kw)
    b = Pair(form=capi.FORM_FAST, sort_every=1000, **kw)
    try:
        v = np.tile(np.array([[20.0, 5.0, -3.0]], np.float32), (a.n, 1))
        a.gpu.write(0, capi.P_VEL, v)
        b.gpu.write(0, capi.P_VEL, v)
        for _ in range(40):
            a.gpu.step(a.dt)
            b.gpu.step(b.dt)
        st = b.gpu.stats()
        assert st["sorts"] >= 3, st
        ga, gb = a.gpu_particles(), b.gpu_particles()
        for k in ("x", "v", "F"):
            assert rel(gb[k], ga[k]) < 1e-5, k
        # the velocity field is uniform, so C is rounding noise around 0:
        # compare it on the scale |v| / h of a velocity gradient
        dC = float(np.abs(gb["C"] - ga["C"]).max())
        assert dC * a.h / float(np.abs(ga["v"]).max()) < 1e-5, dC
    finally:
        a.close()
        b.close()


def test_grid_after_p2g_phase_holds_spills(gpu):
    """Stepping by phases while particles leave their block arenas (0.13 cells
    per step, re-sort cap out of reach): the grid read right after each P2G
    phase holds the spilled particles' share, equal to the plain form's, on
    the steps where the FAST P2G listed particles for the atomic fallback."""
    kw = dict(type_=capi.JELLO, seed=17, randomize=False, dt=1e-4, gravity=(0.0, 0.0, 0.0))
    a = Pair(form=capi.FORM_PLAIN, **kw)
    b = Pair(form=capi.FORM_FAST, sort_every=1000, **kw)
    try:
        v = np.tile(np.array([[20.0, 5.0, -3.0]], np.float32), (a.n, 1))
        a.gpu.write(0, capi.P_VEL, v)
        b.gpu.write(0, capi.P_VEL, v)
        spilled_steps = 0
        for _ in range(30):
            for p in (a, b):
                p.gpu.phase(capi.PHASE_P2G, p.dt)
            for f in (capi.G_MASS, capi.G_MOMENTUM):
                e = rel(b.gpu.grid(f), a.gpu.grid(f))
                assert e < 1e-5, (f, e)
            for p in (a, b):
                p.gpu.phase(capi.PHASE_GRID, p.dt)
                p.gpu.phase(capi.PHASE_G2P, p.dt)
            spilled_steps += b.gpu.stats()["spilled"] > 0
        print("steps with spills:", spilled_steps)
        assert spilled_steps >= 1
    finally:
        a.close()
        b.close()


def test_copy_bandwidth_probe(gpu):
    """The streaming-copy probe the benchmark reports beside the HBM peak:
    a plausible figure (below the 8 TB/s peak) and its argument checks."""
    g = capi.probe_copy_bandwidth(256 << 20, 5)
    print(f"copy {g:.1f} GB/s")
    assert 500.0 < g < 8000.0
    with pytest.raises(capi.MpmError):
        capi.probe_copy_bandwidth(1000, 5)
    with pytest.raises(capi.MpmError):
        capi.probe_copy_bandwidth(1 << 20, 0)


def test_soa_bandwidth_probe(gpu):
    """The structure-of-arrays probe (a G2P-shaped pass: 13 planes read, 31
    written, 1M elements): a plausible figure.  Its 176 MB fit the MALL, so
    the rate may exceed the HBM peak."""
    g = capi.probe_soa_bandwidth(1_000_000, 13, 31, 5)
    print(f"soa 13->31 {g:.1f} GB/s")
    assert 500.0 < g < 40000.0
    with pytest.raises(capi.MpmError):
        capi.probe_soa_bandwidth(1000, 0, 31, 5)


def test_lds_scatter_probe(gpu):
    """The LDS scatter probe behind the P2G design: every mode runs and the
    float atomics (ds_add_f32) are slower than a plain float4
    read-modify-write of the same stencil; its argument checks."""
    ms = {m: min(capi.probe_lds_scatter(m, 64) for _ in range(2)) for m in capi.LDS_PROBE_MODES}
    per = {m: t * 1e-3 * 2.4e9 * 256 / (8192 * 64 * 64) for m, t in ms.items()}
    for m, t in ms.items():
        print(f"mode {m} ({capi.LDS_PROBE_MODES[m]}): {t:.3f} ms, {per[m]:.1f} CU-clk/particle at 2.4 GHz")
    assert all(t > 0.0 for t in ms.values())
    assert ms[0] > ms[2]
    for bad in ((4, 64), (-1, 64), (0, 0), (0, 5000)):
        with pytest.raises(capi.MpmError):
            capi.probe_lds_scatter(*bad)


@pytest.mark.parametrize("type_", [capi.JELLO, capi.SNOW], ids=["jello", "snow"])
def test_stress_paths_against_oracle(gpu, type_):
    """The device stress on both of its paths -- the polar rotation by Newton
    steps for F within a factor 2 of a rotation, the Jacobi SVD outside --
    against the oracle's FP64 fixed-corotated stress (MPMCUDA.cu:229-264)
    with the snow hardening exp(xi (1 - J_p)).  Singular values log-uniform in
    [0.3, 3] cover both paths; bar 2e-5 of max|tau| per matrix."""
    rng = np.random.default_rng(17)
    n = 4000
    mu, lam, xi = O.lame(1.4e5, 0.2)[0], O.lame(1.4e5, 0.2)[1], 10.0
    F = np.empty((n, 3, 3))
    for k in range(n):
        U, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        V, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        U *= np.sign(np.linalg.det(U))
        V *= np.sign(np.linalg.det(V))
        s = np.exp(rng.uniform(np.log(0.3), np.log(3.0), 3))
        F[k] = U @ np.diag(s) @ V.T
    F = F.astype(np.float32)
    Jp = rng.uniform(0.9, 1.05, n).astype(np.float32)
    got = capi.probe_stress(F, type_, mu, lam, xi, Jp if type_ == capi.SNOW else None)
    inside = np.all((np.linalg.svd(F.astype(np.float64), compute_uv=False) > 0.5)
                    & (np.linalg.svd(F.astype(np.float64), compute_uv=False) < 2.0), axis=1)
    assert 0.1 * n < inside.sum() < 0.9 * n          # both paths exercised
    worst = {True: 0.0, False: 0.0}
    for k in range(n):
        h = np.exp(xi * (1.0 - float(Jp[k]))) if type_ == capi.SNOW else 1.0
        t = O.fcr_tau(F[k].astype(np.float64), mu * h, lam * h)
        want = np.array([t[0, 0], t[1, 1], t[2, 2], t[0, 1], t[0, 2], t[1, 2]])
        e = float(np.abs(got[k] - want).max() / np.abs(want).max())
        worst[bool(inside[k])] = max(worst[bool(inside[k])], e)
    print({"polar": f"{worst[True]:.2e}", "svd": f"{worst[False]:.2e}"})
    assert worst[True] < 2e-5 and worst[False] < 2e-5, worst
