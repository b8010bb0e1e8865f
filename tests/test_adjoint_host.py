"""CPU checks of the adjoint workloads' host side and of the oracle's tangent-linear model.

The oracle's tangent (dual-number advection, linear stages applied to the perturbation) is
pinned by central finite differences of the oracle's own forward step, which is itself pinned
by the reference's golden vectors. The synthetic-input generators are pinned by the published
splitmix64 test vector, by the discrete divergence they must annihilate, and by a committed
SHA-256 of the 128^3 perturbation."""
import numpy as np
import pytest

from oracle import ofref
from solver_amd import inputs

C4_SEED = 20261203
# SHA-256 of perturbation(128, C4_SEED, 1e-2) (little-endian float64 storage of u, v, w)
C4_PERTURBATION_SHA256 = "d53bac0ab83f372220055d630efcaba8743af3c780a1fbba957d4ac931aa1a66"


def test_splitmix64_published_vector():
    # splitmix64 with state 0: the first outputs of the reference implementation
    # (Vigna's splitmix64.c) are 0xe220a8397b1dcdaf, 0x6e789e6aa1b965f4, 0x06c45d188009454f
    z = inputs.splitmix64(0, 3)
    assert [int(x) for x in z] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    # the stream is counter based: an offset continues it
    assert int(inputs.splitmix64(0, 1, offset=2)[0]) == 0x06C45D188009454F


def test_normals_moments():
    x = inputs.normals(5, 400_000)
    assert abs(x.mean()) < 5e-3
    assert abs(x.var() - 1.0) < 1e-2
    u = inputs.uniform01(5, 1000)
    assert u.min() > 0.0 and u.max() < 1.0
    # aligned sub-ranges of the deviate stream are generated independently
    assert np.array_equal(inputs.normals(5, 40)[10:30], inputs.normals(5, 20, offset=10))
    with pytest.raises(ValueError):
        inputs.normals(5, 3)


@pytest.mark.parametrize("n", [8, 16, 24])
def test_perturbation_divergence_free(n):
    f = inputs.perturbation(n, 11, 0.3)
    d = inputs.mac_divergence(f, n)
    assert np.max(np.abs(d)) < 1e-13
    act = [f[0][1:n + 1, 1:n + 1, 1:n + 1], f[1][1:n + 1, 1:n + 1, 1:n + 1], f[2][1:n + 1, 1:n + 1, 1:n + 1]]
    rms = np.sqrt(sum(float((a * a).sum()) for a in act) / (3 * n ** 3))
    assert rms == pytest.approx(0.3, rel=1e-13)
    # periodic duplicate faces carry face 1's value; ghosts are zero
    assert np.array_equal(f[0][1:n + 1, 1:n + 1, n + 1], f[0][1:n + 1, 1:n + 1, 1])
    assert np.array_equal(f[2][n + 1, 1:n + 1, 1:n + 1], f[2][1, 1:n + 1, 1:n + 1])
    assert not f[0][0].any() and not f[1][:, :, 0].any()


def test_perturbation_spectrum_is_low_wavenumber():
    n = 32
    f = inputs.perturbation(n, 3, 1.0)
    u = f[0][1:n + 1, 1:n + 1, 1:n + 1]
    e = np.abs(np.fft.fftn(u)) ** 2
    k = np.fft.fftfreq(n, 1.0 / n)
    kz, ky, kx = np.meshgrid(k, k, k, indexing="ij")
    km = np.sqrt(kx * kx + ky * ky + kz * kz)
    # energy per mode follows amplitude^2 = |k|^-4 exp(-|k|^2 / 32): normalised by it, the mean
    # over a band of a few hundred modes is the same in every band (random phases and the
    # projection's 2/3 share of the energy in u are independent of |k|)
    amp2 = np.ones_like(km)
    amp2[km > 0] = km[km > 0] ** -4 * np.exp(-km[km > 0] ** 2 / 32.0)
    flat = e / amp2
    b1 = flat[(km >= 2) & (km <= 4)].mean()
    b2 = flat[(km >= 6) & (km <= 10)].mean()
    assert 0.8 < b1 / b2 < 1.25
    assert e[0, 0, 0] < 1e-20 * e.sum()  # no mean flow


def test_c4_perturbation_checksum():
    f = inputs.perturbation(128, C4_SEED, 1e-2)
    assert inputs.sha256_fields(f) == C4_PERTURBATION_SHA256


def test_forcing_field():
    kx, ky = inputs.forcing_wavenumbers(C4_SEED)
    assert 1 <= kx <= 4 and 1 <= ky <= 4
    n = 16
    f = inputs.forcing(n, C4_SEED, 1e-2)
    h = 2 * np.pi / n
    i, j = 3, 5  # storage indices
    assert f[0][4, j, i] == pytest.approx(1e-2 * np.sin(kx * (i - 1) * h) * np.sin(ky * (j - 0.5) * h), abs=1e-17)
    assert not f[1].any() and not f[2].any()


# ------------------------------------------------------------------ oracle tangent-linear model
def _rand_vel(g, rng, scale=1.0):
    return [rng.standard_normal(ofref.storage_shape(g.nx, g.ny, g.nz, loc)) * scale
            for loc in (ofref.FACE_X, ofref.FACE_Y, ofref.FACE_Z)]


@pytest.mark.parametrize("scheme,eps,tol", [(2, 1e-3, 1e-9), (1, 1e-6, 1e-8), (0, 1e-7, 1e-6)])
def test_oracle_advection_tangent_matches_central_differences(scheme, eps, tol):
    g = ofref.make_grid(6, 5, 4, 0.3, 0.4, 0.5)
    rng = np.random.default_rng(4 + scheme)
    v = _rand_vel(g, rng)
    d = _rand_vel(g, rng)
    tl = ofref.advection_tangent(g, v, d, scheme)
    plus = ofref.advection(g, [a + eps * b for a, b in zip(v, d)], scheme)
    minus = ofref.advection(g, [a - eps * b for a, b in zip(v, d)], scheme)
    for c in range(3):
        fd = (plus[c] - minus[c]) / (2 * eps)
        err = np.max(np.abs(fd - tl[c])) / np.max(np.abs(tl[c]))
        assert err < tol, (c, err)


def test_oracle_advection_tangent_is_linear_in_direction():
    g = ofref.make_grid(5, 5, 5, 0.3, 0.3, 0.3)
    rng = np.random.default_rng(9)
    v, d1, d2 = _rand_vel(g, rng), _rand_vel(g, rng), _rand_vel(g, rng)
    a = ofref.advection_tangent(g, v, [x + 3.0 * y for x, y in zip(d1, d2)], 0)
    b1 = ofref.advection_tangent(g, v, d1, 0)
    b2 = ofref.advection_tangent(g, v, d2, 0)
    for c in range(3):
        assert np.allclose(a[c], b1[c] + 3.0 * b2[c], rtol=1e-12, atol=1e-12)


def _tg_case(n=8, steps=3, nu=0.05):
    dt_probe = ofref.tg_config(n, n, n, viscosity=nu, cfl_limit=0.5, final_time=1.0)
    upper = ofref.tg_dt(dt_probe)
    c = ofref.tg_config(n, n, n, viscosity=nu, cfl_limit=0.5, final_time=steps * upper * 0.999,
                        max_iterations=400, tolerance=1e-13)
    s = ofref.tg_init(c, convention=1)
    pert = inputs.perturbation(n, C4_SEED, 0.05)
    vel = [a + b for a, b in zip(s.vel, pert)]
    return c, vel, s.p


@pytest.mark.parametrize("what", ["velocity", "forcing"])
def test_oracle_step_tangent_matches_central_differences(what):
    n, steps = 8, 3
    c, vel, p = _tg_case(n, steps)
    g = ofref.tg_grid(c)
    force = inputs.forcing(n, 5, 0.2)
    dvel = inputs.perturbation(n, C4_SEED + 1, 1.0)
    rng = np.random.default_rng(2)
    # a forcing direction with a gradient part as well (the projection removes that part)
    dforce = [a + 0.3 * rng.standard_normal(a.shape) for a in inputs.perturbation(n, C4_SEED + 2, 1.0)]
    args = dict(velocity=dict(dvel=dvel), forcing=dict(dforce=dforce))[what]
    _, _, ke, _, dke = ofref.tg_run_tl(c, steps, vel, p, force=force,
                                       dvel=args.get("dvel", [np.zeros_like(a) for a in vel]),
                                       dp=args.get("dp"), dforce=args.get("dforce"))
    eps = 1e-6

    def J(sign):
        v2 = [a + sign * eps * b for a, b in zip(vel, args["dvel"])] if "dvel" in args else vel
        p2 = p + sign * eps * args["dp"] if "dp" in args else p
        f2 = [a + sign * eps * b for a, b in zip(force, args["dforce"])] if "dforce" in args else force
        return ofref.tg_run_tl(c, steps, v2, p2, force=f2)[2]

    fd = (J(1) - J(-1)) / (2 * eps)
    assert abs(dke) > 1e-6
    assert abs(fd - dke) <= 1e-6 * abs(dke), (fd, dke)
    assert g.nx == n


def test_oracle_pressure_perturbation_leaves_velocity_unchanged():
    # Periodic boxes: the factored Helmholtz operator commutes with the MAC gradient, so a
    # pressure perturbation enters the predictor as a pure gradient and the projection removes
    # it; the velocity tangent is zero to rounding after one step.
    n = 8
    c, vel, p = _tg_case(n, 1)
    dp = np.random.default_rng(1).standard_normal(p.shape)
    zero = [np.zeros_like(a) for a in vel]
    _, _, _, du, dke = ofref.tg_run_tl(c, 1, vel, p, dvel=zero, dp=dp)
    _, _, _, du_ref, dke_ref = ofref.tg_run_tl(c, 1, vel, p, dvel=inputs.perturbation(n, 3, 1.0))
    assert max(np.max(np.abs(a)) for a in du) < 1e-12 * max(np.max(np.abs(a)) for a in du_ref)
    assert abs(dke) < 1e-12 * abs(dke_ref)


def test_oracle_forced_run_reduces_to_unforced_steps():
    # with no force and no perturbation the run equals the reference-restating tg_steps
    n, steps = 8, 2
    c, vel, p = _tg_case(n, steps)
    u, pp, ke, _, _ = ofref.tg_run_tl(c, steps, vel, p)
    s = ofref.State(ofref.tg_grid(c))
    for a, b in zip(s.vel, vel):
        a[...] = b
    s.p[...] = p
    # tg_steps reads the state as stored: give it the ghosts tg_run_tl fills
    ofref.velocity_bc(ofref.tg_grid(c), ofref.tg_bcs(n), s.vel)
    ofref.pressure_bc(ofref.tg_grid(c), ofref.derive_pressure_bc(ofref.tg_bcs(n)), s.p)
    ofref.tg_steps(c, s, steps)
    for a, b in zip(u, s.vel):
        assert np.array_equal(a, b)
    assert ke == ofref.kinetic_energy(ofref.tg_grid(c), s.vel)
