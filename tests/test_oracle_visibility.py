"""Pins of the oracle's model visibilities (generate_visibilities_from_local_skymodel,
src/stp/visibility/visibility.cpp:15-41), V = sum_k flux_k n_k exp(+2 pi i (u l_k + v m_k
+ w (n_k - 1))), n = sqrt(1 - l^2 - m^2), by closed forms that a dropped term, a
wrong sign or a swapped coordinate would fail.  The reference's own fixture
(test-data/.../expected_model.npz, visibility_test_ModelGeneration.cpp) is not
available here; its tolerance, 1e-11 absolute, is the bar used below."""
import numpy as np

from tests import simdata

TOL = 1e-11  # visibility_test_ModelGeneration.cpp:15 (10e-12 absolute)


def _uvw(rows):
    return np.asfortranarray(np.array(rows, dtype=np.float64).reshape(-1, 3))


def test_phase_centre_source_is_its_flux(orc):
    """l = m = 0: n = 1, every baseline sees the flux, real."""
    rng = np.random.default_rng(3)
    uvw = np.asfortranarray(rng.uniform(-5e4, 5e4, (1000, 3)))
    v = orc.model_visibilities([[0.0, 0.0, 7.5]], uvw)
    assert np.all(v == 7.5 + 0j)


def test_closed_form_phases(orc):
    """Sources whose n is exact (l = 0.6 -> n = 0.8, m = 0.8 -> n = 0.6) on baselines
    chosen so that each term of the phase is a quarter turn: the sign of each term,
    the n factor on the flux and the w (n - 1) term."""
    f = 2.0
    # u l = 1/4 -> +i ; v m = 1/4 -> +i ; w (n - 1) = -1/4 -> -i
    cases = [
        ([0.6, 0.0, f], [5.0 / 12.0, 0.0, 0.0], 0.8 * f * 1j),
        ([0.0, 0.8, f], [0.0, 0.3125, 0.0], 0.6 * f * 1j),
        ([0.6, 0.0, f], [0.0, 0.0, 1.25], -0.8 * f * 1j),
        ([0.0, 0.8, f], [0.0, 0.0, 0.625], -0.6 * f * 1j),
        # half turns: u l = 1/2 -> -1
        ([0.6, 0.0, f], [5.0 / 6.0, 0.0, 0.0], -0.8 * f),
        # u l + w (n - 1) = 1/4 - 1/4 = 0 -> +1
        ([0.6, 0.0, f], [5.0 / 12.0, 0.0, 1.25], 0.8 * f),
    ]
    for src, bl, want in cases:
        got = orc.model_visibilities([src], _uvw([bl]))[0]
        assert abs(got - want) < TOL, (src, bl, got, want)


def test_conjugate_symmetry_and_linearity(orc):
    """V(-uvw) = conj V(uvw); the model of a union of sky models is the sum."""
    rng = np.random.default_rng(5)
    uvw = np.asfortranarray(rng.uniform(-2e4, 2e4, (3000, 3)))
    sky1 = np.column_stack([rng.uniform(-0.02, 0.02, 7), rng.uniform(-0.02, 0.02, 7), rng.uniform(0.5, 9.0, 7)])
    sky2 = np.column_stack([rng.uniform(-0.02, 0.02, 4), rng.uniform(-0.02, 0.02, 4), rng.uniform(0.5, 9.0, 4)])
    v1 = orc.model_visibilities(sky1, uvw)
    vm = orc.model_visibilities(sky1, np.asfortranarray(-uvw))
    assert np.max(np.abs(vm - np.conj(v1))) < 1e-9
    v12 = orc.model_visibilities(np.vstack([sky1, sky2]), uvw)
    v2 = orc.model_visibilities(sky2, uvw)
    assert np.max(np.abs(v12 - (v1 + v2))) < 1e-9


def test_matches_the_independent_simulator(orc):
    """The test-data simulator's reference-convention model (its own numpy code) with no
    noise equals the oracle's model for the same sky."""
    snap = simdata.snapshot(512, 10.0, 20_000, 6, seed=77, field_frac=0.5, w_max=800.0, noise_sigma=0.0,
                            reference_model=True)
    sky = np.array([[l, m, a] for a, l, m in snap["sources"]])
    v = orc.model_visibilities(sky, snap["uvw"])
    peak = np.max(np.abs(sky[:, 2]))
    assert np.max(np.abs(v - snap["vis"])) < 1e-9 * peak


def test_residual_of_the_model_images_to_zero(orc):
    """Subtracting the model of the very sky that made the data leaves the noise only:
    the residual image's peak drops to the noise level."""
    from fastimaging_amd import capi
    snap = simdata.snapshot(256, 20.0, 20_000, 4, seed=78, field_frac=0.25, noise_sigma=0.0, reference_model=True)
    sky = np.array([[l, m, a] for a, l, m in snap["sources"]])
    pars = capi.imager_pars(256, 20.0)
    img, _, _, _ = orc.image_visibilities(pars, snap["vis"], snap["weights"], snap["uvw"])
    res = snap["vis"] - orc.model_visibilities(sky, snap["uvw"])
    rimg, _, _, _ = orc.image_visibilities(pars, res, snap["weights"], snap["uvw"])
    assert np.max(np.abs(rimg)) < 1e-9 * np.max(np.abs(img))
