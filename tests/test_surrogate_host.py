"""Host side of the learned surrogate (lymph3d_apsc_amd.surrogate): seeded
instance parameters, the material field they define, the common time step,
and the error measure.  No GPU needed."""
import numpy as np

from lymph3d_apsc_amd import surrogate as S


def test_parameters_are_seeded_and_in_range():
    a, b = S.sample_params(5), S.sample_params(5)
    assert np.array_equal(a, b) and a.shape == (8,)
    th = np.stack([S.sample_params(s) for s in range(500)])
    assert (th[:, :3] >= 0.3).all() and (th[:, :3] <= 0.7).all()
    assert (th[:, 3] >= 4).all() and (th[:, 3] <= 8).all()
    assert (np.abs(th[:, 4:]) <= S.CLIP).all()
    assert abs(th[:, 4:].std() - 1) < 0.1


def test_instance_material_field_and_source():
    th = np.array([0.4, 0.5, 0.6, 6.0, 0.3, -1.0, 2.0, 0.5])
    ins = S.instance(th, 3, 2, S.receivers(8))
    m = ins["mesh"]
    bc = m.xyz[m.tet_vert].mean(axis=1)
    assert np.allclose(m.lam, np.exp(0.5 * (0.3 - 1.0 * (2 * bc[:, 0] - 1))), rtol=1e-14)
    assert np.allclose(m.mu, np.exp(0.5 * (2.0 + 0.5 * (2 * bc[:, 2] - 1))), rtol=1e-14)
    assert (m.rho == 1).all()
    assert ins["f0"] == 6.0 and np.array_equal(ins["source"], th[None, :3])
    # the cached cube is shared, its material arrays are not
    other = S.instance(S.sample_params(1), 3, 2, S.receivers(8))
    assert other["mesh"].xyz is m.xyz and other["mesh"].lam is not m.lam


def test_time_step_covers_the_fastest_instance():
    n, p = 16, 2
    dt = S.cfl_dt(n, p)
    # the fastest P wave any clipped parameter set gives
    lam = mu = np.exp(S.SIGMA * 2 * S.CLIP)
    assert dt <= 0.2 * (1 / n) / (np.sqrt(lam + 2 * mu) * p * p) * (1 + 1e-12)


def test_relative_error_per_instance():
    a = np.ones((3, 4, 2))
    b = 2 * np.ones((3, 4, 2))
    assert np.allclose(S.rel_err(a, b), 0.5)
    assert np.allclose(S.rel_err(b, b), 0.0)


def test_receiver_features_travel_times_and_cells():
    """Travel times equal distance x mean slowness along the segment (checked
    by a fine trapezoid rule through the instance's field), the geometry
    entries are what they name, and the cell coordinates lie in [0, 1)."""
    rec = S.receivers(6)
    th = np.stack([S.sample_params(s) for s in range(7)])
    f = S.receiver_features(th, rec, 16)
    assert f.shape == (7, 6, S.N_FEAT) and np.isfinite(f).all()
    k, r = 4, 2
    src, xi = th[k, :3], th[k, 4:]
    t = np.linspace(0, 1, 40001)
    pts = src + t[:, None] * (rec[r] - src)
    lam = np.exp(S.SIGMA * (xi[0] + xi[1] * (2 * pts[:, 0] - 1)))
    mu = np.exp(S.SIGMA * (xi[2] + xi[3] * (2 * pts[:, 2] - 1)))
    d = np.linalg.norm(rec[r] - src)
    assert abs(f[k, r, 9] - d) < 1e-14
    assert np.allclose(f[k, r, 6:9], (rec[r] - src) / d, rtol=1e-14)
    assert abs(f[k, r, 16] / (d * np.trapezoid(1 / np.sqrt(lam + 2 * mu), t)) - 1) < 1e-9
    assert abs(f[k, r, 17] / (d * np.trapezoid(1 / np.sqrt(mu), t)) - 1) < 1e-9
    assert np.allclose(f[k, r, 18], f[k, r, 16] * th[k, 3])
    cells = f[..., 24:30]
    assert (cells >= 0).all() and (cells < 1).all()
    assert np.allclose(f[k, r, 24:27], 16 * src - np.floor(16 * src))
    assert (f[..., 30:] == 0).all()


class _NoNet:
    """Stand-in for the GPU network (predicts zeros) to test the host side."""

    def __init__(self, widths, max_batch=256, seed=0):
        self.widths = widths

    def set_data(self, x, y):
        assert x.shape[1] == self.widths[0] and y.shape[1] == self.widths[-1]

    def train(self, order, batch, lr):
        return np.zeros(1, np.float32)

    def predict(self, x):
        return np.zeros((len(x), self.widths[-1]), np.float32)

    def close(self):
        pass


def test_receiver_model_basis_round_trip(monkeypatch):
    """With as many modes as a receiver trace has entries the SVD basis
    represents the training traces exactly (spreading factor undone); a
    network that predicts zero coefficients returns the mean scaled trace
    divided back by each receiver's spreading factor."""
    monkeypatch.setattr(S, "Mlp", _NoNet)
    rng = np.random.default_rng(1)
    rec = S.receivers(5)
    th = np.stack([S.sample_params(s) for s in range(40)])
    T = 8
    tr = rng.standard_normal((40, T, 5, 3))
    sur = S.ReceiverSurrogate(rec, T, 16, n_nets=2, width=64, n_modes=T * 3)
    sur.fit(th, tr, epochs=1, batch=8)
    assert abs(sur.basis_energy - 1) < 1e-12
    assert np.abs(sur.project(th, tr) - tr.reshape(40, -1)).max() < 1e-12
    mean, spread = sur.predict(th[:3])
    d = np.linalg.norm(rec[None] - th[:3, None, :3], axis=2)
    expect = sur.y_mean.reshape(1, 1, T, 3) / np.maximum(d, S.R_MIN)[:, :, None, None]
    assert np.allclose(mean, expect.transpose(0, 2, 1, 3).reshape(3, -1))
    assert (spread == 0).all()
