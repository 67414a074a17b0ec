"""Checkpoint files (LBM.h:123-126 of the reference declares save/load of
macroscopic and mesoscopic ticks without implementing them): a mesoscopic
checkpoint continues a run bit for bit; a macroscopic one restores the saved
fields and re-initialises at equilibrium like initialize_fields."""
import numpy as np
import pytest

from helpers import SEED, channel_ids, random_state

pytestmark = pytest.mark.gpu


def make_case(bfa, kind, storage="fp32"):
    if kind == "thermal_multiphase_walls":
        X, Y, Z = 24, 18, 10
        n = X * Y * Z
        rng = np.random.default_rng(SEED + 21)
        rho = (1.0 + 0.2 * rng.random(n)).astype(np.float32)
        ids = np.zeros((Z, Y, X), np.uint8)
        ids[:, 0, :] = 1
        ids[:, -1, :] = 2
        s = bfa.Solver((X, Y, Z), 0.9, periodic=(1, 0, 1), velocity_set="D3Q19", storage=storage, multiphase=True)
        s.set_object(1, temperature=1.2)
        s.set_object(2, temperature=0.8)
        s.set_param("interaction_strength", -4.5)
        s.set_param("reference_temperature", 1.0)
        force = np.tile(np.array([0, -1e-5, 0], np.float32), (n, 1))
        s.init_fields(rho, np.zeros((n, 3), np.float32), force=force,
                      temperature=np.full(n, 1.0, np.float32), boundary_id=ids.ravel())
        return s
    X, Y, Z = 20, 16, 12
    n = X * Y * Z
    rho, u = random_state(n, seed=SEED + 22)
    s = bfa.Solver((X, Y, Z), 0.6, velocity_set="D3Q19", storage=storage)
    s.init_fields(rho, u)
    return s


@pytest.mark.parametrize("kind,storage", [("thermal_multiphase_walls", "fp32"), ("periodic", "fp16")])
def test_mesoscopic_checkpoint_continues_bit_for_bit(bfa, tmp_path, kind, storage):
    print(f"case: mesoscopic checkpoint {kind} {storage}", flush=True)
    a = make_case(bfa, kind, storage)
    a.step(17)                       # odd tick count: the AA parity is part of the state
    a.save_mesoscopic(str(tmp_path))
    a.step(13, save=True)
    ref = a.read("populations_raw"), a.read("velocity_density")
    b = make_case(bfa, kind, storage)
    b.set_param("relaxation_time", 0.77)   # overwritten by the file
    b.load_mesoscopic(str(tmp_path), 17)
    assert b.ticks == 17
    b.step(13, save=True)
    assert b.ticks == 30
    got = b.read("populations_raw"), b.read("velocity_density")
    assert np.array_equal(ref[0], got[0]) and np.array_equal(ref[1], got[1])
    if kind != "periodic":
        assert np.array_equal(a.read("thermal_raw"), b.read("thermal_raw"))
        assert np.array_equal(a.read("psi"), b.read("psi"))


def test_macroscopic_checkpoint_restores_fields(bfa, tmp_path):
    print("case: macroscopic checkpoint D2Q9 channel", flush=True)
    X, Y = 48, 20
    n = X * Y
    rho, u = random_state(n, seed=SEED + 23)
    ids = channel_ids(X, Y)
    a = bfa.Solver((X, Y, 1), 0.7, periodic=(1, 0, 1), velocity_set="D2Q9")
    a.init_fields(rho, u, boundary_id=ids)
    a.step(25, save=True)
    saved = a.read("velocity_density")
    a.save_macroscopic(str(tmp_path))
    # a fresh solver initialised from the saved fields directly ...
    fresh = bfa.Solver((X, Y, 1), 0.7, periodic=(1, 0, 1), velocity_set="D2Q9")
    fresh.init_fields(saved[:, 3].copy(), saved[:, :3].copy(), boundary_id=ids)
    fresh.step(5, save=True)
    # ... and one loaded from the checkpoint agree
    b = bfa.Solver((X, Y, 1), 0.7, periodic=(1, 0, 1), velocity_set="D2Q9")
    b.load_macroscopic(str(tmp_path), 25)
    assert b.ticks == 25
    np.testing.assert_array_equal(b.read("boundary"), ids)
    b.step(5, save=True)
    assert np.array_equal(b.read("populations_raw"), fresh.read("populations_raw"))


@pytest.mark.parametrize("kind", ["thermal_constant_force", "multiphase_constant_force", "thermal_force_field"])
def test_macroscopic_checkpoint_restores_the_external_force(bfa, tmp_path, kind):
    """The saved F channel holds the total force of the save tick (external
    force + buoyancy, or + the Shan-Chen force).  Loading a macroscopic file
    must restart from the external force alone: it equals a fresh
    initialisation from the saved u, rho and T with the same external force,
    bit for bit, and a restart from the F channel would differ."""
    print(f"case: macroscopic checkpoint with {kind}", flush=True)
    X, Y, Z = 20, 14, 8
    n = X * Y * Z
    rng = np.random.default_rng(SEED + 24)
    ids = channel_ids(X, Y, Z)
    periodic = (1, 0, 1)
    multiphase = kind.startswith("multiphase")
    if kind == "thermal_force_field":
        force = (1e-5 * rng.standard_normal((n, 3))).astype(np.float32)
    else:
        force = np.tile(np.array([2e-5, -1e-5, 0], np.float32), (n, 1))
    rho = (1.0 + 0.2 * rng.random(n)).astype(np.float32) if multiphase else \
        (1.0 + 0.02 * rng.random(n)).astype(np.float32)
    u = np.zeros((n, 3), np.float32)
    T = (0.5 + 0.1 * rng.random(n)).astype(np.float32) if kind.startswith("thermal") else None

    def solver():
        s = bfa.Solver((X, Y, Z), 0.8, periodic=periodic, velocity_set="D3Q19", multiphase=multiphase)
        if multiphase:
            s.set_param("interaction_strength", -4.5)
        if T is not None:
            s.set_object(1, temperature=0.7)
            s.set_param("reference_temperature", 0.5)
        return s

    a = solver()
    a.init_fields(rho, u, force=force, temperature=T, boundary_id=ids)
    a.step(21, save=True)
    saved, ft = a.read("velocity_density"), a.read("force_temperature")
    total = ft[:, :3][ids == 0]
    assert np.abs(total - force[ids == 0]).max() > 1e-7      # the F channel is not the external force
    a.save_macroscopic(str(tmp_path))
    fresh = solver()
    fresh.init_fields(saved[:, 3].copy(), saved[:, :3].copy(), force=force,
                      temperature=None if T is None else ft[:, 3].copy(), boundary_id=ids)
    b = solver()
    b.load_macroscopic(str(tmp_path), 21)
    assert b.info()["forcing"] == fresh.info()["forcing"] == (2 if kind == "thermal_force_field" else 1)
    assert b.info()["thermal"] == fresh.info()["thermal"]
    fresh.step(6, save=True)
    b.step(6, save=True)
    assert np.array_equal(b.read("populations_raw"), fresh.read("populations_raw"))
    assert np.array_equal(b.read("force_temperature"), fresh.read("force_temperature"))
    if T is not None:
        assert np.array_equal(b.read("thermal_raw"), fresh.read("thermal_raw"))


def test_checkpoint_errors(bfa, tmp_path):
    s = make_case(bfa, "periodic")
    with pytest.raises(bfa.BFError, match="no checkpoint"):
        s.load_mesoscopic(str(tmp_path), 3)
    s.save_mesoscopic(str(tmp_path))
    other = bfa.Solver((20, 16, 12), 0.6, velocity_set="D3Q27")
    other.init_fields(np.ones(20 * 16 * 12, np.float32), np.zeros((20 * 16 * 12, 3), np.float32))
    with pytest.raises(bfa.BFError, match="another configuration"):
        other.load_mesoscopic(str(tmp_path), 0)
