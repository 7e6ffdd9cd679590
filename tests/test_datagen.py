"""The seeded generator against its fixture and against a scalar SplitMix64."""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
MASK = (1 << 64) - 1


def splitmix64_scalar(seed, count):
    state, out = seed, []
    for _ in range(count):
        state = (state + 0x9E3779B97F4A7C15) & MASK
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
        out.append(z ^ (z >> 31))
    return out


def test_splitmix64_matches_scalar_definition():
    from gprat_amd import datagen
    for seed in (0, 1, 12345, MASK):
        assert [int(v) for v in datagen.splitmix64(seed, 8)] == splitmix64_scalar(seed, 8)


def test_fixture():
    from gprat_amd import datagen
    with open(os.path.join(HERE, "golden", "datagen_vectors.json")) as f:
        ref = json.load(f)
    for seed in (0, 1):
        u, y = datagen.msd_series(16, seed)
        assert np.array_equal(u, ref[str(seed)]["u"])
        assert np.array_equal(y, ref[str(seed)]["y"])
        assert [int(w) for w in datagen.splitmix64(seed, 4)] == ref[str(seed)]["splitmix64_first4"]


def test_properties():
    from gprat_amd import datagen
    u, y = datagen.msd_series(20000, 0)
    assert np.all(np.abs(u) <= 1.0) and abs(u.mean()) < 0.02 and abs(u.var() - 1 / 3) < 0.02
    x, yy, z = datagen.gp_problem(100, 10, 8)
    assert x.shape == (107,) and np.all(x[:7] == 0.0) and yy.shape == (100,) and z.shape == (17,)
    # prefix stability: a longer series starts with the shorter one
    u2, _ = datagen.msd_series(50, 3)
    u3, _ = datagen.msd_series(10, 3)
    assert np.array_equal(u2[:10], u3)


def test_written_dataset_reads_back_through_gp_data(tmp_path):
    """The seeded dataset writer produces the reference's data-directory
    files, and GP_data (utils_c.cpp:54-81 semantics) reads them back exactly,
    with the n_reg - 1 leading zeros."""
    import gprat_amd as gprat
    from gprat_amd import datagen
    n_train, n_test, n_reg = 300, 40, 8
    datagen.write_dataset(str(tmp_path), n_train, n_test)
    names = sorted(os.listdir(tmp_path))
    assert names == ["test_input.txt", "test_output.txt", "training_input.txt", "training_output.txt"]
    u, y = datagen.msd_series(n_train, 0)
    ut, yt = datagen.msd_series(n_test, 1)
    for name, ref, n in (("training_input", u, n_train), ("training_output", y, n_train),
                         ("test_input", ut, n_test), ("test_output", yt, n_test)):
        d = gprat.GP_data(os.path.join(str(tmp_path), name + ".txt"), n, n_reg)
        got = np.asarray(d.data)
        assert got.shape == (n + n_reg - 1,)
        assert np.all(got[:n_reg - 1] == 0.0)
        assert np.array_equal(got[n_reg - 1:], ref)
    x, _, z = datagen.gp_problem(n_train, n_test, n_reg)
    assert np.array_equal(np.asarray(gprat.GP_data(os.path.join(str(tmp_path), "training_input.txt"), n_train, n_reg).data), x)
    assert np.array_equal(np.asarray(gprat.GP_data(os.path.join(str(tmp_path), "test_input.txt"), n_test, n_reg).data), z)


def _uniform_scalar(seed, count):
    """[0, 1) doubles from the top 53 bits of SplitMix64, scalar."""
    return [(w >> 11) * (1.0 / 9007199254740992.0) for w in splitmix64_scalar(seed, count)]


def test_mass_spring_damper_against_state_space_recurrence():
    """The semi-implicit Euler step written as a linear state-space map,
    s_t = A s_(t-1) + B u_t with s = (x, v):
        v' = (1 - dt c/m) v - (dt k/m) x + (dt/m) u
        x' = x + dt v' = (1 - dt^2 k/m) x + dt (1 - dt c/m) v + (dt^2/m) u,
    evaluated from scalar SplitMix64 inputs, gives the generator's noise-free
    output to rounding (different operation order), and the noise is the
    Box-Muller draw from the same stream: y - x = 0.01 sqrt(-2 ln(1 - u1))
    cos(2 pi u2) with u1, u2 the stream's second and third thirds."""
    import math
    from gprat_amd import datagen
    n, seed = 400, 0
    m, k, c, dt = 1.0, 1.0, 0.5, 0.05
    A = [[1 - dt * dt * k / m, dt * (1 - dt * c / m)], [-dt * k / m, 1 - dt * c / m]]
    B = [dt * dt / m, dt / m]
    r = _uniform_scalar(seed, 3 * n)
    u = [2.0 * v - 1.0 for v in r[:n]]
    xs, x, v = [], 0.0, 0.0
    for t in range(n):
        x, v = A[0][0] * x + A[0][1] * v + B[0] * u[t], A[1][0] * x + A[1][1] * v + B[1] * u[t]
        xs.append(x)
    noise = [0.01 * math.sqrt(-2.0 * math.log1p(-r[n + t])) * math.cos(2.0 * math.pi * r[2 * n + t]) for t in range(n)]
    gu, gy = datagen.msd_series(n, seed)
    assert np.array_equal(gu, np.array(u))
    assert np.max(np.abs((gy - np.array(noise)) - np.array(xs))) < 1e-12
    assert np.max(np.abs(gy - np.array(xs) - np.array(noise))) < 1e-12


def test_mass_spring_damper_step_response():
    """The oscillator the generator integrates (m = 1, k = 1, c = 0.5,
    dt = 0.05, semi-implicit Euler), driven by a constant u = 1, settles at
    the static deflection u / k = 1 with a bounded overshoot: the
    parameters give a stable, underdamped system."""
    m, k, c, dt = 1.0, 1.0, 0.5, 0.05
    x = v = 0.0
    peak = 0.0
    for _ in range(4000):
        v = v + dt * (1.0 - k * x - c * v) / m
        x = x + dt * v
        peak = max(peak, x)
    assert abs(x - 1.0 / k) < 1e-6 and 1.0 < peak < 2.0
