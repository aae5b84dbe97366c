"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py
with the oracle): each is checked against constructions that do not use the
oracle (numpy lexsort, searchsorted CSR, brute-force distances, the leapfrog
update identities), and the oracle must reproduce it exactly."""
import sys
from pathlib import Path

import numpy as np
import pytest

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE / "golden"))

import make_golden as MG  # noqa: E402


def load(D):
    with np.load(HERE / "golden" / f"golden_{D}d.npz", allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def sorted_x(g):
    return g["x_in"][g["perm"] - 1]


@pytest.mark.parametrize("D", [2, 3])
def test_fixture_permutation_is_stable_lexicographic(D):
    g = load(D)
    c = 2 * float(g["h"])
    keys = np.floor(g["x_in"] / c).astype(np.int64)
    order = np.lexsort(tuple(keys[:, d] for d in reversed(range(D))))
    assert np.array_equal(g["perm"] - 1, order)


@pytest.mark.parametrize("D", [2, 3])
def test_fixture_grid_matches_searchsorted(D):
    g = load(D)
    c = 2 * float(g["h"])
    x = sorted_x(g)
    ming, ng = g["mingridx"], g["ngridx"]
    # origin on the lattice of the cutoff, at least 2 cells below the minimum
    assert np.all(ming <= x.min(0) - 2 * c)
    assert np.allclose(ming / c, np.round(ming / c), rtol=0, atol=1e-9)
    cell = np.floor((x - ming) / c).astype(np.int64)
    flat = cell[:, 0]
    for d in range(1, D):
        flat = flat * int(ng[d]) + cell[:, d]
    ncells = int(np.prod(ng))
    assert np.all(np.diff(flat) >= 0)
    expect = np.searchsorted(flat, np.arange(ncells + 1), side="left") + 1
    assert np.array_equal(g["cell_start"], expect)
    assert g["cell_start"][-1] == len(x) + 1


@pytest.mark.parametrize("D", [2, 3])
def test_fixture_counts_match_brute_force(D):
    g = load(D)
    x = sorted_x(g)
    c2 = (2 * float(g["h"])) ** 2
    dx = x[:, None, :] - x[None, :, :]
    r2 = dx[..., 0] * dx[..., 0]
    for d in range(1, D):
        r2 = r2 + dx[..., d] * dx[..., d]
    np.fill_diagonal(r2, np.inf)
    assert np.array_equal(g["counts"], (r2 < c2).sum(1))


@pytest.mark.parametrize("D", [2, 3])
def test_fixture_step_identities(D):
    """One kick-drift step from dvdt = 0, drhodt = 0 (initial rates are zero,
    src/Particles.jl:84,196): v = v0 + dt·a, x = x0 + dt·v, ρ = fma(dt, ρ̇, ρ0);
    Σ m·a = 0 for a self-interaction with no body force."""
    g = load(D)
    dt = float(g["dt"])
    x0, v0, r0 = g["x_in"], g["v_in"], g["rho_in"]
    a, rd = g["step_dvdt"], g["step_drhodt"]
    v = g["step_v"]
    assert np.array_equal(v, v0 + dt * a)
    assert np.array_equal(g["step_x"], x0 + dt * v)
    assert np.allclose(g["step_rho"], r0 + dt * rd, rtol=1e-15, atol=0)
    assert np.max(np.abs(a.sum(0))) < 1e-10 * np.max(np.abs(a)) * len(a)
    assert np.any(a != 0) and np.any(rd != 0)


@pytest.mark.parametrize("D", [2, 3])
def test_oracle_reproduces_fixture(D):
    g = load(D)
    new = MG.make(D)
    for k, v in g.items():
        if k == "provenance":
            continue
        assert np.array_equal(new[k], v), k
