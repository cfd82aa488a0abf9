"""The whole gray 2-D step as one device-resident sequence (imc2d_step),
launched directly and replayed from a captured HIP graph, against the staged
entry points (imc2d_set_material, imc2d_sample_boundary, imc2d_sample_volume,
imc2d_transport, imc2d_comb with the host callables of IMC/IMC2D.py:1591-1897).

All three draw the same Philox streams, so they differ only by the summation
order of the FP64 tally atomics: temperatures, radiation temperatures,
energies and census sizes agree to 1e-10 relative over several steps.  The
graph runs see the pool grow past a capture's bound, a changed Nmax (a new
capture), and a time-dependent face temperature (new face energies replayed
through the same graph)."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _announce(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _problem(geometry):
    from radtranbook_amd.materials import LinearEOS, PowerLawOpacity

    nx, ny = 24, 20
    e1 = np.linspace(0.0 if geometry == "rz" else 0.1, 0.5, nx + 1)
    e2 = np.linspace(0.0, 0.4, ny + 1)
    g = np.random.default_rng(11).standard_normal((nx, ny))
    sigma0 = np.exp(2.3 + 1.2 * g)
    cvc = 0.3 + 0.1 * np.random.default_rng(12).random((nx, ny))
    return e1, e2, sigma0, cvc, PowerLawOpacity(sigma0, 3.0, 1e-3), LinearEOS(cvc)


def _run(mode, geometry, precision, nsteps, Nmax_of, Tb, reflect=(False, False, False, False)):
    from radtranbook_amd import IMC2D

    e1, e2, sigma0, cvc, opac, eos = _problem(geometry)
    nx, ny = sigma0.shape
    T0 = np.full((nx, ny), 1e-3)
    st = IMC2D.init_simulation(20000, T0, T0, e1, e2, eos.eos, eos.inv_eos, geometry=geometry, precision=precision,
                               seed=3)
    if mode == "staged":
        sf = lambda T: sigma0 / np.maximum(T, 1e-3) ** 3  # noqa: E731
        cvf = lambda T: cvc + 0.0 * T  # noqa: E731
        inv = lambda u: u / cvc  # noqa: E731
    else:
        sf, cvf, inv = opac, eos.cv_func, eos.inv_eos
        st._engine.use_graph = mode == "graph"
    out = []
    for k in range(nsteps):
        st, info = IMC2D.step(st, 10000, 8000, 0, Nmax_of(k), Tb, 0.005, e1, e2, sf, inv, cvf, np.zeros((nx, ny)),
                              geometry=geometry, reflect=reflect)
        assert abs(info["energy_residual"]) <= 1e-12 * info["total_energy"], (mode, k, info["energy_residual"])
        out.append(info)
    census = st._engine.get_census()
    st._engine.close()
    return out, census


def _compare(a_run, b_run, rtol):
    for a, b in zip(a_run, b_run):
        np.testing.assert_allclose(a["temperature"], b["temperature"], rtol=rtol, atol=0)
        np.testing.assert_allclose(a["radiation_temperature"], b["radiation_temperature"], rtol=rtol, atol=0)
        np.testing.assert_allclose(a["deposited"], b["deposited"], rtol=rtol, atol=rtol * np.abs(b["deposited"]).max())
        np.testing.assert_allclose(a["emitted_energies"], b["emitted_energies"], rtol=rtol, atol=0)
        for key in ("total_energy", "total_internal_energy", "total_radiation_energy", "boundary_loss",
                    "boundary_emission", "comb_discrepancy"):
            assert a[key] == pytest.approx(b[key], rel=rtol, abs=rtol * a["total_energy"]), key
        assert a["N_particles"] == b["N_particles"]
        ta, tb = a["profiling"]["transport_events"], b["profiling"]["transport_events"]
        assert ta["n_particles_transported"] == tb["n_particles_transported"]
        assert ta["total"] == tb["total"]


@pytest.mark.parametrize("geometry,precision", [("xy", "fp64"), ("rz", "fp64"), ("xy", "fp32"), ("rz", "fp32")])
def test_device_step_graph_and_direct_match_staged(geometry, precision):
    from radtranbook_amd import IMC2D, _lib

    _lib.load()
    _lib.require_gpu()
    IMC2D.VERBOSE = False
    _announce(f"test_device_step_graph_and_direct_match_staged {geometry} {precision}")
    Tb = (0.5, 0.0, 0.0, 0.0) if geometry == "xy" else (0.0, 0.0, 0.5, 0.0)
    # Nmax grows, then drops: the pool outgrows the first captures' bounds and
    # a changed Nmax needs a capture of its own
    nmax = lambda k: [30000, 40000, 60000, 60000, 60000, 40000, 40000][k]  # noqa: E731
    runs = {m: _run(m, geometry, precision, 7, nmax, Tb) for m in ("graph", "direct", "staged")}
    _compare(runs["graph"][0], runs["direct"][0], 1e-10)
    _compare(runs["graph"][0], runs["staged"][0], 1e-10)
    # the final census is the same particle list (the comb writes in pool order)
    for a, b in zip(runs["graph"][1], runs["staged"][1]):
        if a.dtype.kind == "f":
            np.testing.assert_allclose(a, b, rtol=1e-9, atol=1e-12 * max(np.abs(b).max(), 1e-300))
        else:
            np.testing.assert_array_equal(a, b)
    assert runs["graph"][0][-1]["temperature"].max() > 10 * 1e-3


def test_device_step_graph_replays_time_dependent_faces():
    from radtranbook_amd import IMC2D, _lib

    _lib.load()
    _lib.require_gpu()
    IMC2D.VERBOSE = False
    _announce("test_device_step_graph_replays_time_dependent_faces")
    # left face heats up with time, the top face switches on at the third
    # step; the bottom face reflects
    Tb = (lambda t: 0.2 + 40.0 * t, 0.0, 0.0, lambda t: 0.3 if t > 0.009 else 0.0)
    refl = (False, False, True, False)
    runs = {m: _run(m, "xy", "fp64", 6, lambda k: 40000, Tb, reflect=refl) for m in ("graph", "direct", "staged")}
    _compare(runs["graph"][0], runs["direct"][0], 1e-10)
    _compare(runs["graph"][0], runs["staged"][0], 1e-10)
    be = [i["boundary_emission"] for i in runs["graph"][0]]
    assert be[0] < be[1] < be[2] < be[3]
