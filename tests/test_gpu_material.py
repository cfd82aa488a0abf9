"""The gray 2-D step on the device-held material (power-law opacity, linear
EOS: imc2d_set_material_powerlaw / imc2d_fleck_powerlaw /
imc2d_update_material) against the same step with the host callables
(IMC/IMC2D.py:1646-1651, 1800-1810 evaluated in NumPy).

Both runs draw the same Philox streams, so they differ only by the summation
order of the FP64 tally atomics: temperatures, radiation temperatures and
energies agree to 1e-10 relative over several steps."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _announce(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("geometry,precision", [("xy", "fp64"), ("rz", "fp64"), ("xy", "fp32")])
def test_device_material_step_matches_host_callables(geometry, precision):
    from radtranbook_amd import IMC2D, _lib
    from radtranbook_amd.materials import LinearEOS, PowerLawOpacity

    _lib.load()
    _lib.require_gpu()
    IMC2D.VERBOSE = False
    _announce(f"test_device_material_step_matches_host_callables {geometry} {precision}")
    nx, ny = 20, 16
    e1 = np.linspace(0.0 if geometry == "rz" else 0.1, 0.5, nx + 1)
    e2 = np.linspace(0.0, 0.4, ny + 1)
    g = np.random.default_rng(4).standard_normal((nx, ny))
    sigma0 = np.exp(2.3 + 1.2 * g)
    cvc = 0.3 + 0.1 * np.random.default_rng(5).random((nx, ny))
    opac = PowerLawOpacity(sigma0, 3.0, 1e-3)
    eos = LinearEOS(cvc)
    T0 = np.full((nx, ny), 1e-3)
    Tb = (0.5, 0.0, 0.0, 0.0) if geometry == "xy" else (0.0, 0.0, 0.5, 0.0)
    runs = {}
    for mode in ("device", "host"):
        st = IMC2D.init_simulation(20000, T0, T0, e1, e2, eos.eos, eos.inv_eos, geometry=geometry,
                                   precision=precision, seed=7)
        if mode == "device":
            sf, cvf, inv = opac, eos.cv_func, eos.inv_eos
        else:
            sf = lambda T: sigma0 / np.maximum(T, 1e-3) ** 3  # noqa: E731
            cvf = lambda T: cvc + 0.0 * T  # noqa: E731
            inv = lambda u: u / cvc  # noqa: E731
        out = []
        for _ in range(4):
            st, info = IMC2D.step(st, 10000, 10000, 0, 40000, Tb, 0.005, e1, e2, sf, inv, cvf, np.zeros((nx, ny)),
                                  geometry=geometry)
            assert abs(info["energy_residual"]) <= 1e-12 * info["total_energy"]
            out.append(info)
        runs[mode] = (out, st)
    dev, host = runs["device"][0], runs["host"][0]
    for a, b in zip(dev, host):
        np.testing.assert_allclose(a["temperature"], b["temperature"], rtol=1e-10, atol=0)
        np.testing.assert_allclose(a["radiation_temperature"], b["radiation_temperature"], rtol=1e-10, atol=0)
        for key in ("total_energy", "total_internal_energy", "total_radiation_energy", "boundary_loss"):
            assert a[key] == pytest.approx(b[key], rel=1e-10)
        assert a["N_particles"] == b["N_particles"]
    # the device path keeps the state's arrays in step with the device
    np.testing.assert_allclose(runs["device"][1].internal_energy, runs["host"][1].internal_energy, rtol=1e-10)
    assert dev[-1]["temperature"].max() > 10 * 1e-3
    # a host edit of the state (a new array) is uploaded before the next step
    st = runs["device"][1]
    st.temperature = st.temperature * 1.0
    st.internal_energy = st.internal_energy * 1.0
    st, info = IMC2D.step(st, 10000, 10000, 0, 40000, Tb, 0.005, e1, e2, opac, eos.inv_eos, eos.cv_func,
                          np.zeros((nx, ny)), geometry=geometry)
    assert abs(info["energy_residual"]) <= 1e-12 * info["total_energy"]
