"""Host-side selection of the device-held material path (radtranbook_amd/materials.py)."""
import numpy as np


def test_device_material_detection():
    """Only a PowerLawOpacity with the bound c_v / inverse EOS of one LinearEOS
    runs on the device; any other callable stays on the host."""
    from radtranbook_amd.materials import LinearEOS, PowerLawOpacity, device_material

    op = PowerLawOpacity(np.full((3, 2), 5.0), 3.0, 1e-3)
    eos = LinearEOS(0.3)
    got = device_material(op, eos.cv_func, eos.inv_eos, (3, 2))
    assert got is not None and got[1] == 3.0 and got[2] == 1e-3 and got[3].shape == (3, 2)
    other = LinearEOS(0.3)
    assert device_material(op, eos.cv_func, other.inv_eos, (3, 2)) is None
    assert device_material(lambda T: 5.0 / T ** 3, eos.cv_func, eos.inv_eos, (3, 2)) is None
    assert device_material(op, lambda T: 0.3 + 0 * T, eos.inv_eos, (3, 2)) is None
    assert device_material(op, eos.cv_func, eos.eos, (3, 2)) is None
    # the host form evaluates the same numbers the device kernel does
    T = np.array([[1e-4, 0.2], [0.5, 1.0], [2.0, 1e-3]])
    np.testing.assert_array_equal(op(T), 5.0 / (np.maximum(T, 1e-3) ** 3))
