"""The drop-in boundary from the reference's side (CPU): the GpuSystemForce
adapter compiled against ProtoMol's own headers and registered over the CPU
exemplars (factories/registerForceExemplarsCutoff.cpp:36,
registerForceExemplarsFastElectrostatic.cpp:27, frontend/Factory.h:153-166)
serves every id of the GPU path with the exemplar's id, keyword and parameter
list, and ForceFactory::make reports parameter errors as the CPU exemplar
does (make returns NULL and appends to errMsg)."""
import pytest

from tests import adapter_lib

pytestmark = pytest.mark.skipif(not adapter_lib.available(),
                                reason="oracle/_ref/libpmadapter.so not built (needs the reference tree)")


@pytest.fixture(scope="module")
def adp():
    return adapter_lib.Adapter()


def test_adapters_registered_under_the_cpu_ids(adp):
    adp.register(False)
    cpu = {fid: adp.describe(fid) for fid in adapter_lib.GPU_IDS}
    for fid, d in cpu.items():
        assert d is not None, f"the reference build does not register {fid}"
        assert not d["gpu"]
    n = adp.register(True)
    assert n == len(adapter_lib.GPU_IDS)
    for fid in adapter_lib.GPU_IDS:
        g = adp.describe(fid)
        assert g["gpu"], fid
        assert g["id"] == g["cpu_id"] == cpu[fid]["id"]
        assert g["nparams"] == g["cpu_nparams"] == cpu[fid]["nparams"]
        assert g["keyword"] == cpu[fid]["keyword"]
    # the PME alias resolves to the adapter as well
    assert adp.describe("CoulombPME")["gpu"]


@pytest.mark.parametrize("fid,bad", [
    ("LennardJones -switchingFunction C2 -algorithm NonbondedCutoff", [7.0, -8.0, 8.0]),
    ("Coulomb -switchingFunction C1 -algorithm NonbondedCutoff", [8.0, -1.0]),
    ("Coulomb -algorithm PMEwald -reciprocal -correction -interpolation BSpline", [24, 24, 24, -8.0, 4, 1e-6, -1.0]),
])
def test_make_reports_parameter_errors_as_the_reference(adp, fid, bad):
    adp.register(False)
    ok_cpu, _, err_cpu = adp.make(fid, bad)
    adp.register(True)
    ok_gpu, _, err_gpu = adp.make(fid, bad)
    assert not ok_cpu and not ok_gpu
    assert err_cpu and err_gpu == err_cpu


def test_make_without_a_device_appends_the_reason(adp):
    """On a machine without a HIP device the adapter's doMake returns NULL
    with the reason appended to errMsg (no abort)."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    adp.reset()
    adp.register(True)
    ok, _, err = adp.make("LennardJones -switchingFunction C2 -algorithm NonbondedCutoff", [7.0, 8.0, 8.0])
    assert not ok
    assert "no HIP device" in err
