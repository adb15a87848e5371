"""The MEX gateways compile against the project's mex.h stand-in, export mexFunction, and reject a
wrong argument count with the reference's texts (solvers/functions/mex/*Mex.c gateways, e.g.
calculatedeformedlengthMex.c:57-60) before touching the device, so this runs without a GPU."""
import ctypes as C

import pytest

from bb_pd_amd import build as B

import mexcall

# gateway -> (argument count, error text on another count)
NRHS = {
    "calculatedeformedlengthMex": (8, "Eight input arguments required \n"),
    "calculatebsfexponentialMex": (9, "Nine input arguments required \n"),
    "calculatebondforcesMex": (14, "Fourteen input arguments required \n"),
    "calculatenodalforcesopenmpMex": (8, "Eight input arguments required \n"),
    "calculatenodalforcesMex": (5, "Five input arguments required \n"),
    "timeintegrationeulercromerMex": (10, "Ten input arguments required \n"),
    "calculatebondfailureMex": (6, "Six input arguments required \n"),
    "calculatebondsofteningfactorMex": (6, "Six input arguments required \n"),
    "calculatebsftrilinearMex": (7, "Seven input arguments required \n"),
    "calculatedampingcoefficientMex": (6, "Six input arguments required \n"),
    "timeintegrationADRMex": (9, "Nine input arguments required \n"),
    "calculatecontactforceMex": (9, "Nine input arguments required \n"),
}


def test_every_gateway_is_built_and_exports_mexfunction():
    built = sorted(p.stem for p in B.MEX_OUT.glob("*Mex.so"))
    assert built == sorted(list(NRHS) + ["bbpdStepMex"])
    for name in built:
        lib = C.CDLL(str(B.MEX_OUT / f"{name}.so"))
        assert hasattr(lib, "mexFunction")


@pytest.mark.parametrize("name", sorted(NRHS))
def test_wrong_argument_count(name):
    n, text = NRHS[name]
    for k in (0, n - 1, n + 1):
        with pytest.raises(mexcall.MexError) as e:
            mexcall.call(name, *([0.0] * k))
        assert str(e.value) == text


def test_step_gateway_rejects_bad_calls():
    with pytest.raises(mexcall.MexError, match="sub-command"):
        mexcall.call("bbpdStepMex")
    with pytest.raises(mexcall.MexError, match="usage"):
        mexcall.call("bbpdStepMex", "create")
    with pytest.raises(mexcall.MexError, match="handle"):
        mexcall.call("bbpdStepMex", "step", 1.0, 3.0)
